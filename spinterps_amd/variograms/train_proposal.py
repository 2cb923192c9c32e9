"""Train the variogram proposal network on simulated fields.

    python -m spinterps_amd.variograms.train_proposal --fields 200000 \\
        --out proposal_net.npz

Fields: unconditional Gaussian simulations (FP64 Cholesky on the GPU) at
n_lo..n_hi random stations out of a pool spread over a square domain, each
with a random Sph/Exp/Gau variogram (range 5-80 km, nugget fraction
0-0.3, unit sill).  Labels: the best fit of BatchVariogramFitter with the
settings of variograms/fitvgs.py:187-206 (mdr 0.5, nk 50, robust
estimator, Sph/Exp/Gau with a nugget, permutations of 1 and 2).  The
network is trained by the library's forward/backward/Adam kernels.
"""
from __future__ import annotations

import argparse
import time

import numpy as np
import torch

from ..interp.steps import DeviceProblem
from .fitvgs import BatchVariogramFitter
from .learned import (MODEL_TYPES, VariogramProposalNet, features,
                      labels_from_fits, training_fields)

FIT_KW = dict(mdr=0.5, nk=50, perm_r_list=(1, 2), fit_vgs=MODEL_TYPES,
              fil_nug_vg='Nug', evg='robust', fit_thresh=0.01)


def station_pool(n=1500, span=5e5, seed=0):
    """n stations uniform over a span x span square at a UTM-like
    origin."""
    rng = np.random.default_rng(seed)
    return (350000 + rng.uniform(0, span, n),
            5300000 + rng.uniform(0, span, n))


def build_training_set(n_fields, chunk=128, seed=1, n_pool=1500,
                       span=5e5, n_lo=300, n_hi=1500, fit_chunk=2048,
                       opt_iters=10000, log=print):
    """(features [F x 40] float32, classes [F], targets [F x 3]) of F
    simulated and fitted fields (unfitted or unlabelled fields dropped).

    All fields are simulated first (chunks of `chunk`), then the pair list
    is sorted once and the fields are featurised and fitted in chunks of
    `fit_chunk`."""
    x, y = station_pool(n_pool, span, 0)
    prob = DeviceProblem(x, y, x[:1], y[:1])
    t0 = time.perf_counter()
    vals = np.empty((n_fields, n_pool))
    for c0 in range(0, n_fields, chunk):
        T = min(chunk, n_fields - c0)
        fields, info, _ = training_fields(prob, T, seed=seed * 1000003 + c0,
                                          n_lo=n_lo, n_hi=n_hi)
        vals[c0:c0 + T] = fields.cpu().numpy()
        bad = (info != 0).cpu().numpy()
        vals[c0:c0 + T][bad] = np.nan
        if (c0 // chunk) % 50 == 0:
            log(f'simulated {c0 + T}/{n_fields} '
                f'{time.perf_counter() - t0:.1f}s')
    fitter = BatchVariogramFitter(x, y)
    feats, clss, regs = [], [], []
    for c0 in range(0, n_fields, fit_chunk):
        v = vals[c0:c0 + fit_chunk]
        feat, scale = features(fitter, v, FIT_KW['mdr'])
        _, det = fitter.fit(v, seed=seed + c0, opt_iters=opt_iters,
                            return_details=True, **FIT_KW)
        if det is None:
            continue
        st = det['steps']
        cls, reg = labels_from_fits(det['best'], scale.cpu().numpy()[st])
        keep = cls >= 0
        feats.append(feat.cpu().numpy()[st][keep])
        clss.append(cls[keep])
        regs.append(reg[keep])
        nit = det['nit']
        log(f'fitted {min(c0 + fit_chunk, n_fields)}/{n_fields} labelled '
            f'{sum(map(len, clss))} {time.perf_counter() - t0:.1f}s (DE '
            f'generations mean {nit.mean():.0f} max {nit.max()})')
    return (np.concatenate(feats), np.concatenate(clss),
            np.concatenate(regs))


def train(feat, cls, reg, epochs=60, batch=512, lr=2e-3, seed=0,
          reg_w=1.0, log=print):
    net = VariogramProposalNet(seed=seed)
    hist = net.train(feat, cls, reg, epochs=epochs, batch=batch, lr=lr,
                     reg_w=reg_w, seed=seed,
                     log=lambda ep, v: log(f'epoch {ep} loss {v:.5f}'))
    return net, hist


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--fields', type=int, default=200000)
    ap.add_argument('--chunk', type=int, default=128)
    ap.add_argument('--fit-chunk', type=int, default=2048)
    ap.add_argument('--epochs', type=int, default=60)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--train-seed', type=int, default=0)
    ap.add_argument('--out', default='proposal_net.npz')
    ap.add_argument('--data-out', default=None)
    ap.add_argument('--data-in', default=None, nargs='+',
                    help='train on features/labels saved by --data-out '
                         '(several files are concatenated, so a large set '
                         'can be labelled in parts with distinct --seed)')
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    if a.data_in:
        parts = []
        for fn in a.data_in:
            with np.load(fn, allow_pickle=False) as z:
                parts.append((z['feat'], z['cls'], z['reg']))
        feat = np.concatenate([q[0] for q in parts])
        cls = np.concatenate([q[1] for q in parts])
        reg = np.concatenate([q[2] for q in parts])
    else:
        feat, cls, reg = build_training_set(
            a.fields, a.chunk, a.seed, fit_chunk=a.fit_chunk,
            log=lambda s: print(s, flush=True))
    t_data = time.perf_counter() - t0
    if a.data_out:
        np.savez_compressed(a.data_out, feat=feat, cls=cls, reg=reg)
    if a.epochs <= 0:
        print(f'labelled {feat.shape[0]} of {a.fields} fields in '
              f'{t_data:.1f}s; no training (--epochs 0)', flush=True)
        return
    t1 = time.perf_counter()
    net, hist = train(feat, cls, reg, epochs=a.epochs, seed=a.train_seed,
                      log=lambda s: print(s, flush=True))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    counts = np.bincount(cls, minlength=3)
    np.savez(a.out, params=net.state(), loss=np.asarray(hist),
             n_fields=a.fields, n_labelled=feat.shape[0],
             class_counts=counts, t_data=t_data, t_train=t_train)
    print(f'saved {a.out}: {feat.shape[0]} labelled of {a.fields}, classes '
          f'{dict(zip(MODEL_TYPES, counts.tolist()))}, data {t_data:.1f}s, '
          f'train {t_train:.1f}s', flush=True)


if __name__ == '__main__':
    main()
