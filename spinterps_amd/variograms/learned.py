"""Learned variogram proposal for the step fits of variograms/fitvgs.py.

A small network maps a step's lag-binned empirical variogram (40 features)
to a model type (Sph, Exp, Gau), a sill, a range and a nugget fraction.
The proposal seeds the same differential-evolution fit that
BatchVariogramFitter runs with the reference's rules (vgs.py:661-848), on
the proposed type only; the plain fit over every candidate stays the
fallback and, with ``verify``, replaces any step whose proposal-seeded fit
error exceeds the plain fit's by more than 2 %.

Everything numeric runs in the library's HIP kernels: field simulation
(FP64 Cholesky of the simulated covariance), features, the network's
forward and backward passes and Adam.  Torch only holds device memory.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from .._lib import check, lib
from ..interp.steps import DeviceProblem, vg_device
from .fitvgs import TYPE_ID, BatchVariogramFitter, _ctx, candidate_sets

MODEL_TYPES = ('Sph', 'Exp', 'Gau')
N_LAG = 20
N_FEAT = 2 * N_LAG
N_OUT = 6


# ---- synthetic fields --------------------------------------------------------
def simulate_fields(prob: DeviceProblem, mask_vals, vgs, means, seed=0):
    """Unconditional Gaussian fields [T x N] (NaN where mask_vals is NaN),
    mean[t] + L xi with L L^T the covariance of vgs[t] (FP64 Cholesky).
    Returns (fields, info): info[t] has SPK_ST_FAILED when the Cholesky
    failed (that row stays NaN)."""
    dev = prob.device
    mv = torch.as_tensor(np.ascontiguousarray(mask_vals, np.float64),
                         device=dev)
    T = mv.shape[0]
    assert mv.ndim == 2 and mv.shape[1] == prob.n_stns
    mu = torch.as_tensor(np.ascontiguousarray(means, np.float64), device=dev)
    assert mu.shape == (T,)
    vg_dev = vg_device(vgs, dev)
    out = torch.empty((T, prob.n_stns), dtype=torch.float64, device=dev)
    info = torch.zeros(T, dtype=torch.int32, device=dev)
    ctx = _ctx(dev)
    pst = prob.c_struct()
    check(lib().spk_simulate_fields_dev(
        ctx.h, ctypes.byref(pst), T, mv.data_ptr(), vg_dev.data_ptr(),
        mu.data_ptr(), int(seed) & ((1 << 64) - 1), out.data_ptr(),
        info.data_ptr()))
    return out, info


def features(fitter: BatchVariogramFitter, vals, mdr=0.5, nlag=N_LAG):
    """(features float32 [T x 2 nlag], scale float64 [T x 2] = (lag scale,
    variance)) of every step, on the device."""
    v = torch.as_tensor(vals, dtype=torch.float64, device=fitter.dev)
    v = v.contiguous()
    T = v.shape[0]
    feat = torch.zeros((T, 2 * nlag), dtype=torch.float32, device=fitter.dev)
    scale = torch.zeros((T, 2), dtype=torch.float64, device=fitter.dev)
    ctx = _ctx(fitter.dev)
    check(lib().spk_vg_features_dev(
        ctx.h, fitter.pi.data_ptr(), fitter.pj.data_ptr(),
        fitter.pd.data_ptr(), fitter.M, fitter.N, v.data_ptr(), T,
        float(mdr), int(nlag), feat.data_ptr(), scale.data_ptr()))
    return feat, scale


# ---- network -------------------------------------------------------------------
class VariogramProposalNet:
    """40 -> 128 -> 128 -> 128 -> 6 perceptron (GELU), float32, trained and
    run by the library's kernels."""

    def __init__(self, device=0, seed=0, params=None):
        if not torch.cuda.is_available():
            raise RuntimeError('a ROCm GPU is required')
        self.dev = torch.device('cuda', device)
        self.n_params = int(lib().spk_mlp_n_params())
        self.params = torch.empty(self.n_params, dtype=torch.float32,
                                  device=self.dev)
        if params is None:
            check(lib().spk_mlp_init_dev(_ctx(self.dev).h,
                                         int(seed) & ((1 << 64) - 1),
                                         self.params.data_ptr()))
        else:
            p = np.asarray(params, np.float32).reshape(-1)
            assert p.size == self.n_params
            self.params.copy_(torch.from_numpy(p))
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grads = torch.zeros_like(self.params)
        self.step = 0

    def forward(self, X):
        X = self._x(X)
        out = torch.empty((X.shape[0], N_OUT), dtype=torch.float32,
                          device=self.dev)
        check(lib().spk_mlp_forward_dev(_ctx(self.dev).h,
                                        self.params.data_ptr(), X.data_ptr(),
                                        X.shape[0], out.data_ptr()))
        return out

    def loss_grad(self, X, cls, reg, reg_w=1.0):
        """Loss (device [1]) and gradients (device, layout of params)."""
        X = self._x(X)
        cls = torch.as_tensor(cls, dtype=torch.int32,
                              device=self.dev).contiguous()
        reg = torch.as_tensor(reg, dtype=torch.float32,
                              device=self.dev).contiguous()
        assert cls.shape == (X.shape[0],) and reg.shape == (X.shape[0], 3)
        loss = torch.empty(1, dtype=torch.float32, device=self.dev)
        check(lib().spk_mlp_loss_grad_dev(
            _ctx(self.dev).h, self.params.data_ptr(), X.data_ptr(),
            X.shape[0], cls.data_ptr(), reg.data_ptr(), float(reg_w),
            self.grads.data_ptr(), loss.data_ptr()))
        return loss, self.grads

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        check(lib().spk_mlp_adam_dev(
            _ctx(self.dev).h, self.params.data_ptr(), self.grads.data_ptr(),
            self.m.data_ptr(), self.v.data_ptr(), self.step, float(lr),
            float(b1), float(b2), float(eps)))

    def train(self, X, cls, reg, epochs=100, batch=256, lr=1e-3, reg_w=1.0,
              seed=0, lr_final=None, log=None):
        """Mini-batch Adam over (X, cls, reg); returns the mean loss of each
        epoch.  The learning rate decays by cosine to lr_final."""
        X = self._x(X)
        cls = torch.as_tensor(cls, dtype=torch.int32, device=self.dev)
        reg = torch.as_tensor(reg, dtype=torch.float32, device=self.dev)
        n = X.shape[0]
        lr_final = lr * 0.05 if lr_final is None else lr_final
        g = torch.Generator(device='cpu')
        g.manual_seed(int(seed))
        hist = []
        n_b = max(1, n // batch)
        total = epochs * n_b
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            acc = torch.zeros(1, dtype=torch.float32, device=self.dev)
            for b in range(n_b):
                ix = perm[b * batch:(b + 1) * batch]
                loss, _ = self.loss_grad(X[ix].contiguous(),
                                         cls[ix].contiguous(),
                                         reg[ix].contiguous(), reg_w)
                acc += loss
                lr_k = lr_final + 0.5 * (lr - lr_final) * (
                    1.0 + math.cos(math.pi * k / total))
                self.adam_step(lr_k)
                k += 1
            hist.append(float(acc) / n_b)
            if log is not None:
                log(ep, hist[-1])
        return hist

    def propose(self, feat, scale):
        """Per step (type index into MODEL_TYPES, total sill, range, nugget
        fraction) as host arrays."""
        y = self.forward(feat).double().cpu().numpy()
        sc = scale.cpu().numpy() if isinstance(scale, torch.Tensor) else \
            np.asarray(scale)
        typ = np.argmax(y[:, :3], axis=1)
        sill = sc[:, 1] * np.exp(np.clip(y[:, 3], -10, 10))
        rng = sc[:, 0] * np.exp(np.clip(y[:, 4], -20, 5))
        nugf = np.clip(y[:, 5], 0.0, 1.0)
        return typ, sill, rng, nugf

    def state(self):
        return self.params.cpu().numpy().copy()

    def _x(self, X):
        X = torch.as_tensor(X, dtype=torch.float32, device=self.dev)
        X = X.contiguous()
        assert X.ndim == 2 and X.shape[1] == N_FEAT
        return X


# ---- labels from fits ----------------------------------------------------------
def labels_from_fits(best, scale):
    """Training targets from BatchVariogramFitter.fit details: per fitted
    step the class of the best fit's structured component and (log sill /
    variance, log range / lag scale, nugget fraction).  Steps whose best
    fit has no Sph/Exp/Gau component get class -1."""
    T = len(best)
    cls = np.full(T, -1, dtype=np.int32)
    reg = np.full((T, 3), np.nan, dtype=np.float32)
    for k, b in enumerate(best):
        if not b:
            continue
        names, rngs, sills = b[-1][0], b[-1][1], b[-1][2]
        st = [i for i, nm in enumerate(names) if nm in MODEL_TYPES]
        if len(st) != 1:
            continue
        i = st[0]
        tot = float(np.sum(sills))
        nug = float(np.sum([s for nm, s in zip(names, sills) if nm == 'Nug']))
        if not (tot > 0 and scale[k, 1] > 0 and scale[k, 0] > 0):
            continue
        cls[k] = MODEL_TYPES.index(names[i])
        reg[k] = (math.log(tot / scale[k, 1]), math.log(rngs[i] / scale[k, 0]),
                  nug / tot)
    return cls, reg


# ---- proposal-seeded fit -------------------------------------------------------
def _best_of(o_list, par, fun, npts, fit_thresh=0.01):
    """AIC choice of vgs.py:779-848 over the fitted candidates of one step:
    (names, ranges, sills, objective) of the best, or None."""
    best = None
    min_obj = np.inf
    skip = []
    for c, o in enumerate(o_list):
        if o in skip:
            continue
        D = 2 * len(o)
        x = par[c, :D]
        rngs, sills = x[0::2].copy(), x[1::2].copy()
        with np.errstate(divide='ignore', invalid='ignore'):
            aic = npts * np.log(fun[c]) + 2 * D
        c1 = aic < min_obj * (1.0 - fit_thresh)
        names = list(o)
        if not np.all(np.ediff1d(rngs) >= 0):
            srt = np.argsort(rngs)
            rngs, sills = rngs[srt], sills[srt]
            skip.append(tuple(np.array(o)[srt]))
            names = list(np.array(o)[srt])
        if c1 and np.all(np.ediff1d(rngs) >= 0):
            min_obj = aic
            best = (names, rngs, sills, float(fun[c]))
    return best


def _vg_str(b):
    names, rr, ss = b[0], b[1], b[2]
    return ' + '.join('%0.5f %s(%0.5f)' % (s, nm, r)
                      for nm, r, s in zip(names, rr, ss))


class LearnedVariogramFitter:
    """Proposal-seeded variogram fits of many steps over one station set."""

    def __init__(self, stn_x, stn_y, net: VariogramProposalNet, device=0,
                 fitter: BatchVariogramFitter = None):
        self.fitter = fitter or BatchVariogramFitter(stn_x, stn_y, device)
        self.net = net
        self.dev = self.fitter.dev

    def fit(self, vals, mdr=0.5, nk=50, fil_nug_vg='Nug',
            fit_vgs=MODEL_TYPES, perm_r_list=(1, 2), fit_thresh=0.01,
            opt_iters=10000, evg='robust', min_vld_stns=10, range_ub=1e9,
            seed=0, popmul=10, spread=1.0, verify=True, tol_ratio=1.02,
            return_details=False):
        """Best variogram string per step (None for a step with too few
        values or no spread).

        The network proposes a type t; the candidates (t) and
        (fil_nug_vg, t) are fitted by differential evolution seeded at the
        proposal (population popmul * D^2 in a log-box of half-width
        ``spread`` around it, then the reference's bounds and mutation).
        With ``verify`` the plain fit of every candidate in fit_vgs also
        runs and is kept for each step whose seeded fit error exceeds
        tol_ratio times the plain one.
        """
        f = self.fitter
        vals = np.array(vals, dtype=np.float64, copy=True)
        T = vals.shape[0]
        ok = np.zeros(T, dtype=bool)
        npts = np.zeros(T, dtype=np.int64)
        var = np.zeros(T)
        for t in range(T):
            z = vals[t][~np.isnan(vals[t])]
            npts[t] = z.size
            if z.size < min_vld_stns or np.isclose(z.std(), 0.0):
                continue
            ok[t] = True
            var[t] = np.var(z)
        steps = np.where(ok)[0]
        out = [None] * T
        det = dict(steps=steps, learned_fun=None, plain_fun=None,
                   fallback=None, types=None)
        if steps.size == 0:
            return (out, det) if return_details else out
        sub = vals[steps]
        t0 = time.perf_counter()
        feat, scale = features(f, sub, mdr)
        typ, sill, rng, nugf = self.net.propose(feat, scale)
        bins = f.empirical(sub, mdr, nk, evg)
        bh = bins[0][:, 0].cpu().numpy()
        bg = bins[1][:, 0].cpu().numpy()
        # two candidates per step: (t) and (nugget, t), seeded per step
        S = steps.size
        cand_l = [[(MODEL_TYPES[k],), (fil_nug_vg, MODEL_TYPES[k])]
                  for k in typ]
        init = np.zeros((S, 2, 8))
        init[:, 0, 0] = rng
        init[:, 0, 1] = sill
        init[:, 1, 0] = bh
        init[:, 1, 1] = bg
        init[:, 1, 2] = rng
        init[:, 1, 3] = np.maximum(sill - bg, 1e-5)
        par = np.zeros((S, 2, 8))
        fun = np.zeros((S, 2))
        nit = np.zeros((S, 2), dtype=np.int64)
        # one launch per proposed type (candidate specs are per launch)
        for k, name in enumerate(MODEL_TYPES):
            sel = np.where(typ == k)[0]
            if sel.size == 0:
                continue
            b_sel = tuple(x[torch.as_tensor(sel, device=self.dev)]
                          .contiguous() for x in bins)
            p_, f_, n_ = self._de_seeded(
                b_sel, var[steps][sel], [(name,), (fil_nug_vg, name)],
                fil_nug_vg, init[sel], opt_iters, popmul, range_ub,
                seed + k, spread)
            par[sel], fun[sel], nit[sel] = p_, f_, n_
        learned = [_best_of(cand_l[i], par[i], fun[i], npts[steps[i]],
                            fit_thresh) for i in range(S)]
        t_learned = time.perf_counter() - t0
        lfun = np.array([b[3] if b else np.inf for b in learned])
        final = list(learned)
        fallback = np.zeros(S, dtype=bool)
        pfun = None
        if verify:
            t1 = time.perf_counter()
            _, pdet = f.fit(vals, mdr=mdr, nk=nk, perm_r_list=perm_r_list,
                            fit_vgs=fit_vgs, fil_nug_vg=fil_nug_vg,
                            fit_thresh=fit_thresh, opt_iters=opt_iters,
                            evg=evg, min_vld_stns=min_vld_stns,
                            range_ub=range_ub, seed=seed + 7,
                            return_details=True)
            pfun = np.full(S, np.inf)
            pbest = [None] * S
            pos = {int(t): i for i, t in enumerate(pdet['steps'])}
            for i, t in enumerate(steps):
                j = pos.get(int(t))
                if j is not None and pdet['best'][j]:
                    b = pdet['best'][j][-1]
                    pbest[i] = (b[0], b[1], b[2], float(b[3]))
                    pfun[i] = b[3]
            fallback = lfun > tol_ratio * pfun
            for i in np.where(fallback)[0]:
                final[i] = pbest[i]
            det['t_plain'] = time.perf_counter() - t1
            det['plain_nit'] = pdet['nit']
            det['plain_best'] = pbest
            det['plain_cands'] = pdet['cands']
        for i, t in enumerate(steps):
            if final[i] is not None:
                out[t] = _vg_str(final[i])
        det.update(learned_fun=lfun, plain_fun=pfun, fallback=fallback,
                   types=typ, nit=nit, t_learned=t_learned,
                   proposal=(typ, sill, rng, nugf))
        return (out, det) if return_details else out

    def _de_seeded(self, bins, var, cands, fil_nug_vg, init, opt_iters,
                   popmul, range_ub, seed, spread):
        bh, bg, bn, nb = bins
        T, max_bins = bh.shape
        spec = np.zeros((len(cands), 6), dtype=np.int32)
        for c, o in enumerate(cands):
            spec[c, 0] = len(o)
            for k, nm in enumerate(o):
                spec[c, 1 + k] = TYPE_ID[nm]
            spec[c, 5] = 1 if (len(o) > 1 and o[0] == fil_nug_vg) else 0
        dev = self.dev
        var_d = torch.as_tensor(np.asarray(var, np.float64), device=dev)
        init_d = torch.as_tensor(np.ascontiguousarray(init, np.float64),
                                 device=dev)
        par = torch.zeros((T, len(cands), 8), dtype=torch.float64, device=dev)
        fun = torch.zeros((T, len(cands)), dtype=torch.float64, device=dev)
        nit = torch.zeros((T, len(cands)), dtype=torch.int32, device=dev)
        check(lib().spk_vgfit_de_seeded_dev(
            _ctx(dev).h, bh.data_ptr(), bg.data_ptr(), nb.data_ptr(),
            max_bins, var_d.data_ptr(), T,
            spec.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(cands),
            int(opt_iters), int(popmul), 0.01, float(range_ub),
            int(seed) & ((1 << 64) - 1), init_d.data_ptr(), float(spread),
            par.data_ptr(), fun.data_ptr(), nit.data_ptr()))
        return (par.cpu().numpy(), fun.cpu().numpy(),
                nit.cpu().numpy().astype(np.int64))


# ---- training-set generation -----------------------------------------------------
def random_vgs(rng, T, var=1.0, range_lo=5e3, range_hi=8e4, nug_hi=0.3):
    """Random Sph/Exp/Gau variogram strings with a nugget fraction in
    [0, nug_hi) and the true (type, range, nugget fraction) of each."""
    typ = rng.integers(0, 3, T)
    rr = rng.uniform(range_lo, range_hi, T)
    nf = rng.uniform(0.0, nug_hi, T)
    vgs = [f'{float(nf[t] * var)!r} Nug(0.0) + {float((1 - nf[t]) * var)!r} '
           f'{MODEL_TYPES[typ[t]]}({float(rr[t])!r})' for t in range(T)]
    return vgs, typ, rr, nf


def training_fields(prob: DeviceProblem, T, seed=0, n_lo=300, n_hi=1500,
                    **kw):
    """T simulated fields over prob's stations, each at a random subset of
    n_lo..n_hi stations with a random variogram (random_vgs)."""
    rng = np.random.default_rng(seed)
    N = prob.n_stns
    mask = np.full((T, N), np.nan)
    for t in range(T):
        n = int(rng.integers(n_lo, min(n_hi, N) + 1))
        mask[t, rng.choice(N, n, replace=False)] = 0.0
    vgs, typ, rr, nf = random_vgs(rng, T, **kw)
    fields, info = simulate_fields(prob, mask, vgs, np.zeros(T), seed)
    return fields, info, dict(vgs=vgs, type=typ, range=rr, nug=nf)


def default_net_path():
    from pathlib import Path
    return Path(__file__).with_name('proposal_net.npz')


def load_proposal_net(path=None, device=0):
    """The trained network shipped with the package (or the npz at path,
    as written by train_proposal)."""
    with np.load(path or default_net_path(), allow_pickle=False) as z:
        return VariogramProposalNet(device=device, params=z['params'])
