#!/usr/bin/env python3
"""Headline benchmark: ordinary-kriging cell-estimates/s including the
estimation variance, batched mixed-precision path vs the plain FP64 path.

Workload (hourly): 2000 stations over 500 km x 500 km, a 1000 x 1000 grid at
500 m, per step a random 200-2000 active stations and its own variogram
'<f v> Nug(0.0) + <(1-f) v> {Sph|Exp|Gau}(5-80 km)'; lognormal values.
One bench step = one batch of --batch time steps over the full grid through
the batched mixed path.  Stations and cells are resident in HBM; the timed
region copies each step's values and variogram record to the device and
both fields back as float32 (the reference's output type).  Every timed
step is checked against the batched FP64 path on sampled cells, and a
spread of them against the exact solution of the reference's system.
Synthetic seeded data.

Prints ONE JSON line (rank 0, N = 1).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


# HBM bytes per dispatch of the variance contractions, measured by
# rocprofv3 --pmc (FETCH_SIZE x2 + WRITE_SIZE x1, calibrated with
# k_clip_cast).  The split-BF16 and int8 figures come from the hourly
# workload with Sph/Exp models and the FP32-tier limit at 1e3 (--models
# Sph,Exp --fp32-limit 1000: both contractions, no pseudo-inverse steps,
# whose rocBLAS calls end counter runs in a host fault of the profiler),
# batches of 64 steps, 34 stripes per batch; the int8 figure is its 5-plane
# launch (no step of that run needed six planes).  The FP64 figure is from
# a run in which k_tri64 contracted every step above the FP32 limit
# (stripes of 10 880 cells).
TRI64_TRAFFIC = 2 * 4537366 * 1024 + 22939 * 1024
TRI32_TRAFFIC = 2 * 5403411 * 1024 + 202497 * 1024
TRI8_TRAFFIC = 2 * 14085710 * 1024 + 171559 * 1024
TRIBF_TRAFFIC = 2 * 1603109 * 1024 + 115980 * 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=64,
                    help='time steps per bench step')
    ap.add_argument('--plain-steps', type=int, default=16)
    ap.add_argument('--grid', type=int, default=1000)
    ap.add_argument('--models', default='Sph,Exp,Gau',
                    help='variogram models the hourly steps draw from '
                         '(the workload: Sph,Exp,Gau)')
    ap.add_argument('--cpu-cells', type=int, default=200)
    ap.add_argument('--gate-cells', type=int, default=4096,
                    help='sampled cells of every timed step compared with '
                         'the batched FP64 path')
    ap.add_argument('--gate-oracle-steps', type=int, default=12,
                    help='timed steps (spread) compared with the exact '
                         'oracle, besides the 4 unflagged steps of largest '
                         'condition estimate')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--vg-fields', type=int, default=512,
                    help='simulated fields for the variogram-fit leg '
                         '(0 skips it)')
    ap.add_argument('--cpu-vg-fields', type=int, default=4)
    ap.add_argument('--idw-steps', type=int, default=365,
                    help='IDW leg: daily steps (1000 stations, 500 x 500 '
                         'cells at 1 km, p = 2); 0 skips it')
    ap.add_argument('--cpu-idw-cells', type=int, default=2000)
    ap.add_argument('--cpu-idw-steps', type=int, default=8)
    ap.add_argument('--neb-steps', type=int, default=365,
                    help='daily steps of the nrst neighbourhood leg '
                         '(0 = skip)')
    ap.add_argument('--neb-nebs', type=int, default=50)
    ap.add_argument('--cpu-neb-steps', type=int, default=2)
    ap.add_argument('--cpu-neb-side', type=int, default=60,
                    help='side of the cell window the oracle runs on')
    ap.add_argument('--stress', type=int, default=0,
                    help='also run this many steps of the ill-conditioned '
                         'stress set (daily inputs, Gaussian model, nugget '
                         '1e-6, 10 %% of stations in 1-10 m pairs) and '
                         'report errors vs numpy.linalg.solve')
    ap.add_argument('--stress-cells', type=int, default=200)
    ap.add_argument('--e2e', type=int, default=256,
                    help='also run this many hourly steps end to end: '
                         'proposal-seeded variogram fits of the step '
                         'values, then batched mixed OK with the fitted '
                         'variograms')
    ap.add_argument('--stress-oracle-steps', type=int, default=20)
    ap.add_argument('--precision', default='mixed',
                    choices=('mixed', 'mixed_f32'),
                    help='mixed: the well-conditioned steps\' variance '
                         'contraction on split-BF16 operands; mixed_f32: '
                         'on FP32 operands')
    ap.add_argument('--fp32-limit', type=float, default=0.0,
                    help='condition estimate above which a step\'s variance '
                         'leaves the split-BF16 / FP32 contraction for the '
                         'int8 planes, for every model (0: the library\'s '
                         'policy, 1e5 and 1e3 with a Gaussian component); '
                         'with --models Sph,Exp and 1000 a counter run gets '
                         'both contractions without Gaussian steps, whose '
                         'pseudo-inverse fallback calls rocBLAS')
    ap.add_argument('--calib', action='store_true',
                    help='also run k_clip_cast over 2^28 float64 values '
                         '(2 GiB read, 1 GiB written) as the reference '
                         'kernel for HBM counter scaling')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU: the library has no multi-GPU '
              'path (time steps would shard across GPUs with no exchange '
              'during compute).  Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    import multiprocessing as mp

    import numpy as np
    # CPU baseline workers, forked before anything touches the GPU: one
    # worker per core of this process's share (at most 16), each with one
    # BLAS thread, as the reference runs its chunks (interp/main.py:84-98)
    n_cores = min(16, len(os.sched_getaffinity(0)))
    pool = None if args.no_cpu else mp.get_context('fork').Pool(n_cores)
    import torch

    from spinterps_amd import _lib, configs
    from spinterps_amd.interp.steps import (DeviceProblem, clip_cast,
                                            interp_steps, prof_enable,
                                            prof_get, prof_work, vg_device)

    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    dev = torch.device('cuda', 0)
    B, K, W, P = args.batch, args.steps, args.warmup, args.plain_steps
    w = configs.hourly(n_side=args.grid, cell=5e5 / args.grid)
    C, N = w.n_cells, w.n_stns
    log(f'hourly workload: N={N} C={C} batch={B} steps={K} '
        f'warmup={W} plain_steps={P}')
    n_tot = (W + K) * B
    t = time.time()
    models = tuple(args.models.split(','))
    vals, vgs = configs.hourly_steps(w, 0, n_tot + P, models=models)
    log(f'generated {n_tot + P} steps ({"/".join(models)}) in '
        f'{time.time() - t:.1f}s')
    prob = DeviceProblem.from_workload(w, device=0)
    vrec = _lib.VG_DTYPE.itemsize
    n_act = (~np.isnan(vals)).sum(1)
    # per-step inputs on the host, pinned: the values [T x N] and the parsed
    # variogram records; each run copies its steps to the device and the
    # fields back as the reference's float32 (interp/steps.py:815-831)
    vals_h = torch.from_numpy(vals).pin_memory()
    vg_h = vg_device(vgs, torch.device('cpu')).pin_memory()
    rows = max(K, W) * B
    vals_d = torch.empty((rows, N), dtype=torch.float64, device=dev)
    vg_dd = torch.empty(rows * vrec, dtype=torch.uint8, device=dev)
    est = torch.empty((rows, C), dtype=torch.float64, device=dev)
    var = torch.empty((rows, C), dtype=torch.float64, device=dev)
    e32_h = torch.empty((rows, C), dtype=torch.float32, pin_memory=True)
    v32_h = torch.empty((rows, C), dtype=torch.float32, pin_memory=True)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)

    def run(s0, s1, kap=None):
        """Steps s0..s1 as one call, as a user interpolates a series (batch
        b+1's systems are factorised while batch b's cells contract):
        H2D of the per-step inputs, OK estimate + variance, float32 cast,
        D2H of both fields."""
        m = s1 - s0
        vals_d[:m].copy_(vals_h[s0:s1], non_blocking=True)
        vg_dd[:m * vrec].copy_(vg_h[s0 * vrec:s1 * vrec], non_blocking=True)
        ev0.record()
        # the float32 fields reach the pinned host arrays batch by batch
        # while later batches compute (spk_interp_steps_f32_dev)
        _, _, info = interp_steps(prob, vals_d[:m], None, 'OK',
                                  precision=args.precision, est=est[:m],
                                  var=var[:m], fp64var_kappa=args.fp32_limit,
                                  vg_dev=vg_dd[:m * vrec], kappa=kap,
                                  out32=(e32_h[:m], v32_h[:m]))
        ev1.record()
        return info

    # ---- warmup + timed region (mixed) --------------------------------
    if W:
        run(0, W * B)
        torch.cuda.synchronize()
        log(f'warmup: {W} x {B} steps done')
    TK = K * B
    kap_t = torch.zeros(TK, dtype=torch.float64, device=dev)
    prof_enable(0, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info_d = run(W * B, n_tot, kap_t)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / K
    value = B * C / ((t1 - t0) / K)
    dev_ms = ev0.elapsed_time(ev1)
    log(f'mixed: {ms_per_step:.1f} ms per batch of {B} steps incl. H2D of '
        f'the inputs and D2H of the float32 fields, {value:.3e} '
        f'cell-estimates/s; device only {dev_ms / K:.1f} ms per batch')
    kern = {}
    for nm in ('nearest', 'kgen', 'gemm_tri_bf16', 'gemm_tri_f32',
               'gemm_tri_f64', 'slice_cols', 'gemm_tri_i8', 'var_reduce',
               'potrf', 'trtri', 'kappa', 'dual', 'slice_rows', 'assemble',
               'special',
               'cast_tri', 'split_tri', 'pinv', 'pinv_contract'):
        ms, cnt = prof_get(nm)
        kern[nm] = (ms, cnt)
    macs32 = prof_work('gemm_tri_f32_macs')
    macsbf = prof_work('gemm_tri_bf16_macs')
    ops8x = prof_work('gemm_tri_i8_ops')
    prof_enable(0, False)
    info_t = info_d.cpu().numpy()
    kap_np = kap_t.cpu().numpy()
    timed_n = n_act[W * B:n_tot]
    special_t = (info_t & (_lib.ST_NOSTNS | _lib.ST_MEANFILL |
                           _lib.ST_ONESTN)) != 0
    failed_t = (info_t & _lib.ST_FAILED) != 0
    ill_t = (info_t & _lib.ST_ILLCOND) != 0
    solved_t = ~special_t & ~failed_t
    pinv_t = (info_t & _lib.ST_PINV) != 0
    # steps the batched contractions ran (not the pseudo-inverse fallback)
    batched_t = solved_t & ~pinv_t
    f64_steps = batched_t & ((info_t & _lib.ST_FP64VAR) != 0)
    i8_steps = batched_t & ~f64_steps & ((info_t & _lib.ST_INT8VAR) != 0)
    f32_steps = batched_t & ~f64_steps & ~i8_steps
    # algorithmic flops of W = L^-1 delta: n^2 per cell (n(n+1)/2 MACs)
    nsq = timed_n.astype(np.float64) ** 2
    fl32 = float(np.sum(nsq[f32_steps]) * C)
    fl64 = float(np.sum(nsq[f64_steps]) * C)
    fl8 = float(np.sum(nsq[i8_steps]) * C)
    ms32, c32 = kern['gemm_tri_f32']
    msbf, cbf = kern['gemm_tri_bf16']
    ms64, c64 = kern['gemm_tri_f64']
    ms8, c8 = kern['gemm_tri_i8']
    # the FP32 contraction skips the zero chunks of delta: its MFMA rate
    # counts the multiply-adds it executed, its algorithmic rate n^2 C
    ach32 = 2 * macs32 / (ms32 * 1e-3) / 1e12 if ms32 > 0 else 0.0
    alg32 = fl32 / (ms32 * 1e-3) / 1e12 if ms32 > 0 else 0.0
    # split-BF16 contraction of the same steps: its BF16 MFMA rate counts
    # the executed plane products (six per 32-row chunk), its algorithmic
    # rate n^2 C (the FP32-equivalent L^-1 delta)
    achbf = 2 * macsbf / (msbf * 1e-3) / 1e12 if msbf > 0 else 0.0
    algbf = fl32 / (msbf * 1e-3) / 1e12 if msbf > 0 else 0.0
    ach64 = fl64 / (ms64 * 1e-3) / 1e12 if ms64 > 0 else 0.0
    # int8: it skips the zero chunks of delta too, so its MFMA rate counts
    # the plane products executed; the algorithmic count is 15 plane
    # products of the triangle per step up to condition estimate 1e7, 21
    # above, and the FP64-equivalent rate that of the same L^-1 delta
    prods = np.where(kap_np <= 1e7, 15.0, 21.0)
    ops8 = float(np.sum((prods * nsq)[i8_steps]) * C)
    ach8 = ops8x / (ms8 * 1e-3) / 1e12 if ms8 > 0 else 0.0
    alg8 = ops8 / (ms8 * 1e-3) / 1e12 if ms8 > 0 else 0.0
    eq8 = fl8 / (ms8 * 1e-3) / 1e12 if ms8 > 0 else 0.0
    # kgen: algorithmic HBM bytes = delta written (FP32-tier steps: 4 B,
    # or 6 B of bfloat16 planes; others 8 B) + cell coords
    kb = float(np.sum(timed_n[f32_steps]) * C *
               (6 if args.precision == 'mixed' else 4) +
               np.sum(timed_n[f64_steps | i8_steps]) * C * 8 +
               16 * C * TK)
    mskg, ckg = kern['kgen']
    kg_gbs = kb / (mskg * 1e-3) / 1e9 if mskg > 0 else 0.0
    log('kernel totals (ms, launches): ' +
        ', '.join(f'{k}={v[0]:.1f}/{v[1]}' for k, v in kern.items()))

    # ---- accuracy gate on the timed steps ------------------------------
    # (b) every timed step against the batched FP64 path (FP64 contraction,
    #     the same dual-form estimate) on sampled cells
    rng = np.random.default_rng(5)
    gidx = np.sort(rng.choice(C, args.gate_cells, replace=False))
    gidx_d = torch.as_tensor(gidx, device=dev)
    probs = DeviceProblem(w.stn_x, w.stn_y, w.cell_x[gidx], w.cell_y[gidx],
                          origin=prob.origin)
    vals_d[:TK].copy_(vals_h[W * B:n_tot])
    vg_dd[:TK * vrec].copy_(vg_h[W * B * vrec:n_tot * vrec])
    eb, vb, infob = interp_steps(probs, vals_d[:TK], None, 'OK',
                                 precision='fp64b', vg_dev=vg_dd[:TK * vrec])
    em_s = est[:TK].index_select(1, gidx_d).cpu().numpy()
    vm_s = var[:TK].index_select(1, gidx_d).cpu().numpy()
    eb, vb = eb.cpu().numpy(), vb.cpu().numpy()

    def rel(a, b):
        # relative error, floor 1e-3 x the mean |reference|; equal values
        # count as 0 (steps whose variances are all 0)
        f = 1e-3 * np.abs(b).mean(axis=-1, keepdims=True)
        d = np.abs(a - b)
        r = d / np.maximum(np.abs(b), f)
        return np.where(d == 0, 0.0, r).max(axis=-1)

    with np.errstate(invalid='ignore', divide='ignore'):
        re_b, rv_b = rel(em_s, eb), rel(vm_s, vb)
    ok_b = solved_t & ~ill_t
    fl_b = solved_t & ill_t
    gate_b = {'steps_compared': int(ok_b.sum()), 'cells': int(len(gidx)),
              'est_max': float(re_b[ok_b].max()) if ok_b.any() else None,
              'var_max': float(rv_b[ok_b].max()) if ok_b.any() else None,
              'flagged_steps': int(fl_b.sum()),
              'flagged_est_max': float(re_b[fl_b].max()) if fl_b.any()
              else None,
              'flagged_var_max': float(rv_b[fl_b].max()) if fl_b.any()
              else None,
              'nan_variance_steps_mixed': int(np.isnan(vm_s[solved_t])
                                              .any(axis=1).sum()),
              'nan_variance_steps_fp64b': int(np.isnan(vb[solved_t])
                                              .any(axis=1).sum()),
              'pinv_steps': int(((info_t & _lib.ST_PINV) != 0).sum())}
    log(f'gate (mixed vs batched FP64, all timed steps): {gate_b}')
    # (a) a spread of timed steps, and the unflagged ones of largest
    #     condition estimate, against the exact solution of the reference's
    #     system (oracle.krige_exact) on the first 100 sampled cells; flagged
    #     steps of largest condition estimate are reported with the
    #     reference's own pinv error beside the library's
    from oracle import spk_oracle as O
    oc = gidx[:100]
    cand = np.flatnonzero(solved_t)
    pick = list(cand[np.linspace(0, len(cand) - 1, args.gate_oracle_steps)
                     .astype(int)]) if len(cand) else []
    unfl = cand[~ill_t[cand]]
    pick += list(unfl[np.argsort(kap_np[unfl])[-4:]])
    flg = cand[ill_t[cand]]
    srt = flg[np.argsort(kap_np[flg])]
    pick_ill = sorted(set(list(srt[:2]) + list(srt[-2:])))
    gate_a, ill_rep = [], []
    for t in sorted(set(pick)) + pick_ill:
        st = W * B + t
        act = ~np.isnan(vals[st])
        xs, ys, z = w.stn_x[act], w.stn_y[act], vals[st][act]
        rr = O.dists_2d(xs, ys, xs, ys)
        dr = O.dists_2d(w.cell_x[oc], w.cell_y[oc], xs, ys)
        ex, vx, _ = O.krige_exact('OK', vgs[st], rr, dr, z, want_var=True,
                                  its=1)
        ge = est[t].index_select(0, gidx_d[:100]).cpu().numpy()
        gv = var[t].index_select(0, gidx_d[:100]).cpu().numpy()
        row = {'step': int(st), 'n': int(act.sum()),
               'kappa': float(kap_np[t]), 'flags': int(info_t[t]),
               'est_err': float(rel(ge, ex)), 'var_err': float(rel(gv, vx))}
        if ill_t[t]:
            K_, R_ = O.krig_system('OK', vgs[st], rr, dr, None, None, 0.0)
            ep, vp, _ = O.krige_contract(np.linalg.pinv(K_), R_, z, dr, True)
            row.update({'pinv_est_err': float(rel(ep, ex)),
                        'pinv_var_err': float(rel(vp, vx))})
            ill_rep.append(row)
        else:
            gate_a.append(row)
    ea = max((r['est_err'] for r in gate_a), default=0.0)
    va = max((r['var_err'] for r in gate_a), default=0.0)
    log(f'gate (mixed vs exact, {len(gate_a)} timed steps): est {ea:.2e} '
        f'var {va:.2e}; flagged: {ill_rep}')
    gate_pass = (ea <= 1e-6 and va <= 1e-6 and
                 (gate_b['est_max'] or 0.0) <= 1e-6 and
                 (gate_b['var_max'] or 0.0) <= 1e-6)
    gate = {'passed': bool(gate_pass), 'bound': 1e-6,
            'floors': '1e-3 x the step\'s mean |estimate| and mean variance',
            'timed_steps': int(TK), 'flagged_ill_conditioned': int(ill_t.sum()),
            'failed_nearest_fill': int(failed_t.sum()),
            'special_steps': int(special_t.sum()),
            'vs_batched_fp64_all_timed_steps': gate_b,
            'vs_exact_oracle': {'steps': len(gate_a), 'cells': int(len(oc)),
                                'est_max': ea, 'var_max': va,
                                'per_step': gate_a},
            'flagged_steps_vs_exact': ill_rep}

    # ---- plain FP64 path on P steps (the ratio's denominator) ----------
    # P of the timed steps, evenly spaced, so that the two rates are taken
    # over the same mix of station counts and models
    psel = W * B + np.arange(P) * max(1, TK // P)
    vals_d[:P].copy_(vals_h[psel])
    vg_dd[:P * vrec].copy_(vg_h.view(-1, vrec)[psel].reshape(-1))
    pv, pg = vals_d[:P], vg_dd[:P * vrec]
    e64 = est[:P]
    v64 = var[:P]
    # untimed pass over the same steps: warms the plain kernels and lets
    # the tile autotuner time its variants for every size class here
    interp_steps(prob, pv, None, 'OK', precision='fp64', est=e64, var=v64,
                 vg_dev=pg)
    prof_enable(0, True)
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    _, _, pinfo = interp_steps(prob, pv, None, 'OK', precision='fp64',
                               est=e64, var=v64, vg_dev=pg,
                               out32=(e32_h[:P], v32_h[:P]))
    torch.cuda.synchronize()
    tp1 = time.perf_counter()
    msp, cp = prof_get('gemm_plain_f64')
    prof_enable(0, False)
    plain_rate = P * C / (tp1 - tp0)
    log(f'plain fp64: {(tp1 - tp0) * 1e3 / P:.1f} ms per step, '
        f'{plain_rate:.3e} cell-estimates/s')
    pn = n_act[psel].astype(np.float64)
    flp = float(np.sum(2 * (pn + 1) ** 2 * C))
    achp = flp / (msp * 1e-3) / 1e12 if msp > 0 else 0.0
    pi = pinfo.cpu().numpy()
    plain_err = []
    for s in range(P):
        if pi[s] & (_lib.ST_NOSTNS | _lib.ST_MEANFILL | _lib.ST_ONESTN |
                    _lib.ST_FAILED):
            continue
        st = int(psel[s])
        act = ~np.isnan(vals[st])
        xs, ys, z = w.stn_x[act], w.stn_y[act], vals[st][act]
        rr = O.dists_2d(xs, ys, xs, ys)
        dr = O.dists_2d(w.cell_x[oc], w.cell_y[oc], xs, ys)
        ex, vx, _ = O.krige_exact('OK', vgs[st], rr, dr, z, want_var=True,
                                  its=1)
        ge = e64[s].index_select(0, gidx_d[:100]).cpu().numpy()
        gv = v64[s].index_select(0, gidx_d[:100]).cpu().numpy()
        plain_err.append({'step': int(st), 'flags': int(pi[s]),
                          'est_err': float(rel(ge, ex)),
                          'var_err': float(rel(gv, vx))})
    pe = [r for r in plain_err if not r['flags'] & _lib.ST_ILLCOND]
    plain_gate = {'steps': len(pe),
                  'flagged': len(plain_err) - len(pe),
                  'est_max': max((r['est_err'] for r in pe), default=None),
                  'var_max': max((r['var_err'] for r in pe), default=None)}
    log(f'plain vs exact: {plain_gate}')

    # ---- CPU baseline (reference loop structure, bounded sample) -------
    # The oracle's restatement of interp/steps.py:350-434 (pinv per step,
    # np.matmul per cell) on timed steps: one step per worker, one worker
    # per core, all at once; then one step alone for the 1-core figure.
    cpu = None
    if pool is not None:
        rng = np.random.default_rng(7)
        cidx = rng.choice(C, args.cpu_cells, replace=False)
        steps_c = [W * B + int(t) for t in
                   np.linspace(0, TK - 1, n_cores).astype(int)]
        jobs = [(w.stn_x, w.stn_y, w.cell_x[cidx], w.cell_y[cidx], vals[st],
                 vgs[st]) for st in steps_c]
        tc0 = time.perf_counter()
        res = pool.map(_cpu_step, jobs)
        tc1 = time.perf_counter()
        per_step = [ts + C * tc for ts, tc in res]
        all_rate = n_cores * C / float(np.mean(per_step))
        # the 1-core figure: the same steps one after another on one worker
        # (the first four, to bound the time)
        one = [ts + C * tc for ts, tc in map(_cpu_step, jobs[:4])]
        one_rate = C / float(np.mean(one))
        pool.close()
        cpu = {'value': all_rate, 'unit': 'cell-estimates/s',
               'cores': n_cores, 'kind': 'port',
               'one_core_value': one_rate,
               'one_core_vs_parallel_same_steps': float(
                   np.mean(per_step[:4]) / np.mean(one)),
               'sample': (f'{n_cores} timed hourly steps x '
                          f'{args.cpu_cells} cells, one per worker process '
                          f'with one BLAS thread, all {n_cores} at once '
                          f'({tc1 - tc0:.1f} s); numpy pinv per step + '
                          'np.matmul per cell (interp/steps.py:350-434); '
                          'rate = cores x C / mean(setup + C x per-cell '
                          'time); one_core_value: the first 4 steps alone, '
                          'one after another in this process')}
        log(f'cpu baseline: {cpu}')

    if args.calib:
        x = torch.zeros(1 << 28, dtype=torch.float64, device=dev)
        for _ in range(2):
            clip_cast(x)
        torch.cuda.synchronize()
        del x
        log('calibration copy done: k_clip_cast, 2 GiB read + 1 GiB write')

    vgfit = vgfit_leg(args, log) if args.vg_fields > 0 else None
    idw = idw_leg(args, log) if args.idw_steps > 0 else None
    neb = neb_leg(args, log) if args.neb_steps > 0 else None
    stress = stress_leg(args, log) if args.stress > 0 else None
    e2e = e2e_leg(args, log) if args.e2e > 0 else None

    # the roofline object names the contraction kernel with the larger
    # share of the timed region
    note = ('HBM bytes per dispatch, FETCH_SIZE x2 + WRITE_SIZE x1 '
            '(calibrated), from profiles/kernels_hourly_mixed_b64.md (the '
            'split-BF16 and int8 figures on this workload with Sph/Exp '
            'models and --fp32-limit 1000)')
    r64 = {'bound': 'mfma', 'achieved': ach64, 'peak': 78.6,
           'unit': 'TFLOP/s', 'frac': ach64 / 78.6,
           'traffic': TRI64_TRAFFIC,
           'kernel': 'k_tri64 (L^-1 delta, FP64 MFMA)', 'traffic_note': note}
    r32 = {'bound': 'mfma', 'achieved': ach32, 'peak': 157.3,
           'unit': 'TFLOP/s', 'frac': ach32 / 157.3,
           'algorithmic_tflops': alg32, 'traffic': TRI32_TRAFFIC,
           'kernel': 'k_gemm<float,TRI> (L^-1 delta, FP32 MFMA, FP64 sums '
                     'every 32 rows)', 'traffic_note': note}
    rbf = {'bound': 'mfma', 'achieved': achbf, 'peak': 2500.0,
           'unit': 'TFLOP/s', 'frac': achbf / 2500.0,
           'algorithmic_tflops': algbf, 'traffic': TRIBF_TRAFFIC,
           'kernel': 'k_tri_bf (L^-1 delta on split-BF16 operands: six '
                     'plane products on BF16 MFMA, FP32 sums per 32 rows, '
                     'FP64 across them)', 'traffic_note': note}
    r8 = {'bound': 'mfma', 'achieved': ach8, 'peak': 5000.0,
          'unit': 'TOPS', 'frac': ach8 / 5000.0, 'traffic': TRI8_TRAFFIC,
          'algorithmic_tops': alg8, 'fp64_equivalent_tflops': eq8,
          'kernel': 'k_tri8 (L^-1 delta, exact int8 plane products, int8 '
                    'MFMA; 5- and 6-plane launches together)',
          'traffic_note': note}
    dominant = max(((msbf, rbf), (ms32, r32), (ms64, r64), (ms8, r8)),
                   key=lambda x: x[0])[1]
    out = {
        'metric': 'OK cell-estimates/s incl. variance: batched mixed vs '
                  'plain FP64 @<=1e-6 rel',
        'value': value,
        'unit': 'cell-estimates/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': W,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32+f64',
        'data': 'synthetic (seeded hourly workload generator)',
        'config': {'workload': 'hourly: 2000 stations, 1000x1000 grid at '
                               '500 m, per-step variogram and 200-2000 '
                               'active stations, OK estimate + variance',
                   'models': args.models, 'precision': args.precision,
                   'fp32_limit': args.fp32_limit or 'library policy',
                   'time_steps_per_bench_step': B, 'n_cells': C,
                   'n_stations': w.n_stns},
        'timed_region': 'H2D of the steps\' values and variogram '
                        'records, OK estimate + variance, float32 cast and '
                        'D2H of both fields into pinned host memory (batch '
                        'by batch beside later batches; stations and cells '
                        'resident)',
        'device_only': {'value': TK * C / (dev_ms * 1e-3),
                        'ms_per_step': dev_ms / K},
        'plain_fp64': {'value': plain_rate, 'steps': P,
                       'sample': f'every {max(1, TK // P)}th of the timed '
                                 f'steps, one call, after an untimed pass',
                       'ratio_mixed_over_plain': value / plain_rate,
                       'gemm_tflops': achp, 'vs_exact': plain_gate,
                       'per_step': plain_err},
        'accuracy_gate': gate,
        'roofline': dominant,
        'rooflines': {
            'gemm_tri_bf16': {'bound': 'mfma', 'achieved': achbf,
                              'peak': 2500.0, 'unit': 'TFLOP/s',
                              'frac': achbf / 2500.0,
                              'algorithmic_tflops': algbf},
            'gemm_tri_f32': {'bound': 'mfma', 'achieved': ach32,
                             'peak': 157.3, 'unit': 'TFLOP/s',
                             'frac': ach32 / 157.3,
                             'algorithmic_tflops': alg32},
            'gemm_tri_f64': {'bound': 'mfma', 'achieved': ach64,
                             'peak': 78.6, 'unit': 'TFLOP/s',
                             'frac': ach64 / 78.6},
            'gemm_tri_i8': {'bound': 'mfma', 'achieved': ach8,
                            'peak': 5000.0, 'unit': 'TOPS',
                            'frac': ach8 / 5000.0,
                            'algorithmic_tops': alg8,
                            'fp64_equivalent_tflops': eq8},
            'kgen': {'bound': 'hbm', 'achieved': kg_gbs, 'peak': 8000.0,
                     'unit': 'GB/s', 'frac': kg_gbs / 8000.0},
            'gemm_plain_f64': {'bound': 'mfma', 'achieved': achp,
                               'peak': 78.6, 'unit': 'TFLOP/s',
                               'frac': achp / 78.6}},
        'steps_fp64_variance': int(f64_steps.sum()),
        'steps_int8_variance': int(i8_steps.sum()),
        'steps_fp32_variance': int(f32_steps.sum()),
        # variants the tile autotuner picked (family, size class of 1 024
        # rows, variant, ms of each variant), all timed before the timed
        # regions; the factorisation sequences ran as captured hipGraphs
        'tile_tuning': [list(r) for r in _lib.context(0).tuning_report()],
        'factorisation_graphs': True,
        'cpu_baseline': cpu,
        'variograms_fitted': vgfit,
        'idw': idw,
        'neighbourhood': neb,
    }
    if stress is not None:
        out['stress'] = stress
    if e2e is not None:
        out['end_to_end'] = e2e
    print(json.dumps(out), flush=True)


def _cpu_step(job):
    """One hourly step of the CPU baseline in a worker process, with one
    BLAS thread: (setup seconds, seconds per cell)."""
    from threadpoolctl import threadpool_limits
    from oracle import spk_oracle as O
    sx, sy, cx, cy, z, vg = job
    with threadpool_limits(1):
        ts, tc, _, _ = O.time_ok_step_loop(sx, sy, cx, cy, z, vg)
    return ts, tc


def e2e_leg(args, log):
    """Fit + interpolate: each hourly step's variogram is fitted from its
    own values by the proposal-seeded fit (fitvgs.py settings, Sph/Exp/Gau
    with a nugget), then all steps are kriged (OK + variance) by the
    batched mixed path with those strings."""
    import numpy as np
    import torch

    from spinterps_amd import configs
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    from spinterps_amd.variograms.learned import (LearnedVariogramFitter,
                                                  load_proposal_net)
    T = args.e2e
    w = configs.hourly(n_side=args.grid, cell=5e5 / args.grid)
    vals, _ = configs.hourly_steps(w, 0, T)
    prob = DeviceProblem.from_workload(w)
    lf = LearnedVariogramFitter(w.stn_x, w.stn_y, load_proposal_net())
    lf.fit(vals[:4], verify=False, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    strs = lf.fit(vals, verify=False, seed=2)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    ok = [s is not None for s in strs]
    vgs = [s if s is not None else '1.0 Nug(0.0)' for s in strs]
    est = torch.empty((T, w.n_cells), dtype=torch.float64, device=prob.device)
    var = torch.empty_like(est)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    interp_steps(prob, vals, vgs, 'OK', precision='mixed', est=est, var=var)
    torch.cuda.synchronize()
    t_ok = time.perf_counter() - t1
    res = {'steps': T, 'fitted': int(sum(ok)), 'fit_seconds': t_fit,
           'variograms_fitted_per_s': sum(ok) / t_fit,
           'krige_seconds': t_ok,
           'cell_estimates_per_s_krige': T * w.n_cells / t_ok,
           'cell_estimates_per_s_end_to_end': T * w.n_cells / (t_fit + t_ok),
           'example_vg': strs[0]}
    log(f'end to end: {res}')
    return res


def stress_leg(args, log):
    """Ill-conditioned stress set on the daily inputs: OK estimate and
    variance on the full 500 x 500 grid by the plain FP64 and the mixed
    paths, and errors of both on sampled cells against the oracle's
    numpy.linalg.solve of the reference's bordered system.  Steps a path
    flags ill-conditioned (condition estimate above 1e12, where no FP64
    formulation reproduces another) are reported apart from the rest."""
    import numpy as np
    import torch

    from oracle import spk_oracle as O
    from spinterps_amd import _lib, configs
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.daily()
    sx, sy = configs.stress_stations(w)
    vals, vgs = configs.stress_steps(w, 0, args.stress)
    T = vals.shape[0]
    prob = DeviceProblem(sx, sy, w.cell_x, w.cell_y)
    res = {}
    outs = {}
    for prec in ('fp64', 'mixed'):
        kap = torch.zeros(T, dtype=torch.float64, device=prob.device)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e, v, info = interp_steps(prob, vals, vgs, 'OK', precision=prec,
                                  kappa=kap)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        inf = info.cpu().numpy()
        outs[prec] = (e, v, inf, kap.cpu().numpy())
        res[prec] = {'seconds': dt,
                     'cell_estimates_per_s': T * w.n_cells / dt,
                     'steps_failed_nnb_fill': int(((inf & 4) != 0).sum()),
                     'steps_ill_conditioned': int(((inf & 16) != 0).sum()),
                     'steps_fp64_variance': int(((inf & 32) != 0).sum())}
    rng = np.random.default_rng(11)
    cidx = rng.choice(w.n_cells, args.stress_cells, replace=False)
    cx, cy = w.cell_x[cidx], w.cell_y[cidx]
    errs = {p: ([], []) for p in outs}
    errs_ill = {p: ([], []) for p in outs}
    used = []
    for t in range(min(T, args.stress_oracle_steps)):
        act = ~np.isnan(vals[t])
        xs, ys, z = sx[act], sy[act], vals[t][act]
        rr = O.dists_2d(xs, ys, xs, ys)
        dr = O.dists_2d(cx, cy, xs, ys)
        K, R = O.krig_system('OK', vgs[t], rr, dr, None, None, 0.0)
        est, var, _ = O.krige_solve(K, R, z, dr, True)
        for p, (e, v, inf, kp) in outs.items():
            if inf[t] & 4:      # nearest-station fill: not comparable
                continue
            ge = e[t].cpu().numpy()[cidx]
            gv = v[t].cpu().numpy()[cidx]
            fe = 1e-3 * np.abs(est).mean()
            fv = 1e-3 * np.abs(var).mean()
            dst = errs_ill if inf[t] & 16 else errs
            dst[p][0].append(np.abs(ge - est) / np.maximum(np.abs(est), fe))
            dst[p][1].append(np.abs(gv - var) / np.maximum(np.abs(var), fv))
        used.append(t)
    for p, (le, lv) in errs.items():
        if le:
            ee, ev = np.concatenate(le), np.concatenate(lv)
            res[p].update({'est_max': float(ee.max()),
                           'est_median': float(np.median(ee)),
                           'var_max': float(ev.max()),
                           'var_median': float(np.median(ev)),
                           'steps_compared': len(le)})
        li, lvi = errs_ill[p]
        if li:
            res[p]['ill_conditioned_steps'] = {
                'steps': len(li),
                'est_max': float(np.concatenate(li).max()),
                'var_max': float(np.concatenate(lvi).max())}
    kp = outs['fp64'][3]
    res['kappa_median'] = float(np.median(kp))
    res['kappa_max'] = float(kp.max())
    res['oracle'] = (f'numpy.linalg.solve of [[G,1],[1,0]] on {len(used)} '
                     f'steps x {args.stress_cells} cells; floors 1e-3 x mean '
                     '|estimate| and 1e-3 x mean variance')
    log(f'stress: {res}')
    return res


def idw_leg(args, log):
    """IDW cell-estimates/s on the daily workload (1 000 stations, 0-20 %
    missing per step, 500 x 500 cells at 1 km, p = 2), inputs resident in
    HBM; one k_idw_nnb launch per batch of steps (for p = 2 one pass over
    the active stations per cell, weights 1 / d^2: the reference's
    normalisation by the max distance cancels in the ratio).  The
    oracle's IDW (interp/steps.py:293-313 vectorised with NumPy, one BLAS
    thread) on a cell sample is the CPU baseline and the reference values
    for the error."""
    import numpy as np
    import torch
    from spinterps_amd import configs
    from spinterps_amd.interp.steps import (DeviceProblem, interp_steps,
                                            prof_enable, prof_get)
    w = configs.daily()
    T = min(args.idw_steps, w.n_steps)
    vals, _ = configs.daily_steps(w, 0, T)
    C = len(w.cell_x)
    dev = torch.device('cuda:0')
    prob = DeviceProblem.from_workload(w)
    vals_d = torch.as_tensor(vals, device=dev)
    est = torch.empty((T, C), dtype=torch.float64, device=dev)
    interp_steps(prob, vals_d[:2], None, 'IDW', idw_exp=2.0, est=est[:2])
    prof_enable(0, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    interp_steps(prob, vals_d, None, 'IDW', idw_exp=2.0, est=est)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_k, cnt = prof_get('idw_nnb')
    prof_enable(0, False)
    rate = T * C / (t1 - t0)
    n_act = np.sum(~np.isnan(vals), axis=1).astype(np.float64)
    # 4 n C flops per step: num and den, a multiply-add each per active
    # station; the kernel multiplies over all N stations (masked)
    fl = float(np.sum(4.0 * n_act * C))
    ach = fl / (ms_k * 1e-3) / 1e12 if ms_k > 0 else 0.0
    exe = 4.0 * w.n_stns * C * T / (ms_k * 1e-3) / 1e12 if ms_k > 0 else 0.0
    log(f'idw: {(t1 - t0) * 1e3:.1f} ms for {T} steps, {rate:.3e} '
        f'cell-estimates/s, kernel {ms_k:.1f} ms / {cnt}')
    out = {'value': rate, 'unit': 'cell-estimates/s', 'steps': T,
           'cells': C, 'ms': (t1 - t0) * 1e3,
           'config': 'daily: 1000 stations, 0-20 % missing per step, '
                     '500x500 cells at 1 km, p = 2',
           'roofline': {'bound': 'mfma', 'achieved': ach, 'peak': 78.6,
                        'unit': 'TFLOP/s', 'frac': ach / 78.6,
                        'executed_tflops': exe,
                        'kernel': 'k_idw_mfma (num = (Z.M) W^T, den = M W^T '
                                  'on FP64 MFMA, W = d^-p generated per '
                                  'block)',
                        'note': 'achieved: algorithmic 4 n C flops per step '
                                '(n active stations); executed: 4 N C over '
                                'all N stations (masked)'}}
    if not args.no_cpu:
        from threadpoolctl import threadpool_limits
        from oracle import spk_oracle as O
        rng = np.random.default_rng(9)
        cidx = np.sort(rng.choice(C, args.cpu_idw_cells, replace=False))
        S = min(args.cpu_idw_steps, T)
        with threadpool_limits(1):
            tc0 = time.perf_counter()
            _, f64, _ = O.interp_all(w.stn_x, w.stn_y, w.cell_x[cidx],
                                     w.cell_y[cidx], vals[:S], None,
                                     [('IDW', 'IDW', 2.0)])
            tc1 = time.perf_counter()
        ref = f64['IDW']
        got = est[:S].cpu().numpy()[:, cidx]
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        out['error_vs_oracle'] = {'max': float(rel.max()),
                                  'median': float(np.median(rel)),
                                  'steps': S, 'cells': int(len(cidx))}
        out['cpu_baseline'] = {
            'value': S * len(cidx) / (tc1 - tc0), 'unit': 'cell-estimates/s',
            'cores': 1, 'kind': 'port',
            'sample': f'{S} daily steps x {len(cidx)} cells, the oracle\'s '
                      'NumPy IDW with its distances, one BLAS thread'}
        log(f"idw vs oracle: {out['error_vs_oracle']}, cpu "
            f"{out['cpu_baseline']['value']:.3e}/s")
    return out


def neb_leg(args, log):
    """OK cell-estimates/s incl. variance with nrst neighbourhoods (the
    reference's own test setting, 50 stations per cell) on the daily
    workload (1 000 stations, 0-20 % missing per step, 500 x 500 cells at
    1 km, exponential model), FP64, inputs resident in HBM.  The oracle
    (interp/steps.py:592-831 restated: neighbour sets, groups, pinv per
    group, one BLAS thread) on a window of cells is the CPU baseline and
    the reference values for the error; the stations the full grid keeps
    (interp/steps.py:592-608) come from a k-d tree query over every cell."""
    import numpy as np
    import torch
    from spinterps_amd import configs
    from spinterps_amd.interp.steps import (DeviceProblem, interp_neb_steps,
                                            prof_enable, prof_get, prof_work,
                                            vg_device)
    w = configs.daily()
    T = min(args.neb_steps, w.n_steps)
    K = args.neb_nebs
    vals, vgs = configs.daily_steps(w, 0, T)
    C = len(w.cell_x)
    dev = torch.device('cuda:0')
    prob = DeviceProblem(w.stn_x, w.stn_y, w.cell_x, w.cell_y)
    vals_d = torch.as_tensor(vals, device=dev)
    vg_d = vg_device(vgs, dev)
    est = torch.empty((T, C), dtype=torch.float64, device=dev)
    var = torch.empty((T, C), dtype=torch.float64, device=dev)
    kw = dict(neb_sel_mthd='nrst', n_nebs=K)
    # warm-up over one batch (64 steps) sizes the workspaces
    Tw = min(T, 64)
    interp_neb_steps(prob, vals_d[:Tw], vgs[:Tw], 'OK', est=est[:Tw],
                     var=var[:Tw], vg_dev=vg_d[:Tw * (vg_d.numel() // T)],
                     **kw)
    prof_enable(0, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, info = interp_neb_steps(prob, vals_d, vgs, 'OK', est=est, var=var,
                                  vg_dev=vg_d, **kw)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = {nm: prof_get(nm) for nm in ('neb_select', 'neb_group',
                                        'neb_krige', 'neb_pinv')}
    chunks = prof_work('neb_chunks')
    prof_enable(0, False)
    info = info.cpu().numpy()
    rate = T * C / (t1 - t0)
    # the reference's per-cell work: lambda = K^-1 rhs with K of size
    # k + 1, then the estimate and variance dot products
    # (interp/steps.py:416-434): 2 (k+1)^2 + 4 (k+1) flops per cell
    fl = float(T * C * (2 * (K + 1) ** 2 + 4 * (K + 1)))
    ms_k = prof['neb_krige'][0]
    ach = fl / (ms_k * 1e-3) / 1e12 if ms_k > 0 else 0.0
    log(f'neb: {(t1 - t0) * 1e3:.1f} ms for {T} steps, {rate:.3e} '
        f'cell-estimates/s; ' + ', '.join(
            f'{k} {v[0]:.1f} ms/{v[1]}' for k, v in prof.items()))
    out = {'value': rate, 'unit': 'cell-estimates/s', 'steps': T,
           'cells': C, 'ms': (t1 - t0) * 1e3, 'n_nebs': K,
           'config': 'daily: 1000 stations, 0-20 % missing per step, '
                     '500x500 cells at 1 km, OK estimate + variance, nrst '
                     f'neighbourhoods of {K} stations, FP64',
           'kernel_ms': {k: v[0] for k, v in prof.items()},
           'group_chunks_per_step': chunks / T,
           'steps_pinv': int(np.sum((info & 128) != 0)),
           'steps_ill_conditioned': int(np.sum((info & 16) != 0)),
           'roofline': {'bound': 'mfma', 'achieved': ach, 'peak': 78.6,
                        'unit': 'TFLOP/s', 'frac': ach / 78.6,
                        'kernel': 'k_neb_krige (per group: covariance '
                                  'system, Cholesky and inverse in LDS; per '
                                  'cell: dual-form estimate and variance)',
                        'note': 'achieved: the reference\'s per-cell '
                                '2 (k+1)^2 + 4 (k+1) flops over the '
                                'kernel\'s time (factorisations not '
                                'counted); the kernel runs on the FP64 '
                                'vector units, latency-bound in LDS'}}
    if not args.no_cpu:
        from scipy.spatial import cKDTree
        from threadpoolctl import threadpool_limits
        from oracle import spk_oracle as O
        _, nn = cKDTree(np.c_[w.stn_x, w.stn_y]).query(
            np.c_[w.cell_x, w.cell_y], k=K)
        keep = np.zeros(w.n_stns, dtype=bool)
        keep[np.unique(nn)] = True
        side = w.grid_shape[0]
        s = args.cpu_neb_side
        r0 = (side - s) // 2
        rows = np.arange(r0, r0 + s)
        cidx = (rows[:, None] * w.grid_shape[1] + rows[None, :]).ravel()
        S = min(args.cpu_neb_steps, T)
        with threadpool_limits(1):
            tc0 = time.perf_counter()
            _, f64, _ = O.interp_all(
                w.stn_x, w.stn_y, w.cell_x[cidx], w.cell_y[cidx], vals[:S],
                vgs[:S], [('OK', 'OK', None)], neb_sel_mthd='nrst',
                n_nebs=K, stn_keep=keep)
            tc1 = time.perf_counter()
        ge = est[:S].cpu().numpy()[:, cidx]
        gv = var[:S].cpu().numpy()[:, cidx]
        re, rv = f64['OK'], f64['EST_VARS_OK']
        fe = 1e-3 * np.nanmax(np.abs(re))
        fv = 1e-3 * np.nanmax(np.abs(rv))
        erre = np.abs(ge - re) / np.maximum(np.abs(re), fe)
        errv = np.abs(gv - rv) / np.maximum(np.abs(rv), fv)
        out['error_vs_oracle'] = {
            'est_max': float(np.nanmax(erre)), 'var_max': float(np.nanmax(errv)),
            'floors': 'relative to max(|ref|, 1e-3 max|ref|)',
            'nan_mismatch': int(np.sum(np.isnan(ge) != np.isnan(re))),
            'steps': S, 'cells': int(len(cidx)),
            'kept_stations': int(keep.sum())}
        out['cpu_baseline'] = {
            'value': S * len(cidx) / (tc1 - tc0), 'unit': 'cell-estimates/s',
            'cores': 1, 'kind': 'port',
            'sample': f'{S} daily steps x a {s}x{s} window of cells, the '
                      'oracle\'s neighbour sets, groups and pinv per group, '
                      'one BLAS thread'}
        log(f"neb vs oracle: {out['error_vs_oracle']}, cpu "
            f"{out['cpu_baseline']['value']:.3e}/s")
    return out


def vgfit_leg(args, log):
    """Variograms fitted/s on simulated fields (1500-station pool over
    500 km x 500 km, 300-1500 active stations, random Sph/Exp/Gau with a
    nugget): the plain GPU fit (all 7 candidates of fitvgs.py's settings),
    the proposal-seeded fit with and without the 2 % verification, and the
    SciPy restatement on the host as the CPU baseline."""
    import numpy as np
    import torch

    from spinterps_amd.interp.steps import DeviceProblem
    from spinterps_amd.variograms.fitvgs import BatchVariogramFitter
    from spinterps_amd.variograms.learned import (LearnedVariogramFitter,
                                                  load_proposal_net,
                                                  training_fields)
    from spinterps_amd.variograms.train_proposal import FIT_KW, station_pool
    F = args.vg_fields
    x, y = station_pool()
    prob = DeviceProblem(x, y, x[:1], y[:1])
    parts = []
    for c0 in range(0, F, 128):
        fl, info, _ = training_fields(prob, min(128, F - c0),
                                      seed=777000 + c0)
        parts.append(fl.cpu().numpy())
    vals = np.concatenate(parts)
    fitter = BatchVariogramFitter(x, y)
    lf = LearnedVariogramFitter(x, y, load_proposal_net(), fitter=fitter)
    fitter.fit(vals[:8], seed=1, **FIT_KW)  # warm-up
    lf.fit(vals[:8], verify=False, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fitter.fit(vals, seed=2, **FIT_KW)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    lf.fit(vals, verify=False, seed=2)
    torch.cuda.synchronize()
    t_fast = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, det = lf.fit(vals, verify=True, seed=2, return_details=True)
    torch.cuda.synchronize()
    t_ver = time.perf_counter() - t0
    ratio = det['learned_fun'] / det['plain_fun']
    log(f'vgfit: plain {F / t_plain:.1f}/s, learned {F / t_fast:.1f}/s, '
        f'verified {F / t_ver:.1f}/s, fallback {det["fallback"].mean():.3f}')
    res = {'fields': F, 'unit': 'variograms fitted/s',
           'plain_gpu': F / t_plain, 'learned_gpu': F / t_fast,
           'learned_verified_gpu': F / t_ver,
           'fallback_fraction': float(det['fallback'].mean()),
           'learned_over_plain_error_median': float(np.median(ratio)),
           'learned_over_plain_error_p90': float(np.percentile(ratio, 90))}
    if not args.no_cpu and args.cpu_vg_fields > 0:
        from threadpoolctl import threadpool_limits
        from oracle import vgfit_oracle as V
        t0 = time.perf_counter()
        n = min(args.cpu_vg_fields, F)
        with threadpool_limits(1):
            for t in range(n):
                V.fit_step(x, y, vals[t], seed=t, **FIT_KW)
        dt = time.perf_counter() - t0
        res['cpu_baseline'] = {
            'value': n / dt, 'unit': 'variograms fitted/s', 'cores': 1,
            'kind': 'port',
            'sample': f'{n} fields, scipy differential_evolution as '
                      'vgs.py:758-764 calls it, one BLAS thread'}
        log(f'vgfit cpu: {n / dt:.3f}/s')
    return res


if __name__ == '__main__':
    main()
