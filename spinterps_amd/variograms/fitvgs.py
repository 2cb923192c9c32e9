"""Batched variogram estimation and fitting on the GPU.

Drop-in for the per-step fit of variograms/fitvgs.py:143-278
(FitVariogramsSteps._fit_vgs_step) with the Variogram.fit semantics of
variograms/vgs.py:20-870, for many steps in a few launches: station pairs
are sorted once, every step's empirical variogram is binned on the GPU
and every (step, candidate model) is fitted by its own differential-
evolution workgroup.  The candidate rules, the AIC choice and the output
strings follow vgs.py:700-848 on the host.
"""
from __future__ import annotations

import ctypes
from itertools import permutations

import numpy as np
import torch

from .._lib import check, context, lib

ORDER = ['Nug', 'Sph', 'Exp', 'Lin', 'Gau', 'Pow', 'Hol']
TYPE_ID = {'Rng': 0, 'Nug': 1, 'Sph': 2, 'Exp': 3, 'Lin': 4, 'Gau': 5,
           'Pow': 6, 'Hol': 7}


def candidate_sets(fit_vgs, perm_r_list, fil_nug_vg):
    """The model permutations vgs.py:700-769 fits, in its loop order."""
    if 'all' in fit_vgs:
        mix = list(ORDER)
    else:
        mix = [v for v in ORDER if v in fit_vgs]
    if not mix:
        raise ValueError('No relevant variograms in fit_vgs list.')
    rs = sorted(set(min(int(r), len(mix)) for r in perm_r_list))
    if fil_nug_vg is not None:
        mix.append(fil_nug_vg)
    out = []
    for n in rs:
        for o in permutations(mix, n):
            if n == 1 or (n > 1 and o[0] == fil_nug_vg):
                out.append(o)
    return out


def _ctx(dev):
    ctx = context(dev.index or 0)
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    return ctx


class BatchVariogramFitter:
    """Variogram fits of many steps over one station set (GPU)."""

    def __init__(self, stn_x, stn_y, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError('a ROCm GPU is required')
        self.dev = torch.device('cuda', device)
        self.x = torch.as_tensor(np.asarray(stn_x, np.float64), device=self.dev)
        self.y = torch.as_tensor(np.asarray(stn_y, np.float64), device=self.dev)
        N = self.x.numel()
        self.N = N
        self.M = N * (N - 1) // 2
        self.pi = torch.empty(self.M, dtype=torch.int32, device=self.dev)
        self.pj = torch.empty(self.M, dtype=torch.int32, device=self.dev)
        self.pd = torch.empty(self.M, dtype=torch.float64, device=self.dev)
        ctx = _ctx(self.dev)
        check(lib().spk_vgfit_pairs_dev(
            ctx.h, self.x.data_ptr(), self.y.data_ptr(), N,
            self.pi.data_ptr(), self.pj.data_ptr(), self.pd.data_ptr()))

    def empirical(self, vals, mdr=0.5, nk=50, evg='robust'):
        """Bins of every step: (h, gamma, pairs per bin, n bins), device."""
        v = torch.as_tensor(np.ascontiguousarray(vals, np.float64),
                            device=self.dev)
        T = v.shape[0]
        max_bins = max(1, -(-self.M // nk))
        bh = torch.zeros((T, max_bins), dtype=torch.float64, device=self.dev)
        bg = torch.zeros_like(bh)
        bn = torch.zeros((T, max_bins), dtype=torch.int32, device=self.dev)
        nb = torch.zeros(T, dtype=torch.int32, device=self.dev)
        ctx = _ctx(self.dev)
        check(lib().spk_vgfit_evg_dev(
            ctx.h, self.pi.data_ptr(), self.pj.data_ptr(), self.pd.data_ptr(),
            self.M, self.N, v.data_ptr(), T, float(mdr), int(nk),
            0 if evg == 'robust' else 1, max_bins, bh.data_ptr(),
            bg.data_ptr(), bn.data_ptr(), nb.data_ptr()))
        return bh, bg, bn, nb

    def de(self, bins, var, cands, fil_nug_vg, opt_iters=10000, popmul=50,
           tol=0.01, range_ub=1e9, seed=0):
        bh, bg, bn, nb = bins
        T, max_bins = bh.shape
        spec = np.zeros((len(cands), 6), dtype=np.int32)
        for c, o in enumerate(cands):
            spec[c, 0] = len(o)
            for k, nm in enumerate(o):
                spec[c, 1 + k] = TYPE_ID[nm]
            spec[c, 5] = 1 if (len(o) > 1 and o[0] == fil_nug_vg) else 0
        var_d = torch.as_tensor(np.asarray(var, np.float64), device=self.dev)
        par = torch.zeros((T, len(cands), 8), dtype=torch.float64,
                          device=self.dev)
        fun = torch.zeros((T, len(cands)), dtype=torch.float64,
                          device=self.dev)
        nit = torch.zeros((T, len(cands)), dtype=torch.int32, device=self.dev)
        ctx = _ctx(self.dev)
        check(lib().spk_vgfit_de_dev(
            ctx.h, bh.data_ptr(), bg.data_ptr(), nb.data_ptr(), max_bins,
            var_d.data_ptr(), T,
            spec.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(cands),
            int(opt_iters), int(popmul), float(tol), float(range_ub),
            int(seed) & ((1 << 64) - 1), par.data_ptr(), fun.data_ptr(),
            nit.data_ptr()))
        return par, fun, nit

    def fit(self, vals, mdr=0.5, nk=50, perm_r_list=(1,),
            fit_vgs=('Sph', 'Exp'), fil_nug_vg=None, fit_thresh=0.01,
            n_best=1, opt_iters=10000, evg='robust', min_vld_stns=10,
            drop_zeros=False, range_ub=1e9, seed=0, return_details=False):
        """Per step the variogram strings best first (fitvgs.py:247-276),
        or None for a step with too few values or no spread."""
        vals = np.array(vals, dtype=np.float64, copy=True)
        if drop_zeros:
            vals[vals == 0] = np.nan
        T = vals.shape[0]
        ok = np.zeros(T, dtype=bool)
        var = np.zeros(T)
        npts = np.zeros(T, dtype=np.int64)
        for t in range(T):
            z = vals[t][~np.isnan(vals[t])]
            npts[t] = z.size
            if z.size < min_vld_stns or np.isclose(z.std(), 0.0):
                continue
            ok[t] = True
            var[t] = np.var(z)
        cands = candidate_sets(fit_vgs, perm_r_list, fil_nug_vg)
        steps = np.where(ok)[0]
        out = [None] * T
        if steps.size == 0:
            return (out, None) if return_details else out
        bins = self.empirical(vals[steps], mdr, nk, evg)
        par, fun, nit = self.de(bins, var[steps], cands, fil_nug_vg,
                                opt_iters, 50, 0.01, range_ub, seed)
        par = par.cpu().numpy()
        fun = fun.cpu().numpy()
        details = []
        for k, t in enumerate(steps):
            best = []
            min_obj = np.inf
            skip = []
            for c, o in enumerate(cands):
                if o in skip:
                    continue
                D = 2 * len(o)
                x = par[k, c, :D]
                rngs, sills = x[0::2].copy(), x[1::2].copy()
                with np.errstate(divide='ignore'):
                    aic = npts[t] * np.log(fun[k, c]) + 2 * D
                c1 = aic < min_obj * (1.0 - fit_thresh)
                names = list(o)
                if not np.all(np.ediff1d(rngs) >= 0):
                    srt = np.argsort(rngs)
                    rngs, sills = rngs[srt], sills[srt]
                    skip.append(tuple(np.array(o)[srt]))
                    names = list(np.array(o)[srt])
                c2 = np.all(np.ediff1d(rngs) >= 0)
                if c1 and c2:
                    min_obj = aic
                    best.append((names, rngs, sills, fun[k, c], aic))
                if len(best) > n_best:
                    del best[0]
            strs = [' + '.join('%0.5f %s(%0.5f)' % (s, nm, r)
                               for nm, r, s in zip(nms, rr, ss))
                    for nms, rr, ss, _, _ in best]
            out[t] = strs[::-1] if strs else None
            details.append(best)
        if return_details:
            return out, dict(steps=steps, bins=bins, params=par, fun=fun,
                             nit=nit.cpu().numpy(), cands=cands,
                             best=details)
        return out
