"""CPU oracle (test infrastructure only) for the variogram estimation and fit
that label the step variograms: a NumPy/SciPy restatement of
variograms/vgs.py (Variogram: pairs, distance clip, equal-count bins,
estimators, differential-evolution fit, AIC choice) and
variograms/fitvgs.py:143-278 (_fit_vgs_step).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.

Pinning status: pairs, clip, bins and the estimators are pinned by closed
forms on tiny inputs (tests/test_vgfit_oracle.py).  The differential
evolution is SciPy's own routine, called as vgs.py:758-764 calls it; it is
stochastic (unseeded in the reference), so a fit is pinned by recovering
known parameters from noise-free variograms, not by exact values.  The
model choice and the output strings (select_models, vg_strings) are pinned
on hand-built objective values whose AIC outcomes are worked out by hand
from vgs.py:771-846.
"""
from __future__ import annotations

import math
from itertools import permutations

import numpy as np
from scipy.optimize import differential_evolution


# vgs.py:533-590: the fit models take (h, (range, sill)); Sph/Lin switch
# with ">" (not ">=" as the cyth models do)
def nug(h, a):
    return np.full(h.shape, a[1])


def sph(h, a):
    r, s = a
    out = s * ((1.5 * h) / r - h ** 3 / (2 * r ** 3))
    out[h > r] = s
    return out


def exp_(h, a):
    r, s = a
    return s * (1 - np.exp(-3 * (h / r)))


def lin(h, a):
    r, s = a
    out = s * (h / r)
    out[h > r] = s
    return out


def gau(h, a):
    r, s = a
    return s * (1 - np.exp(-3 * (h ** 2 / r ** 2)))


def pow_(h, a):
    r, s = a
    return s * (h ** r)


def hol(h, a):
    r, s = a
    out = np.zeros(h.shape[0])
    idx = np.where(h > 0)
    x = (math.pi * h[idx]) / r
    out[idx] = s * (1 - (np.sin(x) / x))
    return out


FIT_FTNS = {'Nug': nug, 'Sph': sph, 'Exp': exp_, 'Lin': lin, 'Gau': gau,
            'Pow': pow_, 'Hol': hol}
ORDER = ['Nug', 'Sph', 'Exp', 'Lin', 'Gau', 'Pow', 'Hol']


def evg_robust(xs, ys):
    """vgs.py:414-422."""
    nh = xs.shape[0]
    if nh == 0:
        return 0
    evg = ((np.sum(np.abs(xs - ys) ** 0.5, axis=0)) / nh) ** 4
    return np.divide(evg, (0.457 + (0.494 / nh))) / 2


def evg_classic(xs, ys):
    """vgs.py:399-412: median of squared differences (no 1/2)."""
    if xs.shape[0] == 0:
        return np.nan
    return np.median((np.subtract(xs, ys)) ** 2)


def pairs(x, y, z):
    """vgs.py:345-391: (z_i, z_j, d_ij) for i > j, sorted by distance
    (stable sort; ties have measure zero for continuous coordinates)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    n = x.size
    ii, jj = np.tril_indices(n, -1)
    # the reference loops i (outer) then j < i (inner)
    order = np.lexsort((jj, ii))
    ii, jj = ii[order], jj[order]
    d = np.sqrt((x[ii] - x[jj]) ** 2 + (y[ii] - y[jj]) ** 2)
    srt = np.argsort(d, kind='stable')
    return np.stack([z[ii[srt]], z[jj[srt]], d[srt]], axis=1)


def clip_pairs(pr, mdr):
    """vgs.py:597-659 for a float mdr and uh=None.  Keeps the pairs closer
    than clip_ul, the pair distance nearest to mdr * max distance (or that
    distance itself when the nearest pair is not beyond it)."""
    d = pr[:, 2]
    act_max = d.max()
    max_dist = act_max * mdr
    idx_clip = int(np.argmin(np.abs(d - max_dist)))
    clip_ul = d[idx_clip]
    if not (clip_ul - max_dist) > 0:
        clip_ul = max_dist
    if idx_clip != pr.shape[0] - 1:
        keep = np.clip(d, 0, clip_ul) != clip_ul
        pr = pr[keep]
    return pr


def var_bins(pr, nk, evg='robust'):
    """vgs.py:495-520: bins of nk consecutive pairs -> (h mean, gamma, n)."""
    est = evg_robust if evg == 'robust' else evg_classic
    out = []
    for m in range(0, pr.shape[0], nk):
        b = pr[m:m + nk]
        out.append((np.mean(b[:, 2]), est(b[:, 0], b[:, 1]), b.shape[0]))
    return np.array(out)


def empirical(x, y, z, mdr=0.5, nk=50, evg='robust'):
    pr = clip_pairs(pairs(x, y, z), mdr)
    return var_bins(pr, nk, evg)


def objective(params, names, h, g, wts=None):
    """vgs.py:661-678."""
    vg = np.zeros(h.shape)
    for i, nm in enumerate(names):
        vg += FIT_FTNS[nm](h, params[2 * i:2 * i + 2])
    if wts is not None:
        return np.sum((wts / vg ** 2) * np.square(g - vg))
    return np.sum(np.square(g - vg))


def candidate_sets(fit_vgs, perm_r_list, fil_nug_vg):
    """The permutations vgs.py:700-769 fits, with their bounds rule."""
    mix = [v for v in ORDER if v in fit_vgs]
    # vgs.py:706-713 clamps r to the model count before the nugget joins
    rs = sorted(set(min(r, len(mix)) for r in perm_r_list))
    if fil_nug_vg is not None:
        mix.append(fil_nug_vg)
    out = []
    for n in rs:
        for o in permutations(mix, n):
            if n == 1 or (n > 1 and o[0] == fil_nug_vg):
                out.append(o)
    return out


def bounds_for(o, h0, g0, var, fil_nug_vg, range_ub=1e9):
    """vgs.py:731-750 (range_ub = 1e9 there)."""
    b = []
    cond2 = len(o) > 1 and o[0] == fil_nug_vg
    for p, nm in enumerate(o):
        if nm == 'Pow':
            b += [(1e-5, 2), (1e-5, 2 * var)]
        elif cond2 and p == 0:
            b += [(h0, h0), (g0, g0)]
        else:
            b += [(h0, range_ub), (1e-5, 2 * var)]
    return b


def select_models(cands, solve, tot_pts, fit_thresh=0.01, n_best=1):
    """The model choice of vgs.py:771-812 over the candidates in order.
    solve(o) -> (fun, x) with x = [range_0, sill_0, range_1, sill_1, ...].
    A candidate is accepted when its AIC = tot_pts ln(fun) + 2 len(x) is
    below min_AIC (1 - fit_thresh) (as written, also for a negative AIC)
    and its ranges ascend after sorting; out-of-order ranges are sorted,
    with names and sills, and the sorted permutation is skipped when it
    comes up later.  The last n_best accepted are kept, oldest first.
    Returns (best [(names, ranges, sills, fun, aic)], skipped, funs)."""
    min_obj = np.inf
    best = []
    skip = []
    funs = {}
    for o in cands:
        if o in skip:
            continue
        fun, xs = solve(o)
        xs = np.asarray(xs, np.float64)
        funs[o] = (fun, xs.copy())
        rngs = xs[0::2].copy()
        sills = xs[1::2].copy()
        aic = tot_pts * np.log(fun) + 2 * xs.shape[0]
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
            best.append((names, rngs, sills, fun, aic))
        if len(best) > n_best:
            del best[0]
    return best, skip, funs


def vg_strings(best):
    """vgs.py:826-846: '%0.5f %s(%0.5f)' % (sill, name, range) joined by
    ' + ', one string per kept model."""
    return [' + '.join('%0.5f %s(%0.5f)' % (s, nm, r)
                       for nm, r, s in zip(nms, rr, ss))
            for nms, rr, ss, _, _ in best]


def fit(x, y, z, mdr=0.5, nk=50, perm_r_list=(1,), fit_vgs=('Sph', 'Exp'),
        fil_nug_vg=None, fit_thresh=0.01, n_best=1, opt_iters=10000,
        evg='robust', use_wts=False, seed=None, range_ub=1e9):
    """Variogram(...).fit() of vgs.py:850-870 with the fitvgs.py settings;
    returns (vg_str_list, details)."""
    z = np.asarray(z, np.float64)
    var = np.var(z)
    tot_pts = z.size
    bins = empirical(x, y, z, mdr, nk, evg)
    h, g, npairs = bins[:, 0], bins[:, 1], bins[:, 2]
    wts = npairs if use_wts else None
    rng = np.random.default_rng(seed)

    def solve(o):
        bnds = bounds_for(o, h[0], g[0], var, fil_nug_vg, range_ub)
        res = differential_evolution(objective, bnds, args=(o, h, g, wts),
                                     maxiter=opt_iters,
                                     popsize=len(bnds) * 50, polish=False,
                                     seed=rng)
        return res.fun, res.x

    best, _, funs = select_models(
        candidate_sets(fit_vgs, perm_r_list, fil_nug_vg), solve, tot_pts,
        fit_thresh, n_best)
    return vg_strings(best), dict(bins=bins, best=best, var=var, funs=funs)


def fit_step(x, y, z, min_vld_stns=10, drop_zeros=False, **kw):
    """fitvgs.py:143-278 for one step: None when there are too few values
    or no spread, else the best-first variogram strings."""
    z = np.asarray(z, np.float64)
    ok = ~np.isnan(z)
    x, y, z = np.asarray(x)[ok], np.asarray(y)[ok], z[ok]
    if z.size < min_vld_stns:
        return None
    if drop_zeros:
        nz = z != 0
        x, y, z = x[nz], y[nz], z[nz]
        if nz.sum() < min_vld_stns:
            return None
    if np.isclose(z.std(), 0.0):
        return None
    strs, _ = fit(x, y, z, **kw)
    return strs[::-1]
