"""CPU oracle (test infrastructure only) for the spinterps interpolation path.

A float64 NumPy restatement of the reference's numeric kernels
(``cyth/interpmthds.pyx``) and of the per-step driver semantics
(``interp/steps.py``, ``interp/grps.py``, ``interp/vgclus.py``,
``misc.py:1027-1105``).  Every function cites the reference lines it follows.

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``spinterps_amd``) never does.

Pinning status (how each function is checked, see tests/test_oracle.py):
  vg models / parse / fill_vg_var_arr / theo vg : closed forms + the
      ``test/test_theo_vg.py:26-28`` string as a known-answer case.
  distances / gathers / IDW helpers              : brute-force scalar loops.
  kriging systems (pinv)                         : invariants (sum of weights,
      drift reproduction, nugget-shift invariance, exact interpolation,
      n=2 closed form, scipy solve).
  sel_equidist_refs / nrst neighbours            : brute force + rules read
      from pyx:846-889 and interp/grps.py:147-247.  Degenerate geometry
      (stations on axes, equal distances) is parity unpinned.
"""
from __future__ import annotations

import math

import numpy as np

# ---------------------------------------------------------------------------
# Variogram models, cyth/interpmthds.pyx:38-95
# ---------------------------------------------------------------------------


def rng_vg(h, r, s):
    """pyx:38-39: returns the lag itself."""
    return np.asarray(h, dtype=np.float64).copy()


def nug_vg(h, r, s):
    """pyx:42-43: the sill for every lag, h == 0 included."""
    return np.full(np.shape(h), s, dtype=np.float64)


def sph_vg(h, r, s):
    """pyx:46-55: s for h >= r, else s*((1.5h)/r - h^3/(2 r^3))."""
    h = np.asarray(h, dtype=np.float64)
    a = (1.5 * h) / r
    b = (h * h * h) / (2 * (r * r * r))
    out = s * (a - b)
    return np.where(h >= r, s, out)


def exp_vg(h, r, s):
    """pyx:58-59: s*(1 - exp(-3h/r))."""
    h = np.asarray(h, dtype=np.float64)
    return s * (1 - np.exp((-3 * h) / r))


def lin_vg(h, r, s):
    """pyx:62-66: s for h > r, else s*(h/r)."""
    h = np.asarray(h, dtype=np.float64)
    return np.where(h > r, s, s * (h / r))


def gau_vg(h, r, s):
    """pyx:69-70: s*(1 - exp(-3 h^2/r^2))."""
    h = np.asarray(h, dtype=np.float64)
    return s * (1 - np.exp(-3 * ((h * h) / (r * r))))


def pow_vg(h, r, s):
    """pyx:73-74: s*h^r (C pow, so 0^0 == 1)."""
    h = np.asarray(h, dtype=np.float64)
    return s * np.power(h, r)


def hol_vg(h, r, s):
    """pyx:77-83: 0 at h == 0, else s*(1 - sin(a)/a) with a = pi*h/r."""
    h = np.asarray(h, dtype=np.float64)
    with np.errstate(invalid='ignore', divide='ignore'):
        a = (math.pi * h) / r
        out = s * (1 - (np.sin(a) / a))
    return np.where(h == 0, 0.0, out)


VG_FTNS = {
    'Rng': rng_vg, 'Nug': nug_vg, 'Sph': sph_vg, 'Exp': exp_vg,
    'Lin': lin_vg, 'Gau': gau_vg, 'Pow': pow_vg, 'Hol': hol_vg,
}


def parse_vg(vg_str: str):
    """Parse a variogram string the way fill_vg_var_arr does (pyx:174-184).

    Components are split on '+', stripped, split on exactly one space into
    (sill, rest), then rest.split('(') into (name, range) and the range is
    cut at ')'.  The range is clamped to max(1e-5, range) (pyx:183).
    Returns a list of (name, sill, range).  Unknown names raise KeyError
    (the reference dereferences a null function pointer there).
    """
    comps = []
    for comp in vg_str.split('+'):
        comp = comp.strip()
        sill_s, vg_s = comp.split(' ')
        name, range_s = vg_s.split('(')
        range_s = range_s.split(')')[0]
        if name not in VG_FTNS:
            raise KeyError(name)
        comps.append((name, float(sill_s), max(1e-5, float(range_s))))
    return comps


def vg_sum(d, vg_str, covar_flag=0, min_vg_val=None):
    """Sum of components as fill_vg_var_arr forms it (pyx:162-216)."""
    d = np.asarray(d, dtype=np.float64)
    out = np.zeros(d.shape, dtype=np.float64)
    cov_mult = 1.0 if covar_flag else 0.0
    cov_sign = -1.0 if covar_flag else 1.0
    for name, sill, rng in parse_vg(vg_str):
        out += (cov_mult * sill) + (cov_sign * VG_FTNS[name](d, rng, sill))
    if min_vg_val is not None:
        out[out <= min_vg_val] = 0.0
    return out


def fill_vg_var_arr(dists, in_vars, covar_flag, diag_mat_flag, vg_str,
                    min_vg_val):
    """pyx:146-226.  With diag_mat_flag only g >= h is computed, then the
    lower triangle is mirrored from the upper one."""
    dists = np.asarray(dists, dtype=np.float64)
    assert in_vars.shape == dists.shape
    vals = vg_sum(dists, vg_str, covar_flag, min_vg_val)
    if diag_mat_flag:
        upper = np.triu(np.ones(dists.shape, dtype=bool))
        vals = np.where(upper, vals, vals.T)
    in_vars[...] = vals


def fill_dists_2d_mat(x1s, y1s, x2s, y2s, dists):
    """pyx:123-143: dists[i, j] = sqrt((x1i-x2j)^2 + (y1i-y2j)^2)."""
    dx = x1s[:, None] - x2s[None, :]
    dy = y1s[:, None] - y2s[None, :]
    dists[...] = ((dx * dx) + (dy * dy)) ** 0.5


def dists_2d(x1s, y1s, x2s, y2s):
    out = np.empty((x1s.size, x2s.size), dtype=np.float64)
    fill_dists_2d_mat(x1s, y1s, x2s, y2s, out)
    return out


def copy_2d_arr_at_idxs(arr, row_idxs, col_idxs, subset_arr):
    """pyx:229-248: subset[i, j] = arr[row_idxs[i], col_idxs[j]] on the
    leading (len(row_idxs), len(col_idxs)) block; the rest is untouched."""
    subset_arr[:row_idxs.size, :col_idxs.size] = arr[np.ix_(row_idxs,
                                                            col_idxs)]


def fill_dists_one_pt(x, y, xs, ys, dists):
    """pyx:768-781."""
    dists[...] = (((x - xs) ** 2) + ((y - ys) ** 2)) ** 0.5


def fill_wts_and_sum(dists, wts, idw_exp):
    """pyx:784-795: wts = 1/d^p; returns the running sum (left to right)."""
    with np.errstate(divide='ignore'):
        wts[...] = 1.0 / (dists ** idw_exp)
    s = 0.0
    for w in wts:
        s += w
    return s


def get_mults_sum(wts, data):
    """pyx:798-808: left-to-right sum of wts*data."""
    s = 0.0
    for w, v in zip(wts, data):
        s += w * v
    return s


def fill_theo_vg_vals(vg_str, h_arr, r, s, vg_arr):
    """pyx:98-120: vg_arr += model(h, r, s); no range clamp; asserts."""
    name = vg_str.strip()
    ftn = VG_FTNS[name]
    assert h_arr.shape[0]
    assert h_arr.shape[0] == vg_arr.shape[0]
    assert s >= 0
    assert r >= 0
    vg_arr += ftn(h_arr, r, s)


def get_theo_vg_vals(in_model, h_arr):
    """misc.py:1027-1047."""
    vg_vals = np.zeros_like(h_arr, dtype=np.float64)
    for sub in str(in_model).split('+'):
        sub = sub.strip()
        sill, sub = sub.split(' ')
        sub, rng = sub.split('(')
        rng = rng.split(')')[0]
        fill_theo_vg_vals(sub, h_arr, float(rng), float(sill), vg_vals)
    return vg_vals


def check_full_nuggetness(in_model, min_vg_val):
    """misc.py:1074-1105.  Note: the single-'Nug' test compares the whole
    component text with 'Nug' and never fires; a nugget-only string is
    caught by its zero range instead."""
    in_model = str(in_model)
    if in_model == 'nan':
        return False
    models = in_model.split('+')
    sill = 0.0
    rng = 0.0
    for sub in models:
        sub = sub.strip()
        sill += float(sub.split('(')[0].strip()[:-3].strip())
        rng = max(rng, float(sub.split('(')[1].split(')')[0]))
    nug = bool(np.any(np.array([sill, rng]) <= min_vg_val))
    if (len(models) == 1) and models[0] == 'Nug':
        nug = True
    return nug


def get_nd_dists(pts):
    """pyx:893-923: condensed distances for i > j, row-major over i."""
    pts = np.asarray(pts, dtype=np.float64)
    n = pts.shape[0]
    out = []
    for i in range(n):
        for j in range(i):
            out.append(float(np.sqrt(((pts[i] - pts[j]) ** 2).sum())))
    return np.array(out, dtype=np.float64)


# ---------------------------------------------------------------------------
# Neighbour selection, cyth/interpmthds.pyx:811-890 and interp/grps.py
# ---------------------------------------------------------------------------


def sel_equidist_refs(dst_x, dst_y, ref_xs, ref_ys, n_pies, min_dist_thresh,
                      not_neb_flag, dists, tem_ref_sel_dists,
                      ref_sel_pie_idxs, ref_pie_idxs, ref_pie_cts):
    """pyx:811-890.  Each station gets its distance rank within its angular
    sector.  The angle is atan(dy/dx) with the quadrant fixes of pyx:853-866
    (dx == 0 gives angle 0).  Raises IndexError where the reference would
    write past ref_pie_cts (angle rounding to exactly 2*pi)."""
    n_refs = ref_xs.size
    two_pi = 2 * math.pi
    ref_sel_pie_idxs[:] = not_neb_flag
    fill_dists_one_pt(dst_x, dst_y, ref_xs, ref_ys, dists)
    min_dist = np.inf
    min_idx = -1
    for i in range(n_refs):
        if (dists[i] <= min_dist_thresh) and (dists[i] < min_dist):
            min_idx = i
            min_dist = dists[i]
    if min_dist < np.inf:
        ref_sel_pie_idxs[min_idx] = 0
        return
    ref_pie_cts[:n_pies] = 0
    for j in range(n_refs):
        xd = ref_xs[j] - dst_x
        yd = ref_ys[j] - dst_y
        if not xd:
            ang = 0.0
        else:
            ang = math.atan(yd / xd)
            if (xd < 0) and (yd > 0):
                ang = math.pi + ang
            elif (xd < 0) and (yd < 0):
                ang = math.pi + ang
            elif (xd > 0) and (yd < 0):
                ang = two_pi + ang
        pie = int(ang * n_pies / two_pi)
        if pie >= n_pies:
            raise IndexError('pie index out of range')
        ref_pie_idxs[j] = pie
        ref_pie_cts[pie] += 1
    for j in range(n_pies):
        if not ref_pie_cts[j]:
            continue
        tem_ref_sel_dists[:] = np.where(ref_pie_idxs[:n_refs] == j,
                                        dists[:n_refs], np.inf)
        order = np.argsort(tem_ref_sel_dists, kind='stable')
        for i in range(n_refs):
            if tem_ref_sel_dists[order[i]] == np.inf:
                break
            ref_sel_pie_idxs[order[i]] = i


def neb_idxs_all(n_dst, n_refs):
    """interp/grps.py:141-145."""
    return np.tile(np.arange(n_refs), (n_dst, 1))


def neb_idxs_nrst(dst_xs, dst_ys, ref_xs, ref_ys, n_nebs):
    """interp/grps.py:147-166: sorted indices of the n_nebs nearest.
    Ties (measure zero for continuous coordinates) use a stable order."""
    k = min(n_nebs, ref_xs.size)
    out = np.empty((dst_xs.size, k), dtype=np.int64)
    for i in range(dst_xs.size):
        d = (((dst_xs[i] - ref_xs) ** 2) + ((dst_ys[i] - ref_ys) ** 2)) ** 0.5
        out[i] = np.sort(np.argsort(d, kind='stable')[:n_nebs])
    return out


def neb_idxs_pie(dst_xs, dst_ys, ref_xs, ref_ys, n_nebs, n_pies):
    """interp/grps.py:168-247 (min_dist_thresh = -1)."""
    n_refs = ref_xs.size
    dists = np.zeros(n_refs)
    pie_idxs = np.zeros(n_refs, dtype=np.uint64)
    tem = np.zeros(n_refs)
    sel = np.zeros(n_refs, dtype=np.int64)
    cts = np.zeros(n_pies, dtype=np.uint64)
    out = np.full((dst_xs.size, n_nebs), -1, dtype=np.int64)
    for i in range(dst_xs.size):
        sel_equidist_refs(dst_xs[i], dst_ys[i], ref_xs, ref_ys, n_pies, -1.0,
                          -1, dists, tem, sel, pie_idxs, cts)
        assert np.all(sel != -1)
        full = []
        for rank in np.unique(sel):
            same = np.where(sel == rank)[0]
            full.extend(same[np.argsort(dists[same], kind='stable')].tolist())
        full = np.array(full)
        assert full.size == n_refs
        out[i, :] = np.sort(full[:n_nebs])
    return out


def neb_groups(all_neb_idxs):
    """interp/grps.py:103-139: cells with identical index rows, in order of
    first appearance."""
    groups = {}
    order = []
    for i in range(all_neb_idxs.shape[0]):
        key = all_neb_idxs[i].tobytes()
        if key not in groups:
            groups[key] = []
            order.append(key)
        groups[key].append(i)
    return [np.array(groups[k], dtype=np.int64) for k in order]


def time_groups(data):
    """interp/grps.py:57-101 on a [T x N] array with NaN for missing: steps
    with the same available-station set, in order of first appearance.
    Returns a list of (station_index_array, step_bool_mask)."""
    avail = ~np.isnan(data)
    keys = {}
    order = []
    for t in range(data.shape[0]):
        key = np.where(avail[t])[0].tobytes()
        if key not in keys:
            keys[key] = np.zeros(data.shape[0], dtype=bool)
            order.append((key, np.where(avail[t])[0]))
        keys[key][t] = True
    return [(stns, keys[key]) for key, stns in order]


# ---------------------------------------------------------------------------
# Kriging systems and the step driver, interp/steps.py:170-435
# ---------------------------------------------------------------------------


def krig_system(method, vg_str, ref_ref_dists, dst_ref_dists, ref_drfts,
                dst_drfts, min_vg_val):
    """interp/steps.py:170-243 for one step group: returns (K, R) where K is
    the (n+add) square system and R the [C x (n+add)] right-hand sides.

    OK : K = [[G, 1], [1^T, 0]], R = [g | 1]          (steps.py:212-216)
    SK : K = Cov, R = cov                            (steps.py:218-219)
    SKM: the SK system (simple kriging with the step mean as known mean)
    EDK: K = [[G, 1, S], [1^T, 0, 0], [S^T, 0, 0]], R = [g | 1 | s]
                                                     (steps.py:221-234)
    G and g are semivariances (covar_flag 0), Cov and cov covariances
    (covar_flag 1), both from fill_vg_var_arr (pyx:146-226).
    """
    n = ref_ref_dists.shape[0]
    ncell = dst_ref_dists.shape[0]
    if method == 'OK':
        add, cflag = 1, 0
    elif method in ('SK', 'SKM'):
        add, cflag = 0, 1
    elif method == 'EDK':
        add, cflag = 1 + ref_drfts.shape[1], 0
    else:
        raise NotImplementedError(method)
    K = np.full((n + add, n + add), np.nan)
    R = np.full((ncell, n + add), np.nan)
    tmp = np.empty((n, n))
    fill_vg_var_arr(ref_ref_dists, tmp, cflag, 1, vg_str, min_vg_val)
    K[:n, :n] = tmp
    tmp2 = np.empty((ncell, n))
    fill_vg_var_arr(dst_ref_dists, tmp2, cflag, 0, vg_str, min_vg_val)
    R[:, :n] = tmp2
    if method == 'OK':
        K[n, :n] = 1.0
        K[:, n] = 1.0
        K[n, n] = 0.0
        R[:, n] = 1.0
    elif method == 'EDK':
        K[n, :n] = 1.0
        K[:n, n] = 1.0
        R[:, n] = 1.0
        K[n:, n:] = 0.0
        for k in range(ref_drfts.shape[1]):
            K[n + 1 + k, :n] = ref_drfts[:, k]
            K[:n, n + 1 + k] = ref_drfts[:, k]
            R[:, n + 1 + k] = dst_drfts[k, :]
    return K, R


def krige_contract(Kinv, R, z, dst_ref_dists, want_var, known_mean=None):
    """interp/steps.py:403-435 for one step: per cell lmds = Kinv @ rhs; if
    not isclose(sum(lmds[:n]), 1) the cell takes the nearest station's
    value (first index on ties) and variance 0; else estimate = lmds[:n].z
    and variance = sum(lmds * rhs) + lmds[n].  Vectorised over cells with
    one GEMM; the per-cell arithmetic is the same.

    known_mean m (simple kriging with a known mean, no counterpart in the
    reference): estimate = m + lmds.(z - m) for every cell, no sum test."""
    n = z.size
    L = R @ Kinv.T                                   # [C x (n+add)]
    if known_mean is not None:
        est = known_mean + L[:, :n] @ (z - known_mean)
        return est, None, np.ones(est.shape, dtype=bool)
    lsum = L[:, :n].sum(axis=1)
    ok = np.isclose(lsum, 1.0)
    nnb = np.argmin(dst_ref_dists, axis=1)
    est = np.where(ok, (L[:, :n] * z[None, :]).sum(axis=1), z[nnb])
    var = None
    if want_var:
        v = (L * R).sum(axis=1) + L[:, n]
        var = np.where(ok, v, 0.0)
    return est, var, ok


def krige_solve(K, R, z, dst_ref_dists, want_var):
    """krige_contract with lambda from numpy.linalg.solve(K, rhs) instead of
    the pseudo-inverse: the FP64 reference for ill-conditioned stress
    systems, where pinv's rcond cut and a direct solve differ."""
    n = z.size
    L = np.linalg.solve(K, R.T).T
    lsum = L[:, :n].sum(axis=1)
    ok = np.isclose(lsum, 1.0)
    nnb = np.argmin(dst_ref_dists, axis=1)
    est = np.where(ok, (L[:, :n] * z[None, :]).sum(axis=1), z[nnb])
    var = None
    if want_var:
        v = (L * R).sum(axis=1) + L[:, n]
        var = np.where(ok, v, 0.0)
    return est, var, ok


def _refined_solve(C, B, its=2):
    """X = C^-1 B by LU (float64) with `its` steps of iterative refinement
    whose residuals B - C X are formed in long double (64-bit mantissa), so
    that X is accurate to float64 rounding while kappa(C) u_64 << 1."""
    import scipy.linalg as sla
    lu = sla.lu_factor(C)
    X = sla.lu_solve(lu, B)
    Cl = C.astype(np.longdouble)
    Bl = B.astype(np.longdouble)
    for _ in range(its):
        r = (Bl - Cl @ X.astype(np.longdouble)).astype(np.float64)
        X = (X.astype(np.longdouble) +
             sla.lu_solve(lu, r).astype(np.longdouble)).astype(np.float64)
    return X


def krige_exact(method, vg_str, ref_ref_dists, dst_ref_dists, z,
                ref_drfts=None, dst_drfts=None, min_vg_val=0.0,
                want_var=False, known_mean=None, its=2):
    """The exact solution of the reference's kriging system of one step
    (interp/steps.py:181-234, 403-435) without its per-cell sum-of-weights
    test: the float64-rounded answer that pinv (steps.py:351) approximates.

    For OK / EDK the semivariogram matrix G is turned into C = c0 - G
    (c0 = the sill sum for bounded models, else 2 max|G| + |sill sum|).
    Because the weights sum to one, G lambda = c0 1 - C lambda, so the
    bordered system [[G, F], [F^T, 0]] [lambda; mu] = [g; f] (F = [1 | S],
    f = [1; s_cell]) becomes C lambda = k + F mu, F^T lambda = f with
    k = c0 - g: lambda = C^-1 k + W mu, mu = B^-1 (f - F^T C^-1 k),
    W = C^-1 F, B = F^T W.  OK variance (steps.py:432-434):
    lambda.g + 2 mu = c0 - lambda.k + 2 mu.  SK: lambda = Cov^-1 cov
    (steps.py:218-219); with known_mean m: m + lambda.(z - m).
    Every C^-1 is _refined_solve (long-double residuals), so the result is
    exact to float64 rounding for kappa(C) up to about 1e15, where float64
    pinv / solve lose kappa(C) u digits.  Returns (est [C], var [C] or
    None, lambda_sum [C])."""
    n = z.size
    if method in ('SK', 'SKM'):
        Cm = np.empty((n, n))
        fill_vg_var_arr(ref_ref_dists, Cm, 1, 1, vg_str, min_vg_val)
        k = np.empty(dst_ref_dists.shape)
        fill_vg_var_arr(dst_ref_dists, k, 1, 0, vg_str, min_vg_val)
        lam = _refined_solve(Cm, k.T, its).T
        if method == 'SKM' or known_mean is not None:
            m = z.mean() if known_mean is None else known_mean
            return m + lam @ (z - m), None, lam.sum(1)
        return lam @ z, None, lam.sum(1)
    G = np.empty((n, n))
    fill_vg_var_arr(ref_ref_dists, G, 0, 1, vg_str, min_vg_val)
    g = np.empty(dst_ref_dists.shape)
    fill_vg_var_arr(dst_ref_dists, g, 0, 0, vg_str, min_vg_val)
    comps = parse_vg(vg_str)
    bounded = all(nm not in ('Rng', 'Pow', 'Hol') for nm, _, _ in comps)
    ssum = sum(s for _, s, _ in comps)
    c0 = ssum if bounded else 2 * np.abs(G).max() + abs(ssum)
    C = c0 - G
    k = c0 - g                                       # [cells x n]
    F = np.ones((n, 1))
    f = np.ones((k.shape[0], 1))
    if method == 'EDK':
        F = np.concatenate([F, np.asarray(ref_drfts, np.float64)], axis=1)
        f = np.concatenate([f, np.asarray(dst_drfts, np.float64).T], axis=1)
    elif method != 'OK':
        raise NotImplementedError(method)
    X = _refined_solve(C, np.concatenate([k.T, F], axis=1), its)
    Ck, W = X[:, :k.shape[0]], X[:, k.shape[0]:]
    B = F.T @ W
    mu = np.linalg.solve(B, (f - Ck.T @ F).T).T      # [cells x p]
    lam = Ck.T + mu @ W.T                            # [cells x n]
    est = lam @ z
    var = None
    if want_var and method == 'OK':
        var = c0 - (lam * k).sum(1) + 2.0 * mu[:, 0]
    return est, var, lam.sum(1)


def get_interp(ref_data, interp_type, ref_drfts, dst_drfts, idw_exp, models,
               interp_steps_flags, nuggetness_flags, dst_ref_dists,
               ref_ref_dists, min_vg_val, want_var, prblm=None,
               sub_steps=None):
    """interp/steps.py:245-401 for one (time group, cell group).

    ref_data [T x n] (no NaN), dst_ref_dists [C x n], ref_ref_dists [n x n].
    Returns (dst_data [T x C] float64, est_vars [T x C] float64 or None).
    The reference stores est_vars as float32 (steps.py:269-271); callers
    cast.  IDW normalises each cell's distances by their maximum in place
    (steps.py:297-301); that changes no argmin here, so the copy is kept.
    """
    n_dsts = dst_ref_dists.shape[0]
    n_refs = ref_data.shape[1]
    n_time = ref_data.shape[0]
    dst_data = np.full((n_time, n_dsts), np.nan)
    est_vars = None
    if want_var and interp_type == 'OK':
        est_vars = np.full((n_time, n_dsts), np.nan)
    ref_means = ref_data.mean(axis=1)
    if n_refs == 1:
        for j in range(n_time):
            dst_data[j, :] = ref_data[j, :]
    elif interp_type == 'NNB':
        nnb = np.argmin(dst_ref_dists, axis=1)
        dst_data[:, :] = ref_data[:, nnb]
    elif interp_type == 'IDW':
        d = dst_ref_dists.copy()
        dmax = d.max(axis=1)
        pos = dmax > 0
        d[pos] /= dmax[pos, None]
        with np.errstate(divide='ignore', invalid='ignore'):
            w = 1.0 / (d ** idw_exp)
            wsum = w.sum(axis=1)
            assert np.all(wsum >= 1e-14)
            num = ref_data @ w.T
            vals = num / wsum[None, :]
        for j in range(n_time):
            dst_data[j, :] = vals[j] if interp_steps_flags[j] else ref_means[j]
    elif interp_type in ('OK', 'SK', 'SKM', 'EDK'):
        order = np.argsort(np.asarray(models, dtype=object).astype(str),
                           kind='stable')
        old_model = ''
        failed = False
        Kinv = None
        for j in order:
            model = models[j]
            if (not interp_steps_flags[j]) or nuggetness_flags[j]:
                dst_data[j, :] = ref_means[j]
                if est_vars is not None:
                    est_vars[j, :] = 0.0
                continue
            if model == 'nan':
                raise ValueError('NaN VG!')
            if model != old_model:
                K, R = krig_system(interp_type, model, ref_ref_dists,
                                   dst_ref_dists, ref_drfts, dst_drfts,
                                   min_vg_val)
                old_model = model
                try:
                    Kinv = np.linalg.pinv(K)
                    failed = False
                except Exception:
                    failed = True
                    if prblm is not None and sub_steps is not None:
                        if sub_steps[j] not in prblm:
                            prblm.append(sub_steps[j])
            if failed:
                nnb = np.argmin(dst_ref_dists, axis=1)
                dst_data[j, :] = ref_data[j, nnb]
                if est_vars is not None:
                    est_vars[j, :] = 0.0
            else:
                est, var, _ = krige_contract(
                    Kinv, R, ref_data[j], dst_ref_dists,
                    est_vars is not None,
                    known_mean=ref_means[j] if interp_type == 'SKM' else None)
                dst_data[j, :] = est
                if est_vars is not None:
                    est_vars[j, :] = var
    else:
        raise NotImplementedError(interp_type)
    return dst_data, est_vars


def mod_min_max(fld, min_var_cut, max_var_cut):
    """interp/steps.py:466-476 (clip estimates, NaN untouched)."""
    with np.errstate(invalid='ignore'):
        if min_var_cut is not None:
            fld[fld < min_var_cut] = min_var_cut
        if max_var_cut is not None:
            fld[fld > max_var_cut] = max_var_cut


def interp_all(stn_x, stn_y, cell_x, cell_y, data, vgs, methods, *,
               stn_drfts=None, cell_drfts=None, min_var_thr=-np.inf,
               min_var_cut=None, max_var_cut=None, min_vg_val=0.0,
               neb_sel_mthd='all', n_nebs=None, n_pies=None,
               est_vars_flag=True, stn_keep=None):
    """The numeric core of interp/steps.py:478-835 without files.

    data [T x N] with NaN for missing stations, vgs a length-T list of
    variogram strings (or None without kriging), methods a list of
    (type, label, idw_exp).  Returns (fields, fields64, prblm_steps):
    fields[label] is the [T x C] float32 output (steps.py:659-663,
    815-831), fields64 the same before the float32 cast (variances in
    float64 too).  Steps without stations stay NaN (steps.py:677-688).
    """
    T, N = data.shape
    C = cell_x.size
    types = [m[0] for m in methods]
    labels = [m[1] for m in methods]
    edk = 'EDK' in types
    want_var = est_vars_flag and ('OK' in types)
    fields = {lab: np.full((T, C), np.nan, dtype=np.float32) for lab in labels}
    f64 = {lab: np.full((T, C), np.nan) for lab in labels}
    if want_var:
        fields['EST_VARS_OK'] = np.full((T, C), np.nan, dtype=np.float32)
        f64['EST_VARS_OK'] = np.full((T, C), np.nan)
    prblm = []
    steps = np.arange(T)
    if neb_sel_mthd != 'all':
        # stations that no cell selects with every station present are
        # dropped before the time grouping (interp/steps.py:592-608); a
        # dropped station behaves as one without values.  stn_keep: the
        # kept stations of a larger grid whose cells these are a sample of
        if stn_keep is not None:
            neb0 = np.where(np.asarray(stn_keep, dtype=bool))[0]
        elif neb_sel_mthd == 'nrst':
            neb0 = neb_idxs_nrst(cell_x, cell_y, stn_x, stn_y, n_nebs)
        elif neb_sel_mthd == 'pie':
            neb0 = neb_idxs_pie(cell_x, cell_y, stn_x, stn_y, n_nebs, n_pies)
        else:
            raise NotImplementedError(neb_sel_mthd)
        drop = np.ones(N, dtype=bool)
        drop[np.unique(neb0)] = False
        if drop.any():
            data = data.copy()
            data[:, drop] = np.nan
    dst_ref_all = dists_2d(cell_x, cell_y, stn_x, stn_y)
    ref_ref_all = dists_2d(stn_x, stn_y, stn_x, stn_y)
    for stns, tmask in time_groups(data):
        sub_steps = steps[tmask]
        if not stns.size:
            for s in sub_steps:
                if s not in prblm:
                    prblm.append(s)
            continue
        vals = data[np.ix_(tmask, stns)]
        tidx = np.where(tmask)[0]
        if vgs is not None:
            models = [vgs[t] for t in tidx]
            nugf = np.array([check_full_nuggetness(m, min_vg_val)
                             for m in models])
        else:
            models, nugf = None, None
        sx, sy = stn_x[stns], stn_y[stns]
        if neb_sel_mthd == 'all':
            neb = neb_idxs_all(C, stns.size)
        elif neb_sel_mthd == 'nrst':
            neb = neb_idxs_nrst(cell_x, cell_y, sx, sy, n_nebs)
        elif neb_sel_mthd == 'pie':
            neb = neb_idxs_pie(cell_x, cell_y, sx, sy, n_nebs, n_pies)
        else:
            raise NotImplementedError(neb_sel_mthd)
        for cells in neb_groups(neb):
            sub = neb[cells[0]]
            sub_vals = vals[:, sub]
            gstns = stns[sub]
            dst_ref = dst_ref_all[np.ix_(cells, gstns)]
            ref_ref = ref_ref_all[np.ix_(gstns, gstns)]
            flags = np.array([np.any(sub_vals[j] >= min_var_thr)
                              for j in range(sub_vals.shape[0])])
            if edk:
                s_rd = stn_drfts[gstns, :]
                s_dd = cell_drfts[:, cells]
            else:
                s_rd = s_dd = None
            for (typ, lab, idw_exp) in methods:
                if lab == 'EST_VARS_OK':
                    continue
                ivals, evars = get_interp(
                    sub_vals, typ, s_rd, s_dd,
                    idw_exp if typ == 'IDW' else 5, models, flags, nugf,
                    dst_ref, ref_ref, min_vg_val, want_var, prblm, sub_steps)
                mod_min_max(ivals, min_var_cut, max_var_cut)
                for j, t in enumerate(tidx):
                    f64[lab][t, cells] = ivals[j]
                    fields[lab][t, cells] = ivals[j].astype(np.float32)
                    if evars is not None:
                        f64['EST_VARS_OK'][t, cells] = evars[j]
                        fields['EST_VARS_OK'][t, cells] = evars[j]
    return fields, f64, prblm


# ---------------------------------------------------------------------------
# Class API, cyth/interpmthds.pyx:251-719
# ---------------------------------------------------------------------------


def _class_parse(model):
    comps = []
    for sub in model.split('+'):
        sub = sub.strip()
        sill, vg = sub.split(' ')
        vg, rng = vg.split('(')
        rng = rng.split(')')[0]
        comps.append((vg, float(sill), max(1e-5, float(rng))))
    return comps


def _class_in_vars(xi, yi, comps, sign, start):
    n = xi.size
    d = dists_2d(xi, yi, xi, yi)
    M = np.full((n, n), start, dtype=np.float64)
    off = ~np.eye(n, dtype=bool)
    for name, sill, rng in comps:
        M[off] += sign * VG_FTNS[name](d[off], rng, sill)
    return M


def _class_out_vars(xi, yi, xk, yk, comps, sign, start):
    d = dists_2d(xk, yk, xi, yi)
    out = np.full(d.shape, start, dtype=np.float64)
    nz = d != 0.0
    for name, sill, rng in comps:
        out[nz] += sign * VG_FTNS[name](d[nz], rng, sill)
    return out


def ordinary_kriging(xi, yi, zi, xk, yk, model='1.0 Sph(2)'):
    """pyx:251-362: zero-diagonal Gamma, rhs skips d == 0, pinv, variance
    max(0, lambda.gamma + mu).  Returns dict of the class attributes."""
    comps = _class_parse(model)
    n = xi.size
    K = np.zeros((n + 1, n + 1))
    K[:n, :n] = _class_in_vars(xi, yi, comps, 1.0, 0.0)
    K[n, :n] = 1.0
    K[:, n] = 1.0
    K[n, n] = 0.0
    Kinv = np.linalg.pinv(K)
    g = _class_out_vars(xi, yi, xk, yk, comps, 1.0, 0.0)
    R = np.concatenate([g, np.ones((xk.size, 1))], axis=1)
    L = R @ Kinv.T
    mus = L[:, n]
    lam = L[:, :n]
    zk = (lam * zi[None, :]).sum(axis=1)
    ev = np.maximum(0.0, (lam * g).sum(axis=1) + mus)
    return dict(zk=zk, est_vars=ev, mus=mus, lambdas=lam, rhss=R,
                in_vars=K, in_vars_inv=Kinv)


def simple_kriging(xi, yi, zi, xk, yk, model='1.0 Sph(2)'):
    """pyx:365-471.  in_covars starts at 0 (self.covar before the sills are
    summed), so K = -Gamma off the diagonal and 0 on it; rhs starts at the
    summed sill.  est_covars = max(0, covar - lambda.rhs)."""
    comps = _class_parse(model)
    covar = sum(c[1] for c in comps)
    K = _class_in_vars(xi, yi, comps, -1.0, 0.0)
    Kinv = np.linalg.pinv(K)
    rhs = _class_out_vars(xi, yi, xk, yk, comps, -1.0, covar)
    lam = rhs @ Kinv.T
    zk = (lam * zi[None, :]).sum(axis=1)
    ec = np.maximum(0.0, covar - (lam * rhs).sum(axis=1))
    return dict(zk=zk, est_covars=ec, lambdas=lam, rhss=rhs, out_covars=rhs,
                in_covars=K, in_covars_inv=Kinv)


def external_drift_kriging_md(xi, yi, zi, si, xk, yk, sk, model='1.0 Sph(2)'):
    """pyx:588-719: zero-diagonal Gamma, a [1 | S] constraint block with a
    zero corner, rhs skips d == 0; mus_arr[(m+1) x C]."""
    assert si.ndim == 2 and xi.ndim == 1
    assert xi.shape[0] == yi.shape[0] == zi.shape[0] == si.shape[1]
    assert xk.shape[0] == yk.shape[0] == sk.shape[1]
    assert si.shape[0] == sk.shape[0]
    comps = _class_parse(model)
    n, m = xi.size, si.shape[0]
    K = np.zeros((n + m + 1, n + m + 1))
    K[:n, :n] = _class_in_vars(xi, yi, comps, 1.0, 0.0)
    K[n, :n] = 1.0
    K[:n, n] = 1.0
    for k in range(m):
        K[n + 1 + k, :n] = si[k]
        K[:n, n + 1 + k] = si[k]
    Kinv = np.linalg.pinv(K)
    g = _class_out_vars(xi, yi, xk, yk, comps, 1.0, 0.0)
    R = np.concatenate([g, np.ones((xk.size, 1)), sk.T], axis=1)
    L = R @ Kinv.T
    lam = L[:, :n]
    zk = (lam * zi[None, :]).sum(axis=1)
    return dict(zk=zk, mus_arr=L[:, n:].T.copy(), lambdas=lam, rhss=R,
                in_vars=K, in_vars_inv=Kinv)


# ---------------------------------------------------------------------------
# CPU baseline timing: the reference's per-step cost structure
# ---------------------------------------------------------------------------


def time_ok_step_loop(stn_x, stn_y, cell_x, cell_y, z_t, vg_str,
                      min_vg_val=0.0):
    """Time one OK step (estimate + variance) the way interp/steps.py runs it
    for a step with its own variogram: semivariances of the cells against
    the step's stations (steps.py:639-650 via pyx:146-226), the station
    matrix and its pinv (steps.py:336-352), then one np.matmul per cell
    and the isclose / estimate / variance of steps.py:415-434.

    Returns (t_step_setup, t_per_cell, est, var) for the given cells; the
    caller extrapolates to a full grid as t_setup + C * t_per_cell.
    """
    import time
    act = ~np.isnan(z_t)
    sx, sy, z = stn_x[act], stn_y[act], z_t[act]
    n = z.size
    t0 = time.perf_counter()
    rr = dists_2d(sx, sy, sx, sy)
    K = np.empty((n + 1, n + 1))
    tmp = np.empty((n, n))
    fill_vg_var_arr(rr, tmp, 0, 1, vg_str, min_vg_val)
    K[:n, :n] = tmp
    K[n, :n] = 1.0
    K[:, n] = 1.0
    K[n, n] = 0.0
    Kinv = np.linalg.pinv(K)
    t1 = time.perf_counter()
    dr = dists_2d(cell_x, cell_y, sx, sy)
    R = np.empty((cell_x.size, n + 1))
    tmp2 = np.empty((cell_x.size, n))
    fill_vg_var_arr(dr, tmp2, 0, 0, vg_str, min_vg_val)
    R[:, :n] = tmp2
    R[:, n] = 1.0
    est = np.empty(cell_x.size)
    var = np.empty(cell_x.size)
    for i in range(cell_x.size):
        lmds = np.matmul(Kinv, R[i])
        if not np.isclose(lmds[:n].sum(), 1.0):
            est[i] = z[np.argmin(dr[i])]
            var[i] = 0.0
        else:
            est[i] = (lmds[:n] * z).sum()
            var[i] = (lmds * R[i]).sum() + lmds[n]
    t2 = time.perf_counter()
    return t1 - t0, (t2 - t1) / cell_x.size, est, var


# ---------------------------------------------------------------------------
# vgclus weight heuristic, vgclus/tvgs_clus.py:92-137
# ---------------------------------------------------------------------------

def threshold_wts(wts, min_kg_abs_wt):
    """tvgs_clus.py:109-130 on sorted weights: weights whose share of
    sum |w| is at most min_kg_abs_wt become 0 (unless that is all of them)
    and the rest are rescaled to sum to one."""
    wts = np.array(wts, dtype=np.float64)
    if min_kg_abs_wt:
        abs_wts = np.abs(wts)
        rel_wts = abs_wts / abs_wts.sum()
        zero_idxs = rel_wts <= min_kg_abs_wt
        if zero_idxs.sum() == zero_idxs.size:
            zero_idxs = np.zeros_like(zero_idxs, dtype=bool)
        sclr = 1.0 / wts[~zero_idxs].sum()
        assert zero_idxs.sum() < zero_idxs.size
        wts[zero_idxs] = 0.0
        wts[~zero_idxs] *= sclr
    assert (wts[1:] - wts[:-1]).min() >= 0.0
    assert np.isclose(wts.sum(), 1.0)
    return wts


def srtd_krg_wts(tvg_str, rnd_pts, min_kg_abs_wt=None):
    """tvgs_clus.py:92-137: the ordinary-kriging weights (class API,
    pyx:251-362) of rnd_pts[1:] for the point rnd_pts[0], sorted ascending,
    then thresholded."""
    rnd_pts = np.asarray(rnd_pts, dtype=np.float64)
    out = ordinary_kriging(rnd_pts[1:, 0], rnd_pts[1:, 1],
                           np.zeros(rnd_pts.shape[0] - 1),
                           rnd_pts[:1, 0], rnd_pts[:1, 1], tvg_str)
    wts = out['lambdas'].ravel().copy()
    wts.sort()
    return threshold_wts(wts, min_kg_abs_wt)
