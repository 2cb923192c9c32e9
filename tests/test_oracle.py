"""Pins the CPU oracle (oracle/spk_oracle.py) against closed forms,
library routines, invariants and the reference's own known-answer string.

Nothing here retypes an oracle formula: each check comes from a property
a plausible mistake would break (a dropped term, a swapped range/sill, a
wrong sign, a missing Lagrange multiplier).
"""
import math

import numpy as np
import pytest
from scipy.linalg import solve
from scipy.spatial.distance import cdist, pdist

from oracle import spk_oracle as O


# ---- variogram models (pyx:38-83) ----------------------------------------

def test_models_closed_form_points():
    s, r = 2.0, 1000.0
    assert O.sph_vg(0.0, r, s) == 0.0
    assert O.sph_vg(r, r, s) == s                      # h >= r -> sill
    assert math.isclose(O.sph_vg(r / 2, r, s), s * 0.6875, rel_tol=1e-15)
    assert math.isclose(O.exp_vg(r, r, s), s * (1 - math.exp(-3)), rel_tol=1e-15)
    assert math.isclose(O.gau_vg(r, r, s), s * (1 - math.exp(-3)), rel_tol=1e-15)
    assert math.isclose(O.gau_vg(r / 2, r, s), s * (1 - math.exp(-0.75)),
                        rel_tol=1e-15)
    assert O.lin_vg(r / 2, r, s) == s / 2
    assert O.lin_vg(r, r, s) == s                      # h == r: s*(h/r)
    assert O.lin_vg(3 * r, r, s) == s
    assert math.isclose(O.pow_vg(4.0, 1.5, s), s * 8.0, rel_tol=1e-15)
    assert O.pow_vg(0.0, 0.0, s) == s                  # C pow(0, 0) == 1
    assert O.hol_vg(0.0, r, s) == 0.0
    assert math.isclose(O.hol_vg(r / 2, r, s), s * (1 - 2 / math.pi),
                        rel_tol=1e-15)
    assert O.nug_vg(0.0, 0.0, s) == s                  # nugget at h == 0 too
    assert O.rng_vg(7.5, 1.0, s) == 7.5


def test_theo_vg_known_answer():
    """test/test_theo_vg.py:26-28: '100 Sph(10000) + 10 Exp(1000000)' on
    linspace(0, 1e6, 10): h_k = k e6/9, the Sph part is at its sill for
    every k >= 1 (h > 1e4) and the Exp part is 10 (1 - e^{-k/3})."""
    h = np.linspace(0, 1e6, 10)
    got = O.get_theo_vg_vals('100 Sph(10000) + 10 Exp(1000000)', h)
    exp = np.array([0.0] + [100.0 + 10.0 * (1 - math.exp(-k / 3.0))
                            for k in range(1, 10)])
    np.testing.assert_allclose(got, exp, rtol=1e-14, atol=0)


def test_parse_clamps_range_and_rejects_unknown():
    comps = O.parse_vg('1.5 Sph(0.0) + 2 Exp(10)')
    assert comps == [('Sph', 1.5, 1e-5), ('Exp', 2.0, 10.0)]
    with pytest.raises(KeyError):
        O.parse_vg('1 Foo(2)')
    with pytest.raises(ValueError):
        O.parse_vg('1  Sph(2)')  # two spaces: not exactly (sill, rest)


def test_fill_vg_var_arr_invariants():
    rng = np.random.default_rng(1)
    x, y = rng.uniform(0, 1e4, 12), rng.uniform(0, 1e4, 12)
    d = O.dists_2d(x, y, x, y)
    vg = '0.3 Nug(0.0) + 1.7 Exp(3000.0)'
    g = np.empty_like(d)
    c = np.empty_like(d)
    O.fill_vg_var_arr(d, g, 0, 1, vg, -np.inf)
    O.fill_vg_var_arr(d, c, 1, 1, vg, -np.inf)
    # covariance + semivariance = total sill everywhere
    np.testing.assert_allclose(g + c, 2.0, rtol=0, atol=1e-15)
    # the nugget reaches the diagonal of the semivariogram matrix
    np.testing.assert_allclose(np.diag(g), 0.3, rtol=0, atol=0)
    np.testing.assert_array_equal(g, g.T)
    # SK covariance diagonal: the nugget contributes s - s = 0
    np.testing.assert_allclose(np.diag(c), 1.7, rtol=0, atol=1e-15)
    # min_vg_val zeroes small entries
    g2 = np.empty_like(d)
    O.fill_vg_var_arr(d, g2, 0, 0, vg, 1.0)
    assert np.all((g2 == 0) | (g2 > 1.0))
    assert np.all(g2[g > 1.0] == g[g > 1.0])


def test_dists_match_cdist_and_pdist():
    rng = np.random.default_rng(2)
    a = rng.uniform(3.5e5, 4.5e5, (30, 2))
    b = rng.uniform(3.5e5, 4.5e5, (17, 2))
    d = O.dists_2d(a[:, 0], a[:, 1], b[:, 0], b[:, 1])
    np.testing.assert_allclose(d, cdist(a, b), rtol=1e-15)
    pts = rng.normal(size=(9, 3))
    cond = O.get_nd_dists(pts)
    # pdist is ordered i < j row-major; the reference orders i > j by i
    full = cdist(pts, pts)
    exp = [full[i, j] for i in range(9) for j in range(i)]
    np.testing.assert_allclose(cond, exp, rtol=1e-15)
    assert cond.size == pdist(pts).size


def test_idw_helpers_closed_form():
    d = np.array([1.0, 2.0, 4.0])
    w = np.empty(3)
    s = O.fill_wts_and_sum(d, w, 2.0)
    np.testing.assert_array_equal(w, [1.0, 0.25, 0.0625])
    assert s == 1.3125
    assert O.get_mults_sum(w, np.array([2.0, 4.0, 16.0])) == 4.0


# ---- kriging systems (interp/steps.py:170-435) -------------------------

def _system_case(n=6, seed=3, vg='0.2 Nug(0.0) + 1.0 Sph(6000.0)'):
    rng = np.random.default_rng(seed)
    sx, sy = rng.uniform(0, 1e4, n), rng.uniform(0, 1e4, n)
    cx, cy = rng.uniform(0, 1e4, 40), rng.uniform(0, 1e4, 40)
    rr = O.dists_2d(sx, sy, sx, sy)
    dr = O.dists_2d(cx, cy, sx, sy)
    z = rng.gamma(0.8, 5.0, n)
    return sx, sy, cx, cy, rr, dr, z, vg


def test_ok_two_station_closed_form():
    """Cell equidistant from two stations: lambda = (1/2, 1/2) by symmetry,
    mu = g_c - (gamma_0 + gamma_12) / 2 with gamma_0 = nugget on the
    diagonal, variance = lambda.g + 2 mu = 3 g_c - gamma_0 - gamma_12
    (interp/steps.py:432-434 counts mu twice)."""
    vg = '0.1 Nug(0.0) + 1.0 Exp(500.0)'
    sx, sy = np.array([0.0, 100.0]), np.array([0.0, 0.0])
    cx, cy = np.array([50.0]), np.array([80.0])
    rr = O.dists_2d(sx, sy, sx, sy)
    dr = O.dists_2d(cx, cy, sx, sy)
    K, R = O.krig_system('OK', vg, rr, dr, None, None, 0.0)
    Kinv = np.linalg.pinv(K)
    z = np.array([3.0, 7.0])
    est, var, ok = O.krige_contract(Kinv, R, z, dr, True)
    gc = 0.1 + (1 - math.exp(-3 * math.hypot(50, 80) / 500.0))
    g12 = 0.1 + (1 - math.exp(-3 * 100 / 500.0))
    assert ok[0]
    assert math.isclose(est[0], 5.0, rel_tol=1e-12)
    assert math.isclose(var[0], 3 * gc - 0.1 - g12, rel_tol=1e-12)


def test_ok_invariants():
    sx, sy, cx, cy, rr, dr, z, vg = _system_case()
    K, R = O.krig_system('OK', vg, rr, dr, None, None, 0.0)
    L = R @ np.linalg.pinv(K).T
    n = z.size
    np.testing.assert_allclose(L[:, :n].sum(1), 1.0, atol=1e-12)
    # weights do not depend on the nugget (constant shift of all entries)
    K2, R2 = O.krig_system('OK', '1.0 Sph(6000.0)', rr, dr, None, None, 0.0)
    L2 = R2 @ np.linalg.pinv(K2).T
    np.testing.assert_allclose(L[:, :n], L2[:, :n], atol=1e-11)
    # same solution as a direct solve
    L3 = solve(K, R.T).T
    np.testing.assert_allclose(L, L3, atol=1e-10)


def test_ok_exact_at_station_and_variance_is_nugget():
    sx, sy, cx, cy, rr, dr, z, vg = _system_case(vg='0.25 Nug(0.0) + 1 Exp(4000.0)')
    dr_st = O.dists_2d(sx, sy, sx, sy)    # cells placed on the stations
    K, R = O.krig_system('OK', vg, rr, dr_st, None, None, 0.0)
    est, var, ok = O.krige_contract(np.linalg.pinv(K), R, z, dr_st, True)
    np.testing.assert_allclose(est, z, rtol=1e-10)
    np.testing.assert_allclose(var, 0.25, rtol=1e-9)


def test_edk_reproduces_drift_and_sk_matches_solve():
    sx, sy, cx, cy, rr, dr, z, vg = _system_case(n=9)
    sd = (0.3 * sx + 0.1 * sy)[:, None]
    cd = (0.3 * cx + 0.1 * cy)[None, :]
    K, R = O.krig_system('EDK', vg, rr, dr, sd, cd, 0.0)
    L = R @ np.linalg.pinv(K).T
    n = z.size
    np.testing.assert_allclose(L[:, :n] @ sd[:, 0], cd[0], rtol=1e-9)
    np.testing.assert_allclose(L[:, :n].sum(1), 1.0, atol=1e-11)
    # a field that is exactly linear in the drift is reproduced exactly
    est, _, ok = O.krige_contract(np.linalg.pinv(K), R, 2.0 + 3.0 * sd[:, 0],
                                  dr, False)
    np.testing.assert_allclose(est[ok], (2.0 + 3.0 * cd[0])[ok], rtol=1e-9)
    Ks, Rs = O.krig_system('SK', vg, rr, dr, None, None, 0.0)
    lam = solve(Ks, Rs.T).T
    np.testing.assert_allclose(Rs @ np.linalg.pinv(Ks).T, lam, atol=1e-10)


def test_sk_known_mean_invariants():
    """Simple kriging with the step mean as known mean (labelled SKM): a
    cell beyond the range of every station has zero covariances and takes
    the mean; a cell on a station (no nugget in the SK covariance) takes
    that station's value; shifting all values by t shifts every estimate
    by t (the mean shifts with them)."""
    sx, sy, cx, cy, rr, dr, z, vg = _system_case(n=7, vg='1.0 Sph(3000.0)')
    m = z.mean()
    far = np.array([[1e6, 1e6]])
    dr_far = O.dists_2d(far[:, 0], far[:, 1], sx, sy)
    K, R = O.krig_system('SKM', vg, rr, dr_far, None, None, 0.0)
    est, var, ok = O.krige_contract(np.linalg.pinv(K), R, z, dr_far, False,
                                    known_mean=m)
    assert var is None and ok.all()
    np.testing.assert_allclose(est, m, rtol=1e-15)
    dr_st = O.dists_2d(sx, sy, sx, sy)
    K, R = O.krig_system('SKM', vg, rr, dr_st, None, None, 0.0)
    est, _, _ = O.krige_contract(np.linalg.pinv(K), R, z, dr_st, False,
                                 known_mean=m)
    np.testing.assert_allclose(est, z, rtol=1e-10)
    K, R = O.krig_system('SKM', vg, rr, dr, None, None, 0.0)
    Ki = np.linalg.pinv(K)
    e1, _, _ = O.krige_contract(Ki, R, z, dr, False, known_mean=m)
    e2, _, _ = O.krige_contract(Ki, R, z + 7.5, dr, False, known_mean=m + 7.5)
    np.testing.assert_allclose(e2 - e1, 7.5, rtol=1e-12)
    # the reference's SK on the same system differs (no mean term)
    e_sk, _, _ = O.krige_contract(Ki, R, z, dr, False)
    assert np.abs(e_sk - e1).max() > 1e-3


def test_get_interp_idw_and_special_steps():
    # two stations at distances 1 and 2: normalised 0.5, 1 -> w = 4, 1
    dr = np.array([[1.0, 2.0]])
    data = np.array([[10.0, 20.0], [1.0, 1.0]])
    out, _ = O.get_interp(data, 'IDW', None, None, 2.0, None,
                          np.array([True, False]), None, dr, None, 0.0, False)
    assert math.isclose(out[0, 0], (4 * 10.0 + 20.0) / 5, rel_tol=1e-15)
    assert out[1, 0] == 1.0                      # flag off -> step mean
    # single station: its value everywhere, variance stays NaN
    out, ev = O.get_interp(np.array([[5.0]]), 'OK', None, None, 5, ['1 Sph(1)'],
                           np.array([True]), np.array([False]),
                           np.array([[3.0], [4.0]]), np.zeros((1, 1)), 0.0,
                           True)
    assert np.all(out == 5.0) and np.all(np.isnan(ev))
    # nugget-only step -> mean fill and zero variance
    rr = np.array([[0.0, 1.0], [1.0, 0.0]])
    out, ev = O.get_interp(np.array([[1.0, 3.0]]), 'OK', None, None, 5,
                           ['1.0 Nug(0.0)'], np.array([True]),
                           np.array([O.check_full_nuggetness('1.0 Nug(0.0)',
                                                             0.0)]),
                           np.array([[0.5, 0.7]]), rr, 0.0, True)
    assert out[0, 0] == 2.0 and ev[0, 0] == 0.0


def test_nuggetness_parse():
    assert O.check_full_nuggetness('1.0 Nug(0.0)', 0.0)
    assert not O.check_full_nuggetness('0.1 Nug(0.0) + 0.9 Sph(25.0)', 0.0)
    assert O.check_full_nuggetness('0.0 Sph(25.0)', 0.0)
    assert not O.check_full_nuggetness('nan', 0.0)


# ---- neighbour selection (interp/grps.py) ----------------------------------

def test_nrst_matches_brute_force():
    rng = np.random.default_rng(5)
    sx, sy = rng.uniform(0, 1, 40), rng.uniform(0, 1, 40)
    cx, cy = rng.uniform(0, 1, 25), rng.uniform(0, 1, 25)
    got = O.neb_idxs_nrst(cx, cy, sx, sy, 7)
    d = cdist(np.c_[cx, cy], np.c_[sx, sy])
    for i in range(25):
        exp = np.sort(np.lexsort((np.arange(40), d[i]))[:7])
        np.testing.assert_array_equal(got[i], exp)


def test_pie_quadrants_and_ranks():
    """pyx:850-870: x offset 0 -> angle 0 whatever y; west (y == 0, x < 0)
    gives atan(-0.0) and stays in pie 0; NW/SW add pi, SE adds 2 pi."""
    sx = np.array([1.0, 0.0, -3.0, -1.0, -3.0, 2.0, 5.0])
    sy = np.array([0.0, 2.0, 0.0, 2.0, -3.0, -2.0, 0.0])
    n = sx.size
    dists = np.zeros(n)
    tem = np.zeros(n)
    sel = np.zeros(n, dtype=np.int64)
    pie = np.zeros(n, dtype=np.uint64)
    cts = np.zeros(4, dtype=np.uint64)
    O.sel_equidist_refs(0.0, 0.0, sx, sy, 4, -1.0, -1, dists, tem, sel, pie,
                        cts)
    # angles: 0, (x=0) 0, -0.0, pi-atan2 -> 2.03, 5pi/4, 7pi/4, 0
    np.testing.assert_array_equal(pie, [0, 0, 0, 1, 2, 3, 0])
    np.testing.assert_array_equal(cts, [4, 1, 1, 1])
    # distance ranks inside each pie: pie 0 has d = 1, 2, 3, 5
    np.testing.assert_array_equal(sel, [0, 1, 2, 0, 0, 0, 3])
    # tem holds the last non-empty pie's distances, inf elsewhere
    assert np.isinf(tem[[0, 1, 2, 3, 4, 6]]).all()
    assert tem[5] == math.hypot(2, 2)
    out = O.neb_idxs_pie(np.array([0.0]), np.array([0.0]), sx, sy, 5, 4)
    # rank 0 by distance: 0 (1), 3 (2.24), 5 (2.83), 4 (4.24); rank 1: 1
    np.testing.assert_array_equal(out[0], [0, 1, 3, 4, 5])


def test_time_groups_first_appearance():
    d = np.array([[1.0, np.nan, 2.0], [3.0, 4.0, 5.0], [6.0, np.nan, 7.0]])
    g = O.time_groups(d)
    assert len(g) == 2
    np.testing.assert_array_equal(g[0][0], [0, 2])
    np.testing.assert_array_equal(g[0][1], [True, False, True])
    np.testing.assert_array_equal(g[1][0], [0, 1, 2])


def test_interp_all_small_runs():
    from spinterps_amd import configs
    w = configs.small(n_cells_side=8)
    vals, vgs = configs.small_steps(w, 0, 3)
    f, f64, prb = O.interp_all(w.stn_x, w.stn_y, w.cell_x, w.cell_y, vals,
                               vgs, [('IDW', 'IDW', 2.0), ('OK', 'OK', None)])
    assert f['OK'].dtype == np.float32 and f['OK'].shape == (3, 64)
    assert np.isfinite(f64['OK']).all() and np.isfinite(f64['EST_VARS_OK']).all()
    assert not prb


def _neb_case(seed=11):
    """30 stations on [0, 10]^2 plus one far away (never among the 6
    nearest of any cell: dropped), 49 cells, 4 steps with missing
    stations; low values in the west so that some neighbourhoods fall
    entirely below min_var_thr while their step does not."""
    rng = np.random.default_rng(seed)
    sx = np.r_[rng.uniform(0, 10, 30), 60.0]
    sy = np.r_[rng.uniform(0, 10, 30), 60.0]
    g = np.linspace(0.5, 9.5, 7)
    cx, cy = [a.ravel() for a in np.meshgrid(g, g[::-1])]
    T = 4
    z = rng.gamma(2.0, 2.0, (T, sx.size)) + 0.1
    z[:, sx < 3.5] = rng.uniform(0.0, 0.4, (T, int((sx < 3.5).sum())))
    z[1, rng.choice(30, 8, replace=False)] = np.nan
    z[2, :25] = np.nan          # 5 left + the far one: only it is dropped
    z[3, rng.choice(30, 3, replace=False)] = np.nan
    vgs = ['0.1 Nug(0.0) + 2.0 Sph(6.0)', '1.5 Exp(4.0)',
           '0.05 Nug(0.0) + 1.0 Gau(5.0)', '2.0 Sph(8.0)']
    return sx, sy, cx, cy, z, vgs


def _brute_neb(sx, sy, cx, cy, z, vgs, k, thr):
    """Per cell, without cell groups: the k nearest active stations that
    survive the never-selected drop, then OK by numpy.linalg.solve of the
    bordered system, IDW and nearest station over them."""
    d_all = cdist(np.c_[cx, cy], np.c_[sx, sy])
    keep = np.zeros(sx.size, dtype=bool)
    for i in range(cx.size):
        keep[np.lexsort((np.arange(sx.size), d_all[i]))[:k]] = True
    T, C = z.shape[0], cx.size
    out = {lab: np.full((T, C), np.nan) for lab in ('OK', 'V', 'IDW', 'NNB')}
    for t in range(T):
        act = np.where(~np.isnan(z[t]) & keep)[0]
        for i in range(C):
            s = act[np.lexsort((act, d_all[i, act]))[:k]]
            zs, d = z[t, s], d_all[i, s]
            out['NNB'][t, i] = zs[np.argmin(d)]
            if s.size == 1:
                out['OK'][t, i] = out['IDW'][t, i] = zs[0]
                continue
            if not np.any(zs >= thr):
                out['OK'][t, i] = out['IDW'][t, i] = zs.mean()
                out['V'][t, i] = 0.0
                continue
            w = 1.0 / (d / d.max()) ** 2
            out['IDW'][t, i] = (w * zs).sum() / w.sum()
            n = s.size
            K = np.zeros((n + 1, n + 1))
            K[:n, :n] = O.vg_sum(cdist(np.c_[sx[s], sy[s]],
                                       np.c_[sx[s], sy[s]]), vgs[t])
            K[n, :n] = K[:n, n] = 1.0
            r = np.r_[O.vg_sum(d, vgs[t]), 1.0]
            lam = np.linalg.solve(K, r)
            out['OK'][t, i] = lam[:n] @ zs
            out['V'][t, i] = lam @ r + lam[n]
    return out


def test_interp_all_nrst_matches_per_cell_brute_force():
    """interp/steps.py:592-608 (never-selected stations dropped) and
    :732-831 (per cell group: flags, mean fill, system) against a per-cell
    restatement without groups."""
    sx, sy, cx, cy, z, vgs = _neb_case()
    methods = [('OK', 'OK', None), ('IDW', 'IDW', 2.0), ('NNB', 'NNB', None)]
    f, f64, prb = O.interp_all(sx, sy, cx, cy, z, vgs, methods,
                               min_var_thr=0.5, neb_sel_mthd='nrst',
                               n_nebs=6)
    exp = _brute_neb(sx, sy, cx, cy, z, vgs, 6, 0.5)
    assert not prb
    for lab, key in (('OK', 'OK'), ('IDW', 'IDW'), ('NNB', 'NNB'),
                     ('EST_VARS_OK', 'V')):
        np.testing.assert_allclose(f64[lab], exp[key], rtol=1e-10,
                                   atol=1e-12, err_msg=lab)
    # mean-filled neighbourhoods inside steps that are interpolated
    assert (exp['V'] == 0.0).any() and (exp['V'] > 0.0).any()
    # the drop matters: step 2 keeps five stations and would otherwise
    # also use the far one
    f2, f642, _ = O.interp_all(sx, sy[:], cx, cy, z[2:3], vgs[2:3],
                               methods, min_var_thr=0.5,
                               neb_sel_mthd='nrst', n_nebs=6)
    np.testing.assert_array_equal(f642['OK'], f64['OK'][2:3])
    zz = z[2:3].copy()
    zz[:, -1] = np.nan
    _, f64all, _ = O.interp_all(sx, sy, cx, cy, zz, vgs[2:3], methods,
                                min_var_thr=0.5)
    np.testing.assert_allclose(f642['OK'], f64all['OK'], rtol=1e-10)
    _, f64far, _ = O.interp_all(sx, sy, cx, cy, z[2:3], vgs[2:3], methods,
                                min_var_thr=0.5)
    assert np.abs(f64far['OK'] - f642['OK']).max() > 1e-3


def test_interp_all_nrst_with_every_station_equals_all():
    """n_nebs at least the station count: one neighbourhood of every active
    station per step, so the results are those of 'all' bit for bit."""
    sx, sy, cx, cy, z, vgs = _neb_case(3)
    sx, sy, z = sx[:30], sy[:30], z[:, :30]
    methods = [('OK', 'OK', None), ('IDW', 'IDW', 2.0)]
    _, a, _ = O.interp_all(sx, sy, cx, cy, z, vgs, methods, min_var_thr=0.5)
    _, b, _ = O.interp_all(sx, sy, cx, cy, z, vgs, methods, min_var_thr=0.5,
                           neb_sel_mthd='nrst', n_nebs=40)
    for lab in a:
        np.testing.assert_array_equal(a[lab], b[lab])


# ---- exact solution of the kriging system (ill-conditioned steps) --------

def _bordered_solve_mp(K, rhs, dps=40):
    """The reference's bordered system K x = rhs solved with 40 significant
    digits (mpmath inverse by LU): an algorithm independent of krige_exact's
    covariance-form shift and of float64 rounding."""
    import mpmath
    with mpmath.workdps(dps):
        A = mpmath.matrix(K.tolist())
        B = mpmath.matrix(rhs.tolist())
        X = mpmath.inverse(A) * B
        return np.array([[float(X[i, j]) for j in range(rhs.shape[1])]
                         for i in range(K.shape[0])])


def _ill_step(n=60, ncell=12, seed=5):
    """A Gaussian-model OK step with a long range and four stations 2-5 m
    from another one: condition number of C about 1e11."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1e4, n)
    y = rng.uniform(0, 1e4, n)
    x[:4] = x[4:8] + rng.uniform(2, 5, 4)
    y[:4] = y[4:8]
    cx = rng.uniform(0, 1e4, ncell)
    cy = rng.uniform(0, 1e4, ncell)
    z = rng.lognormal(0, 1, n)
    vg = '0.01 Nug(0.0) + 1.0 Gau(5000.0)'
    return x, y, cx, cy, z, vg


def test_krige_exact_matches_pinv_when_well_conditioned():
    rng = np.random.default_rng(2)
    n = 60
    x, y = rng.uniform(0, 1e4, n), rng.uniform(0, 1e4, n)
    cx, cy = rng.uniform(0, 1e4, 30), rng.uniform(0, 1e4, 30)
    z = rng.gamma(0.8, 5.0, n)
    rr, dr = O.dists_2d(x, y, x, y), O.dists_2d(cx, cy, x, y)
    for vg in ('0.2 Nug(0.0) + 1.5 Sph(3000.0)', '1.0 Exp(2500.0)',
               '0.5 Lin(4000.0) + 0.1 Nug(0.0)', '2.0 Pow(1.2)'):
        K, R = O.krig_system('OK', vg, rr, dr, None, None, 0.0)
        e_p, v_p, ok = O.krige_contract(np.linalg.pinv(K), R, z, dr, True)
        assert ok.all()
        e_x, v_x, lsum = O.krige_exact('OK', vg, rr, dr, z, want_var=True)
        np.testing.assert_allclose(e_x, e_p, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(v_x, v_p, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(lsum, 1.0, rtol=0, atol=1e-13)


def test_krige_exact_against_long_double_bordered_solve():
    """On an ill-conditioned step (kappa(C) about 1e11) the float64 pinv
    loses digits; krige_exact agrees with a 40-digit solve of the bordered
    system [[G, 1], [1^T, 0]] to 5e-9 where pinv does not (7e-6)."""
    x, y, cx, cy, z, vg = _ill_step()
    rr, dr = O.dists_2d(x, y, x, y), O.dists_2d(cx, cy, x, y)
    K, R = O.krig_system('OK', vg, rr, dr, None, None, 0.0)
    n = z.size
    ev = np.linalg.eigvalsh(1.01 - K[:n, :n])
    assert ev[-1] / ev[0] > 1e10
    L = _bordered_solve_mp(K, R.T).T                 # [cells x (n+1)]
    e_ref = L[:, :n] @ z
    v_ref = (L * R).sum(1) + L[:, n]
    e_x, v_x, lsum = O.krige_exact('OK', vg, rr, dr, z, want_var=True)
    fe = 1e-3 * np.abs(e_ref).mean()
    fv = 1e-3 * np.abs(v_ref).mean()
    err_e = np.abs(e_x - e_ref) / np.maximum(np.abs(e_ref), fe)
    err_v = np.abs(v_x - v_ref) / np.maximum(np.abs(v_ref), fv)
    assert err_e.max() < 5e-9 and err_v.max() < 5e-9
    e_p, v_p, _ = O.krige_contract(np.linalg.pinv(K), R, z, dr, True)
    err_p = np.abs(e_p - e_ref) / np.maximum(np.abs(e_ref), fe)
    assert err_p.max() > 10 * err_e.max()       # pinv is the less exact one


def test_krige_exact_edk_reproduces_drift_and_is_shift_invariant():
    """EDK: sum(lambda s) = s_cell exactly (drift reproduction) and the
    estimate does not depend on the covariance shift c0 (a nugget-only
    component adds a constant to every entry of G and changes c0)."""
    rng = np.random.default_rng(4)
    n = 50
    x, y = rng.uniform(0, 1e4, n), rng.uniform(0, 1e4, n)
    cx, cy = rng.uniform(0, 1e4, 25), rng.uniform(0, 1e4, 25)
    s = (0.3 * x + 0.1 * y)[:, None] / 1e3
    sc = (0.3 * cx + 0.1 * cy)[None, :] / 1e3
    z = 2.0 + 0.5 * s[:, 0] + rng.normal(0, 0.1, n)
    rr, dr = O.dists_2d(x, y, x, y), O.dists_2d(cx, cy, x, y)
    vg = '1.0 Sph(4000.0)'
    e1, _, l1 = O.krige_exact('EDK', vg, rr, dr, z, s, sc)
    e2, _, _ = O.krige_exact('EDK', '5.0 Nug(0.0) + ' + vg, rr, dr, z, s, sc)
    np.testing.assert_allclose(l1, 1.0, atol=1e-13)
    np.testing.assert_allclose(e1, e2, rtol=1e-11, atol=1e-12)
    # a field exactly linear in the drift is reproduced exactly
    e3, _, _ = O.krige_exact('EDK', vg, rr, dr, 2.0 + 0.5 * s[:, 0], s, sc)
    np.testing.assert_allclose(e3, 2.0 + 0.5 * sc[0], rtol=1e-11)
    K, R = O.krig_system('EDK', vg, rr, dr, s, sc, 0.0)
    e_p, _, ok = O.krige_contract(np.linalg.pinv(K), R, z, dr, False)
    assert ok.all()
    np.testing.assert_allclose(e1, e_p, rtol=1e-10, atol=1e-12)


# ---- class API (pyx:251-719) ------------------------------------------------

def test_ok_class_two_stations_closed_form():
    """OrdinaryKriging with two stations: Gamma = [[0, g12], [g12, 0]] (zero
    diagonal), so for a point with semivariances (g1, g2) the system solves
    by hand: l1 = (1 - (g1 - g2) / g12) / 2, l2 = 1 - l1,
    mu = g1 - g12 l2; est_vars = max(0, l1 g1 + l2 g2 + mu)."""
    xi, yi = np.array([0.0, 3000.0]), np.array([0.0, 0.0])
    zi = np.array([2.0, 5.0])
    xk, yk = np.array([1000.0, 2500.0, -400.0]), np.array([0.0, 700.0, 300.0])
    model = '0.3 Nug(0.0) + 1.2 Sph(5000.0)'
    out = O.ordinary_kriging(xi, yi, zi, xk, yk, model)

    def gam(d):
        return 0.3 + 1.2 * (1.5 * d / 5000 - 0.5 * (d / 5000) ** 3)
    g12 = gam(3000.0)
    for c in range(3):
        d1 = math.hypot(xk[c] - xi[0], yk[c] - yi[0])
        d2 = math.hypot(xk[c] - xi[1], yk[c] - yi[1])
        g1, g2 = gam(d1), gam(d2)
        l1 = (1 - (g1 - g2) / g12) / 2
        l2 = 1 - l1
        mu = g1 - g12 * l2
        assert math.isclose(out['zk'][c], l1 * 2.0 + l2 * 5.0, rel_tol=1e-12)
        assert math.isclose(out['mus'][c], mu, rel_tol=1e-11, abs_tol=1e-13)
        assert math.isclose(out['est_vars'][c],
                            max(0.0, l1 * g1 + l2 * g2 + mu), rel_tol=1e-11)


def test_ok_class_exact_at_stations_and_weights_sum_to_one():
    """Zero-diagonal Gamma and the rhs skipping d == 0 make lambda = e_j,
    mu = 0 the solution at station j even with a nugget: the value is
    reproduced and est_vars is 0 there; elsewhere sum(lambda) = 1."""
    sx, sy, cx, cy, rr, dr, z, vg = _system_case(
        vg='0.25 Nug(0.0) + 1 Exp(4000.0)')
    xk = np.concatenate([sx[:3], cx])
    yk = np.concatenate([sy[:3], cy])
    out = O.ordinary_kriging(sx, sy, z, xk, yk, '0.25 Nug(0.0) + 1 Exp(4000.0)')
    np.testing.assert_allclose(out['lambdas'][:3], np.eye(z.size)[:3],
                               atol=1e-10)
    np.testing.assert_allclose(out['zk'][:3], z[:3], rtol=1e-10)
    np.testing.assert_allclose(out['est_vars'][:3], 0.0, atol=1e-10)
    np.testing.assert_allclose(out['lambdas'].sum(1), 1.0, atol=1e-12)
    assert (out['est_vars'][3:] > 0.25).all()   # nugget + spatial part


def test_sk_class_init_order_quirk():
    """SimpleKriging fills in_covars from self.covar before the sills are
    summed (pyx:395-399, 413, 439), so K = -Gamma with a zero diagonal
    (Gamma from fill_vg_var_arr's semivariances); the weights are then the
    minimum-norm least-squares solution of K lambda = rhs (lstsq)."""
    sx, sy, cx, cy, rr, dr, z, vg = _system_case(vg='1.5 Sph(6000.0)')
    out = O.simple_kriging(sx, sy, z, cx, cy, '1.5 Sph(6000.0)')
    G = np.empty(rr.shape)
    O.fill_vg_var_arr(rr, G, 0, 1, '1.5 Sph(6000.0)', -np.inf)
    np.fill_diagonal(G, 0.0)
    np.testing.assert_array_equal(out['in_covars'], -G)
    lam = np.linalg.lstsq(-G, out['rhss'].T, rcond=None)[0].T
    np.testing.assert_allclose(out['lambdas'], lam, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out['zk'], lam @ z, rtol=1e-9)
    cov = np.empty(dr.shape)
    O.fill_vg_var_arr(dr, cov, 1, 0, '1.5 Sph(6000.0)', -np.inf)
    np.testing.assert_allclose(out['rhss'], cov, rtol=1e-15)


def test_edk_md_class_reproduces_drifts():
    """ExternalDriftKriging_MD: sum(lambda) = 1 and sum(lambda s_k) = s_k at
    every point for each drift, a field linear in the drifts is reproduced
    exactly, and lambda = e_j at station j."""
    sx, sy, cx, cy, rr, dr, z, vg = _system_case(n=14)
    si = np.stack([(0.3 * sx + 0.1 * sy) / 1e3, np.sin(sx / 2e3)])
    sk = np.stack([(0.3 * cx + 0.1 * cy) / 1e3, np.sin(cx / 2e3)])
    zl = 1.0 + 2.0 * si[0] - 0.5 * si[1]
    out = O.external_drift_kriging_md(sx, sy, zl, si, cx, cy, sk, vg)
    np.testing.assert_allclose(out['lambdas'].sum(1), 1.0, atol=1e-10)
    np.testing.assert_allclose(out['lambdas'] @ si.T, sk.T, rtol=1e-9,
                               atol=1e-10)
    np.testing.assert_allclose(out['zk'], 1.0 + 2.0 * sk[0] - 0.5 * sk[1],
                               rtol=1e-9)
    at = O.external_drift_kriging_md(sx, sy, z, si, sx[:2], sy[:2],
                                     si[:, :2], vg)
    np.testing.assert_allclose(at['lambdas'], np.eye(sx.size)[:2],
                               atol=1e-9)
    assert at['mus_arr'].shape == (3, 2)


# ---- step-driver helpers -------------------------------------------------------

def test_neb_groups_partition_cells_by_identical_rows():
    rng = np.random.default_rng(0)
    idx = rng.integers(0, 3, (40, 2))          # few distinct rows
    groups = O.neb_groups(idx)
    allc = np.concatenate(groups)
    assert sorted(allc.tolist()) == list(range(40))
    uniq, inv = np.unique(idx, axis=0, return_inverse=True)
    assert len(groups) == len(uniq)
    for g in groups:
        assert (idx[g] == idx[g[0]]).all()
        assert len(set(inv[g].tolist())) == 1
    firsts = [g[0] for g in groups]
    assert firsts == sorted(firsts)             # order of first appearance
    assert all((np.diff(g) > 0).all() for g in groups)


def test_mod_min_max_clips_and_keeps_nan():
    f = np.array([[-3.0, 0.5, np.nan], [7.0, 2.0, -0.0]])
    g = f.copy()
    O.mod_min_max(g, 0.0, 5.0)
    np.testing.assert_array_equal(g, np.where(np.isnan(f), np.nan,
                                              np.clip(f, 0.0, 5.0)))
    g = f.copy()
    O.mod_min_max(g, None, None)
    np.testing.assert_array_equal(g, f)
    g = f.copy()
    O.mod_min_max(g, None, 1.0)
    np.testing.assert_array_equal(g, np.where(np.isnan(f), np.nan,
                                              np.minimum(f, 1.0)))


def test_krige_solve_matches_pinv_when_well_conditioned():
    sx, sy, cx, cy, rr, dr, z, vg = _system_case()
    K, R = O.krig_system('OK', vg, rr, dr, None, None, 0.0)
    e1, v1, ok1 = O.krige_contract(np.linalg.pinv(K), R, z, dr, True)
    e2, v2, ok2 = O.krige_solve(K, R, z, dr, True)
    assert ok1.all() and ok2.all()
    np.testing.assert_allclose(e2, e1, rtol=1e-10)
    np.testing.assert_allclose(v2, v1, rtol=1e-10)


# ---- vgclus weight heuristic (vgclus/tvgs_clus.py:92-137) ----------------

def test_threshold_wts_by_hand():
    w = np.array([-0.05, 0.05, 0.1, 0.3, 0.6])   # sum |w| = 1.1
    # shares 0.045, 0.045, 0.091, 0.27, 0.55: a cut at 0.05 drops the two
    # smallest and rescales the rest by 1 / 1.0
    got = O.threshold_wts(w, 0.05)
    np.testing.assert_allclose(got, [0.0, 0.0, 0.1, 0.3, 0.6])
    # a cut at 0.1 also drops 0.1 and rescales 0.3, 0.6 by 1 / 0.9
    got = O.threshold_wts(w, 0.1)
    np.testing.assert_allclose(got, [0.0, 0.0, 0.0, 1 / 3, 2 / 3])
    # a cut above every share keeps them all (the reference's guard)
    got = O.threshold_wts(np.array([0.2, 0.2, 0.2, 0.2, 0.2]), 0.5)
    np.testing.assert_allclose(got, [0.2] * 5)
    # no cut
    np.testing.assert_array_equal(O.threshold_wts(w, None), w)


def test_srtd_krg_wts_symmetric_and_sum_one():
    """Four stations on a circle around the point: equal weights 1/4; a
    random set: weights sorted and summing to one, a station at the point
    takes all the weight (the class API's zero-diagonal Gamma)."""
    ang = np.arange(4) * np.pi / 2
    pts = np.r_[[[0.0, 0.0]], np.c_[np.cos(ang), np.sin(ang)] * 3.0]
    w = O.srtd_krg_wts('1.0 Sph(10.0)', pts)
    np.testing.assert_allclose(w, 0.25, rtol=1e-12)
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 10, (12, 2))
    w = O.srtd_krg_wts('0.1 Nug(0.0) + 1.0 Exp(5.0)', pts)
    assert np.all(np.diff(w) >= 0) and abs(w.sum() - 1) < 1e-12
    pts[5] = pts[0]
    w = O.srtd_krg_wts('1.0 Exp(5.0)', pts)
    np.testing.assert_allclose(w[-1], 1.0, atol=1e-10)
    np.testing.assert_allclose(w[:-1], 0.0, atol=1e-10)
