"""GPU parity: libspk against the CPU oracle on seeded inputs.

Tolerances (stated per test): distances and indices bit-exact; FP64
variogram values 1e-13 relative; FP64 kriging estimates and variances
1e-9 relative (the plain FP64 path and the batched FP64 path); the mixed
path's estimates 1e-9 and its variances within the bound in the test.
"""
import math

import numpy as np
import pytest

from oracle import spk_oracle as O
from spinterps_amd import configs

pytestmark = pytest.mark.gpu

VGS_ALL = ['0.1 Nug(0.0) + 0.9 Sph(25000.0)', '2 Exp(3000)', '1.5 Gau(8000.0)',
           '1 Lin(5000) + 0.2 Nug(0)', '0.5 Pow(1.2)', '1 Hol(4000)',
           '1 Rng(0)', '3 Sph(0)']


def _rel(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


def test_cyth_distances_bit_exact(gpu):
    from spinterps_amd import cyth
    rng = np.random.default_rng(11)
    x1 = configs.ORIGIN_X + rng.uniform(0, 5e5, 300)
    y1 = configs.ORIGIN_Y + rng.uniform(0, 5e5, 300)
    x2 = configs.ORIGIN_X + rng.uniform(0, 5e5, 77)
    y2 = configs.ORIGIN_Y + rng.uniform(0, 5e5, 77)
    out = np.full((300, 77), np.nan)
    cyth.fill_dists_2d_mat(x1, y1, x2, y2, out)
    np.testing.assert_array_equal(out, O.dists_2d(x1, y1, x2, y2))
    d1 = np.full(77, np.nan)
    cyth.fill_dists_one_pt(x1[3], y1[3], x2, y2, d1)
    np.testing.assert_array_equal(d1, out[3])
    pts = rng.normal(size=(40, 3))
    np.testing.assert_allclose(cyth.get_nd_dists(pts), O.get_nd_dists(pts),
                               rtol=1e-15)


@pytest.mark.parametrize('vg', VGS_ALL)
@pytest.mark.parametrize('covar,diag,mvv', [(0, 1, 0.0), (1, 0, 0.0),
                                            (0, 0, 0.3), (1, 1, -np.inf)])
def test_cyth_vg_var_arr(gpu, vg, covar, diag, mvv):
    from spinterps_amd import cyth
    rng = np.random.default_rng(12)
    x, y = rng.uniform(0, 2e4, 50), rng.uniform(0, 2e4, 50)
    d = O.dists_2d(x, y, x, y)
    got = np.full_like(d, np.nan)
    exp = np.full_like(d, np.nan)
    cyth.fill_vg_var_arr(d, got, covar, diag, vg, mvv)
    O.fill_vg_var_arr(d, exp, covar, diag, vg, mvv)
    np.testing.assert_allclose(got, exp, rtol=1e-13, atol=1e-300)


def test_cyth_small_helpers(gpu):
    from spinterps_amd import cyth
    rng = np.random.default_rng(13)
    arr = rng.normal(size=(20, 30))
    ri = np.array([3, 0, 19, 7], dtype=np.int64)
    ci = np.array([29, 1, 1, 5, 0], dtype=np.int64)
    sub = np.full((6, 7), np.nan)
    cyth.copy_2d_arr_at_idxs(arr, ri, ci, sub)
    np.testing.assert_array_equal(sub[:4, :5], arr[np.ix_(ri, ci)])
    assert np.isnan(sub[4:]).all() and np.isnan(sub[:, 5:]).all()
    with pytest.raises(IndexError):
        cyth.copy_2d_arr_at_idxs(arr, np.array([20]), ci, sub)
    d = rng.uniform(0.1, 1.0, 100)
    w = np.empty(100)
    we = np.empty(100)
    s = cyth.fill_wts_and_sum(d, w, 2.0)
    se = O.fill_wts_and_sum(d, we, 2.0)
    np.testing.assert_allclose(w, we, rtol=1e-15)
    assert math.isclose(s, se, rel_tol=1e-13)
    z = rng.normal(size=100)
    assert math.isclose(cyth.get_mults_sum(w, z), O.get_mults_sum(w, z),
                        rel_tol=1e-12)
    h = np.linspace(0, 1e6, 10)
    vals = np.zeros(10)
    for nm, r, sl in [('Sph', 1e4, 100.0), ('Exp', 1e6, 10.0)]:
        cyth.fill_theo_vg_vals(nm, h, r, sl, vals)
    np.testing.assert_allclose(
        vals, O.get_theo_vg_vals('100 Sph(10000) + 10 Exp(1000000)', h),
        rtol=1e-14)
    with pytest.raises(AssertionError):
        cyth.fill_theo_vg_vals('Sph', h, -1.0, 1.0, vals)


def test_neighbours_bit_exact(gpu):
    from spinterps_amd import _lib
    from spinterps_amd._lib import check, context, dptr, i64ptr, lib
    rng = np.random.default_rng(14)
    sx, sy = rng.uniform(0, 1e5, 150), rng.uniform(0, 1e5, 150)
    cx, cy = rng.uniform(0, 1e5, 400), rng.uniform(0, 1e5, 400)
    for k in (1, 7, 64, 150, 200):
        kk = min(k, 150)
        out = np.full((400, kk), -1, dtype=np.int64)
        check(lib().spk_neb_idxs_nrst(context().h, dptr(cx), dptr(cy), 400,
                                      dptr(sx), dptr(sy), 150, k, i64ptr(out)))
        np.testing.assert_array_equal(out, O.neb_idxs_nrst(cx, cy, sx, sy, k))
    for k, p in ((8, 4), (20, 6), (9, 9)):
        out = np.full((400, k), -1, dtype=np.int64)
        check(lib().spk_neb_idxs_pie(context().h, dptr(cx), dptr(cy), 400,
                                     dptr(sx), dptr(sy), 150, k, p,
                                     i64ptr(out)))
        np.testing.assert_array_equal(out, O.neb_idxs_pie(cx, cy, sx, sy, k, p))


def test_sel_equidist_single_point(gpu):
    from spinterps_amd import cyth
    sx = np.array([1.0, 0.0, -3.0, -1.0, -3.0, 2.0, 5.0])
    sy = np.array([0.0, 2.0, 0.0, 2.0, -3.0, -2.0, 0.0])
    res = []
    for impl in (cyth.sel_equidist_refs, O.sel_equidist_refs):
        a = [np.zeros(7), np.zeros(7), np.zeros(7, np.int64),
             np.zeros(7, np.uint64), np.zeros(4, np.uint64)]
        impl(0.0, 0.0, sx, sy, 4, -1.0, -1, *a)
        res.append(a)
    for g, e in zip(*res):
        np.testing.assert_array_equal(g, e)


# ---- step interpolation ---------------------------------------------------------

def _small_case(n_side=20, steps=6, mask=False, seed=20240601):
    w = configs.small(seed=seed, n_cells_side=n_side, cell=1e5 / n_side)
    vals, vgs = configs.small_steps(w, 0, steps)
    if mask:
        rng = np.random.default_rng(99)
        for t in range(steps):
            vals[t, rng.choice(w.n_stns, 10 + 3 * t, replace=False)] = np.nan
    return w, vals, vgs


def _oracle_fields(w, vals, vgs, methods, **kw):
    return O.interp_all(w.stn_x, w.stn_y, w.cell_x, w.cell_y, vals, vgs,
                        methods, stn_drfts=w.stn_drift, cell_drfts=w.cell_drift,
                        **kw)


@pytest.mark.parametrize('method,p', [('IDW', 2.0), ('IDW', 1.5),
                                      ('NNB', 2.0)])
def test_idw_nnb_steps(gpu, method, p):
    """IDW (p = 2: one pass with 1/d^2 weights; other p: the reference's
    two passes) and NNB against the oracle."""
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w, vals, vgs = _small_case(mask=True)
    prob = DeviceProblem.from_workload(w)
    est, _, info = interp_steps(prob, vals, None, method, idw_exp=p)
    _, f64, _ = _oracle_fields(w, vals, None, [(method, method, p)])
    got = est.cpu().numpy()
    if method == 'NNB':
        np.testing.assert_array_equal(got, f64[method])
    else:
        np.testing.assert_allclose(got, f64[method], rtol=1e-12)


@pytest.mark.parametrize('precision', ['fp64', 'fp64b', 'mixed'])
def test_ok_small_all_precisions(gpu, precision):
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w, vals, vgs = _small_case(mask=True)
    prob = DeviceProblem.from_workload(w)
    est, var, info = interp_steps(prob, vals, vgs, 'OK', precision=precision)
    _, f64, _ = _oracle_fields(w, vals, vgs, [('OK', 'OK', None)])
    e, v = est.cpu().numpy(), var.cpu().numpy()
    assert ((info.cpu().numpy() & 15) == 0).all()
    scale = np.nanmax(np.abs(f64['OK']))
    np.testing.assert_allclose(e, f64['OK'], rtol=1e-9, atol=1e-12 * scale)
    vref = f64['EST_VARS_OK']
    vfloor = 1e-3 * np.nanmax(vref)
    err = _rel(v, vref, vfloor)
    print(f'{precision}: est max rel {_rel(e, f64["OK"], 1e-12 * scale):.2e} '
          f'var max rel (floor {vfloor:.3g}) {err:.2e}')
    assert err <= (1e-9 if precision != 'mixed' else 1e-6)


@pytest.mark.parametrize('method', ['SK', 'SKM', 'EDK'])
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_sk_edk_daily_subset(gpu, method, precision):
    """SK (the reference's), SKM (simple kriging with the step mean as
    known mean) and EDK against the oracle, 1e-9 relative."""
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.daily(n_side=40, cell=12500.0, n_stns=300)
    vals, vgs = configs.daily_steps(w, 0, 4)
    prob = DeviceProblem.from_workload(w)
    est, _, info = interp_steps(prob, vals, vgs, method, precision=precision,
                                want_var=False)
    _, f64, _ = _oracle_fields(w, vals, vgs, [(method, method, None)])
    e = est.cpu().numpy()
    scale = np.nanmax(np.abs(f64[method]))
    np.testing.assert_allclose(e, f64[method], rtol=1e-9, atol=1e-12 * scale)


def test_ok_daily_subset_plain_vs_oracle(gpu):
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.daily(n_side=30, cell=16000.0, n_stns=400)
    vals, vgs = configs.daily_steps(w, 10, 13)
    prob = DeviceProblem.from_workload(w)
    _, f64, _ = _oracle_fields(w, vals, vgs, [('OK', 'OK', None)])
    for prec in ('fp64', 'fp64b'):
        est, var, info = interp_steps(prob, vals, vgs, 'OK', precision=prec)
        e, v = est.cpu().numpy(), var.cpu().numpy()
        np.testing.assert_allclose(e, f64['OK'], rtol=1e-9,
                                   atol=1e-12 * np.nanmax(np.abs(f64['OK'])))
        np.testing.assert_allclose(v, f64['EST_VARS_OK'], rtol=1e-9,
                                   atol=1e-12 * np.nanmax(f64['EST_VARS_OK']))


def test_special_steps(gpu):
    """No station (NaN out), one station (its value, NaN variance), all
    values below min_var_thr and nugget-only variograms (mean, variance
    0): interp/steps.py:280-283, 325-331, 677-688."""
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w, vals, vgs = _small_case(steps=5)
    vals = vals.copy()
    vals[0, :] = np.nan
    vals[1, 1:] = np.nan
    vals[2] = np.minimum(vals[2], 0.5)
    vgs = list(vgs)
    vgs[3] = '1.0 Nug(0.0)'
    prob = DeviceProblem.from_workload(w)
    for prec in ('fp64', 'mixed'):
        est, var, info = interp_steps(prob, vals, vgs, 'OK', precision=prec,
                                      min_var_thr=1.0)
        _, f64, prb = _oracle_fields(w, vals, vgs, [('OK', 'OK', None)],
                                     min_var_thr=1.0)
        e, v = est.cpu().numpy(), var.cpu().numpy()
        np.testing.assert_array_equal(info.cpu().numpy()[:4], [1, 8, 2, 2])
        np.testing.assert_allclose(e, f64['OK'], rtol=1e-9, equal_nan=True)
        np.testing.assert_allclose(v[1:], f64['EST_VARS_OK'][1:],
                                   rtol=1e-9 if prec == 'fp64' else 1e-6,
                                   atol=1e-9, equal_nan=True)
        assert np.isnan(e[0]).all() and np.isnan(v[0]).all()
        assert prb == [0]


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_sk_kriging_branch(gpu, precision):
    """SK keeps the reference's sum-of-weights test (interp/steps.py:418):
    cells 1-5 cm from a station pass it (|sum lambda - 1| < 1e-5) while
    their estimate still differs from the station value by ~1e-7 relative,
    so both the kriging branch and the nearest-station fallback are
    exercised."""
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.small(n_cells_side=6, cell=1e5 / 6)
    rng = np.random.default_rng(21)
    pick = rng.choice(w.n_stns, 40, replace=False)
    ang = rng.uniform(0, 2 * np.pi, 40)
    r = rng.uniform(0.01, 0.05, 40)
    cx = np.concatenate([w.cell_x, w.stn_x[pick] + r * np.cos(ang)])
    cy = np.concatenate([w.cell_y, w.stn_y[pick] + r * np.sin(ang)])
    vals, _ = configs.small_steps(w, 0, 3)
    vgs = ['0.2 Nug(0.0) + 1.0 Exp(30000.0)', '1.0 Gau(20000.0)',
           '0.5 Sph(40000.0)']
    prob = DeviceProblem(w.stn_x, w.stn_y, cx, cy)
    est, _, info = interp_steps(prob, vals, vgs, 'SK', precision=precision,
                                want_var=False)
    f, f64, _ = O.interp_all(w.stn_x, w.stn_y, cx, cy, vals, vgs,
                             [('SK', 'SK', None)])
    e = est.cpu().numpy()
    # the oracle takes the kriging branch for the near-station cells
    near = slice(w.n_cells, None)
    nnb = vals[:, pick]
    assert (np.abs(f64['SK'][:, near] - nnb) > 1e-12 * np.abs(nnb)).any()
    scale = np.nanmax(np.abs(f64["SK"]))
    np.testing.assert_allclose(e, f64["SK"], rtol=1e-9, atol=1e-12 * scale)


def test_multi_batch_overlap_hourly_small(gpu):
    """More steps than one batch (64): systems of batch b+1 are factorised
    on the second stream while batch b is contracted.  Every step of the
    batched FP64 path must match the plain FP64 path (1e-9), the mixed
    path must stay within its bound, and sampled steps match the oracle."""
    import torch
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.hourly(n_side=24, cell=5e5 / 24, n_stns=160)
    vals, vgs = configs.hourly_steps(w, 0, 150, min_active=20)
    prob = DeviceProblem.from_workload(w)
    out = {}
    for prec in ('fp64', 'fp64b', 'mixed'):
        kap = torch.zeros(150, dtype=torch.float64, device='cuda')
        e, v, info = interp_steps(prob, vals, vgs, 'OK', precision=prec,
                                  kappa=kap)
        out[prec] = (e.cpu().numpy(), v.cpu().numpy(), info.cpu().numpy(),
                     kap.cpu().numpy())
    kap = out['fp64b'][3]
    good = (kap < 1e8) & ((out['fp64'][2] & (4 | 128)) == 0)
    assert good.sum() > 100
    e64, v64 = out['fp64'][0][good], out['fp64'][1][good]
    # two FP64 formulations differ by ~1e-16 x condition: 1e-9 for
    # kappa < 1e5, 1e-7 up to 1e8; the mixed path stays within 1e-6
    # (variances relative with a floor of 1e-3 x the step's mean variance)
    well = kap[good] < 1e5
    for prec in ('fp64b', 'mixed'):
        e, v = out[prec][0][good], out[prec][1][good]
        floor = 1e-3 * np.abs(e64).mean(axis=1, keepdims=True)
        assert np.max(np.abs(e - e64) / np.maximum(np.abs(e64), floor)) < 1e-8
        vfloor = 1e-3 * np.abs(v64).mean(axis=1, keepdims=True)
        rv = np.abs(v - v64) / np.maximum(np.abs(v64), vfloor)
        rw, ri = rv[well].max(), rv[~well].max(initial=0.0)
        print(f'{prec}: var max rel vs fp64 {rw:.2e} (kappa < 1e5), '
              f'{ri:.2e} (1e5 <= kappa < 1e8)')
        if prec == 'fp64b':
            assert rw < 1e-9 and ri < 1e-7
        else:
            worst = np.where(good)[0][np.argmax(rv.max(axis=1))]
            assert max(rw, ri) < 1e-6, (worst, kap[worst], vgs[worst],
                                        out[prec][2][worst])
    for t in np.where(good)[0][[0, 50, 100]]:
        _, ref, _ = O.interp_all(w.stn_x, w.stn_y, w.cell_x, w.cell_y,
                                 vals[t:t + 1], vgs[t:t + 1],
                                 [('OK', 'OK', None)])
        np.testing.assert_allclose(out['fp64b'][0][t], ref['OK'][0],
                                   rtol=1e-9, atol=1e-12 * np.abs(ref['OK']).max())
        np.testing.assert_allclose(out['fp64b'][1][t], ref['EST_VARS_OK'][0],
                                   rtol=1e-9)


_SOLVED_MASK = 1 | 2 | 4 | 8 | 16   # no stations, mean fill, failed, one station, ill-conditioned


@pytest.mark.parametrize('mode', ['int8', 'fp32', 'int8_limit'])
def test_mixed_precision_policy_500m_grid(gpu, mode):
    """The mixed path's precision policy on the hourly workload's geometry
    (2 000 stations, 500 m cells).  Default ('int8'): the variance
    contraction runs in FP32 below the FP32 limit (condition estimate 1e5,
    1e3 for Gaussian models) and by exact int8 plane products above it.
    'fp32' (int8 switched off): FP64 above the limit.  'int8_limit'
    (caller's int8 limit 1e10): FP64 above that.  Against the batched FP64
    path (the same FP64 estimate, the variance contraction in FP64) every
    step the mixed path does not flag ill-conditioned stays within 1e-6
    (relative, floor 1e-3 x the step's mean)."""
    import torch
    from spinterps_amd import _lib
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.hourly(n_side=64, cell=500.0)
    T = 96
    vals, vgs = configs.hourly_steps(w, 0, T)
    prob = DeviceProblem.from_workload(w)
    kap = torch.zeros(T, dtype=torch.float64, device='cuda')
    k8 = {'int8': 0.0, 'fp32': -1.0, 'int8_limit': 1e10}[mode]
    e, v, info = interp_steps(prob, vals, vgs, 'OK', precision='mixed',
                              kappa=kap, int8var_kappa=k8)
    e64, v64, i64 = interp_steps(prob, vals, vgs, 'OK', precision='fp64b')
    e, v, inf = e.cpu().numpy(), v.cpu().numpy(), info.cpu().numpy()
    e64, v64, i64 = e64.cpu().numpy(), v64.cpu().numpy(), i64.cpu().numpy()
    k = kap.cpu().numpy()
    gau = np.array(['Gau' in g for g in vgs])
    fp64 = (inf & _lib.ST_FP64VAR) != 0
    int8 = ((inf & _lib.ST_INT8VAR) != 0) & ~fp64
    # steps solved by the batched kernels (not the pseudo-inverse fallback)
    runs = (inf & (1 | 2 | 4 | 8 | 128)) == 0
    solved = runs & ((inf & _lib.ST_ILLCOND) == 0)
    assert runs.sum() > 80 and solved.sum() > 70
    above = runs & np.where(gau, k > 1e3, k > 1e5)
    if mode == 'int8':
        assert int8[above].all() and not int8[runs & ~above].any()
        assert not fp64[runs].any()
        # the FP32 limit above 1e3 is exercised
        assert (runs & ~above & ~gau & (k > 1e3)).sum() >= 5
    elif mode == 'fp32':
        assert not int8[runs].any()
        assert fp64[above].all() and not fp64[runs & ~above].any()
        # the FP32 limit above 1e3 is exercised
        assert (runs & ~fp64 & ~gau & (k > 1e3)).sum() >= 5
    else:
        assert fp64[runs & (k > 1e10)].all() and (runs & (k > 1e10)).any()
        assert int8[above & (k <= 1e10)].all()
    cmp = solved
    floor_e = 1e-3 * np.abs(e64).mean(axis=1, keepdims=True)
    floor_v = 1e-3 * np.abs(v64).mean(axis=1, keepdims=True)
    with np.errstate(invalid='ignore'):   # steps without variance are NaN
        re = (np.abs(e - e64) / np.maximum(np.abs(e64), floor_e)).max(axis=1)
        rv = (np.abs(v - v64) / np.maximum(np.abs(v64), floor_v)).max(axis=1)
    print(f'{mode}: int8 {int8[runs].sum()}, fp64 {fp64[runs].sum()} of '
          f'{runs.sum()} steps; est max {re[cmp].max():.2e}, var max below '
          f'the FP32 limit {rv[cmp & ~above].max():.2e}, above '
          f'{rv[cmp & above].max(initial=0):.2e}')
    worst = np.where(cmp)[0][np.argmax(rv[cmp])]
    assert re[cmp].max() < 1e-6
    assert rv[cmp].max() < 1e-6, (worst, k[worst], vgs[worst])


def test_split_bf16_contraction_matches_fp32_operands(gpu):
    """The FP32-tier steps' variance contraction on split-BF16 operands
    ('mixed': k_split_tri, k_kgen's planes, k_tri_bf) against the same
    contraction on FP32 operands ('mixed_f32') and the batched FP64 path,
    on the hourly workload's geometry (2 000 stations, 500 m cells, 96
    steps, every model): estimates are identical (the contraction only
    feeds the variance), the steps' flags are identical, and on every
    FP32-tier step neither is flagged on, the variance stays within 1e-6 of
    FP64 (relative, floor 1e-3 x the step's mean) and within 4x the FP32
    operands' error (plus 1e-10).  The work counters show which kernel
    ran."""
    import torch
    from spinterps_amd import _lib
    from spinterps_amd.interp.steps import (DeviceProblem, interp_steps,
                                            prof_enable, prof_get, prof_work)
    w = configs.hourly(n_side=64, cell=500.0)
    T = 96
    vals, vgs = configs.hourly_steps(w, 0, T)
    prob = DeviceProblem.from_workload(w)
    out = {}
    for prec in ('mixed', 'mixed_f32', 'fp64b'):
        prof_enable(0, True)
        e, v, info = interp_steps(prob, vals, vgs, 'OK', precision=prec)
        torch.cuda.synchronize()
        out[prec] = (e.cpu().numpy(), v.cpu().numpy(), info.cpu().numpy(),
                     prof_get('gemm_tri_bf16')[1], prof_get('gemm_tri_f32')[1],
                     prof_work('gemm_tri_bf16_macs'))
        prof_enable(0, False)
    eb, vb, ib, nb, nf, wb = out['mixed']
    e3, v3, i3, nb3, nf3, _ = out['mixed_f32']
    e64, v64, i64 = out['fp64b'][:3]
    assert nb > 0 and nf == 0 and wb > 0, (nb, nf, wb)
    assert nb3 == 0 and nf3 > 0, (nb3, nf3)
    np.testing.assert_array_equal(ib, i3)
    np.testing.assert_array_equal(eb, e3)
    tier = (ib & (1 | 2 | 4 | 8 | 16 | 128 | _lib.ST_FP64VAR
                  | _lib.ST_INT8VAR)) == 0
    assert tier.sum() > 40, tier.sum()
    floor = 1e-3 * np.abs(v64).mean(axis=1, keepdims=True)
    with np.errstate(invalid='ignore'):
        rb = (np.abs(vb - v64) / np.maximum(np.abs(v64), floor)).max(axis=1)
        r3 = (np.abs(v3 - v64) / np.maximum(np.abs(v64), floor)).max(axis=1)
    print(f'split-BF16 var max {rb[tier].max():.2e}, FP32 operands '
          f'{r3[tier].max():.2e} over {tier.sum()} steps')
    assert rb[tier].max() < 1e-6 and r3[tier].max() < 1e-6
    bad = np.where(tier & (rb > 4 * r3 + 1e-10))[0]
    assert bad.size == 0, [(int(t), rb[t], r3[t], vgs[t]) for t in bad]


def test_mixed_slot_overflow_promotes(gpu):
    """On a coarse grid a 256-cell block has far more than 16 distinct
    nearest stations, so not every cell gets a staged local fit: the
    step's variance then runs in FP64 (SPK_ST_FP64VAR) at any condition
    estimate and matches the plain FP64 path to 1e-9."""
    import torch
    from spinterps_amd import _lib
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.hourly(n_side=48, cell=5e5 / 48)   # 2 304 cells of 10.4 km
    T = 6
    vals, _ = configs.hourly_steps(w, 0, T, min_active=1500)
    vgs = ['0.2 Nug(0.0) + 1.0 Exp(20000.0)'] * T
    prob = DeviceProblem.from_workload(w)
    kap = torch.zeros(T, dtype=torch.float64, device='cuda')
    _, v, info = interp_steps(prob, vals, vgs, 'OK', precision='mixed',
                              kappa=kap)
    _, v64, _ = interp_steps(prob, vals, vgs, 'OK', precision='fp64')
    k, inf = kap.cpu().numpy(), info.cpu().numpy()
    assert (k < 1e5).all(), k        # FP32 by condition estimate alone
    assert ((inf & _lib.ST_FP64VAR) != 0).all(), inf
    v, v64 = v.cpu().numpy(), v64.cpu().numpy()
    floor = 1e-3 * np.abs(v64).mean(axis=1, keepdims=True)
    assert np.max(np.abs(v - v64) / np.maximum(np.abs(v64), floor)) < 1e-9


def test_mixed_results_independent_of_call_length(gpu):
    """The same 16 steps alone and inside a 64-step call give bit-identical
    estimates and variances.  The cell stripes depend on the batch length
    (the workspace budget is shared by the batch's steps): on a 200 x 200
    grid the 64-step call splits the cells into stripes, the 16-step call
    does not; stripes are whole 256-cell blocks, so every cell keeps its
    block of nearest stations and its local fit."""
    import torch
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    w = configs.hourly(n_side=200, cell=500.0)
    vals, vgs = configs.hourly_steps(w, 0, 64)
    prob = DeviceProblem.from_workload(w)
    e64, v64, _ = interp_steps(prob, vals, vgs, 'OK', precision='mixed')
    e16, v16, _ = interp_steps(prob, vals[:16], vgs[:16], 'OK',
                               precision='mixed')
    torch.cuda.synchronize()
    np.testing.assert_array_equal(e16.cpu().numpy(), e64[:16].cpu().numpy())
    np.testing.assert_array_equal(v16.cpu().numpy(), v64[:16].cpu().numpy())


@pytest.mark.parametrize('model', ['Sph', 'Exp', 'Gau'])
def test_mixed_culled_station_chunks_bit_exact(gpu, model):
    """kgen culls the 16-station sub-chunks that are beyond the model's
    exact-zero distance from a block's cells: their covariances and delta are
    exactly 0, so the culled path gives bit-identical estimates and
    variances.  The reference run appends a zero-sill component of the same
    family ('+ 0.0 M(1.0)'), which leaves every value unchanged but takes
    the general evaluation, which culls nothing.  Between the two runs a
    long-range exponential call fills the delta workspace with nonzero
    values, so a consumer that read a chunk the culled run left unwritten
    would differ."""
    import torch
    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    # 96 x 96 cells at 500 m in one corner of the 500 km station field
    w = configs.hourly(n_side=96, cell=500.0)
    T = 12
    vals, _ = configs.hourly_steps(w, 0, T)
    rng = np.random.default_rng(31)
    vgs, vgs0, vgs_far = [], [], []
    for t in range(T):
        rg = float(rng.uniform(5e3, 8e4))
        f = float(rng.uniform(0.0, 0.3))
        v = float(np.nanvar(vals[t]))
        base = f'{f * v!r} Nug(0.0) + {(1 - f) * v!r} {model}({rg!r})'
        vgs.append(base)
        vgs0.append(base + f' + 0.0 {model}(1.0)')
        vgs_far.append(f'{v!r} Exp(1000000.0)')
    prob = DeviceProblem.from_workload(w)
    e0, v0, i0 = interp_steps(prob, vals, vgs0, 'OK', precision='mixed')
    interp_steps(prob, vals, vgs_far, 'OK', precision='mixed')
    e1, v1, i1 = interp_steps(prob, vals, vgs, 'OK', precision='mixed')
    torch.cuda.synchronize()
    np.testing.assert_array_equal(i1.cpu().numpy(), i0.cpu().numpy())
    np.testing.assert_array_equal(e1.cpu().numpy(), e0.cpu().numpy())
    np.testing.assert_array_equal(v1.cpu().numpy(), v0.cpu().numpy())


def test_int8_plane_contraction_matches_fp64(gpu):
    """Exact int8 plane products (every step forced above the FP32 limit)
    against the batched FP64 path, whose contraction is the same
    |L^-1 delta|^2 in FP64 (relative, floor 1e-3 x the step's mean
    variance): 3e-8 up to condition estimate 1e5, 1e-7 up to 1e6, 1e-5 up
    to 1e8 and 1e-4 up to 1e12, where cancellation between the FP64 terms
    amplifies the planes' truncation (2^-35 of the row and column
    maxima)."""
    import torch
    from spinterps_amd import _lib
    from spinterps_amd.interp.steps import (DeviceProblem, interp_steps,
                                            prof_enable, prof_work)
    w = configs.hourly(n_side=48, cell=500.0)
    T = 40
    vals, vgs = configs.hourly_steps(w, 0, T)
    prob = DeviceProblem.from_workload(w)
    kap = torch.zeros(T, dtype=torch.float64, device='cuda')
    prof_enable(0, True)
    _, v8, inf8 = interp_steps(prob, vals, vgs, 'OK', precision='mixed',
                               kappa=kap, fp64var_kappa=1e-300)
    ops = prof_work('gemm_tri_i8_ops')
    prof_enable(0, False)
    _, v64, i64 = interp_steps(prob, vals, vgs, 'OK', precision='fp64b')
    k = kap.cpu().numpy()
    inf8, i64 = inf8.cpu().numpy(), i64.cpu().numpy()
    # steps whose systems were solved (flagged ones too: the contraction's
    # precision is what is compared, not the conditioning)
    solved = (i64 & (1 | 2 | 4 | 8)) == 0
    int8 = ((inf8 & _lib.ST_INT8VAR) != 0) & ((inf8 & _lib.ST_FP64VAR) == 0)
    sel = solved & int8
    assert sel.sum() >= 20, (k, inf8)
    # the cells cover 24 km x 24 km of the 500 km station field, so most
    # 64-station chunks of delta are exactly zero and the contraction skips
    # them: its executed plane products stay well below those of every
    # chunk of the triangle (128-row x 64-cell tiles, 64-k chunks)
    C = len(w.cell_x)
    full = 0.0
    for t in np.nonzero(sel)[0]:
        n = int(np.sum(~np.isnan(vals[t])))
        prods = 15 if k[t] <= 1e7 else 21
        nch = sum(-(-min(n, r0 + 128) // 64) for r0 in range(0, n, 128))
        full += nch * -(-C // 64) * prods * 2.0 * 64 * 128 * 64
    print(f'int8 ops executed {ops:.3e} of {full:.3e} without skipping')
    assert 0 < ops < 0.8 * full
    v8, v64 = v8.cpu().numpy(), v64.cpu().numpy()
    floor = 1e-3 * np.abs(v64).mean(axis=1, keepdims=True)
    with np.errstate(invalid='ignore'):
        rv = (np.abs(v8 - v64) / np.maximum(np.abs(v64), floor)).max(axis=1)
    print(f'int8 steps {sel.sum()}: max rel {rv[sel].max():.2e} (kappa up to '
          f'{k[sel].max():.1e})')
    assert rv[sel & (k <= 1e5)].max() < 3e-8
    assert rv[sel & (k <= 1e6)].max(initial=0.0) < 1e-7
    assert rv[sel & (k <= 1e8)].max(initial=0.0) < 1e-5
    assert rv[sel].max() < 1e-4
