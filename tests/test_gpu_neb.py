"""Neighbourhood interpolation (nrst / pie, interp/grps.py:141-288,
interp/steps.py:592-608, 732-831) on the GPU against oracle.interp_all
(GPU)."""
import numpy as np
import pytest

from oracle import spk_oracle as O
from spinterps_amd import _lib

pytestmark = pytest.mark.gpu

OX, OY = 350000.0, 5300000.0


def _case(seed=7, n=120, side=24, T=8, dup=False):
    """n stations on 100 km x 100 km plus two far ones (never selected:
    dropped), side x side cells, T steps: missing stations, one step with a
    single station, one without any, low values in the west so that some
    neighbourhoods are mean-filled; optionally two stations at one place
    (singular covariance blocks: the pseudo-inverse fallback)."""
    rng = np.random.default_rng(seed)
    sx = np.r_[OX + rng.uniform(0, 1e5, n), OX + 9e5, OX - 6e5]
    sy = np.r_[OY + rng.uniform(0, 1e5, n), OY + 8e5, OY - 7e5]
    if dup:
        sx[1], sy[1] = sx[0], sy[0]
    N = sx.size
    cell = 1e5 / side
    g = np.arange(side) * cell + cell / 2
    cx, cy = [a.ravel() for a in np.meshgrid(OX + g, OY + g[::-1])]
    z = rng.gamma(0.8, 5.0, (T, N)) + 0.05
    west = sx < OX + 3e4
    z[:, west] = rng.uniform(0.0, 0.3, (T, int(west.sum())))
    z[1, rng.choice(n, n // 5, replace=False)] = np.nan
    z[2, :] = np.nan
    z[2, 17] = 3.5                      # one station
    z[3, :] = np.nan                    # none
    z[5, rng.choice(n, 40, replace=False)] = np.nan
    models = ['Sph', 'Exp', 'Gau', 'Sph', 'Exp', 'Gau', 'Sph', 'Exp']
    vgs = []
    for t in range(T):
        v = float(np.nanvar(z[t])) if np.isfinite(z[t]).sum() > 1 else 1.0
        vgs.append(f'{0.1 * v!r} Nug(0.0) + {0.9 * v!r} '
                   f'{models[t]}({float(rng.uniform(2e4, 6e4))!r})')
    drift = np.sin(sx / 2e4) * 300.0 + sy / 1e3 - OY / 1e3
    cdrift = np.sin(cx / 2e4) * 300.0 + cy / 1e3 - OY / 1e3
    return sx, sy, cx, cy, z, vgs, drift, cdrift


def _run(gpu, case, method, sel, n_nebs, n_pies=None, thr=0.5):
    from spinterps_amd.interp.steps import DeviceProblem, interp_neb_steps
    sx, sy, cx, cy, z, vgs, drift, cdrift = case
    edk = method == 'EDK'
    prob = DeviceProblem(sx, sy, cx, cy,
                         stn_drift=drift[:, None] if edk else None,
                         cell_drift=cdrift[None, :] if edk else None)
    e, v, info = interp_neb_steps(prob, z, vgs, method, neb_sel_mthd=sel,
                                  n_nebs=n_nebs, n_pies=n_pies,
                                  min_var_thr=thr)
    _, f64, prb = O.interp_all(
        sx, sy, cx, cy, z, vgs, [(method, method, 2.0)],
        stn_drfts=drift[:, None] if edk else None,
        cell_drfts=cdrift[None, :] if edk else None, min_var_thr=thr,
        neb_sel_mthd=sel, n_nebs=n_nebs, n_pies=n_pies)
    return (e.cpu().numpy(), None if v is None else v.cpu().numpy(),
            info.cpu().numpy(), f64, prb)


def _close(got, ref, scale, rtol, what):
    assert np.array_equal(np.isnan(got), np.isnan(ref)), what
    ok = ~np.isnan(ref)
    err = np.abs(got[ok] - ref[ok]).max() / scale
    assert err <= rtol, (what, err)


@pytest.mark.parametrize('sel,n_nebs,n_pies', [('nrst', 10, None),
                                                ('nrst', 50, None),
                                                ('pie', 12, 4)])
def test_ok_matches_oracle(gpu, sel, n_nebs, n_pies):
    case = _case()
    e, v, info, f64, prb = _run(gpu, case, 'OK', sel, n_nebs, n_pies)
    z = case[4]
    scale = np.nanmax(np.abs(z))
    # FP64 Cholesky against the oracle's pinv on well-conditioned sets
    _close(e, f64['OK'], scale, 1e-9, 'est')
    _close(v, f64['EST_VARS_OK'], np.nanmax(f64['EST_VARS_OK']), 1e-8, 'var')
    assert sorted(prb) == [3]
    assert info[3] == _lib.ST_NOSTNS
    if sel == 'nrst':
        assert info[2] & _lib.ST_ONESTN    # one station: copied, var NaN
        assert np.all(e[2] == 3.5) and np.all(np.isnan(v[2]))
    else:
        # pie repeats the one station n_nebs times (numpy's broadcast in
        # interp/grps.py:245): a singular set, solved by the pseudo-inverse
        assert info[2] & _lib.ST_PINV
        np.testing.assert_allclose(e[2], 3.5, rtol=1e-12)
    # mean-filled groups (variance 0) flagged where the oracle has them
    fill = [bool((f64['EST_VARS_OK'][t] == 0.0).any()) for t in (0, 1, 4)]
    assert [bool(info[t] & 2) for t in (0, 1, 4)] == fill
    if n_nebs == 10:
        assert any(fill)
    assert not any(info[t] & (4 | 16 | 128) for t in range(len(info))
                   if t != 2)


@pytest.mark.parametrize('method', ['IDW', 'NNB', 'SK', 'SKM', 'EDK'])
def test_other_methods_match_oracle(gpu, method):
    case = _case(seed=3)
    e, v, info, f64, prb = _run(gpu, case, method, 'nrst', 12)
    scale = np.nanmax(np.abs(case[4]))
    # EDK's bordered system holds drifts of a few hundred against
    # covariances of about ten: the oracle's pinv carries ~1e-9 there
    tol = {'NNB': 0.0, 'EDK': 1e-8}.get(method, 1e-9)
    _close(e, f64[method], scale, tol, method)
    assert v is None


def test_edk_reproduces_the_drift(gpu):
    """Values that are a linear function of the drift are reproduced at
    every cell (the weights sum to one and reproduce the drift)."""
    from spinterps_amd.interp.steps import DeviceProblem, interp_neb_steps
    sx, sy, cx, cy, z, vgs, drift, cdrift = _case(seed=3)
    z = np.where(np.isnan(z), np.nan, 2.0 + 0.5 * drift[None, :])
    z[:, -2:] = np.nan           # the far stations
    prob = DeviceProblem(sx, sy, cx, cy, stn_drift=drift[:, None],
                         cell_drift=cdrift[None, :])
    e, _, info = interp_neb_steps(prob, z, vgs, 'EDK', neb_sel_mthd='nrst',
                                  n_nebs=12)
    e = e.cpu().numpy()
    for t in (0, 1, 4, 5, 6, 7):
        np.testing.assert_allclose(e[t], 2.0 + 0.5 * cdrift, rtol=1e-9,
                                   atol=1e-9 * np.abs(cdrift).max())


def test_every_station_equals_all_path(gpu):
    """n_nebs at least the active count: one neighbourhood per step, so the
    results are the 'all' path's (FP64 plain) within rounding."""
    from spinterps_amd.interp.steps import (DeviceProblem, interp_neb_steps,
                                            interp_steps)
    sx, sy, cx, cy, z, vgs, _, _ = _case(n=40)
    sx, sy, z = sx[:40], sy[:40], z[:, :40]
    prob = DeviceProblem(sx, sy, cx, cy)
    e1, v1, i1 = interp_neb_steps(prob, z, vgs, 'OK', neb_sel_mthd='nrst',
                                  n_nebs=64, min_var_thr=0.5)
    e2, v2, i2 = interp_steps(prob, z, vgs, 'OK', precision='fp64',
                              min_var_thr=0.5)
    e1, v1, e2, v2 = [a.cpu().numpy() for a in (e1, v1, e2, v2)]
    _close(e1, e2, np.nanmax(np.abs(z)), 1e-9, 'est')
    _close(v1, v2, np.nanmax(v2), 1e-8, 'var')


def test_singular_sets_use_the_pseudo_inverse(gpu):
    """Two stations at one place make every set holding both singular: the
    Cholesky fails and the group is solved by pinv of the bordered system,
    as the reference solves it."""
    case = _case(seed=5, dup=True)
    z = case[4]
    z[:, 1] = z[:, 0]            # one value per place, as a real duplicate
    e, v, info, f64, prb = _run(gpu, case, 'OK', 'nrst', 10)
    scale = np.nanmax(np.abs(z))
    # both sides drop the exact null vector of the duplicate; the kept
    # spectrum of such sets spans about 1e8, so the two pseudo-inverses
    # (an SVD here, a symmetric eigendecomposition there) agree to ~1e-8
    _close(e, f64['OK'], scale, 1e-7, 'est')
    _close(v, f64['EST_VARS_OK'], np.nanmax(f64['EST_VARS_OK']), 1e-7, 'var')
    # with a dropped eigenvalue the sets are flagged ill-conditioned
    assert info[0] & _lib.ST_PINV and info[0] & _lib.ST_ILLCOND


def test_pie_with_too_few_stations_is_refused(gpu):
    from spinterps_amd.interp.steps import DeviceProblem, interp_neb_steps
    sx, sy, cx, cy, z, vgs, _, _ = _case()
    z = z.copy()
    z[0, 5:] = np.nan            # five active stations, twelve wanted
    prob = DeviceProblem(sx, sy, cx, cy)
    with pytest.raises(Exception, match='fewer active stations'):
        interp_neb_steps(prob, z[:1], vgs[:1], 'OK', neb_sel_mthd='pie',
                         n_nebs=12, n_pies=4)
