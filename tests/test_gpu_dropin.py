"""The drop-in step driver interpolate() (the method loop, clipping to
[min_var_cut, max_var_cut], the float32 cast and the problem-step list of
interp/steps.py:466-476, 478-835) against oracle.interp_all (GPU)."""
import numpy as np
import pytest

from oracle import spk_oracle as O
from spinterps_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_interpolate_matches_reference_driver(gpu, precision):
    import torch

    from spinterps_amd.interp.steps import DeviceProblem, clip_cast, interpolate
    w = configs.small(n_cells_side=30, cell=1e5 / 30)
    vals, vgs = configs.small_steps(w, 0, 10)
    vals[2, :] = np.nan                  # no station: a problem step
    vals[3, 1:] = np.nan                 # one station: its value everywhere
    vals[4, :] = np.abs(vals[4, :]) * 0.0 + 0.25   # all below min_var_thr
    vals[5, ::3] = np.nan                # a different station set
    methods = [('OK', 'OK', None), ('IDW', 'IDW', 2.0), ('NNB', 'NNB', None)]
    kw = dict(min_var_thr=0.5, min_var_cut=0.8, max_var_cut=9.0)
    f_ref, f64_ref, prb_ref = O.interp_all(
        w.stn_x, w.stn_y, w.cell_x, w.cell_y, vals, vgs, methods, **kw)
    prob = DeviceProblem.from_workload(w)
    f, f64, prb = interpolate(prob, vals, vgs, methods, precision=precision,
                              **kw)
    assert prb == sorted(int(s) for s in prb_ref) == [2]
    for lab in ('OK', 'IDW', 'NNB', 'EST_VARS_OK'):
        got = f[lab].cpu().numpy()
        ref = f_ref[lab]
        assert got.dtype == np.float32 and got.shape == ref.shape
        assert np.array_equal(np.isnan(got), np.isnan(ref)), lab
        ok = ~np.isnan(ref)
        if lab != 'EST_VARS_OK':
            # clipped like mod_min_max (estimates only, not variances)
            assert got[ok].min() >= 0.8 and got[ok].max() <= 9.0
        # float32 of values within 1e-9 (1e-6 for mixed variances): at most
        # one float32 rounding step apart
        tol = 1e-6 if (precision == 'mixed' and lab == 'EST_VARS_OK') else 0
        np.testing.assert_allclose(got[ok], ref[ok], rtol=1.2e-7 + tol,
                                   atol=1e-6 * np.abs(ref[ok]).max())
    assert np.all(f['OK'].cpu().numpy()[3] == np.float32(vals[3, 0]).clip(
        0.8, 9.0))
    assert np.all(f['OK'].cpu().numpy()[4] == np.float32(0.8))  # mean fill
    # clip_cast alone: the reference's cut and float32 cast on the device
    x = torch.as_tensor(np.array([-1.0, 0.3, np.nan, 12.0, 4.25]),
                        device=gpu)
    y = clip_cast(x, 0.0, 10.0).cpu().numpy()
    np.testing.assert_array_equal(
        y, np.array([0.0, 0.3, np.nan, 10.0, 4.25], dtype=np.float32))
    y = clip_cast(x).cpu().numpy()
    np.testing.assert_array_equal(y, x.cpu().numpy().astype(np.float32))
