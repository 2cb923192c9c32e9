"""A step series in chunks, written to a NetCDF file as chunks finish,
against one call over every step (GPU)."""
import numpy as np
import pytest

from spinterps_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('neb', [None, 'nrst'])
def test_chunked_series_file_equals_one_call(gpu, tmp_path, neb):
    from spinterps_amd.interp.ncfile import NcWriter, read_nc
    from spinterps_amd.interp.series import interpolate_series
    from spinterps_amd.interp.steps import DeviceProblem, interpolate
    w = configs.small(n_cells_side=20, cell=5000.0)
    vals, vgs = configs.small_steps(w)
    vals[3, :] = np.nan                      # a step without stations
    methods = [('OK', 'OK', None), ('IDW', 'IDW', 2.0)]
    kw = dict(min_var_thr=0.5, min_var_cut=0.0, max_var_cut=30.0)
    if neb:
        kw.update(neb_sel_mthd=neb, n_nebs=12)
    prob = DeviceProblem.from_workload(w)
    ny, nx = w.grid_shape
    xs = np.unique(w.cell_x)
    ys = np.unique(w.cell_y)[::-1]
    wr = NcWriter(tmp_path / 'out.nc', xs, ys, np.arange(w.n_steps),
                  ['OK', 'IDW', 'EST_VARS_OK'])
    plan, prb, kept = interpolate_series(
        prob, vals, vgs, methods, writer=wr, nmrl_prcn=3,
        max_steps_per_chunk=5, first_chunk=7, keep_fields=True, **kw)
    np.testing.assert_array_equal(plan.step_idxs, [0, 7, 12, 17, 22, 24])
    f, _, prb1 = interpolate(prob, vals, vgs, methods, **kw)
    assert prb == prb1 == [3]
    _, _, v = read_nc(tmp_path / 'out.nc')
    for lab in ('OK', 'IDW', 'EST_VARS_OK'):
        one = f[lab].cpu().numpy()
        np.testing.assert_array_equal(kept[lab], one)
        got = v[lab][2].reshape(w.n_steps, -1)
        np.testing.assert_array_equal(got, np.round(one, 3))
