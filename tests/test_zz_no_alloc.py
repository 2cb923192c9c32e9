"""The allocation-free IDW / nearest-neighbour entry point
(spk_idw_steps_noalloc_dev) against spk_interp_steps_dev, through the C
ABI: two calls in a row, the second with a smaller input, must give
bit-identical fields and flags (same kernels, scratch from the library's
static device array instead of allocated workspaces)."""
import ctypes

import numpy as np
import pytest

from spinterps_amd import _lib, configs

pytestmark = pytest.mark.gpu


def _noalloc(prob, vals_d, method, p):
    import torch

    from spinterps_amd.interp.steps import _ctx_for
    T, C = vals_d.shape[0], prob.n_cells
    opts = _lib.SpkOpts()
    opts.method = _lib.METHODS[method]
    opts.precision = 0
    opts.idw_exp = float(p)
    opts.min_vg_val = 0.0
    opts.min_var_thr = float('-inf')
    est = torch.full((T, C), float('nan'), dtype=torch.float64,
                     device=vals_d.device)
    info = torch.zeros(T, dtype=torch.int32, device=vals_d.device)
    pst = prob.c_struct()
    ctx = _ctx_for(vals_d.device)
    _lib.check(_lib.lib().spk_idw_steps_noalloc_dev(
        ctx.h, ctypes.byref(pst), ctypes.byref(opts), T, vals_d.data_ptr(),
        est.data_ptr(), info.data_ptr(), None))
    return est, info


def test_noalloc_matches_allocating_entry(gpu):
    import torch

    from spinterps_amd.interp.steps import DeviceProblem, interp_steps
    cases = [
        # 1000 stations (0-20 % missing), 40 000 cells, 70 steps: two
        # launches of the 64-step scratch
        (configs.daily(n_side=200, cell=2500.0), 0, 70),
        # then a smaller problem: 300 stations, 2 500 cells, 5 steps
        (configs.daily(n_side=50, cell=1e4, n_stns=300, seed=9), 3, 8),
    ]
    for w, t0, t1 in cases:
        vals, _ = configs.daily_steps(w, t0, t1)
        prob = DeviceProblem.from_workload(w)
        vd = torch.as_tensor(vals, device=gpu)
        for method, p in (('IDW', 2.0), ('IDW', 1.5), ('NNB', 2.0)):
            e_ref, _, i_ref = interp_steps(prob, vd, None, method, idw_exp=p)
            e_na, i_na = _noalloc(prob, vd, method, p)
            torch.cuda.synchronize()
            a, b = e_na.cpu().numpy(), e_ref.cpu().numpy()
            assert np.array_equal(a, b, equal_nan=True), (w.n_stns, method, p)
            assert np.array_equal(i_na.cpu().numpy(), i_ref.cpu().numpy())
            assert np.isfinite(a).all()


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    import torch

    from spinterps_amd.interp.steps import DeviceProblem
    rng = np.random.default_rng(0)
    N = 8193
    x, y = rng.uniform(0, 1e4, N), rng.uniform(0, 1e4, N)
    prob = DeviceProblem(x, y, x[:10], y[:10])
    vd = torch.zeros((1, N), dtype=torch.float64, device=gpu)
    with pytest.raises(AssertionError, match='exceed the scratch'):
        _noalloc(prob, vd, 'IDW', 2.0)
