"""Float32 host fields delivered by the library batch by batch
(spk_interp_steps_f32_dev): equal, bit for bit, to casting the call's
float64 device outputs afterwards (estimates clipped to the cuts as
interp/steps.py:466-476 does), for several batches, the plain path and
IDW, into page-locked and ordinary host memory."""
import numpy as np
import pytest

from spinterps_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('method,precision,T,pinned', [
    ('OK', 'mixed', 130, True),     # three batches: overlapped copies
    ('OK', 'mixed', 40, False),
    ('OK', 'fp64', 3, True),
    ('IDW', 'fp64', 10, True),
])
def test_f32_host_fields_match_cast(gpu, method, precision, T, pinned):
    import torch
    from spinterps_amd.interp.steps import (DeviceProblem, clip_cast,
                                            interp_steps)
    w = configs.hourly(n_side=64, cell=500.0)
    vals, vgs = configs.hourly_steps(w, 0, T)
    prob = DeviceProblem.from_workload(w)
    C = prob.n_cells
    e32 = torch.full((T, C), -1.0, dtype=torch.float32, pin_memory=pinned)
    v32 = torch.full((T, C), -1.0, dtype=torch.float32, pin_memory=pinned)
    lo, hi = 0.05, 8.0
    kw = {} if method == 'IDW' else {'vgs': vgs}
    est, var, info = interp_steps(prob, vals, method=method,
                                  precision=precision, out32=(e32, v32),
                                  min_var_cut=lo, max_var_cut=hi, **kw)
    torch.cuda.synchronize()
    ref_e = clip_cast(est, lo, hi).cpu().numpy()
    np.testing.assert_array_equal(e32.numpy(), ref_e)
    if var is not None:
        np.testing.assert_array_equal(v32.numpy(),
                                      clip_cast(var).cpu().numpy())
    else:
        assert (v32.numpy() == -1.0).all()
    assert np.nanmin(e32.numpy()) >= lo and np.nanmax(e32.numpy()) <= hi
