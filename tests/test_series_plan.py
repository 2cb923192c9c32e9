"""Chunk planning of a step series (the role of interp/main.py:652-859)."""
import numpy as np
import pytest

from spinterps_amd.interp.series import output_bytes_per_step, plan_steps

METHODS = [('OK', 'OK', None), ('IDW', 'IDW', 2.0)]


def test_bytes_per_step():
    dev, host = output_bytes_per_step(1000, METHODS)
    # OK, IDW and the OK variance: FP64 + float32 on the device, float32
    # on the host
    assert dev == 1000 * 3 * 12 and host == 1000 * 3 * 4
    dev, host = output_bytes_per_step(1000, METHODS, est_vars_flag=False)
    assert dev == 1000 * 2 * 12


def test_chunks_from_the_tighter_budget():
    C = 10 ** 6
    dev_b, host_b = output_bytes_per_step(C, METHODS)
    # device: 100 steps after the 10 % reserve; host: 2 chunks of 40
    p = plan_steps(1000, C, 1000, METHODS, device_free=dev_b * 100 / 0.9,
                   host_free=host_b * 80)
    assert p.steps_per_chunk == 40
    np.testing.assert_array_equal(p.step_idxs[:3], [0, 40, 80])
    assert p.step_idxs[-1] == 1000 and len(p.step_idxs) == 26
    p = plan_steps(1000, C, 1000, METHODS, device_free=dev_b * 100 / 0.9,
                   host_free=host_b * 10 ** 6)
    assert p.steps_per_chunk == 100
    p = plan_steps(1000, C, 1000, METHODS, device_free=dev_b * 100 / 0.9,
                   host_free=host_b * 10 ** 6, max_steps_per_chunk=30)
    assert p.steps_per_chunk == 30
    np.testing.assert_array_equal(p.grid_idxs, [0, 1000])


def test_no_step_fits():
    with pytest.raises(AssertionError, match='Not enough memory'):
        plan_steps(10, 10 ** 6, 1000, METHODS, device_free=1000,
                   host_free=10 ** 12)
