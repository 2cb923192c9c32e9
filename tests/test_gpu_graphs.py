"""Captured factorisation sequences (hipGraph): the batched path with its
Cholesky, condition-estimate and dual-vector launch sequences replayed
from graphs gives the same results as launching them one by one, on every
replay, with other work and synchronisations between the replays."""
import numpy as np
import pytest

from spinterps_amd import configs

pytestmark = pytest.mark.gpu


def _run(prob, vals, vgs, method, precision):
    import torch
    from spinterps_amd.interp.steps import interp_steps
    e, v, info = interp_steps(prob, vals, vgs, method, precision=precision)
    torch.cuda.synchronize()
    # SK has no variance output
    return (e.cpu().numpy(), None if v is None else v.cpu().numpy(),
            info.cpu().numpy())


def _same(a, b):
    for x, y in zip(a, b):
        assert (x is None) == (y is None)
        if x is not None:
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize('method,precision,T', [
    ('OK', 'mixed', 130),     # batches of 64, 64 and 2: both buffer sets
    ('OK', 'fp64b', 70),
    ('SK', 'mixed', 40),
    ('OK', 'fp64', 3),        # the plain path's Cholesky sequence
])
def test_graph_replays_match_launches(gpu, method, precision, T):
    import torch
    from spinterps_amd import _lib
    from spinterps_amd.interp.steps import DeviceProblem
    w = configs.hourly(n_side=64, cell=500.0)
    vals, vgs = configs.hourly_steps(w, 0, T)
    prob = DeviceProblem.from_workload(w)
    ctx = _lib.context(0)
    try:
        ctx.set_graphs(False)
        ref = _run(prob, vals, vgs, method, precision)
        ctx.set_graphs(True)
        for rep in range(3):
            got = _run(prob, vals, vgs, method, precision)
            _same(ref, got)
            # eager work of the program between replays
            x = torch.randn(1 << 20, device='cuda', dtype=torch.float64)
            (x * 2.0).sum().item()
            # and the library's own launches, one by one, on other steps
            ctx.set_graphs(False)
            _run(prob, vals[-5:], vgs[-5:], method, precision)
            ctx.set_graphs(True)
    finally:
        ctx.set_graphs(True)


def test_graph_replays_after_other_shapes(gpu):
    """Replays stay correct after calls with other step counts and grids
    (other graphs, reallocated buffers) in between."""
    from spinterps_amd import _lib
    from spinterps_amd.interp.steps import DeviceProblem
    ctx = _lib.context(0)
    w1 = configs.hourly(n_side=64, cell=500.0)
    w2 = configs.hourly(n_side=40, cell=800.0)
    v1, g1 = configs.hourly_steps(w1, 0, 70)
    v2, g2 = configs.hourly_steps(w2, 200, 210)
    p1 = DeviceProblem.from_workload(w1)
    p2 = DeviceProblem.from_workload(w2)
    try:
        ctx.set_graphs(False)
        r1 = _run(p1, v1, g1, 'OK', 'mixed')
        r2 = _run(p2, v2, g2, 'OK', 'mixed')
        ctx.set_graphs(True)
        for _ in range(2):
            _same(r1, _run(p1, v1, g1, 'OK', 'mixed'))
            _same(r2, _run(p2, v2, g2, 'OK', 'mixed'))
    finally:
        ctx.set_graphs(True)
