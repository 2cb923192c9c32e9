"""Class API (OrdinaryKriging, SimpleKriging, ExternalDriftKriging_MD of
pyx:251-719) on the GPU against the oracle's restatement (which inverts
with numpy's pinv, as the reference does).  The GPU inverts through the
Cholesky factor of the covariance form, so values agree to the systems'
rounding: 1e-9 relative to each array's largest magnitude."""
import numpy as np
import pytest

from oracle import spk_oracle as O

pytestmark = pytest.mark.gpu


def _pts(seed, n, C, span=1e4):
    rng = np.random.default_rng(seed)
    xi = 350000 + rng.uniform(0, span, n)
    yi = 5300000 + rng.uniform(0, span, n)
    zi = rng.gamma(0.8, 5.0, n)
    xk = 350000 + rng.uniform(0, span, C)
    yk = 5300000 + rng.uniform(0, span, C)
    # two output points on stations: their right-hand sides skip d == 0
    xk[:2], yk[:2] = xi[[3, 7]], yi[[3, 7]]
    return xi, yi, zi, xk, yk


def _close(a, b, tol=1e-9):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    scale = max(np.abs(b).max(), 1e-300)
    np.testing.assert_allclose(a, b, rtol=0, atol=tol * scale)


@pytest.mark.parametrize('model', ['0.2 Nug(0.0) + 1.5 Exp(3000.0)',
                                   '1.0 Sph(4000.0)',
                                   '0.1 Nug(0.0) + 2.0 Gau(2500.0)',
                                   '0.05 Nug(0.0) + 0.01 Pow(1.2)'])
def test_ordinary_kriging_class(gpu, model):
    from spinterps_amd.cyth import OrdinaryKriging
    xi, yi, zi, xk, yk = _pts(1, 40, 300)
    ok = OrdinaryKriging(xi, yi, zi, xk, yk, model)
    ok.krige()
    ref = O.ordinary_kriging(xi, yi, zi, xk, yk, model)
    _close(ok.in_vars, ref['in_vars'], 1e-15)
    _close(ok.rhss, ref['rhss'], 1e-15)
    _close(ok.in_vars_inv, ref['in_vars_inv'])
    _close(ok.lambdas, ref['lambdas'])
    _close(ok.mus, ref['mus'])
    _close(ok.zk, ref['zk'])
    _close(ok.est_vars, ref['est_vars'])
    assert np.all(ok.rhss[:2, [3, 7]][[0, 1], [0, 1]] == 0.0)
    d = np.hypot(xi[:, None] - xi[None], yi[:, None] - yi[None])
    np.testing.assert_allclose(ok.in_dists, d, rtol=1e-15)


def test_simple_kriging_class(gpu):
    from spinterps_amd.cyth import SimpleKriging
    xi, yi, zi, xk, yk = _pts(2, 30, 200)
    model = '0.3 Nug(0.0) + 1.0 Sph(3000.0)'
    sk = SimpleKriging(xi, yi, zi, xk, yk, model)
    sk.krige()
    ref = O.simple_kriging(xi, yi, zi, xk, yk, model)
    _close(sk.in_covars, ref['in_covars'], 1e-15)
    assert np.all(np.diag(sk.in_covars) == 0.0)     # filled before the sum
    _close(sk.out_covars, ref['out_covars'], 1e-15)
    _close(sk.in_covars_inv, ref['in_covars_inv'], 1e-8)
    _close(sk.lambdas, ref['lambdas'], 1e-8)
    _close(sk.zk, ref['zk'], 1e-8)
    _close(sk.est_covars, ref['est_covars'], 1e-8)


def test_external_drift_kriging_md_class(gpu):
    from spinterps_amd.cyth import ExternalDriftKriging_MD
    xi, yi, zi, xk, yk = _pts(3, 50, 250)
    rng = np.random.default_rng(4)
    si = rng.uniform(0, 2500, (2, 50))
    sk = rng.uniform(0, 2500, (2, 250))
    model = '0.1 Nug(0.0) + 1.0 Exp(5000.0)'
    ed = ExternalDriftKriging_MD(xi, yi, zi, si, xk, yk, sk, model)
    ed.krige()
    ref = O.external_drift_kriging_md(xi, yi, zi, si, xk, yk, sk, model)
    _close(ed.in_vars, ref['in_vars'], 1e-15)
    _close(ed.rhss, ref['rhss'], 1e-15)
    _close(ed.in_vars_inv, ref['in_vars_inv'], 1e-8)
    _close(ed.lambdas, ref['lambdas'], 1e-8)
    _close(ed.mus_arr, ref['mus_arr'], 1e-8)
    _close(ed.zk, ref['zk'], 1e-8)
    # the drift is reproduced: sum lambda s_i = s_k
    np.testing.assert_allclose(ed.lambdas @ si.T, sk.T, rtol=1e-7)


def test_class_errors(gpu):
    from spinterps_amd.cyth import ExternalDriftKriging_MD, OrdinaryKriging
    xi, yi, zi, xk, yk = _pts(5, 10, 5)
    with pytest.raises(ValueError):
        OrdinaryKriging(xi, yi, zi, xk, yk, '1.0 Foo(2.0)').krige()
    with pytest.raises(AssertionError):
        ExternalDriftKriging_MD(xi, yi, zi, np.zeros((1, 9)), xk, yk,
                                np.zeros((1, 5)))
