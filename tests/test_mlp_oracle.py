"""Pins the learned-proposal oracle (oracle/mlp_oracle.py): gradients by
central finite differences of the loss, the loss by closed forms, the
features by an independent pdist/histogram computation and the field
generator by its moments (standard normals; sample covariance of many
simulated steps against the model covariance)."""
import math

import numpy as np
from scipy.spatial.distance import pdist
from scipy.stats import kstest, norm

from oracle import mlp_oracle as M
from oracle import spk_oracle as O


def _net(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    _, n = M.offsets()
    return rng.normal(0, scale, n)


def _batch(seed, B=7):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (B, M.N_IN))
    cls = rng.integers(0, 3, B)
    reg = rng.normal(0, 1, (B, 3))
    reg[1, 2] = np.nan
    return X, cls, reg


def test_param_count():
    _, n = M.offsets()
    assert n == 40 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 6 + 6


def test_zero_network_closed_form():
    _, n = M.offsets()
    X, cls, reg = _batch(1)
    loss, g = M.mlp_loss_grad(np.zeros(n), X, cls, reg, 0.5)
    r = np.where(np.isnan(reg), 0.0, reg)
    assert math.isclose(loss, math.log(3) + 0.5 * np.sum(r * r) / 7,
                        rel_tol=1e-14)
    # output bias gradient: mean of (softmax - onehot) and -2 w r
    offs, _ = M.offsets()
    bo, bn = offs[3][1]
    gb = g[bo:bo + bn]
    oh = np.eye(3)[cls]
    np.testing.assert_allclose(gb[:3], np.mean(1 / 3 - oh, axis=0),
                               rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(gb[3:], -2 * 0.5 * r.sum(axis=0) / 7,
                               rtol=1e-13, atol=1e-15)


def test_gelu_definition():
    x = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(M.gelu(x), x * norm.cdf(x), rtol=1e-13,
                               atol=1e-15)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(M.gelu_d(x), fd, rtol=1e-7, atol=1e-9)


def test_gradients_by_finite_differences():
    p = _net(2)
    X, cls, reg = _batch(3)
    _, g = M.mlp_loss_grad(p, X, cls, reg, 0.7)
    rng = np.random.default_rng(4)
    offs, n = M.offsets()
    # every layer's weights and biases are sampled
    pick = np.concatenate(
        [rng.integers(wo, wo + wn, 12) for (wo, wn), _ in offs] +
        [rng.integers(bo, bo + bn, 4) for _, (bo, bn) in offs])
    h = 1e-6
    for k in pick:
        e = np.zeros(n)
        e[k] = h
        fd = (M.mlp_loss(p + e, X, cls, reg, 0.7)
              - M.mlp_loss(p - e, X, cls, reg, 0.7)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_adam_first_step_is_signed_lr():
    g = np.array([0.5, -2.0, 1e-3])
    p, m, v = M.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01,
                     eps=0.0)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-12)


def test_features_against_histogram():
    rng = np.random.default_rng(5)
    n = 40
    x = rng.uniform(0, 1e4, n)
    y = rng.uniform(0, 1e4, n)
    z = rng.normal(size=n)
    z[[3, 17]] = np.nan
    f, (md, var) = M.features(x, y, z, mdr=0.6, nlag=20)
    ok = ~np.isnan(z)
    d = pdist(np.c_[x[ok], y[ok]])
    dz = pdist(z[ok][:, None])
    assert math.isclose(md, 0.6 * d.max(), rel_tol=1e-15)
    assert math.isclose(var, np.var(z[ok]), rel_tol=1e-14)
    keep = d <= md
    cnt, _ = np.histogram(d[keep], bins=20, range=(0, md))
    sm, _ = np.histogram(d[keep], bins=20, range=(0, md),
                         weights=0.5 * dz[keep] ** 2)
    np.testing.assert_array_equal(f[20:] * keep.sum() / 20, cnt)
    np.testing.assert_allclose(f[:20], sm / np.maximum(cnt, 1) / var,
                               rtol=1e-12)


def test_generator_normals_moments():
    z = M.normals(123, 5, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
    assert kstest(z, 'norm').pvalue > 1e-3
    # distinct steps and seeds give different streams
    assert not np.allclose(M.normals(123, 6, 10), z[:10])
    assert not np.allclose(M.normals(124, 5, 10), z[:10])


def test_simulation_covariance():
    rng = np.random.default_rng(6)
    n, T = 5, 6000
    x = rng.uniform(0, 3e4, n)
    y = rng.uniform(0, 3e4, n)
    vg = '0.3 Nug(0.0) + 2.0 Exp(20000.0)'
    mask = np.zeros((T, n))
    out = M.simulate(x, y, mask, [vg] * T, np.full(T, 4.0), seed=9)
    d = O.dists_2d(x, y, x, y)
    c = 2.0 * np.exp(-3 * d / 2e4)
    c[np.diag_indices(n)] = 2.3 + 1e-6 * 2.3
    emp = np.cov(out.T, bias=True)
    # standard error of a covariance estimate ~ sqrt((c_ii c_jj + c_ij^2)/T)
    se = np.sqrt((np.outer(np.diag(c), np.diag(c)) + c ** 2) / T)
    assert np.all(np.abs(emp - c) < 4.5 * se)
    assert abs(out.mean() - 4.0) < 4.5 * math.sqrt(c.mean() / T)
