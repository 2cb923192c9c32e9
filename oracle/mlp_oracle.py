"""CPU oracle (test infrastructure only) for the learned variogram proposal:
the synthetic-field generator, the lag-bin features and the multilayer
perceptron's forward pass, loss and gradients, in NumPy float64.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.

The proposal has no counterpart in the reference; what it serves is the
fit of variograms/fitvgs.py:143-278.  Pinning status:
  * mlp_loss_grad: gradients pinned by central finite differences of
    mlp_loss (tests/test_mlp_oracle.py); the forward pass and loss by closed
    forms (an all-zero network gives log 3 and the bias outputs, GELU
    against its integral definition).
  * features: pinned by a brute-force restatement over all station pairs
    of tiny inputs (no sorting, no pair list).
  * simulate: pinned by the covariance identity L L^T = C of the Cholesky
    factor it uses and by the standard-normal moments of the counter-based
    generator.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

from . import spk_oracle as O

N_IN, N_H, N_OUT = 40, 128, 6
DIMS = [N_IN, N_H, N_H, N_H, N_OUT]
M64 = (1 << 64) - 1


def offsets():
    """Flat layout [W1 b1 W2 b2 W3 b3 W4 b4], W_l row-major [out x in]."""
    out, o = [], 0
    for l in range(1, 5):
        w = (o, DIMS[l] * DIMS[l - 1])
        o += w[1]
        b = (o, DIMS[l])
        o += b[1]
        out.append((w, b))
    return out, o


def unpack(p):
    offs, _ = offsets()
    Ws, bs = [], []
    for l, ((wo, wn), (bo, bn)) in enumerate(offs, start=1):
        Ws.append(p[wo:wo + wn].reshape(DIMS[l], DIMS[l - 1]))
        bs.append(p[bo:bo + bn])
    return Ws, bs


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_d(x):
    return (0.5 * (1.0 + erf(x / math.sqrt(2.0)))
            + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi))


def mlp_forward(p, X):
    """Returns the output [B x 6] and the pre-activations of every layer."""
    Ws, bs = unpack(np.asarray(p, np.float64))
    H = np.asarray(X, np.float64)
    Zs, Hs = [], [H]
    for l in range(4):
        Z = H @ Ws[l].T + bs[l]
        H = gelu(Z) if l < 3 else Z
        Zs.append(Z)
        Hs.append(H)
    return H, Zs, Hs


def mlp_loss(p, X, cls, reg, reg_w):
    Y, _, _ = mlp_forward(p, X)
    return _loss_and_dy(Y, np.asarray(cls), np.asarray(reg, np.float64),
                        reg_w)[0]


def _loss_and_dy(Y, cls, reg, reg_w):
    B = Y.shape[0]
    lg = Y[:, :3]
    mx = lg.max(axis=1, keepdims=True)
    e = np.exp(lg - mx)
    se = e.sum(axis=1, keepdims=True)
    lse = mx[:, 0] + np.log(se[:, 0])
    has = (cls >= 0) & (cls < 3)
    cc = np.where(has, cls, 0)
    loss = np.sum(np.where(has, lse - lg[np.arange(B), cc], 0.0))
    dY = np.zeros_like(Y)
    dY[:, :3] = np.where(has[:, None], e / se, 0.0)
    dY[np.arange(B)[has], cls[has]] -= 1.0
    r = Y[:, 3:] - reg
    ok = ~np.isnan(reg)
    r = np.where(ok, r, 0.0)
    loss += reg_w * np.sum(r * r)
    dY[:, 3:] = 2.0 * reg_w * r
    return loss / B, dY / B


def mlp_loss_grad(p, X, cls, reg, reg_w):
    """Mean softmax cross-entropy over (Sph, Exp, Gau) + reg_w * squared
    error of the 3 regression outputs (NaN targets ignored; a class outside
    0..2 drops that sample's cross-entropy term)."""
    p = np.asarray(p, np.float64)
    Y, Zs, Hs = mlp_forward(p, X)
    loss, dA = _loss_and_dy(Y, np.asarray(cls), np.asarray(reg, np.float64),
                            reg_w)
    Ws, _ = unpack(p)
    offs, n = offsets()
    g = np.zeros(n)
    for l in range(3, -1, -1):
        if l < 3:
            dA = dA * gelu_d(Zs[l])
        (wo, wn), (bo, bn) = offs[l]
        g[wo:wo + wn] = (dA.T @ Hs[l]).reshape(-1)
        g[bo:bo + bn] = dA.sum(axis=0)
        dA = dA @ Ws[l]
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step))
                                           + eps)
    return p, m, v


# ---- features ------------------------------------------------------------
def features(x, y, z, mdr=0.5, nlag=20):
    """Lag-bin features of one step from its active stations: nlag equal
    bins over (0, mdr * largest pair distance]; half mean squared
    difference / variance, then each bin's share of the pairs times nlag
    (1 for pairs spread evenly over the bins).  Returns
    (features [2 nlag], (lag scale, variance))."""
    ok = ~np.isnan(z)
    x, y, z = np.asarray(x)[ok], np.asarray(y)[ok], np.asarray(z)[ok]
    var = np.var(z)
    n = z.size
    sums = np.zeros(nlag)
    cnts = np.zeros(nlag, dtype=np.int64)
    md = 0.0
    for i in range(n):
        for j in range(i):
            md = max(md, math.sqrt((x[i] - x[j]) ** 2 + (y[i] - y[j]) ** 2))
    md *= mdr
    for i in range(n):
        for j in range(i):
            d = math.sqrt((x[i] - x[j]) ** 2 + (y[i] - y[j]) ** 2)
            if md <= 0 or d > md:
                continue
            b = min(int(d / md * nlag), nlag - 1)
            sums[b] += 0.5 * (z[i] - z[j]) ** 2
            cnts[b] += 1
    f = np.zeros(2 * nlag)
    g = np.where(cnts > 0, sums / np.maximum(cnts, 1), 0.0)
    f[:nlag] = g / var if var > 0 else 0.0
    tot = cnts.sum()
    f[nlag:] = cnts * nlag / tot if tot else 0.0
    return f, (md, var)


# ---- synthetic fields -------------------------------------------------------
def _smix(x):
    x = (x + np.uint64(0x9e3779b97f4a7c15)) & np.uint64(M64)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    return x ^ (x >> np.uint64(31))


def normals(seed, t, n):
    """The generator's standard normals for step t (counter-based
    splitmix64 hashing + Box-Muller)."""
    with np.errstate(over='ignore'):
        sd = _smix(np.uint64(seed) ^ _smix(np.array([t], np.uint64)))[0]
        k = np.arange(n, dtype=np.uint64)
        a = _smix(sd ^ _smix(np.uint64(2) * k))
        b = _smix(sd ^ _smix(np.uint64(2) * k + np.uint64(1)))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0
    u2 = (b >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def sim_covariance(x, y, vg):
    """Covariance of the simulation: sum over components of sill - gamma
    (the nugget's sill on the diagonal) + 1e-6 of the sill sum."""
    d = O.dists_2d(x, y, x, y)
    c = O.vg_sum(d, vg, covar_flag=1)
    comps = O.parse_vg(vg)
    nug = sum(sl for nm, sl, r in comps if nm == 'Nug')
    tot = sum(sl for nm, sl, r in comps)
    c[np.diag_indices_from(c)] += nug + 1e-6 * tot
    return c


def simulate(x, y, mask_vals, vgs, means, seed):
    """[T x N] fields with NaN where mask_vals is NaN."""
    mask_vals = np.asarray(mask_vals, np.float64)
    T, N = mask_vals.shape
    out = np.full((T, N), np.nan)
    for t in range(T):
        act = np.where(~np.isnan(mask_vals[t]))[0]
        if act.size < 2:
            continue
        c = sim_covariance(np.asarray(x)[act], np.asarray(y)[act], vgs[t])
        L = np.linalg.cholesky(c)
        out[t, act] = means[t] + L @ normals(seed, t, act.size)
    return out
