"""Operand precision of the mixed path's variance contraction: the
contraction q = |L^-1 delta|^2 (delta = k - C_J beta, the covariances
deflated by a local fit on the 4 nearest stations as k_kgen does it, so
delta_J = 0 and q(k) = |L^-1 delta|^2 + beta^T C_JJ beta) emulated on the
CPU on hourly steps at 500 m (256 random cells).  Plain BF16 operands
(FP32 products and sums, added to FP64 every 32 rows) miss the 1e-6 bound
by up to three orders of magnitude, FP32 operands stay inside it, and so
do split-BF16 operands in the arithmetic of k_tri_bf (three bfloat16 terms
per value, six plane products, FP32 sums per 32-row chunk, FP64 across
chunks), whose error stays at FP32's."""
import numpy as np
from scipy.linalg import cho_factor, solve_triangular

from oracle import spk_oracle as O
from spinterps_amd import configs


def _bf16(a):
    """Round to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + np.uint32(0x7FFF)
    return ((u + r) & np.uint32(0xFFFF0000)).view(np.float32)


def _split3(a):
    """x = hi + mid + lo in bfloat16 terms, each the nearest to what the
    previous ones leave, from FP32 values (k_kgen's split3f) or FP64 values
    with FP64 remainders (k_split_tri's split3d)."""
    a = np.asarray(a)
    if a.dtype == np.float32:
        h = _bf16(a)
        r = (a - h).astype(np.float32)
        m = _bf16(r)
        return h, m, _bf16((r - m).astype(np.float32))
    h = _bf16(a.astype(np.float32))
    r = a - h.astype(np.float64)
    m = _bf16(r.astype(np.float32))
    return h, m, _bf16((r - m.astype(np.float64)).astype(np.float32))


# the plane products of k_tri_bf, in its order (small terms first)
_PAIRS = ((2, 0), (1, 1), (0, 2), (1, 0), (0, 1), (0, 0))


def _contract_split(Li, D):
    """k_tri_bf's arithmetic: both operands split from their FP32 values
    (L^-1 as k_tri_bf stages it, delta as k_kgen writes it), per 32-row
    chunk of k the six plane products (exact in FP32; each MFMA's sum
    rounded to FP32 once), the chunk sums added in FP64, q = sum of
    squares in FP64."""
    A = [x.astype(np.float64) for x in _split3(np.float32(Li))]
    B = [x.astype(np.float64) for x in _split3(np.float32(D))]
    n = Li.shape[0]
    acc = np.zeros((n, D.shape[1]))
    for k0 in range(0, n, 32):
        c = np.zeros((n, D.shape[1]), np.float32)
        for i, j in _PAIRS:
            t = A[i][:, k0:k0 + 32] @ B[j][k0:k0 + 32]
            c = (c.astype(np.float64) + t).astype(np.float32)
        acc += c
    return np.sum(acc * acc, axis=0)


def _contract(Li, D, cast):
    """sum_i (L^-1 D)_i^2 per column, operands cast, FP32 products and sums
    over 32-row blocks, the blocks added in FP64."""
    A, B = cast(Li), cast(D)
    u = (A @ B).astype(np.float32)
    q = np.zeros(D.shape[1])
    for r0 in range(0, u.shape[0], 32):
        q += np.sum(u[r0:r0 + 32] * u[r0:r0 + 32], axis=0,
                    dtype=np.float32).astype(np.float64)
    return q


def test_bf16_operands_miss_the_variance_bound():
    w = configs.hourly()
    vals, vgs = configs.hourly_steps(w, 0, 40)
    side = w.grid_shape[0]
    cidx = np.random.default_rng(1).choice(w.n_cells, 256, replace=False)
    cx, cy = w.cell_x[cidx], w.cell_y[cidx]
    err32, err16, used = [], [], 0
    for t in range(vals.shape[0]):
        act = np.where(~np.isnan(vals[t]))[0]
        if 'Gau' in vgs[t] or act.size > 900 or used == 6:
            continue
        used += 1
        sx, sy = w.stn_x[act], w.stn_y[act]
        c0 = sum(s for _, s, _ in O.parse_vg(vgs[t]))
        C = c0 - O.vg_sum(O.dists_2d(sx, sy, sx, sy), vgs[t])
        K = c0 - O.vg_sum(O.dists_2d(cx, cy, sx, sy), vgs[t]).T   # [n x C]
        L = np.linalg.cholesky(C)
        Li = solve_triangular(L, np.eye(act.size), lower=True)
        q_ex = np.sum((Li @ K) ** 2, axis=0)
        # local fit on the 4 nearest stations of each cell
        D = np.empty_like(K)
        qJ = np.empty(K.shape[1])
        d2 = O.dists_2d(cx, cy, sx, sy)
        for c in range(K.shape[1]):
            J = np.argsort(d2[c])[:4]
            beta = np.linalg.solve(C[np.ix_(J, J)], K[J, c])
            D[:, c] = K[:, c] - C[:, J] @ beta
            qJ[c] = beta @ C[np.ix_(J, J)] @ beta
        w1 = solve_triangular(L, np.ones(act.size), lower=True)
        a = w1 @ (Li @ K)
        b = w1 @ w1
        var = c0 - q_ex + (a - 1) * (a - 2) / b
        floor = 1e-3 * np.mean(var)
        for cast, errs in ((lambda x: np.float32(x), err32), (_bf16, err16)):
            q = _contract(Li, D, cast) + qJ
            v = c0 - q + (a - 1) * (a - 2) / b
            errs.append(np.max(np.abs(v - var) / np.maximum(np.abs(var),
                                                            floor)))
    assert used == 6
    # FP32 operands: at most 2e-8 on these steps; BF16: up to 8e-4, above
    # 1e-5 on four of the six
    assert max(err32) <= 1e-6, err32
    assert max(err16) >= 1e-4, err16
    assert sum(e > 1e-5 for e in err16) >= 3, err16


def test_split_bf16_operands_keep_fp32_accuracy():
    """Split-BF16 operands in k_tri_bf's arithmetic on the first eight
    hourly steps with at most 1 200 active stations (spherical,
    exponential and Gaussian models): within 1e-6 wherever FP32 operands
    are, and never more than 4x FP32's error (plus 1e-10).  Three plane
    products instead of six leave 10-100x FP32's error: the check fails
    if a product is dropped."""
    w = configs.hourly()
    vals, vgs = configs.hourly_steps(w, 0, 60)
    cidx = np.random.default_rng(1).choice(w.n_cells, 256, replace=False)
    cx, cy = w.cell_x[cidx], w.cell_y[cidx]
    used = 0
    res = []
    for t in range(vals.shape[0]):
        act = np.where(~np.isnan(vals[t]))[0]
        if act.size > 1200 or used == 8:
            continue
        used += 1
        sx, sy = w.stn_x[act], w.stn_y[act]
        c0 = sum(s for _, s, _ in O.parse_vg(vgs[t]))
        C = c0 - O.vg_sum(O.dists_2d(sx, sy, sx, sy), vgs[t])
        K = c0 - O.vg_sum(O.dists_2d(cx, cy, sx, sy), vgs[t]).T
        L = np.linalg.cholesky(C)
        Li = solve_triangular(L, np.eye(act.size), lower=True)
        d2 = O.dists_2d(cx, cy, sx, sy)
        D = np.empty_like(K)
        qJ = np.empty(K.shape[1])
        for c in range(K.shape[1]):
            J = np.argsort(d2[c])[:4]
            beta = np.linalg.solve(C[np.ix_(J, J)], K[J, c])
            D[:, c] = K[:, c] - C[:, J] @ beta
            qJ[c] = beta @ C[np.ix_(J, J)] @ beta
        w1 = solve_triangular(L, np.ones(act.size), lower=True)
        a = w1 @ (Li @ K)
        b = w1 @ w1
        var = c0 - np.sum((Li @ K) ** 2, axis=0) + (a - 1) * (a - 2) / b
        floor = 1e-3 * np.mean(var)

        def err(q):
            v = c0 - (q + qJ) + (a - 1) * (a - 2) / b
            return np.max(np.abs(v - var) / np.maximum(np.abs(var), floor))

        e32 = err(_contract(Li, D, lambda x: np.float32(x)))
        e6 = err(_contract_split(Li, D))
        res.append((t, e32, e6))
    assert used == 8
    for t, e32, e6 in res:
        if e32 <= 1e-6:
            assert e6 <= 1e-6, res
        assert e6 <= 4 * e32 + 1e-10, res
