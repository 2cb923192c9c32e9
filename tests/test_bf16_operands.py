"""Why the mixed path's variance contraction has no BF16-operand variant:
the contraction q = |L^-1 delta|^2 (delta = k - C_J beta, the covariances
deflated by a local fit on the 4 nearest stations as k_kgen does it, so delta_J = 0 and
q(k) = |L^-1 delta|^2 + beta^T C_JJ beta) emulated on the CPU with FP32 and
with BF16 operands (FP32 products and sums, added to FP64 every 32 rows)
on six hourly steps at 500 m with spherical / exponential models (256
random cells).  FP32 stays inside the 1e-6 bound; BF16 misses it by up to
three orders of magnitude."""
import numpy as np
from scipy.linalg import cho_factor, solve_triangular

from oracle import spk_oracle as O
from spinterps_amd import configs


def _bf16(a):
    """Round to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + np.uint32(0x7FFF)
    return ((u + r) & np.uint32(0xFFFF0000)).view(np.float32)


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
