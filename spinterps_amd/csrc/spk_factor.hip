// Per-step kriging systems: station compaction, covariance assembly,
// Cholesky, triangular inverse, the dual-form vectors of the batched path
// and the explicit inverse operator of the plain path.
//
// Every kernel here runs one workgroup per time step (blockIdx.x = step),
// so a launch factorises a whole batch of uneven systems at once.  Matrices
// are column-major with a common leading dimension `ld`.
//
// Reference: interp/steps.py:170-243 (system layout), :350-371 (inverse),
// misc.py:1074-1105 (nugget-only check), interp/steps.py:276, :760-765
// (step means and interpolation flags).
#include "spk_krige.h"

namespace spk {

// ---- step compaction -------------------------------------------------------
// Active stations (non-NaN values) of each step, in station order (or in
// the order sorder gives, a spatially compact one for the batched path), as
// interp/grps.py:57-101 forms a step's station set; the step mean over them
// (interp/steps.py:276) and the "any value >= min_var_thr" flag
// (interp/steps.py:760-765); then the step status.
__global__ __launch_bounds__(256) void k_compact(
    const double *__restrict__ vals, int64_t N, int32_t *__restrict__ idx,
    double *__restrict__ zact, int64_t ldz, StepSc *__restrict__ sc,
    const spk_vg *__restrict__ vgs, int method, double thr,
    double min_vg_val, const int32_t *__restrict__ sorder) {
  const int s = blockIdx.x;
  const double *row = vals + (int64_t)s * N;
  int32_t *my_idx = idx + (int64_t)s * N;
  double *my_z = zact + (int64_t)s * ldz;
  __shared__ int wave_cnt[4];
  __shared__ double red[256], red_abs[256];
  __shared__ int red_any[256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int base_off = 0;
  double sum = 0.0, sabs = 0.0;
  int any = 0;
  for (int64_t b = 0; b < N; b += 256) {
    const int64_t i = (b + tid < N && sorder) ? (int64_t)sorder[b + tid]
                                               : b + tid;
    const double v = (b + tid < N) ? row[i] : NAN;
    const bool act = !isnan(v);
    const unsigned long long m = __ballot(act);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_cnt[wave] = __popcll(m);
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wave; ++w) woff += wave_cnt[w];
    if (act) {
      const int pos = base_off + woff + pre;
      my_idx[pos] = (int32_t)i;
      my_z[pos] = v;
      sum += v;
      sabs += fabs(v);
      any |= (v >= thr) ? 1 : 0;
    }
    const int tot = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    __syncthreads();
    base_off += tot;
  }
  red[tid] = sum;
  red_abs[tid] = sabs;
  red_any[tid] = any;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      red[tid] += red[tid + w];
      red_abs[tid] += red_abs[tid + w];
      red_any[tid] |= red_any[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int n = base_off;
    StepSc &o = sc[s];
    o.n = n;
    o.mean = n > 0 ? red[0] / n : NAN;
    o.zabs = n > 0 ? red_abs[0] / n : 0.0;
    int st = 0;
    if (n == 0) {
      st = SPK_ST_NOSTNS;
    } else if (n == 1) {
      st = SPK_ST_ONESTN;
    } else if (method == SPK_M_IDW) {
      if (!red_any[0]) st = SPK_ST_MEANFILL;
    } else if (method == SPK_M_NNB) {
      st = 0;
    } else {
      const spk_vg &vg = vgs[s];
      const bool nug =
          (vg.sill_sum <= min_vg_val) || (vg.range_max_raw <= min_vg_val);
      if (!red_any[0] || nug) st = SPK_ST_MEANFILL;
    }
    o.status = st;
    o.flags = 0;
    o.planes = 0;
    o.pinv = 0;
    o.kappa = 0.0;
    o.c0 = 0.0;
    o.b = 0.0;
    o.a0 = 0.0;
    for (int k = 0; k < kMaxP; ++k) {
      o.h[k] = 0.0;
      o.dh[k] = 0.0;
    }
  }
}

// ---- system assembly ---------------------------------------------------------
// A[i, j] for i, j < n: the semivariogram sum G (OK, EDK) or the covariance
// sum (SK) of fill_vg_var_arr (pyx:146-226) at the distance of active
// stations i and j.  Coordinates are already re-centred.
__global__ __launch_bounds__(256) void k_assemble(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, int64_t N, const StepSc *__restrict__ sc,
    const spk_vg *__restrict__ vgs, int covar_flag, double min_vg_val,
    double *__restrict__ A, int64_t ld, int64_t strideA,
    double *__restrict__ Gfull) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int tiles = (n + 63) / 64;
  const int t = blockIdx.x;
  if (t >= tiles * tiles) return;
  const int bi = t % tiles, bj = t / tiles;
  if (bi < bj) return;  // lower block triangle, mirrored on write
  __shared__ double rx[64], ry[64], cx[64], cy[64];
  __shared__ spk_vg vg;
  const int tid = threadIdx.x;
  const int32_t *my = idx + (int64_t)s * N;
  if (tid < 64) {
    const int i = bi * 64 + tid;
    rx[tid] = i < n ? sx[my[i]] : 0.0;
    ry[tid] = i < n ? sy[my[i]] : 0.0;
  } else if (tid < 128) {
    const int j = bj * 64 + tid - 64;
    cx[tid - 64] = j < n ? sx[my[j]] : 0.0;
    cy[tid - 64] = j < n ? sy[my[j]] : 0.0;
  }
  if (tid == 0) vg = vgs[s];
  __syncthreads();
  double *As = A + (int64_t)s * strideA;
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e & 63, c = e >> 6;
    const int i = bi * 64 + r, j = bj * 64 + c;
    if (i >= n || j >= n || (bi == bj && i < j)) continue;
    const double d = dist2(rx[r], ry[r], cx[c], cy[c]);
    const double v = vg_entry(vg, d, covar_flag, min_vg_val);
    As[i + (int64_t)j * ld] = v;
    As[j + (int64_t)i * ld] = v;
    if (Gfull) {
      double *Gs = Gfull + (int64_t)s * strideA;
      Gs[i + (int64_t)j * ld] = v;
      Gs[j + (int64_t)i * ld] = v;
    }
  }
}

// ---- Cholesky ------------------------------------------------------------------
// For OK/EDK the semivariogram matrix G is first shifted to covariance form
// C = c0 - G (c0 = sum of sills for bounded models, otherwise twice the
// largest entry plus the sill sum).  The constant shift leaves the kriging
// weights unchanged because they sum to one, which is also why the
// reference's nugget (added to every entry, pyx:192-201) cancels.  Then
// C = L L^T, right-looking, 32-wide panels, 64x64 trailing tiles.  A failed
// pivot marks the step SPK_ST_FAILED, which interp/steps.py:363-384 answers
// with the nearest station's value.
__global__ __launch_bounds__(256) void k_potrf(double *__restrict__ A,
                                               int64_t ld, int64_t strideA,
                                               StepSc *__restrict__ sc,
                                               const spk_vg *__restrict__ vgs,
                                               int shift) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  double *As = A + (int64_t)s * strideA;
  const int tid = threadIdx.x;
  __shared__ double D[32][33];
  __shared__ double Pa[64][33];
  __shared__ double Pb[64][33];
  __shared__ int fail;
  __shared__ double red[256];
  __shared__ double c0s;

  if (shift) {
    const spk_vg &vg = vgs[s];
    bool bounded = true;
    for (int k = 0; k < vg.n; ++k) {
      const int ty = vg.type[k];
      if (ty == SPK_VG_RNG || ty == SPK_VG_POW || ty == SPK_VG_HOL)
        bounded = false;
    }
    if (bounded) {
      if (tid == 0) c0s = vg.sill_sum;
    } else {
      double mx = 0.0;
      for (int64_t e = tid; e < (int64_t)n * n; e += 256) {
        const int i = (int)(e % n), j = (int)(e / n);
        if (i >= j) mx = fmax(mx, fabs(As[i + (int64_t)j * ld]));
      }
      red[tid] = mx;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
        __syncthreads();
      }
      if (tid == 0) c0s = 2.0 * red[0] + fabs(vg.sill_sum);
    }
    __syncthreads();
    const double c0 = c0s;
    for (int64_t e = tid; e < (int64_t)n * n; e += 256) {
      const int i = (int)(e % n), j = (int)(e / n);
      if (i >= j) As[i + (int64_t)j * ld] = c0 - As[i + (int64_t)j * ld];
    }
    if (tid == 0) sc[s].c0 = c0;
  }
  if (tid == 0) fail = 0;
  __syncthreads();

  for (int k0 = 0; k0 < n; k0 += 32) {
    const int jb = min(32, n - k0);
    for (int e = tid; e < 32 * 32; e += 256) {
      const int i = e & 31, j = e >> 5;
      D[i][j] = (i < jb && j < jb && i >= j)
                    ? As[(k0 + i) + (int64_t)(k0 + j) * ld]
                    : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < jb; ++j) {
      if (tid == 0) {
        const double d = D[j][j];
        if (!(d > 0.0) || !isfinite(d))
          fail = k0 + j + 1;
        else
          D[j][j] = sqrt(d);
      }
      __syncthreads();
      if (fail) break;
      if (tid > j && tid < jb) D[tid][j] /= D[j][j];
      __syncthreads();
      for (int e = tid; e < 32 * 32; e += 256) {
        const int i = e & 31, c = e >> 5;
        if (c > j && i >= c && i < jb) D[i][c] -= D[i][j] * D[c][j];
      }
      __syncthreads();
    }
    if (fail) break;
    for (int e = tid; e < 32 * 32; e += 256) {
      const int i = e & 31, j = e >> 5;
      if (i < jb && j < jb && i >= j)
        As[(k0 + i) + (int64_t)(k0 + j) * ld] = D[i][j];
    }
    // panel: rows r >= k0 + jb solve x L_kk^T = a
    for (int r = k0 + jb + tid; r < n; r += 256) {
      double x[32];
#pragma unroll
      for (int c = 0; c < 32; ++c)
        x[c] = c < jb ? As[r + (int64_t)(k0 + c) * ld] : 0.0;
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        if (c < jb) {
          double v = x[c];
#pragma unroll
          for (int q = 0; q < c; ++q) v -= x[q] * D[c][q];
          x[c] = v / D[c][c];
        }
      }
#pragma unroll
      for (int c = 0; c < 32; ++c)
        if (c < jb) As[r + (int64_t)(k0 + c) * ld] = x[c];
    }
    __syncthreads();
    // trailing update of rows/cols >= m0
    const int m0 = k0 + jb;
    const int mt = n - m0;
    if (mt <= 0) break;
    const int nt = (mt + 63) / 64;
    const int ty = tid >> 4, tx = tid & 15;
    // tiles (bi, bj <= bi) in order; the next tile's panel rows are loaded
    // into registers while this tile is updated (the panel columns are not
    // written by the trailing update)
    double qa[8], qb[8], qc[4][4];
    auto fetch = [&](int bi, int bj) {
      const int i0 = m0 + bi * 64, j0 = m0 + bj * 64;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = tid + q * 256;
        const int r = e & 63, c = e >> 6;
        qa[q] = (i0 + r < n && c < jb) ? As[(i0 + r) + (int64_t)(k0 + c) * ld]
                                       : 0.0;
        qb[q] = (j0 + r < n && c < jb) ? As[(j0 + r) + (int64_t)(k0 + c) * ld]
                                       : 0.0;
      }
      // the tile's own entries (tiles are disjoint)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int i = i0 + ty * 4 + a, j = j0 + tx * 4 + b;
          qc[a][b] = (i < n && j < n && i >= j) ? As[i + (int64_t)j * ld] : 0.0;
        }
    };
    fetch(0, 0);
    for (int bi = 0; bi < nt; ++bi) {
      for (int bj = 0; bj <= bi; ++bj) {
        const int i0 = m0 + bi * 64, j0 = m0 + bj * 64;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int e = tid + q * 256;
          Pa[e & 63][e >> 6] = qa[q];
          Pb[e & 63][e >> 6] = qb[q];
        }
        double cur[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) cur[a][b] = qc[a][b];
        __syncthreads();
        if (bj < bi)
          fetch(bi, bj + 1);
        else if (bi + 1 < nt)
          fetch(bi + 1, 0);
        double acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
        for (int c = 0; c < jb; ++c) {
          double pa[4], pb[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) pa[a] = Pa[ty * 4 + a][c];
#pragma unroll
          for (int b = 0; b < 4; ++b) pb[b] = Pb[tx * 4 + b][c];
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += pa[a] * pb[b];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int i = i0 + ty * 4 + a;
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const int j = j0 + tx * 4 + b;
            if (i < n && j < n && i >= j)
              As[i + (int64_t)j * ld] = cur[a][b] - acc[a][b];
          }
        }
        __syncthreads();
      }
    }
  }
  __syncthreads();
  if (tid == 0 && fail) sc[s].status = SPK_ST_FAILED;
}

// ---- Cholesky over (block, step) workgroups ------------------------------------
// The same factorisation as k_potrf, with the same operations in the same
// order (bit-identical results), spread over many workgroups: k_chol_shift
// forms C = c0 - G (one workgroup per step); then per 32-wide panel k0,
// k_chol_diag factors the panel's diagonal block, k_chol_solve solves 256
// of the rows below it per workgroup, and k_chol_trail updates the
// trailing triangle, one 64 x 64 tile per workgroup.  A failed pivot sets
// SPK_ST_FAILED, after which the step's later launches return at once.
// (Panels of 32 rows only: a panel shorter than 32 is the last one and
// leaves nothing below it.)
__global__ __launch_bounds__(256) void k_chol_shift(double *__restrict__ A,
                                                    int64_t ld,
                                                    int64_t strideA,
                                                    StepSc *__restrict__ sc,
                                                    const spk_vg *__restrict__ vgs) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  double *As = A + (int64_t)s * strideA;
  const int tid = threadIdx.x;
  __shared__ double red[256];
  __shared__ double c0s;
  const spk_vg &vg = vgs[s];
  bool bounded = true;
  for (int k = 0; k < vg.n; ++k) {
    const int ty = vg.type[k];
    if (ty == SPK_VG_RNG || ty == SPK_VG_POW || ty == SPK_VG_HOL)
      bounded = false;
  }
  if (bounded) {
    if (tid == 0) c0s = vg.sill_sum;
  } else {
    double mx = 0.0;
    for (int64_t e = tid; e < (int64_t)n * n; e += 256) {
      const int i = (int)(e % n), j = (int)(e / n);
      if (i >= j) mx = fmax(mx, fabs(As[i + (int64_t)j * ld]));
    }
    red[tid] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
      __syncthreads();
    }
    if (tid == 0) c0s = 2.0 * red[0] + fabs(vg.sill_sum);
  }
  __syncthreads();
  const double c0 = c0s;
  for (int64_t e = tid; e < (int64_t)n * n; e += 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i >= j) As[i + (int64_t)j * ld] = c0 - As[i + (int64_t)j * ld];
  }
  if (tid == 0) sc[s].c0 = c0;
}

// the panel's 32 x 32 diagonal block, factored in LDS and written back;
// grid (1, steps)
__global__ __launch_bounds__(256) void k_chol_diag(double *__restrict__ A,
                                                   int64_t ld, int64_t strideA,
                                                   StepSc *__restrict__ sc,
                                                   int k0) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  if (k0 >= n) return;
  const int jb = min(32, n - k0);
  double *As = A + (int64_t)s * strideA;
  const int tid = threadIdx.x;
  __shared__ double D[32][33];
  __shared__ int fail;
  if (tid == 0) fail = 0;
  for (int e = tid; e < 32 * 32; e += 256) {
    const int i = e & 31, j = e >> 5;
    D[i][j] = (i < jb && j < jb && i >= j)
                  ? As[(k0 + i) + (int64_t)(k0 + j) * ld]
                  : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < jb; ++j) {
    if (tid == 0) {
      const double d = D[j][j];
      if (!(d > 0.0) || !isfinite(d))
        fail = k0 + j + 1;
      else
        D[j][j] = sqrt(d);
    }
    __syncthreads();
    if (fail) break;
    if (tid > j && tid < jb) D[tid][j] /= D[j][j];
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += 256) {
      const int i = e & 31, c = e >> 5;
      if (c > j && i >= c && i < jb) D[i][c] -= D[i][j] * D[c][j];
    }
    __syncthreads();
  }
  if (fail) {
    if (tid == 0) sc[s].status = SPK_ST_FAILED;
    return;
  }
  for (int e = tid; e < 32 * 32; e += 256) {
    const int i = e & 31, j = e >> 5;
    if (i < jb && j < jb && i >= j)
      As[(k0 + i) + (int64_t)(k0 + j) * ld] = D[i][j];
  }
}

// rows r >= k0 + 32 solve x L_kk^T = a, 256 rows per workgroup;
// grid (rows below the panel / 256, steps)
__global__ __launch_bounds__(256) void k_chol_solve(double *__restrict__ A,
                                                    int64_t ld,
                                                    int64_t strideA,
                                                    const StepSc *__restrict__ sc,
                                                    int k0) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  if (k0 + 32 >= n) return;
  const int jb = 32, m0 = k0 + 32;
  const int rbase = m0 + (int)blockIdx.x * 256;
  if (rbase >= n) return;
  double *As = A + (int64_t)s * strideA;
  const int tid = threadIdx.x;
  __shared__ double D[32][33];
  for (int e = tid; e < 32 * 32; e += 256) {
    const int i = e & 31, j = e >> 5;
    D[i][j] = (i >= j) ? As[(k0 + i) + (int64_t)(k0 + j) * ld] : 0.0;
  }
  __syncthreads();
  const int r = rbase + tid;
  if (r < n) {
    double x[32];
#pragma unroll
    for (int c = 0; c < 32; ++c)
      x[c] = c < jb ? As[r + (int64_t)(k0 + c) * ld] : 0.0;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
      if (c < jb) {
        double v = x[c];
#pragma unroll
        for (int q = 0; q < c; ++q) v -= x[q] * D[c][q];
        x[c] = v / D[c][c];
      }
    }
#pragma unroll
    for (int c = 0; c < 32; ++c)
      if (c < jb) As[r + (int64_t)(k0 + c) * ld] = x[c];
  }
}

// grid (tiles of the largest trailing triangle, steps): tile t = (bi, bj),
// bj <= bi, t = bi (bi + 1) / 2 + bj, of the rows / columns >= k0 + 32
__global__ __launch_bounds__(256) void k_chol_trail(double *__restrict__ A,
                                                    int64_t ld,
                                                    int64_t strideA,
                                                    const StepSc *__restrict__ sc,
                                                    int k0) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  if (k0 + 32 >= n) return;
  const int jb = 32, m0 = k0 + 32;
  const int nt = (n - m0 + 63) / 64;
  const int t = blockIdx.x;
  if (t >= nt * (nt + 1) / 2) return;
  int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (bi * (bi + 1) / 2 > t) --bi;
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  const int bj = t - bi * (bi + 1) / 2;
  double *As = A + (int64_t)s * strideA;
  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  __shared__ double Pa[64][33];
  __shared__ double Pb[64][33];
  const int i0 = m0 + bi * 64, j0 = m0 + bj * 64;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int e = tid + q * 256;
    const int r = e & 63, c = e >> 6;
    Pa[r][c] = (i0 + r < n && c < jb) ? As[(i0 + r) + (int64_t)(k0 + c) * ld]
                                      : 0.0;
    Pb[r][c] = (j0 + r < n && c < jb) ? As[(j0 + r) + (int64_t)(k0 + c) * ld]
                                      : 0.0;
  }
  double cur[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = i0 + ty * 4 + a, j = j0 + tx * 4 + b;
      cur[a][b] = (i < n && j < n && i >= j) ? As[i + (int64_t)j * ld] : 0.0;
    }
  __syncthreads();
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int c = 0; c < jb; ++c) {
    double pa[4], pb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) pa[a] = Pa[ty * 4 + a][c];
#pragma unroll
    for (int b = 0; b < 4; ++b) pb[b] = Pb[tx * 4 + b][c];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += pa[a] * pb[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty * 4 + a;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + tx * 4 + b;
      if (i < n && j < n && i >= j)
        As[i + (int64_t)j * ld] = cur[a][b] - acc[a][b];
    }
  }
}

// ---- triangular inverse ----------------------------------------------------------
// X = L^-1 (lower, zeros above the diagonal) by 64 x 64 blocks: the diagonal
// blocks are the inverses of L's, X[r, r] = inv(L_rr), and below them
// X[r, c] = X[r, r] (- sum_{c <= k < r} L[r, k] X[k, c]).  Column block c is
// a chain over r that needs only its own earlier blocks, so the column
// blocks of every step run in parallel: k_trtri_diag inverts the diagonal
// blocks (grid (blocks, steps)), then k_trtri_cols runs one column block's
// chain per workgroup (grid (blocks, steps)).  The n x n region of X is
// fully written (upper part zero).
__global__ __launch_bounds__(256) void k_trtri_diag(
    const double *__restrict__ L, double *__restrict__ X, int64_t ld,
    int64_t stride, const StepSc *__restrict__ sc) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int r0 = blockIdx.x * 64;
  if (r0 >= n) return;
  const int rn = min(64, n - r0);
  const double *Ls = L + (int64_t)s * stride;
  double *Xs = X + (int64_t)s * stride;
  const int tid = threadIdx.x;
  __shared__ double Di[64][65];
  for (int e = tid; e < 64 * 64; e += 256) {
    const int i = e & 63, j = e >> 6;
    Di[i][j] = (i < rn && j < rn && i >= j)
                   ? Ls[(r0 + i) + (int64_t)(r0 + j) * ld]
                   : 0.0;
  }
  __syncthreads();
  // in-place inverse, last column first:
  // X_jj = 1/L_jj, X[i, j] = -X_jj sum_{j<k<=i} X[i, k] L[k, j]
  for (int j = rn - 1; j >= 0; --j) {
    double v = 0.0;
    const double djj = 1.0 / Di[j][j];
    if (tid > j && tid < rn)
      for (int k = j + 1; k <= tid; ++k) v += Di[tid][k] * Di[k][j];
    __syncthreads();
    if (tid > j && tid < rn) Di[tid][j] = -v * djj;
    if (tid == j) Di[j][j] = djj;
    __syncthreads();
  }
  for (int e = tid; e < 64 * 64; e += 256) {
    const int i = e & 63, j = e >> 6;
    if (i < rn && j < rn)
      Xs[(r0 + i) + (int64_t)(r0 + j) * ld] = (i >= j) ? Di[i][j] : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_trtri_cols(
    const double *__restrict__ L, double *__restrict__ X, int64_t ld,
    int64_t stride, const StepSc *__restrict__ sc) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int cb = blockIdx.x;
  const int c0 = cb * 64;
  if (c0 >= n) return;
  const double *Ls = L + (int64_t)s * stride;
  double *Xs = X + (int64_t)s * stride;
  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  __shared__ double Di[64][65];  // X[r, r]
  __shared__ union {
    struct {
      double a[64][33];
      double b[32][65];
    } t;
    double acc[64][65];
  } sm;
  const int nb = (n + 63) / 64;
  const int cn = min(64, n - c0);
  // the strictly-upper blocks of this column block
  for (int64_t e = tid; e < (int64_t)c0 * cn; e += 256) {
    const int i = (int)(e % c0), j = (int)(e / c0);
    Xs[i + (int64_t)(c0 + j) * ld] = 0.0;
  }
  for (int rb = cb + 1; rb < nb; ++rb) {
    const int r0 = rb * 64;
    const int rn = min(64, n - r0);
    for (int e = tid; e < 64 * 64; e += 256) {
      const int i = e & 63, j = e >> 6;
      Di[i][j] = (i < rn && j < rn) ? Xs[(r0 + i) + (int64_t)(r0 + j) * ld]
                                    : 0.0;
    }
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    // acc -= L[r, k] X[k, c] for k in [c0, r0); the next chunk's tiles
    // are loaded into registers while this one is multiplied
    double pa_[8], pb_[8];
    auto fetch = [&](int k0) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = tid + q * 256;
        const int r = e & 63, k = e >> 6;
        pa_[q] = (r < rn) ? Ls[(r0 + r) + (int64_t)(k0 + k) * ld] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = tid + q * 256;
        const int k = e & 31, c = e >> 5;
        pb_[q] = (c0 + c < n) ? Xs[(k0 + k) + (int64_t)(c0 + c) * ld] : 0.0;
      }
    };
    fetch(c0);
    for (int k0 = c0; k0 < r0; k0 += 32) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = tid + q * 256;
        sm.t.a[e & 63][e >> 6] = pa_[q];
        sm.t.b[e & 31][e >> 5] = pb_[q];
      }
      __syncthreads();
      if (k0 + 32 < r0) fetch(k0 + 32);
      for (int k = 0; k < 32; ++k) {
        double pa[4], pb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) pa[a] = sm.t.a[ty * 4 + a][k];
#pragma unroll
        for (int b = 0; b < 4; ++b) pb[b] = sm.t.b[k][tx * 4 + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] -= pa[a] * pb[b];
      }
      __syncthreads();
    }
    // X[r, c] = X[r, r] acc
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) sm.acc[ty * 4 + a][tx * 4 + b] = acc[a][b];
    __syncthreads();
    double out[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) out[a][b] = 0.0;
    for (int k = 0; k < 64; ++k) {
      double pa[4], pb[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) pa[a] = Di[ty * 4 + a][k];
#pragma unroll
      for (int b = 0; b < 4; ++b) pb[b] = sm.acc[k][tx * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) out[a][b] += pa[a] * pb[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = r0 + ty * 4 + a;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = c0 + tx * 4 + b;
        if (i < n && j < n) Xs[i + (int64_t)j * ld] = out[a][b];
      }
    }
    __syncthreads();
  }
}

// ---- dual-form vectors (batched path) -----------------------------------------
// With Li = L^-1:  C^-1 x = Li^T (Li x).
//   OK : w = C^-1 1, b = 1^T w, az = C^-1 z, sz = 1^T az,
//        alpha = az - w sz / b, a0 = sz / b  ->  est = k.alpha + a0
//        (the Schur complement of the Lagrange row, interp/steps.py:212-216)
//   SK : alpha = C^-1 z, w = C^-1 1 (sum of weights for the isclose test of
//        interp/steps.py:418)
//   EDK: F = [1 | S], U = Li F, W = Li^T U, B = U^T U, h = B^-1 U^T (Li z),
//        alpha = C^-1 z - W h  ->  est = k.alpha + h.[1; s_cell]
// alpha and h solve the bordered system [C F; F^T 0] [alpha; h] = [z; 0].
// Two steps of iterative refinement follow, with FP64 residuals against
// the covariances the Cholesky factorised (recomputed from the
// coordinates): r1 = z - C alpha - F h, r2 = -F^T alpha, rw = 1 - C w;
// dh = B^-1 (F^T C^-1 r1 - r2), dalpha = C^-1 r1 - W dh, dw = C^-1 rw.
// The corrections are added; the second step's are kept: k.dalpha + f.dh
// is the last change of a cell's estimate, of the size of the error left,
// which the cell kernel uses as its error indicator.
// Vectors are stored per step in V[s][0:ld] (alpha), V[s][ld:2 ld] (w),
// V[s][2 ld:3 ld] (dalpha) and V[s][3 ld:4 ld] (dw); dh in the StepSc.
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the 256-thread block (all threads receive the result).
__device__ inline double block_sum(double v, double *red4) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = v;
  __syncthreads();
  return red4[0] + red4[1] + red4[2] + red4[3];
}

__device__ void small_inverse(double *B, int p, double *Binv) {
  // Gauss-Jordan with partial pivoting on a p x p (p <= kMaxP) row-major.
  double a[kMaxP][2 * kMaxP];
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < 2 * p; ++j)
      a[i][j] = j < p ? B[i * p + j] : (j - p == i ? 1.0 : 0.0);
  for (int c = 0; c < p; ++c) {
    int piv = c;
    for (int r = c + 1; r < p; ++r)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    if (piv != c)
      for (int j = 0; j < 2 * p; ++j) {
        double t = a[c][j];
        a[c][j] = a[piv][j];
        a[piv][j] = t;
      }
    const double d = a[c][c];
    for (int j = 0; j < 2 * p; ++j) a[c][j] /= d;
    for (int r = 0; r < p; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int j = 0; j < 2 * p; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < p; ++j) Binv[i * p + j] = a[i][p + j];
}

// Y[:, r] = Li X[:, r] (r < nr), X and Y column-major with leading dim ld.
// The k loops are unrolled so that a thread's loads of the next entries are
// in flight while it adds up the current ones (in the same order).
__device__ void tri_lower_mv(const double *__restrict__ Li, int64_t ld, int n,
                             const double *__restrict__ X,
                             double *__restrict__ Y, int nr) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double acc[kMaxP + 1];
    for (int r = 0; r < nr; ++r) acc[r] = 0.0;
#pragma unroll 8
    for (int k = 0; k <= i; ++k) {
      const double l = Li[i + (int64_t)k * ld];
      for (int r = 0; r < nr; ++r) acc[r] += l * X[k + (int64_t)r * ld];
    }
    for (int r = 0; r < nr; ++r) Y[i + (int64_t)r * ld] = acc[r];
  }
}

// Y[:, r] = Li^T X[:, r]: one wave per output entry, lanes over rows.
__device__ void tri_lower_tmv(const double *__restrict__ Li, int64_t ld,
                              int n, const double *__restrict__ X,
                              double *__restrict__ Y, int nr) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  for (int k = wave; k < n; k += nw) {
    double acc[kMaxP + 1];
    for (int r = 0; r < nr; ++r) acc[r] = 0.0;
#pragma unroll 4
    for (int i = k + lane; i < n; i += 64) {
      const double l = Li[i + (int64_t)k * ld];
      for (int r = 0; r < nr; ++r) acc[r] += l * X[i + (int64_t)r * ld];
    }
    for (int r = 0; r < nr; ++r) {
      const double v = wave_sum(acc[r]);
      if (lane == 0) Y[k + (int64_t)r * ld] = v;
    }
  }
}

// ---- condition estimate over (block, step) workgroups --------------------------
// kappa = lambda_max(C) lambda_max(C^-1) by power iteration on C = Lc Lc^T
// (which 0, Lc the Cholesky factor) and C^-1 = Li^T Li (which 1), 12 rounds
// each, Rayleigh quotients; the iterate x and u = T^T x or T x are the first
// two scratch columns of the step's T0 (X and U of k_dual).  Launch order
// per which: k_kap_start, then 12 x (k_kap_norm, product into u,
// k_kap_rq, product back into x); then k_kap_flags.  The products run over
// (row blocks, steps), one row per thread (k_kap_lmv, y = T x) or one
// column per wave (k_kap_umv, y = T^T x), with the sums in the order of
// tri_lower_mv / tri_lower_tmv, and the norms as one workgroup per step with
// block_sum, so the estimate is the one a single workgroup per step gives.
__device__ inline double *kap_x(double *T0, int64_t ldt, int s) {
  return T0 + (int64_t)s * ldt * kT0Cols;
}
__device__ inline double *kap_u(double *T0, int64_t ldt, int s) {
  return kap_x(T0, ldt, s) + ldt * (kMaxP + 1);
}

// grid (steps)
__global__ __launch_bounds__(256) void k_kap_start(const StepSc *__restrict__ sc,
                                                   double *__restrict__ T0,
                                                   int64_t ldt, int which) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  double *X = kap_x(T0, ldt, s);
  for (int i = threadIdx.x; i < n; i += 256) {
    // fixed start vector with both smooth and oscillating parts
    const unsigned h = (unsigned)i * 2654435761u;
    X[i] = 1.0 + 0.5 * (double)((h >> 8) & 0xffff) / 65536.0 -
           (which ? 0.5 * (double)(i & 1) : 0.0);
  }
}

// x /= |x|; grid (steps)
__global__ __launch_bounds__(256) void k_kap_norm(const StepSc *__restrict__ sc,
                                                  double *__restrict__ T0,
                                                  int64_t ldt) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  double *X = kap_x(T0, ldt, s);
  __shared__ double red4[4];
  const int tid = threadIdx.x;
  double nv = 0.0;
  for (int i = tid; i < n; i += 256) nv += X[i] * X[i];
  nv = block_sum(nv, red4);
  const double inv = 1.0 / sqrt(nv);
  for (int i = tid; i < n; i += 256) X[i] *= inv;
}

// lam[2 s + which] = |u|^2; grid (steps)
__global__ __launch_bounds__(256) void k_kap_rq(const StepSc *__restrict__ sc,
                                                double *__restrict__ T0,
                                                int64_t ldt,
                                                double *__restrict__ lam,
                                                int which) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const double *U = kap_u(T0, ldt, s);
  __shared__ double red4[4];
  double q = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) q += U[i] * U[i];
  q = block_sum(q, red4);
  if (threadIdx.x == 0) lam[2 * s + which] = q;
}

// y = T v (T lower, column-major, ld), v and y the step's x / u columns
// (to_u: v = x, y = u; else v = u, y = x); grid (ceil(n_max / 256), steps)
__global__ __launch_bounds__(256) void k_kap_lmv(const double *__restrict__ T,
                                                 int64_t ld, int64_t stride,
                                                 const StepSc *__restrict__ sc,
                                                 double *__restrict__ T0,
                                                 int64_t ldt, int to_u) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double *Ts = T + (int64_t)s * stride;
  const double *v = to_u ? kap_x(T0, ldt, s) : kap_u(T0, ldt, s);
  double *y = to_u ? kap_u(T0, ldt, s) : kap_x(T0, ldt, s);
  double acc = 0.0;
#pragma unroll 8
  for (int k = 0; k <= i; ++k) {
    const double l = Ts[i + (int64_t)k * ld];
    acc += l * v[k];
  }
  y[i] = acc;
}

// y = T^T v, one wave per entry k (lanes over rows i >= k);
// grid (ceil(n_max / 4), steps)
__global__ __launch_bounds__(256) void k_kap_umv(const double *__restrict__ T,
                                                 int64_t ld, int64_t stride,
                                                 const StepSc *__restrict__ sc,
                                                 double *__restrict__ T0,
                                                 int64_t ldt, int to_u) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n) return;
  const double *Ts = T + (int64_t)s * stride;
  const double *v = to_u ? kap_x(T0, ldt, s) : kap_u(T0, ldt, s);
  double *y = to_u ? kap_u(T0, ldt, s) : kap_x(T0, ldt, s);
  double acc = 0.0;
#pragma unroll 4
  for (int i = k + lane; i < n; i += 64) {
    const double l = Ts[i + (int64_t)k * ld];
    acc += l * v[i];
  }
  const double r = wave_sum(acc);
  if (lane == 0) y[k] = r;
}

// kappa and the step's precision flags; grid (steps), one thread
__global__ void k_kap_flags(StepSc *__restrict__ sc,
                            const double *__restrict__ lam,
                            const spk_vg *__restrict__ vgs,
                            double fp64var_thr, double fp64var_thr_gau,
                            double int8var_thr) {
  const int s = blockIdx.x;
  if (threadIdx.x != 0 || sc[s].status) return;
  const double kap = lam[2 * s] * lam[2 * s + 1];
  sc[s].kappa = kap;
  int fl = 0;
  if (!(kap <= 1e12)) fl |= SPK_ST_ILLCOND;
  bool gau = false;
  for (int k = 0; k < vgs[s].n; ++k) gau |= vgs[s].type[k] == SPK_VG_GAU;
  // above the FP32 limit: exact int8 plane products up to the int8
  // limit, FP64 beyond it
  int planes = 0;
  if (!(kap <= (gau ? fp64var_thr_gau : fp64var_thr))) {
    if (kap <= int8var_thr)
      planes = kap <= kInt8Deep ? 5 : 6;
    else
      fl |= SPK_ST_FP64VAR;
  }
  if (planes) fl |= SPK_ST_INT8VAR;
  sc[s].flags = fl;
  sc[s].planes = planes;
}

// ---- dual vectors and refinement over (block, step) workgroups ----------------
// Per step, with T0 holding X, U (kMaxP + 1 columns each), R2, Y2, Z2 (two
// each) and the per-step scratch ds (kDualSc doubles: G, Bi, hcoef, hcur,
// dhs):
//   k_dual_rhs   X = [F z]                                    (steps)
//   k_dual_lmv   U = Li X            (p + 1 columns)          (row blocks, steps)
//   k_dual_umv   X = Li^T U = C^-1 [F z]                      (column groups, steps)
//   k_dual_mid   G = U^T U, Bi = (U_F^T U_F)^-1, h, alpha, w  (steps)
//   then twice: k_dual_resid (r1 = z - C alpha - F h, rw = 1 - C w;
//   row blocks, steps), k_dual_lmv / k_dual_umv on [r1 rw], k_dual_upd
//   (the corrections of h, alpha and w; steps); then k_dual_fin (steps).
// Every sum runs in the order of the one-workgroup-per-step form (a row per
// thread, a column per wave, one workgroup per step for the small
// reductions), so the results are the ones that form gives.
constexpr int kDsG = 0;                                  // (p+1)^2
constexpr int kDsBi = kDsG + (kMaxP + 1) * (kMaxP + 1);  // p^2
constexpr int kDsH = kDsBi + kMaxP * kMaxP;              // hcoef
constexpr int kDsHc = kDsH + kMaxP;                      // hcur
constexpr int kDsDh = kDsHc + kMaxP;                     // dhs

struct DualPtrs {
  double *X, *U, *R2, *Y2, *Z2, *alpha, *w, *dalpha, *dw, *ds;
};
__device__ inline DualPtrs dual_ptrs(double *T0, int64_t ldt, double *V,
                                     int64_t ldv, double *DS, int s) {
  DualPtrs d;
  d.X = T0 + (int64_t)s * ldt * kT0Cols;
  d.U = d.X + ldt * (kMaxP + 1);
  d.R2 = d.U + ldt * (kMaxP + 1);  // residuals r1, rw
  d.Y2 = d.R2 + 2 * ldt;           // Li [r1 rw]
  d.Z2 = d.Y2 + 2 * ldt;           // C^-1 [r1 rw]
  d.alpha = V + (int64_t)s * ldv;
  d.w = d.alpha + ldt;
  d.dalpha = d.w + ldt;
  d.dw = d.dalpha + ldt;
  d.ds = DS + (int64_t)s * kDualSc;
  return d;
}

// X columns: 0..p-1 = F (ones, drifts), p = z; grid (steps)
__global__ __launch_bounds__(256) void k_dual_rhs(
    const StepSc *__restrict__ sc, const double *__restrict__ zact,
    int64_t ldz, const int32_t *__restrict__ idx, int64_t N,
    const double *__restrict__ stn_drift, int m, int method,
    double *__restrict__ T0, int64_t ldt) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const double *z = zact + (int64_t)s * ldz;
  const int32_t *my = idx + (int64_t)s * N;
  double *X = T0 + (int64_t)s * ldt * kT0Cols;
  const int p = (method == SPK_M_EDK) ? 1 + m : 1;
  for (int i = threadIdx.x; i < n; i += 256) {
    X[i] = 1.0;
    if (method == SPK_M_EDK)
      for (int k = 0; k < m; ++k)
        X[i + (int64_t)(1 + k) * ldt] = stn_drift[(int64_t)my[i] * m + k];
    X[i + (int64_t)p * ldt] = z[i];
  }
}

// column offsets in T0 of the products' operands (col: 0 X, 1 U, 2 R2,
// 3 Y2, 4 Z2)
__device__ inline double *t0_col(double *T0, int64_t ldt, int s, int col) {
  double *X = T0 + (int64_t)s * ldt * kT0Cols;
  switch (col) {
    case 0: return X;
    case 1: return X + ldt * (kMaxP + 1);
    case 2: return X + 2 * ldt * (kMaxP + 1);
    case 3: return X + 2 * ldt * (kMaxP + 1) + 2 * ldt;
    default: return X + 2 * ldt * (kMaxP + 1) + 4 * ldt;
  }
}

// Y[:, r] = Li X[:, r] (r < nr), one row per thread; grid (ceil(n_max /
// 256), steps)
__global__ __launch_bounds__(256) void k_dual_lmv(
    const double *__restrict__ Li, int64_t ld, int64_t stride,
    const StepSc *__restrict__ sc, double *__restrict__ T0, int64_t ldt,
    int src, int dst, int nr) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double *L = Li + (int64_t)s * stride;
  const double *X = t0_col(T0, ldt, s, src);
  double *Y = t0_col(T0, ldt, s, dst);
  double acc[kMaxP + 1];
  for (int r = 0; r < nr; ++r) acc[r] = 0.0;
#pragma unroll 8
  for (int k = 0; k <= i; ++k) {
    const double l = L[i + (int64_t)k * ld];
    for (int r = 0; r < nr; ++r) acc[r] += l * X[k + (int64_t)r * ldt];
  }
  for (int r = 0; r < nr; ++r) Y[i + (int64_t)r * ldt] = acc[r];
}

// Y[:, r] = Li^T X[:, r], one wave per output entry (lanes over rows);
// grid (ceil(n_max / 4), steps)
__global__ __launch_bounds__(256) void k_dual_umv(
    const double *__restrict__ Li, int64_t ld, int64_t stride,
    const StepSc *__restrict__ sc, double *__restrict__ T0, int64_t ldt,
    int src, int dst, int nr) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n) return;
  const double *L = Li + (int64_t)s * stride;
  const double *X = t0_col(T0, ldt, s, src);
  double *Y = t0_col(T0, ldt, s, dst);
  double acc[kMaxP + 1];
  for (int r = 0; r < nr; ++r) acc[r] = 0.0;
#pragma unroll 4
  for (int i = k + lane; i < n; i += 64) {
    const double l = L[i + (int64_t)k * ld];
    for (int r = 0; r < nr; ++r) acc[r] += l * X[i + (int64_t)r * ldt];
  }
  for (int r = 0; r < nr; ++r) {
    const double v = wave_sum(acc[r]);
    if (lane == 0) Y[k + (int64_t)r * ldt] = v;
  }
}

// G = U^T U (a, b <= p: B = U_F^T U_F, g = U_F^T U_z), Bi = B^-1, the
// Schur coefficients h = Bi g, alpha = C^-1 z - C^-1 F h, w = C^-1 1;
// grid (steps)
__global__ __launch_bounds__(256) void k_dual_mid(
    const StepSc *__restrict__ sc, int m, int method, double *__restrict__ T0,
    int64_t ldt, double *__restrict__ V, int64_t ldv, double *__restrict__ DS) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const DualPtrs d = dual_ptrs(T0, ldt, V, ldv, DS, s);
  const int p = (method == SPK_M_EDK) ? 1 + m : 1;
  const int tid = threadIdx.x;
  __shared__ double G[(kMaxP + 1) * (kMaxP + 1)];
  for (int e = tid >> 6; e < (p + 1) * (p + 1); e += 4) {
    const int a = e / (p + 1), b = e % (p + 1);
    double acc = 0.0;
    for (int i = (tid & 63); i < n; i += 64)
      acc += d.U[i + (int64_t)a * ldt] * d.U[i + (int64_t)b * ldt];
    acc = wave_sum(acc);
    if ((tid & 63) == 0) G[e] = acc;
  }
  __syncthreads();
  __shared__ double hcoef[kMaxP];
  __shared__ double Bi[kMaxP * kMaxP];
  if (tid == 0) {
    double B[kMaxP * kMaxP];
    for (int a = 0; a < p; ++a)
      for (int b = 0; b < p; ++b) B[a * p + b] = G[a * (p + 1) + b];
    if (method == SPK_M_SK) {
      // no constraint rows: alpha = C^-1 z, w = C^-1 1
      hcoef[0] = 0.0;
    } else {
      small_inverse(B, p, Bi);
      for (int a = 0; a < p; ++a) {
        double v = 0.0;
        for (int b = 0; b < p; ++b) v += Bi[a * p + b] * G[b * (p + 1) + p];
        hcoef[a] = v;
      }
    }
    for (int e = 0; e < (p + 1) * (p + 1); ++e) d.ds[kDsG + e] = G[e];
    if (method != SPK_M_SK)
      for (int e = 0; e < p * p; ++e) d.ds[kDsBi + e] = Bi[e];
    for (int a = 0; a < kMaxP; ++a) {
      d.ds[kDsH + a] = a < p ? hcoef[a] : 0.0;
      d.ds[kDsHc + a] = a < p ? hcoef[a] : 0.0;
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += 256) {
    double a = d.X[i + (int64_t)p * ldt];
    if (method != SPK_M_SK)
      for (int k = 0; k < p; ++k) a -= d.X[i + (int64_t)k * ldt] * hcoef[k];
    d.alpha[i] = a;
    d.w[i] = d.X[i];
  }
}

// ---- iterative refinement with FP64 residuals ------------------------------
// r1 = z - C alpha - F h, rw = 1 - C w, with C the covariances the
// Cholesky factorised: c0 - gamma (OK, EDK) or the covariance sum (SK),
// recomputed from the re-centred coordinates with the same arithmetic.
// Two steps: the first removes the error of the explicit-inverse solve,
// the second's correction is of the size of the error left, which the
// cell kernel turns into each cell's error indicator.
// One row per thread; grid (ceil(n_max / 256), steps).
__global__ __launch_bounds__(256) void k_dual_resid(
    const StepSc *__restrict__ sc, const double *__restrict__ zact,
    int64_t ldz, const int32_t *__restrict__ idx, int64_t N,
    const double *__restrict__ stn_drift, int m, int method,
    double *__restrict__ T0, int64_t ldt, double *__restrict__ V, int64_t ldv,
    double *__restrict__ DS, const spk_vg *__restrict__ vgs, double min_vg_val,
    const double *__restrict__ sx, const double *__restrict__ sy) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const int i0 = blockIdx.x * 256;
  if (i0 >= n) return;
  const DualPtrs d = dual_ptrs(T0, ldt, V, ldv, DS, s);
  const double *z = zact + (int64_t)s * ldz;
  const int32_t *my = idx + (int64_t)s * N;
  const int p = (method == SPK_M_EDK) ? 1 + m : 1;
  const int tid = threadIdx.x;
  __shared__ double stx[256], sty[256], sal[256], sw[256];
  __shared__ spk_vg vg;
  if (tid == 0) vg = vgs[s];
  const bool sk = method == SPK_M_SK;
  const double c0 = sc[s].c0;
  const int i = i0 + tid;
  double xi = 0.0, yi = 0.0;
  if (i < n) {
    xi = sx[my[i]];
    yi = sy[my[i]];
  }
  double ca = 0.0, cw = 0.0;
  for (int j0 = 0; j0 < n; j0 += 256) {
    __syncthreads();
    if (j0 + tid < n) {
      const int g = my[j0 + tid];
      stx[tid] = sx[g];
      sty[tid] = sy[g];
      sal[tid] = d.alpha[j0 + tid];
      sw[tid] = d.w[j0 + tid];
    }
    __syncthreads();
    const int cnt = min(256, n - j0);
    if (i < n)
      for (int j = 0; j < cnt; ++j) {
        const double dd = dist2(xi, yi, stx[j], sty[j]);
        const double e = vg_entry(vg, dd, sk ? 1 : 0, min_vg_val);
        const double c = sk ? e : c0 - e;
        ca += c * sal[j];
        cw += c * sw[j];
      }
  }
  if (i < n) {
    double r1 = z[i] - ca;
    if (!sk) {
      r1 -= d.ds[kDsHc + 0];
      for (int k = 1; k < p; ++k)
        r1 -= stn_drift[(int64_t)my[i] * m + (k - 1)] * d.ds[kDsHc + k];
    }
    d.R2[i] = r1;
    d.R2[i + ldt] = 1.0 - cw;
  }
}

// gF[a] = U_F[:, a] . (Li r1) = F^T C^-1 r1, r2[a] = -F[:, a] . alpha, the
// corrections dh = Bi (gF - r2), alpha += Z2[:, 0] - C^-1 F dh,
// w += Z2[:, 1]; grid (steps)
__global__ __launch_bounds__(256) void k_dual_upd(
    const StepSc *__restrict__ sc, const int32_t *__restrict__ idx, int64_t N,
    const double *__restrict__ stn_drift, int m, int method,
    double *__restrict__ T0, int64_t ldt, double *__restrict__ V, int64_t ldv,
    double *__restrict__ DS) {
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const DualPtrs d = dual_ptrs(T0, ldt, V, ldv, DS, s);
  const int32_t *my = idx + (int64_t)s * N;
  const int p = (method == SPK_M_EDK) ? 1 + m : 1;
  const int tid = threadIdx.x;
  const bool sk = method == SPK_M_SK;
  __shared__ double gF[kMaxP], r2s[kMaxP], dhs[kMaxP];
  if (!sk) {
    for (int a = tid >> 6; a < 2 * p; a += 4) {
      double acc = 0.0;
      const int col = a < p ? a : a - p;
      for (int i = (tid & 63); i < n; i += 64) {
        if (a < p) {
          acc += d.U[i + (int64_t)col * ldt] * d.Y2[i];
        } else {
          const double f =
              col == 0 ? 1.0 : stn_drift[(int64_t)my[i] * m + (col - 1)];
          acc -= f * d.alpha[i];
        }
      }
      acc = wave_sum(acc);
      if ((tid & 63) == 0) {
        if (a < p)
          gF[col] = acc;
        else
          r2s[col] = acc;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    for (int a = 0; a < kMaxP; ++a) dhs[a] = 0.0;
    if (!sk)
      for (int a = 0; a < p; ++a) {
        double v = 0.0;
        for (int b = 0; b < p; ++b)
          v += d.ds[kDsBi + a * p + b] * (gF[b] - r2s[b]);
        dhs[a] = v;
      }
  }
  __syncthreads();
  for (int i = tid; i < n; i += 256) {
    double da = d.Z2[i];
    if (!sk)
      for (int k = 0; k < p; ++k) da -= d.X[i + (int64_t)k * ldt] * dhs[k];
    const double dwi = d.Z2[i + ldt];
    d.alpha[i] += da;
    d.w[i] += dwi;
    d.dalpha[i] = da;
    d.dw[i] = dwi;
  }
  if (tid < p) d.ds[kDsHc + tid] += dhs[tid];
  if (tid < kMaxP) d.ds[kDsDh + tid] = dhs[tid];
}

// the step record's dual-form terms; grid (steps)
__global__ void k_dual_fin(StepSc *__restrict__ sc, int m, int method,
                           const double *__restrict__ DS) {
  const int s = blockIdx.x;
  if (threadIdx.x != 0 || sc[s].status) return;
  const double *ds = DS + (int64_t)s * kDualSc;
  const int p = (method == SPK_M_EDK) ? 1 + m : 1;
  const bool sk = method == SPK_M_SK;
  StepSc &o = sc[s];
  o.b = ds[kDsG + 0];  // 1^T C^-1 1
  for (int a = 0; a < p; ++a) {
    o.h[a] = ds[kDsHc + a];
    o.dh[a] = ds[kDsDh + a];
  }
  o.a0 = sk ? 0.0 : ds[kDsHc + 0];
}

// ---- explicit K^-1 for the plain path -------------------------------------------
// Cinv = Li^T Li; W = Cinv F; B = F^T W; Q = W B^-1; P = Cinv - Q W^T.
// The reference's K = [[G, F], [F^T, 0]] with G = c0 11^T - C has the
// inverse [[-P, Q], [Q^T, B^-1 - c0 e0 e0^T]], written to M (ld rows).
// SK has no constraint rows: M = Cinv.
__global__ __launch_bounds__(256) void k_minv(
    const double *__restrict__ Li, int64_t ld, int64_t stride,
    StepSc *__restrict__ sc, const int32_t *__restrict__ idx, int64_t N,
    const double *__restrict__ stn_drift, int m, int method,
    double *__restrict__ M, double *__restrict__ T0,
    const spk_vg *__restrict__ vgs, double min_vg_val) {
  const int64_t ldt = ld;
  const int s = blockIdx.x;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const double *L = Li + (int64_t)s * stride;
  double *Ms = M + (int64_t)s * stride;
  const int32_t *my = idx + (int64_t)s * N;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  __shared__ double Ta[32][65];
  __shared__ double Tb[32][65];
  const int nb = (n + 63) / 64;
  // Cinv tiles (bi >= bj): sum over k >= bi*64 of L[k, i] L[k, j]
  for (int bi = 0; bi < nb; ++bi) {
    for (int bj = 0; bj <= bi; ++bj) {
      double acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
      for (int k0 = bi * 64; k0 < n; k0 += 32) {
        for (int e = tid; e < 32 * 64; e += 256) {
          const int k = e & 31, c = e >> 5;
          const int kk = k0 + k;
          Ta[k][c] = (kk < n && bi * 64 + c < n)
                         ? L[kk + (int64_t)(bi * 64 + c) * ld]
                         : 0.0;
          Tb[k][c] = (kk < n && bj * 64 + c < n)
                         ? L[kk + (int64_t)(bj * 64 + c) * ld]
                         : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < 32; ++k) {
          double pa[4], pb[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) pa[a] = Ta[k][ty * 4 + a];
#pragma unroll
          for (int b = 0; b < 4; ++b) pb[b] = Tb[k][tx * 4 + b];
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += pa[a] * pb[b];
        }
        __syncthreads();
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int i = bi * 64 + ty * 4 + a;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int j = bj * 64 + tx * 4 + b;
          if (i < n && j < n) {
            Ms[i + (int64_t)j * ld] = acc[a][b];
            Ms[j + (int64_t)i * ld] = acc[a][b];
          }
        }
      }
    }
  }
  __syncthreads();
  {
    __shared__ double red[256];
    double mx = 0.0;
    for (int i = tid; i < n; i += 256) mx = fmax(mx, Ms[i + (int64_t)i * ld]);
    red[tid] = mx;
    __syncthreads();
    for (int w2 = 128; w2 > 0; w2 >>= 1) {
      if (tid < w2) red[tid] = fmax(red[tid], red[tid + w2]);
      __syncthreads();
    }
    if (tid == 0) {
      const spk_vg &vg = vgs[s];
      const double cii = (method == SPK_M_SK)
                             ? vg_entry(vg, 0.0, 1, min_vg_val)
                             : sc[s].c0 - vg_entry(vg, 0.0, 0, min_vg_val);
      const double kap = red[0] * fabs(cii);
      sc[s].kappa = kap;
      sc[s].flags = (kap <= 1e12) ? 0 : SPK_ST_ILLCOND;
    }
    __syncthreads();
  }
  if (method == SPK_M_SK) return;
  const int p = (method == SPK_M_EDK) ? 1 + m : 1;
  // per-step work area: kT0Cols columns of ld (W, then Q)
  double *W = T0 + (int64_t)s * ldt * kT0Cols;
  // W = Cinv F
  for (int i = tid; i < n; i += 256) {
    double acc[kMaxP];
    for (int r = 0; r < p; ++r) acc[r] = 0.0;
    for (int k = 0; k < n; ++k) {
      const double c = Ms[i + (int64_t)k * ld];
      acc[0] += c;
      for (int r = 1; r < p; ++r)
        acc[r] += c * stn_drift[(int64_t)my[k] * m + (r - 1)];
    }
    for (int r = 0; r < p; ++r) W[i + (int64_t)r * ldt] = acc[r];
  }
  __syncthreads();
  __shared__ double Bs[kMaxP * kMaxP], Bi[kMaxP * kMaxP];
  for (int e = tid >> 6; e < p * p; e += 4) {
    const int a = e / p, b = e % p;
    double acc = 0.0;
    for (int i = (tid & 63); i < n; i += 64) {
      const double fa = a == 0 ? 1.0 : stn_drift[(int64_t)my[i] * m + a - 1];
      acc += fa * W[i + (int64_t)b * ldt];
    }
    acc = wave_sum(acc);
    if ((tid & 63) == 0) Bs[e] = acc;
  }
  __syncthreads();
  if (tid == 0) small_inverse(Bs, p, Bi);
  __syncthreads();
  // Q = W B^-1 (into columns p..2p-1 of the work array)
  double *Q = W + (int64_t)p * ldt;
  for (int i = tid; i < n; i += 256)
    for (int a = 0; a < p; ++a) {
      double v = 0.0;
      for (int b = 0; b < p; ++b) v += W[i + (int64_t)b * ldt] * Bi[b * p + a];
      Q[i + (int64_t)a * ldt] = v;
    }
  __syncthreads();
  const double c0 = sc[s].c0;
  for (int64_t e = tid; e < (int64_t)n * n; e += 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    double v = Ms[i + (int64_t)j * ld];
    for (int a = 0; a < p; ++a)
      v -= Q[i + (int64_t)a * ldt] * W[j + (int64_t)a * ldt];
    Ms[i + (int64_t)j * ld] = -v;
  }
  for (int i = tid; i < n; i += 256)
    for (int a = 0; a < p; ++a) {
      const double q = Q[i + (int64_t)a * ldt];
      Ms[i + (int64_t)(n + a) * ld] = q;
      Ms[(n + a) + (int64_t)i * ld] = q;
    }
  if (tid < p * p) {
    const int a = tid / p, b = tid % p;
    Ms[(n + a) + (int64_t)(n + b) * ld] =
        Bi[a * p + b] - ((a == 0 && b == 0) ? c0 : 0.0);
  }
}

// Li as float (column-major) for the FP32 contraction.
__global__ void k_cast_tri(const double *__restrict__ Li, float *__restrict__ L32,
                           int64_t ld, int64_t stride,
                           const StepSc *__restrict__ sc) {
  const int s = blockIdx.y;
  // only the steps of the FP32 contraction read L32
  if (sc[s].status || (sc[s].flags & (SPK_ST_FP64VAR | SPK_ST_INT8VAR)))
    return;
  const int n = sc[s].n;
  const int64_t tot = (int64_t)n * n;
  const double *src = Li + (int64_t)s * stride;
  float *dst = L32 + (int64_t)s * stride;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    dst[i + (int64_t)j * ld] = (float)src[i + (int64_t)j * ld];
  }
}

}  // namespace spk
