// Pseudo-inverse fallback for steps whose covariance block does not
// factorise by Cholesky (indefinite, e.g. a bounded linear model over long
// distances in 2-D, or numerically singular, e.g. Gaussian models with long
// ranges).  The reference inverts every step's bordered system with
// numpy.linalg.pinv (interp/steps.py:350-371) and contracts each cell's
// right-hand side with it (interp/steps.py:403-435); for such steps this
// file builds the same bordered K, forms pinv(K) from a symmetric
// eigendecomposition (rocSOLVER dsyevd: for a symmetric matrix the singular
// values are the |eigenvalues| and the singular vectors the eigenvectors),
// with numpy's cutoff (singular values not above 1e-15 x the largest are
// dropped), and leaves it where the plain FP64 contraction reads its K^-1.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "spk_krige.h"

namespace spk {

// The reference's bordered system of one step, column-major with leading
// dimension ld (interp/steps.py:181-234): OK [[G, 1], [1^T, 0]], EDK
// [[G, 1, S], [1^T, 0, 0], [S^T, 0, 0]], SK the covariance matrix, with
// G / Cov the entries fill_vg_var_arr gives (pyx:146-226) at the distances
// of the step's active stations, zero-padded to Pw x Pw (Pw >= n + p).
// Grid (tiles, tiles) of 64 x 64 tiles covering Pw.
__global__ __launch_bounds__(256) void k_pinv_assemble(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, const StepSc *__restrict__ sc,
    const spk_vg *__restrict__ vgp, int method, int m,
    const double *__restrict__ stn_drift, double min_vg_val,
    double *__restrict__ K, int64_t ld, int Pw) {
  const int n = sc->n;
  const int p = method == SPK_M_OK ? 1 : method == SPK_M_EDK ? 1 + m : 0;
  const int P = n + p;
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  if (i0 >= Pw || j0 >= Pw) return;
  __shared__ double rx[64], ry[64], cx[64], cy[64];
  __shared__ spk_vg vg;
  const int tid = threadIdx.x;
  if (tid < 64) {
    const int i = i0 + tid;
    rx[tid] = i < n ? sx[idx[i]] : 0.0;
    ry[tid] = i < n ? sy[idx[i]] : 0.0;
  } else if (tid < 128) {
    const int j = j0 + tid - 64;
    cx[tid - 64] = j < n ? sx[idx[j]] : 0.0;
    cy[tid - 64] = j < n ? sy[idx[j]] : 0.0;
  }
  if (tid == 0) vg = *vgp;
  __syncthreads();
  const int cflag = method == SPK_M_SK ? 1 : 0;
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e & 63, c = e >> 6;
    const int i = i0 + r, j = j0 + c;
    if (i >= Pw || j >= Pw) continue;
    double v;
    if (i >= P || j >= P) {
      v = 0.0;  // padding
    } else if (i < n && j < n) {
      // the upper triangle's value, mirrored (fill_vg_var_arr with the
      // diagonal flag computes g >= h and copies it below)
      const int a = min(i, j), b = max(i, j);
      const double ax = a == i ? rx[r] : cx[c], ay = a == i ? ry[r] : cy[c];
      const double bx = a == i ? cx[c] : rx[r], by = a == i ? cy[c] : ry[r];
      v = vg_entry(vg, dist2(ax, ay, bx, by), cflag, min_vg_val);
    } else if (i < n || j < n) {
      // border: row / column n of ones, then the drifts
      const int st = i < n ? i : j;
      const int k = (i < n ? j : i) - n;
      v = k == 0 ? 1.0 : stn_drift[(int64_t)idx[st] * m + (k - 1)];
    } else {
      v = 0.0;
    }
    K[i + (int64_t)j * ld] = v;
  }
}

// W[:, k] = V[:, k] / D[k] where |D[k]| > 1e-15 max |D| (numpy.linalg.pinv's
// rcond), else 0, for the Pw eigenpairs of a system zero-padded from P to
// Pw (the padding adds Pw - P zero eigenvalues, which the cutoff drops).
// D ascending (dsyevd), so max |D| is at an end.  Block k per column.
// Block 0 also records the outcome in the step record: pinv = 1 when the
// eigendecomposition converged (2 when eigenvalues of the system itself
// were dropped), and kappa = the ratio of the largest to the smallest kept
// |eigenvalue|.
__global__ __launch_bounds__(256) void k_pinv_scale(
    const double *__restrict__ V, int64_t ld, int Pw, int P,
    const double *__restrict__ D, const int *__restrict__ info,
    double *__restrict__ W, StepSc *__restrict__ sc) {
  const int n = sc->n;
  const int k = blockIdx.x;
  if (k >= Pw) return;
  const double mx = fmax(fabs(D[0]), fabs(D[Pw - 1]));
  const double cut = 1e-15 * mx;
  const double dk = D[k];
  const double wk = (fabs(dk) > cut) ? 1.0 / dk : 0.0;
  for (int i = threadIdx.x; i < P; i += 256)
    W[i + (int64_t)k * ld] = V[i + (int64_t)k * ld] * wk;
  if (k == 0 && threadIdx.x == 0) {
    double mn = INFINITY;
    int ncut = 0;
    for (int q = 0; q < Pw; ++q) {
      if (fabs(D[q]) > cut)
        mn = fmin(mn, fabs(D[q]));
      else
        ++ncut;
    }
    // 1: full rank, 2: eigenvalues under the cutoff were dropped
    const bool dropped = ncut > Pw - P;
    sc->pinv = (*info == 0 && n > 0) ? (dropped ? 2 : 1) : 0;
    sc->kappa = mx / mn;
  }
}

// On the contraction stream, after the batched kernels (which skip failed
// steps): a step with a pseudo-inverse leaves SPK_ST_FAILED and is
// contracted from pinv(K).  It is ill-conditioned when eigenvalues were
// dropped (the result then depends on where rounding puts them against
// the cutoff) or the kept ones span more than 1e12.
__global__ void k_mark_pinv(StepSc *__restrict__ sc) {
  if (threadIdx.x == 0 && sc->status == SPK_ST_FAILED && sc->pinv > 0) {
    sc->status = 0;
    // the batched variance contraction did not run for the step
    sc->flags = (sc->flags & ~(SPK_ST_FP64VAR | SPK_ST_INT8VAR)) | SPK_ST_PINV;
    if (sc->pinv == 2 || !(sc->kappa <= 1e12)) sc->flags |= SPK_ST_ILLCOND;
  }
}

// ---- contraction of a step solved by the pseudo-inverse --------------------
// The reference computes lambda = M r per cell (interp/steps.py:416) and
// falls back to the nearest station where sum(lambda[:n]) is not close to
// one (:418-427).  With M symmetric, lambda.z = r.beta and
// sum(lambda[:n]) = r.omega for beta = M [z; 0] and omega = M [1; 0], so
// estimates and the test cost O(n) per cell; only OK cells that pass the
// test and want a variance need the O(n^2) quadratic form, and only those
// are handed to the plain contraction (compacted into a list).

// bw[0:P] = beta, bw[ld:ld + P] = omega.  One thread per row of M.
__global__ __launch_bounds__(256) void k_pinv_vec(
    const double *__restrict__ M, int64_t ld, int p,
    const StepSc *__restrict__ sc, const double *__restrict__ z,
    double *__restrict__ bw) {
  if (sc->status) return;
  const int n = sc->n, P = n + p;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  double b = 0.0, w = 0.0;
  for (int j = 0; j < n; ++j) {
    const double v = M[i + (int64_t)j * ld];
    b += v * z[j];
    w += v;
  }
  bw[i] = b;
  bw[ld + i] = w;
}

// np.isclose(x, 1.0): rtol 1e-5, atol 1e-8 (interp/steps.py:418)
__device__ inline bool close1(double x) { return fabs(x - 1.0) <= 1e-8 + 1e-5; }

// One staged chunk of stations for k_pinv_cells: the nearest station and
// the sums r.beta, r.omega.  MT: the step's single bounded model (VgOne),
// or 0 for the general form; both give vg_fast_entry's values.
template <int MT>
__device__ inline void pinv_chunk(const VgFast &vf, const VgOne &vo,
                                  int cflag, double min_vg_val, double x,
                                  double y, const double *stx,
                                  const double *sty, const double *sb,
                                  const double *sw,
                                  const int32_t *__restrict__ idx, int base,
                                  int cnt, double &e, double &a, double &best,
                                  int &bi) {
  for (int j = 0; j < cnt; ++j) {
    const double d = dist2(x, y, stx[j], sty[j]);
    if (d < best || (d == best && idx[base + j] < idx[bi])) {
      best = d;
      bi = base + j;
    }
    double g;
    if constexpr (MT != 0)
      g = vg_one<MT>(vo, d, min_vg_val);
    else
      g = vg_fast_entry(vf, d, cflag, min_vg_val);
    e += g * sb[j];
    a += g * sw[j];
  }
}

// Per cell of a stripe: the estimate and the sum test from beta / omega.
// Cells that fail take the nearest active station (ties to the lower
// station index) and variance 0; OK cells that pass are appended to list
// (their variance, and their estimate in the reference's arithmetic, come
// from the plain contraction) when want_list is set.  The right-hand side
// takes the batched path's variogram values (vg_fast_entry: within a few
// ulp of fill_vg_var_arr's), with the station loop specialised on a single
// bounded model.
__global__ __launch_bounds__(256) void k_pinv_cells(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, const StepSc *__restrict__ sc,
    const spk_vg *__restrict__ vgp, const double *__restrict__ cx,
    const double *__restrict__ cy, int64_t Cs, int64_t C, int64_t cb,
    const double *__restrict__ cell_drift, int method, int m, int skm,
    double min_vg_val, const double *__restrict__ bw, int64_t ld,
    const double *__restrict__ z, double *__restrict__ est,
    double *__restrict__ var, int want_list, int32_t *__restrict__ list,
    int *__restrict__ cnt) {
  if (sc->status) return;
  const int n = sc->n;
  __shared__ double stx[256], sty[256], sb[256], sw[256];
  __shared__ VgFast vf;
  __shared__ VgOne vo;
  __shared__ int mt;
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x * (int64_t)256 + tid;
  const bool valid = c < Cs;
  const double x = valid ? cx[c] : 0.0, y = valid ? cy[c] : 0.0;
  const int cflag = method == SPK_M_SK ? 1 : 0;
  if (tid == 0) {
    vg_fast_init(*vgp, vf);
    // the single-model form is the semivariogram (OK / EDK)
    const int t = vg_one_of(vf, vo);
    mt = (cflag == 0 && t > 0) ? t : 0;
  }
  double e = 0.0, a = 0.0, best = INFINITY;
  int bi = 0;
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    if (base + tid < n) {
      stx[tid] = sx[idx[base + tid]];
      sty[tid] = sy[idx[base + tid]];
      sb[tid] = bw[base + tid];
      sw[tid] = bw[ld + base + tid];
    }
    __syncthreads();
    const int cnt2 = min(256, n - base);
    if (valid) {
      if (mt == SPK_VG_GAU)
        pinv_chunk<SPK_VG_GAU>(vf, vo, cflag, min_vg_val, x, y, stx, sty, sb,
                               sw, idx, base, cnt2, e, a, best, bi);
      else if (mt == SPK_VG_EXP)
        pinv_chunk<SPK_VG_EXP>(vf, vo, cflag, min_vg_val, x, y, stx, sty, sb,
                               sw, idx, base, cnt2, e, a, best, bi);
      else if (mt == SPK_VG_SPH)
        pinv_chunk<SPK_VG_SPH>(vf, vo, cflag, min_vg_val, x, y, stx, sty, sb,
                               sw, idx, base, cnt2, e, a, best, bi);
      else
        pinv_chunk<0>(vf, vo, cflag, min_vg_val, x, y, stx, sty, sb, sw, idx,
                      base, cnt2, e, a, best, bi);
    }
  }
  if (!valid) return;
  if (method != SPK_M_SK) {
    e += bw[n];
    a += bw[ld + n];
    if (method == SPK_M_EDK)
      for (int k = 0; k < m; ++k) {
        const double f = cell_drift[(int64_t)k * C + cb + c];
        e += f * bw[n + 1 + k];
        a += f * bw[ld + n + 1 + k];
      }
  }
  const int64_t oc = cb + c;
  if (skm) {
    est[oc] = e + sc->mean * (1.0 - a);
  } else if (close1(a)) {
    est[oc] = e;
    if (want_list) list[atomicAdd(cnt, 1)] = (int32_t)oc;
  } else {
    est[oc] = z[bi];
    if (var) var[oc] = 0.0;
  }
}

// Coordinates of the listed cells, and the count of each stripe of the
// list (stripe k holds entries k Csp .. (k + 1) Csp - 1).
__global__ void k_pinv_gather(const int32_t *__restrict__ list,
                              const int *__restrict__ cnt,
                              const double *__restrict__ cx,
                              const double *__restrict__ cy,
                              double *__restrict__ gx, double *__restrict__ gy,
                              int64_t Csp, int nstripes,
                              int *__restrict__ scnt) {
  const int total = *cnt;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    gx[i] = cx[list[i]];
    gy[i] = cy[list[i]];
  }
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < nstripes; k += 256)
      scnt[k] = (int)max((int64_t)0,
                         min(Csp, (int64_t)total - (int64_t)k * Csp));
}

static int blas_status(rocblas_status st, const char *what) {
  if (st != rocblas_status_success)
    return set_error(SPK_EHIP, "%s failed (rocblas status %d)", what,
                     (int)st);
  return SPK_OK;
}

int pinv_steps(spk_ctx *ctx, hipStream_t st, const double *sx,
               const double *sy, const int32_t *idx, int64_t N, StepSc *sc,
               const StepSc *h_sc, const std::vector<int> &steps,
               const spk_vg *vgs, int method, int m, const double *stn_drift,
               double min_vg_val, double *out, int64_t stride, int64_t ld) {
  const int p = method == SPK_M_OK ? 1 : method == SPK_M_EDK ? 1 + m : 0;
  const int nb = (int)steps.size();
  int Pw = 0;
  for (int s : steps) Pw = std::max(Pw, h_sc[s].n + p);
  if (nb == 0 || Pw < 1) return SPK_OK;
  if (!ctx->blas) {
    rocblas_handle h = nullptr;
    SPK_TRY(blas_status(rocblas_create_handle(&h), "rocblas_create_handle"));
    ctx->blas = h;
  }
  rocblas_handle h = (rocblas_handle)ctx->blas;
  SPK_TRY(blas_status(rocblas_set_stream(h, st), "rocblas_set_stream"));
  // the systems side by side (stride ld x ld), their eigenvalues and the
  // off-diagonals of the tridiagonal forms, one scaled eigenvector matrix
  double *Ab, *D, *E, *W;
  int *info;
  SPK_TRY(ws_get(ctx, "pinv_A", (size_t)nb * ld * ld * sizeof(double),
                 (void **)&Ab));
  SPK_TRY(ws_get(ctx, "pinv_D", (size_t)2 * nb * ld * sizeof(double),
                 (void **)&D));
  E = D + (int64_t)nb * ld;
  SPK_TRY(ws_get(ctx, "pinv_info", (size_t)nb * sizeof(int) + 256,
                 (void **)&info));
  SPK_TRY(ws_get(ctx, "pinv_W", (size_t)ld * ld * sizeof(double),
                 (void **)&W));
  const int tiles = (int)ceil_div(Pw, 64);
  for (int k = 0; k < nb; ++k) {
    const int s = steps[k];
    hipLaunchKernelGGL(k_pinv_assemble, dim3(tiles, tiles), dim3(256), 0, st,
                       sx, sy, idx + (int64_t)s * N, sc + s, vgs + s, method,
                       m, stn_drift, min_vg_val, Ab + (int64_t)k * ld * ld,
                       ld, Pw);
  }
  SPK_TRY(check_launch("k_pinv_assemble"));
  SPK_TRY(blas_status(
      rocsolver_dsyevd_strided_batched(
          h, rocblas_evect_original, rocblas_fill_lower, Pw, Ab, (int)ld,
          (rocblas_stride)(ld * ld), D, (rocblas_stride)ld, E,
          (rocblas_stride)ld, info, nb),
      "rocsolver_dsyevd_strided_batched"));
  const double one = 1.0, zero = 0.0;
  for (int k = 0; k < nb; ++k) {
    const int s = steps[k];
    const int P = h_sc[s].n + p;
    const double *V = Ab + (int64_t)k * ld * ld;
    hipLaunchKernelGGL(k_pinv_scale, dim3(Pw), dim3(256), 0, st, V, ld, Pw,
                       P, D + (int64_t)k * ld, info + k, W, sc + s);
    SPK_TRY(check_launch("k_pinv_scale"));
    // M = W V^T = V diag(1 / D) V^T over the kept eigenvalues; the rows
    // and columns of the system itself (P of them) are formed
    SPK_TRY(blas_status(rocblas_dgemm(h, rocblas_operation_none,
                                      rocblas_operation_transpose, P, P, Pw,
                                      &one, W, (int)ld, V, (int)ld, &zero,
                                      out + (int64_t)s * stride, (int)ld),
                        "rocblas_dgemm"));
  }
  return SPK_OK;
}

void pinv_release(spk_ctx *ctx) {
  if (ctx->blas) {
    rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
  }
}

}  // namespace spk
