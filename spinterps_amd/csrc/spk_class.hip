// Class API of cyth/interpmthds.pyx on the GPU: OrdinaryKriging
// (pyx:251-362), SimpleKriging (pyx:365-471) and ExternalDriftKriging_MD
// (pyx:588-719).
//
// These differ from the step driver: the semivariogram matrix has a zero
// diagonal (only g > h is filled, pyx:320-329), right-hand sides skip
// stations at distance 0 (pyx:343-349), OK's variance is
// max(0, lambda.gamma + mu) with a single mu (pyx:358-359), and
// SimpleKriging fills its covariance matrix before the sills are summed
// (pyx:397-399, 413), so its system is K = -Gamma.
//
// The systems are solved with the kriging kernels: Gamma is shifted to the
// SPD covariance C = c0 11^T - Gamma (the zero diagonal becomes c0, which
// includes the nugget), factorised by Cholesky, and K^-1 is formed from
// C^-1 by the Schur complement of the [1 | drifts] border (OK, EDK_MD) or
// by Sherman-Morrison for K = C - c0 11^T (SimpleKriging).  The reference
// uses pinv; both give the inverse of a nonsingular K.  lambda = K^-1 rhs
// for all points is one FP64 GEMM.
#include <algorithm>
#include <vector>

#include "spk_krige.h"

namespace spk {

__global__ void k_cls_gamma(const double *__restrict__ x,
                            const double *__restrict__ y, int n,
                            spk_vg vg, double *__restrict__ dists,
                            double *__restrict__ A, int64_t ld) {
  for (int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
       e < (int64_t)n * n; e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / n), j = (int)(e % n);
    double d = 0.0, g = 0.0;
    if (i != j) {
      d = dist2(x[i], y[i], x[j], y[j]);
      for (int k = 0; k < vg.n; ++k)
        g += vg_comp(vg.type[k], d, vg.range[k], vg.sill[k]);
    }
    dists[e] = d;
    A[i + (int64_t)j * ld] = g;
  }
}

// rhs rows [C x P]: OK/EDK: gamma (0 at distance 0), 1, drifts;
// SK: covar - gamma (covar at distance 0)
__global__ void k_cls_rhs(const double *__restrict__ xi,
                          const double *__restrict__ yi, int n,
                          const double *__restrict__ xk,
                          const double *__restrict__ yk, int64_t C,
                          const double *__restrict__ sk, int m, spk_vg vg,
                          int sk_flag, double covar, double *__restrict__ R,
                          int P) {
  for (int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x; e < C * P;
       e += (int64_t)gridDim.x * 256) {
    const int64_t c = e / P;
    const int i = (int)(e % P);
    double v;
    if (i < n) {
      const double d = dist2(xi[i], yi[i], xk[c], yk[c]);
      double g = 0.0;
      if (d != 0.0)
        for (int k = 0; k < vg.n; ++k)
          g += vg_comp(vg.type[k], d, vg.range[k], vg.sill[k]);
      v = sk_flag ? covar - g : g;
    } else if (i == n) {
      v = 1.0;
    } else {
      v = sk[(int64_t)(i - n - 1) * C + c];
    }
    R[e] = v;
  }
}

// K as the reference assembles it, [P x P] row-major
__global__ void k_cls_kmat(const double *__restrict__ A, int64_t ld, int n,
                           int P, const double *__restrict__ si, int sk_flag,
                           double *__restrict__ K) {
  for (int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
       e < (int64_t)P * P; e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / P), j = (int)(e % P);
    double v = 0.0;
    if (i < n && j < n) {
      v = sk_flag ? -A[i + (int64_t)j * ld] : A[i + (int64_t)j * ld];
    } else if (i < n && j == n) {
      v = 1.0;
    } else if (i == n && j < n) {
      v = 1.0;
    } else if (i < n && j > n) {
      v = si[(int64_t)(j - n - 1) * n + i];
    } else if (i > n && j < n) {
      v = si[(int64_t)(i - n - 1) * n + j];
    }
    K[e] = v;
  }
}

// SimpleKriging: K = C - c0 11^T, so K^-1 = C^-1 + c0 w w^T / (1 - c0 b)
// with w = C^-1 1, b = 1^T w (Sherman-Morrison).  M holds C^-1.
__global__ __launch_bounds__(256) void k_cls_sk_fix(
    double *__restrict__ M, int64_t ld, int n, const StepSc *__restrict__ sc,
    double *__restrict__ w, int *__restrict__ bad) {
  __shared__ double red[256];
  __shared__ double fac;
  const int tid = threadIdx.x;
  double bs = 0.0;
  for (int i = tid; i < n; i += 256) {
    double a = 0.0;
    for (int j = 0; j < n; ++j) a += M[i + (int64_t)j * ld];
    w[i] = a;
    bs += a;
  }
  red[tid] = bs;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  if (tid == 0) {
    const double c0 = sc[0].c0;
    const double den = 1.0 - c0 * red[0];
    if (!(fabs(den) > 1e-300) || !isfinite(den)) *bad = 1;
    fac = c0 / den;
  }
  __syncthreads();
  const double f = fac;
  for (int64_t e = tid; e < (int64_t)n * n; e += 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    M[i + (int64_t)j * ld] += f * w[i] * w[j];
  }
}

// lam [C x P] = R [C x P] * M (M symmetric, column-major with ld)
__global__ __launch_bounds__(256) void k_cls_gemm(const double *__restrict__ R,
                                                  int64_t C, int P,
                                                  const double *__restrict__ M,
                                                  int64_t ld,
                                                  double *__restrict__ lam) {
  __shared__ double As[32][33], Bs[32][33];
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  const int64_t r0 = blockIdx.y * (int64_t)32;
  const int c0 = blockIdx.x * 32;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int k0 = 0; k0 < P; k0 += 32) {
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
      const int i = e >> 5, k = e & 31;
      const int64_t r = r0 + i;
      const int kk = k0 + k;
      As[i][k] = (r < C && kk < P) ? R[r * P + kk] : 0.0;
      const int col = c0 + i;
      Bs[i][k] = (col < P && kk < P) ? M[kk + (int64_t)col * ld] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < 32; ++k) {
      const double a0 = As[ty * 2][k], a1 = As[ty * 2 + 1][k];
      const double b0 = Bs[tx * 2][k], b1 = Bs[tx * 2 + 1][k];
      acc[0][0] += a0 * b0;
      acc[0][1] += a0 * b1;
      acc[1][0] += a1 * b0;
      acc[1][1] += a1 * b1;
    }
    __syncthreads();
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int64_t r = r0 + ty * 2 + a;
      const int col = c0 + tx * 2 + b;
      if (r < C && col < P) lam[r * P + col] = acc[a][b];
    }
}

// zk, variances, mus
__global__ void k_cls_out(const double *__restrict__ lam,
                          const double *__restrict__ R, int64_t C, int n,
                          int P, const double *__restrict__ z, int kind,
                          double covar, double *__restrict__ zk,
                          double *__restrict__ ev) {
  const int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (c >= C) return;
  const double *l = lam + c * P;
  const double *r = R + c * P;
  double e = 0.0, q = 0.0;
  for (int i = 0; i < n; ++i) {
    e += l[i] * z[i];
    q += l[i] * r[i];
  }
  zk[c] = e;
  if (ev) {
    if (kind == 0)
      ev[c] = fmax(0.0, q + l[n]);
    else if (kind == 1)
      ev[c] = fmax(0.0, covar - q);
  }
}

__global__ void k_iota(int32_t *__restrict__ v, int n) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    v[i] = i;
}

}  // namespace spk

using namespace spk;

extern "C" {

int spk_krige_class(spk_ctx *ctx, int kind, const char *model,
                    const double *xi, const double *yi, const double *zi,
                    int64_t n_in, const double *si, int m, const double *xk,
                    const double *yk, const double *sk, int64_t C,
                    double *zk, double *est_vars, double *lambdas,
                    double *mus, double *rhss, double *in_vars,
                    double *in_vars_inv, double *in_dists) {
  if (!ctx || !model || !xi || !yi || !zi || !xk || !yk || !zk)
    return set_error(SPK_EARG, "null argument");
  if (kind < 0 || kind > 2) return set_error(SPK_EARG, "kind must be 0..2");
  if (n_in < 1 || C < 0) return set_error(SPK_ESHAPE, "need input points");
  if (n_in > 16384) return set_error(SPK_ESHAPE, "at most 16384 points");
  if (C > ((int64_t)1 << 21))
    return set_error(SPK_ESHAPE, "at most 2^21 interpolation points");
  if (kind == 2 && (m < 1 || m > kMaxP - 1 || !si || !sk))
    return set_error(SPK_EARG, "EDK_MD needs 1..%d drifts", kMaxP - 1);
  if (kind != 2) m = 0;
  const int n = (int)n_in;
  const int p = kind == 1 ? 0 : 1 + m;
  const int P = n + p;
  spk_vg vg;
  SPK_TRY(parse_vg(model, &vg));
  double covar = 0.0;
  for (int k = 0; k < vg.n; ++k) covar += vg.sill[k];
  SPK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t ld = ((P + 63) / 64) * 64;
  const int64_t slab = ld * ld;
  const int64_t Cr = C > 0 ? C : 1;
  double *dxi, *dyi, *dzi, *dxk, *dyk, *dsi = nullptr, *dsk = nullptr;
  double *A, *Li, *Mm, *T0, *R, *lam, *K, *dist, *dzk, *dev_;
  StepSc *sc;
  spk_vg *dvg;
  int32_t *idx;
  int *bad;
  SPK_TRY(ws_get(ctx, "cls_in", 3 * n * 8, (void **)&dxi));
  dyi = dxi + n;
  dzi = dyi + n;
  SPK_TRY(ws_get(ctx, "cls_out", 2 * Cr * 8, (void **)&dxk));
  dyk = dxk + Cr;
  if (kind == 2) {
    SPK_TRY(ws_get(ctx, "cls_si", (int64_t)m * n * 8 * 2, (void **)&dsi));
    SPK_TRY(ws_get(ctx, "cls_sk", (int64_t)m * Cr * 8, (void **)&dsk));
  }
  SPK_TRY(ws_get(ctx, "cls_A", slab * 8, (void **)&A));
  SPK_TRY(ws_get(ctx, "cls_Li", slab * 8, (void **)&Li));
  SPK_TRY(ws_get(ctx, "cls_M", slab * 8, (void **)&Mm));
  SPK_TRY(ws_get(ctx, "cls_T0", ld * kT0Cols * 8, (void **)&T0));
  SPK_TRY(ws_get(ctx, "cls_R", Cr * P * 8, (void **)&R));
  SPK_TRY(ws_get(ctx, "cls_lam", Cr * P * 8, (void **)&lam));
  SPK_TRY(ws_get(ctx, "cls_K", (int64_t)P * P * 8, (void **)&K));
  SPK_TRY(ws_get(ctx, "cls_dist", (int64_t)n * n * 8, (void **)&dist));
  SPK_TRY(ws_get(ctx, "cls_res", 2 * Cr * 8, (void **)&dzk));
  dev_ = dzk + Cr;
  SPK_TRY(ws_get(ctx, "cls_sc", sizeof(StepSc), (void **)&sc));
  SPK_TRY(ws_get(ctx, "cls_vg", sizeof(spk_vg), (void **)&dvg));
  SPK_TRY(ws_get(ctx, "cls_idx", n * 4, (void **)&idx));
  SPK_TRY(ws_get(ctx, "cls_bad", 64, (void **)&bad));
  SPK_TRY(h2d(ctx, dxi, xi, n * 8));
  SPK_TRY(h2d(ctx, dyi, yi, n * 8));
  SPK_TRY(h2d(ctx, dzi, zi, n * 8));
  if (C > 0) {
    SPK_TRY(h2d(ctx, dxk, xk, C * 8));
    SPK_TRY(h2d(ctx, dyk, yk, C * 8));
  }
  // si [m x n] as given, and transposed [n x m] for the Schur border
  std::vector<double> sit((size_t)n * m);
  if (kind == 2) {
    for (int k = 0; k < m; ++k)
      for (int i = 0; i < n; ++i) sit[(size_t)i * m + k] = si[(size_t)k * n + i];
    SPK_TRY(h2d(ctx, dsi, si, (int64_t)m * n * 8));
    SPK_TRY(h2d(ctx, dsi + (int64_t)m * n, sit.data(), (int64_t)m * n * 8));
    if (C > 0) SPK_TRY(h2d(ctx, dsk, sk, (int64_t)m * C * 8));
  }
  StepSc hs{};
  hs.n = n;
  SPK_TRY(h2d(ctx, sc, &hs, sizeof(StepSc)));
  SPK_TRY(h2d(ctx, dvg, &vg, sizeof(spk_vg)));
  SPK_HIP(hipMemsetAsync(bad, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_iota, dim3(ceil_div(n, 256)), dim3(256), 0, st, idx, n);
  const unsigned gn = (unsigned)std::min<int64_t>(ceil_div((int64_t)n * n, 256),
                                                  65536);
  hipLaunchKernelGGL(k_cls_gamma, dim3(gn), dim3(256), 0, st, dxi, dyi, n, vg,
                     dist, A, ld);
  const unsigned gp = (unsigned)std::min<int64_t>(ceil_div((int64_t)P * P, 256),
                                                  65536);
  hipLaunchKernelGGL(k_cls_kmat, dim3(gp), dim3(256), 0, st, A, ld, n, P, dsi,
                     kind == 1 ? 1 : 0, K);
  SPK_TRY(check_launch("class assembly"));
  // C = c0 - Gamma (lower triangle), Cholesky, triangular inverse
  hipLaunchKernelGGL(k_potrf, dim3(1), dim3(256), 0, st, A, ld, slab, sc, dvg,
                     1);
  {
    const dim3 gt((unsigned)ceil_div(n, 64), 1);
    hipLaunchKernelGGL(k_trtri_diag, gt, dim3(256), 0, st, A, Li, ld, slab, sc);
    hipLaunchKernelGGL(k_trtri_cols, gt, dim3(256), 0, st, A, Li, ld, slab, sc);
  }
  hipLaunchKernelGGL(k_minv, dim3(1), dim3(256), 0, st, Li, ld, slab, sc, idx,
                     (int64_t)n, dsi ? dsi + (int64_t)m * n : nullptr, m,
                     kind == 1 ? SPK_M_SK : (kind == 2 ? SPK_M_EDK : SPK_M_OK),
                     Mm, T0, dvg, -INFINITY);
  if (kind == 1)
    hipLaunchKernelGGL(k_cls_sk_fix, dim3(1), dim3(256), 0, st, Mm, ld, n, sc,
                       T0, bad);
  SPK_TRY(check_launch("class inverse"));
  StepSc hs2;
  int hbad = 0;
  SPK_TRY(d2h(ctx, &hs2, sc, sizeof(StepSc)));
  SPK_TRY(d2h(ctx, &hbad, bad, sizeof(int)));
  SPK_TRY(sync(ctx));
  if (hs2.status & SPK_ST_FAILED)
    return set_error(SPK_ENOTSPD,
                     "covariance form of the system is not positive definite");
  if (hbad) return set_error(SPK_ENOTSPD, "simple-kriging system is singular");
  if (C > 0) {
    const unsigned gr = (unsigned)std::min<int64_t>(ceil_div(C * P, 256),
                                                    65536);
    hipLaunchKernelGGL(k_cls_rhs, dim3(gr), dim3(256), 0, st, dxi, dyi, n, dxk,
                       dyk, C, dsk, m, vg, kind == 1 ? 1 : 0, covar, R, P);
    const dim3 gg((unsigned)ceil_div(P, 32), (unsigned)ceil_div(C, 32));
    hipLaunchKernelGGL(k_cls_gemm, gg, dim3(256), 0, st, R, C, P, Mm, ld, lam);
    hipLaunchKernelGGL(k_cls_out, dim3((unsigned)ceil_div(C, 256)), dim3(256),
                       0, st, lam, R, C, n, P, dzi, kind, covar, dzk,
                       kind == 2 ? nullptr : dev_);
    SPK_TRY(check_launch("class contraction"));
  }
  // host outputs
  if (C > 0) {
    SPK_TRY(d2h(ctx, zk, dzk, C * 8));
    if (est_vars && kind != 2) SPK_TRY(d2h(ctx, est_vars, dev_, C * 8));
  }
  std::vector<double> hl;
  if (C > 0 && (lambdas || mus)) {
    hl.resize((size_t)C * P);
    SPK_TRY(d2h(ctx, hl.data(), lam, C * P * 8));
  }
  std::vector<double> hr;
  if (C > 0 && rhss) SPK_TRY(d2h(ctx, rhss, R, C * P * 8));
  std::vector<double> hm;
  if (in_vars_inv) hm.resize((size_t)ld * P);
  if (in_vars_inv) SPK_TRY(d2h(ctx, hm.data(), Mm, ld * P * 8));
  if (in_vars) SPK_TRY(d2h(ctx, in_vars, K, (int64_t)P * P * 8));
  if (in_dists) SPK_TRY(d2h(ctx, in_dists, dist, (int64_t)n * n * 8));
  SPK_TRY(sync(ctx));
  if (lambdas)
    for (int64_t c = 0; c < C; ++c)
      std::copy(hl.begin() + c * P, hl.begin() + c * P + n, lambdas + c * n);
  if (mus && p > 0)
    for (int64_t c = 0; c < C; ++c)
      for (int a = 0; a < p; ++a) mus[(int64_t)a * C + c] = hl[c * P + n + a];
  if (in_vars_inv)
    for (int i = 0; i < P; ++i)
      for (int j = 0; j < P; ++j)
        in_vars_inv[(int64_t)i * P + j] = hm[(size_t)j * ld + i];
  return SPK_OK;
}

}  // extern "C"
