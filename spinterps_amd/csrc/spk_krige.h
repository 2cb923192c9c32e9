// Shared declarations of the kriging kernels.
#pragma once

#include <vector>

#include "spk_internal.h"

namespace spk {

constexpr int kMaxP = 8;  // constraint rows: 1 + up to 7 drifts
// Columns of ld doubles per step in the factor work array T0: the
// right-hand sides [F z] and their images (2 (kMaxP + 1)), then six for
// the refinement's residuals and their images.
constexpr int kT0Cols = 2 * (kMaxP + 1) + 6;

// Logical block id of linear block id lin in a grid of G blocks, so that
// the blocks dispatched to one XCD (lin mod 8: the hardware deals blocks
// to the eight XCDs in turn) take a contiguous range of logical ids.  XCD
// x holds q + (x < r) of the blocks (q = G / 8, r = G mod 8); the map is a
// bijection of [0, G) for any G.
__device__ inline int xcd_remap(int lin, int G) {
  const int x = lin & 7, q = G >> 3, r = G & 7;
  return x * q + min(x, r) + (lin >> 3);
}

// The steps of a batch one contraction launch takes (blockIdx.z indexes
// s[]), chosen on the host from the step records after the factorisation:
// blocks of the other steps would only exit.
constexpr int kMaxStepsPerBatch = 64;
struct StepList {
  int cnt;
  int s[kMaxStepsPerBatch];
};

// Per-step scalars kept on the device.
struct StepSc {
  int n;          // active stations
  int status;     // SPK_ST_* that skip the step's regular kernels (0 = run)
  int flags;      // informational SPK_ST_* bits (ill-conditioned, FP64 var)
  int planes;     // int8 planes of the variance contraction (0, 5 or 6)
  int pinv;       // 1: a failed step's pseudo-inverse is ready (spk_pinv)
  double kappa;   // max diag(C^-1) * max diag(C), a condition estimate
  double mean;    // mean of the active values (interp/steps.py:276)
  double c0;      // covariance shift of the OK/EDK system
  double b;       // 1^T C^-1 1
  double a0;      // constant term of the dual-form estimate
  double zabs;    // mean |value| of the active stations
  double h[kMaxP];
  double dh[kMaxP];  // last refinement correction of h (dual form)
};

// ---- IDW as a masked contraction on FP64 MFMA (spk_idw.hip) ----------------
// grid (ceil(C / 128), ceil(S / 32)), 256 threads.
__global__ void k_idw_mfma(const double *sx, const double *sy, int64_t N,
                           const double *vals, int S, const StepSc *sc,
                           const int32_t *idx, const double *zact,
                           int64_t ldz, const double *cx, const double *cy,
                           int64_t C, double p, double *est, int *bad);

// ---- pseudo-inverse fallback of failed steps (spk_pinv.hip) ----------------
// For the steps `steps` of a batch (step records sc on the device, their
// host copies h_sc): builds each bordered system, zero-padded to the
// largest, eigendecomposes them in one batched call and writes pinv(K) of
// step s to out + s stride (leading dimension ld), all on stream st;
// records sc[s].pinv and sc[s].kappa on the device.
int pinv_steps(spk_ctx *ctx, hipStream_t st, const double *sx,
               const double *sy, const int32_t *idx, int64_t N, StepSc *sc,
               const StepSc *h_sc, const std::vector<int> &steps,
               const spk_vg *vgs, int method, int m, const double *stn_drift,
               double min_vg_val, double *out, int64_t stride, int64_t ld);
// The step leaves SPK_ST_FAILED for the plain contraction when its
// pseudo-inverse is ready (SPK_ST_PINV | SPK_ST_ILLCOND).
__global__ void k_mark_pinv(StepSc *sc);
__global__ void k_pinv_vec(const double *M, int64_t ld, int p,
                           const StepSc *sc, const double *z, double *bw);
__global__ void k_pinv_cells(const double *sx, const double *sy,
                             const int32_t *idx, const StepSc *sc,
                             const spk_vg *vgp, const double *cx,
                             const double *cy, int64_t Cs, int64_t C,
                             int64_t cb, const double *cell_drift, int method,
                             int m, int skm, double min_vg_val,
                             const double *bw, int64_t ld, const double *z,
                             double *est, double *var, int want_list,
                             int32_t *list, int *cnt);
__global__ void k_pinv_gather(const int32_t *list, const int *cnt,
                              const double *cx, const double *cy, double *gx,
                              double *gy, int64_t Csp, int nstripes,
                              int *scnt);
void pinv_release(spk_ctx *ctx);

// ---- scratch of the allocation-free IDW / nearest-neighbour entry point ----
// spk_idw_steps_noalloc_dev takes every buffer beyond its inputs and outputs
// from one file-scope device array (spk_core.hip), laid out as below: the
// re-centred station and cell coordinates, the weight-sum flag, and the
// per-step compaction of kNaSteps steps at a time.
constexpr int64_t kNaMaxStns = 8192;
constexpr int64_t kNaMaxCells = (int64_t)1 << 20;
constexpr int64_t kNaSteps = 64;
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaOffStn = 0;
constexpr size_t kNaOffCell = kNaOffStn + na_align(2 * kNaMaxStns * 8);
constexpr size_t kNaOffBad = kNaOffCell + na_align(2 * kNaMaxCells * 8);
constexpr size_t kNaOffIdx = kNaOffBad + 256;
constexpr size_t kNaOffZ = kNaOffIdx + na_align(kNaSteps * kNaMaxStns * 4);
constexpr size_t kNaOffSc = kNaOffZ + na_align(kNaSteps * kNaMaxStns * 8);
constexpr size_t kNaBytes = kNaOffSc + na_align(kNaSteps * sizeof(StepSc));
// Device address of that array on `device`, resolved once per device by
// spk_ctx_create; SPK_EHIP if the resolution failed.
int noalloc_scratch(int device, unsigned char **out);

__global__ void k_compact(const double *vals, int64_t N, int32_t *idx,
                          double *zact, int64_t ldz, StepSc *sc,
                          const spk_vg *vgs, int method, double thr,
                          double min_vg_val, const int32_t *sorder);
__global__ void k_assemble(const double *sx, const double *sy,
                           const int32_t *idx, int64_t N, const StepSc *sc,
                           const spk_vg *vgs, int covar_flag, double min_vg_val,
                           double *A, int64_t ld, int64_t strideA,
                           double *Gfull);
__global__ void k_potrf(double *A, int64_t ld, int64_t strideA, StepSc *sc,
                        const spk_vg *vgs, int shift);
// Cholesky over (block, step) workgroups: k_chol_shift (grid (steps)),
// then per 32-wide panel k0: k_chol_diag (grid (1, steps)), k_chol_solve
// (grid (rows below / 256, steps)), k_chol_trail (grid (trailing 64 x 64
// tiles, steps)).  chol_steps launches the sequence.
__global__ void k_chol_shift(double *A, int64_t ld, int64_t strideA,
                             StepSc *sc, const spk_vg *vgs);
__global__ void k_chol_diag(double *A, int64_t ld, int64_t strideA,
                            StepSc *sc, int k0);
__global__ void k_chol_solve(double *A, int64_t ld, int64_t strideA,
                             const StepSc *sc, int k0);
__global__ void k_chol_trail(double *A, int64_t ld, int64_t strideA,
                             const StepSc *sc, int k0);
// X = L^-1: k_trtri_diag then k_trtri_cols, each with grid
// (ceil(n_max / 64), steps)
__global__ void k_trtri_diag(const double *L, double *X, int64_t ld,
                             int64_t stride, const StepSc *sc);
__global__ void k_trtri_cols(const double *L, double *X, int64_t ld,
                             int64_t stride, const StepSc *sc);
// condition estimate: k_kap_start, 12 x (k_kap_norm, k_kap_umv / k_kap_lmv,
// k_kap_rq, k_kap_lmv / k_kap_umv) per power iteration, then k_kap_flags
// (spk_factor.hip)
__global__ void k_kap_start(const StepSc *sc, double *T0, int64_t ldt,
                            int which);
__global__ void k_kap_norm(const StepSc *sc, double *T0, int64_t ldt);
__global__ void k_kap_rq(const StepSc *sc, double *T0, int64_t ldt,
                         double *lam, int which);
__global__ void k_kap_lmv(const double *T, int64_t ld, int64_t stride,
                          const StepSc *sc, double *T0, int64_t ldt, int to_u);
__global__ void k_kap_umv(const double *T, int64_t ld, int64_t stride,
                          const StepSc *sc, double *T0, int64_t ldt, int to_u);
__global__ void k_kap_flags(StepSc *sc, const double *lam, const spk_vg *vgs,
                            double fp64var_thr, double fp64var_thr_gau,
                            double int8var_thr);
// dual vectors and refinement (spk_factor.hip): k_dual_rhs, k_dual_lmv,
// k_dual_umv, k_dual_mid, then twice k_dual_resid, k_dual_lmv, k_dual_umv,
// k_dual_upd, then k_dual_fin.  DS: kDualSc doubles per step.
constexpr int kDualSc = 256;
__global__ void k_dual_rhs(const StepSc *sc, const double *zact, int64_t ldz,
                           const int32_t *idx, int64_t N,
                           const double *stn_drift, int m, int method,
                           double *T0, int64_t ldt);
__global__ void k_dual_lmv(const double *Li, int64_t ld, int64_t stride,
                           const StepSc *sc, double *T0, int64_t ldt, int src,
                           int dst, int nr);
__global__ void k_dual_umv(const double *Li, int64_t ld, int64_t stride,
                           const StepSc *sc, double *T0, int64_t ldt, int src,
                           int dst, int nr);
__global__ void k_dual_mid(const StepSc *sc, int m, int method, double *T0,
                           int64_t ldt, double *V, int64_t ldv, double *DS);
__global__ void k_dual_resid(const StepSc *sc, const double *zact,
                             int64_t ldz, const int32_t *idx, int64_t N,
                             const double *stn_drift, int m, int method,
                             double *T0, int64_t ldt, double *V, int64_t ldv,
                             double *DS, const spk_vg *vgs, double min_vg_val,
                             const double *sx, const double *sy);
__global__ void k_dual_upd(const StepSc *sc, const int32_t *idx, int64_t N,
                           const double *stn_drift, int m, int method,
                           double *T0, int64_t ldt, double *V, int64_t ldv,
                           double *DS);
__global__ void k_dual_fin(StepSc *sc, int m, int method, const double *DS);
__global__ void k_minv(const double *Li, int64_t ld, int64_t stride,
                       StepSc *sc, const int32_t *idx, int64_t N,
                       const double *stn_drift, int m, int method, double *M,
                       double *T0, const spk_vg *vgs, double min_vg_val);
__global__ void k_recenter(const double *in, double *out, int64_t n, double o);
__global__ void k_status_out(const StepSc *sc, int S, int32_t *info,
                             double *kappa);
// Spatially compact order of the cells (Morton keys over the bounding box,
// radix-sorted): perm[k] = original index of the k-th cell, pcx/pcy the
// coordinates in that order.  Runs on ctx->stream.
int cell_order(spk_ctx *ctx, const double *cx, const double *cy, int64_t C,
               int32_t *perm, double *pcx, double *pcy);
__global__ void k_cast_tri(const double *Li, float *L32, int64_t ld,
                           int64_t stride, const StepSc *sc);

// ---- split-BF16 operands of the FP32-tier variance contraction -----------
// (spk_bf16.hip).  A value x is held as three bfloat16 terms hi + mid + lo,
// each the nearest bfloat16 (ties to even) to what the previous ones
// leave; the remainders are exact in FP32.  The planes store 8 consecutive
// k of one row (L^-1) or one cell (delta) in 16 bytes.
__device__ inline uint16_t bf16_bits(float x) {
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ inline float bf16_float(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
__device__ inline void split3f(float x, uint16_t &h, uint16_t &m,
                               uint16_t &l) {
  h = bf16_bits(x);
  const float r = x - bf16_float(h);
  m = bf16_bits(r);
  l = bf16_bits(r - bf16_float(m));
}
// two values at once (one packed conversion per term): the low halves of
// h, m, l hold x0's terms, the high halves x1's
__device__ inline void split3f2(float x0, float x1, uint32_t &h, uint32_t &m,
                                uint32_t &l) {
  typedef float f2_t __attribute__((ext_vector_type(2)));
  typedef __bf16 b2_t __attribute__((ext_vector_type(2)));
  auto cvt = [](float a, float b) {
    return __builtin_bit_cast(uint32_t,
                              __builtin_convertvector(f2_t{a, b}, b2_t));
  };
  h = cvt(x0, x1);
  const float r0 = x0 - __uint_as_float(h << 16);
  const float r1 = x1 - __uint_as_float(h & 0xffff0000u);
  m = cvt(r0, r1);
  l = cvt(r0 - __uint_as_float(m << 16), r1 - __uint_as_float(m & 0xffff0000u));
}
// from FP64: the remainders are taken in FP64
__device__ inline void split3d(double x, uint16_t &h, uint16_t &m,
                               uint16_t &l) {
  h = bf16_bits((float)x);
  const double r = x - (double)bf16_float(h);
  m = bf16_bits((float)r);
  l = bf16_bits((float)(r - (double)bf16_float(m)));
}
// Largest system (ld) the split contraction takes: its chunk list holds
// ld / 32 entries.
constexpr int64_t kBfMaxLd = 8192;
template <int AF>
__global__ void k_split_tri(const double *Li, int64_t ld, int64_t stride,
                            const StepSc *sc, uint4 *Lb);
template <int PF, int AF>
__global__ void k_tri_bf(const uint4 *Lb, int64_t ld, const uint4 *Db,
                         int64_t ldk, const StepSc *sc, StepList lst,
                         int64_t Cs, double *qpart, int nrb,
                         const uint8_t *nzb, int64_t nzk,
                         unsigned long long *nexec);
// exact int8 plane products (spk_int8.hip): up to 6 planes per operand;
// steps up to condition estimate kInt8Deep use 5 planes and the plane
// products up to the fourth diagonal, steps above it 6 planes and the
// fifth diagonal
constexpr int kInt8Planes = 6;
constexpr double kInt8Deep = 1e7;
__global__ void k_slice_rows(const double *Li, int64_t ld, int64_t stride,
                             const StepSc *sc, int8_t *A8, int64_t KS,
                             int64_t R, int *ea);
template <typename T>
__global__ void k_slice_cols(const T *K, int64_t ld, int64_t ldk, int64_t Cs,
                             const StepSc *sc, const double *dmax, int8_t *B8,
                             int64_t KS, int *eb, const uint8_t *nzb,
                             int64_t nzk, uint8_t *nz8);
template <int S, int P>
__global__ void k_tri8(const int8_t *A8, int64_t R, const int *ea,
                       const int8_t *B8, int64_t ldk, const int *eb,
                       int64_t KS, const StepSc *sc, StepList lst,
                       int64_t Cs, double *qpart, int nrb,
                       const uint8_t *nz8, unsigned long long *nexec);

}  // namespace spk
