// Cell-side kernels of the step interpolation and the batch driver behind
// spk_interp_steps_dev.
//
//  k_rhs_plain   station-cell distance -> semivariogram/covariance, written
//                as the reference's right-hand sides R [(n+p) x cells]
//                (interp/steps.py:203-234), plus the nearest active station
//  k_gemm<..>    MFMA contraction over all cells of a step:
//                PLAIN: Lambda = K^-1 R (interp/steps.py:416) with the
//                       estimate / sum-of-weights / variance epilogue of
//                       interp/steps.py:418-434 fused in
//                TRI  : W = L^-1 K, q = |W|^2 per cell (variance term)
//  k_tri64       the TRI contraction in FP64, 128 x 128 tiles
//  k_nearest     each cell's nearest active station (batched variance)
//  k_kgen<..>    fused distance -> covariance for the batched path, with
//                the dual-form estimate k.alpha accumulated in FP64 and the
//                locally deflated covariances delta for the variance
//  k_var_reduce  variance from the FP64 terms and the partial norms
//  (the int8 plane-product contraction is in spk_int8.hip)
//  k_idw_nnb     interp/steps.py:285-313 (nearest neighbour; IDW runs as
//                a masked FP64 MFMA contraction, k_idw_mfma in spk_idw.hip)
//  k_special     one-station, mean-fill and failed-system steps
//                (interp/steps.py:280-283, 325-331, 372-384)
#include <algorithm>
#include <type_traits>
#include <vector>

#include "spk_krige.h"
#include "spk_prof.h"

namespace spk {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Result row held in slot r by lane group g = lane / 16 (column lane % 16):
// f64 16x16x4 interleaves rows (g + 4 r), f32 16x16x4 keeps blocks (4 g + r)
// (checked on gfx950 against a host matmul).
template <>
struct Mfma<double> {
  using V = d4_t;
  __device__ static inline V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int g, int r) { return g + 4 * r; }
};
template <>
struct Mfma<float> {
  using V = f4_t;
  __device__ static inline V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int g, int r) { return 4 * g + r; }
};

// |sum of weights - 1| of a cell above which the plain path flags the step
// SPK_ST_ILLCOND (the sum is exactly one for OK / EDK).  On hourly steps
// the plain path's error against an exact solution stays below 450 times
// this deviation, so unflagged steps are within 5e-7.
constexpr double kSumTol = 1e-9;
// Condition estimate above which a batched step whose Cholesky succeeded is
// solved by the pseudo-inverse of its bordered system instead: there the
// Cholesky solve is rounding noise (the estimate can be off by orders of
// magnitude), while pinv drops the eigenvalues under 1e-15 of the largest
// as the reference does.
constexpr double kPinvKappa = 1e14;
// Relative change of a cell's estimate by the second refinement step above
// which the step is flagged SPK_ST_ILLCOND (the change is of the size of
// the error left by the first step; the refined estimate is at least as
// good), so unflagged steps stay well inside 1e-6.
constexpr double kRefineTol = 1e-7;
// L^-1 of the split-BF16 contraction: 1 = FP32 octets split into the
// bfloat16 planes as k_tri_bf stages them (4 bytes per value read from
// memory instead of 6), 0 = the planes themselves.
constexpr int kBfA32 = 1;

// np.isclose(x, 1.0) with rtol 1e-5, atol 1e-8 (interp/steps.py:418)
__device__ inline bool isclose1(double x) {
  const double tol = 1e-8 + 1e-5 * 1.0;
  return fabs(x - 1.0) <= tol;
}

__global__ void k_recenter(const double *__restrict__ in,
                           double *__restrict__ out, int64_t n, double o) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i] - o;
}

// ---- plain path right-hand sides ---------------------------------------------
__global__ __launch_bounds__(256) void k_rhs_plain(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, const StepSc *__restrict__ sc,
    const spk_vg *__restrict__ vgp, const double *__restrict__ cx,
    const double *__restrict__ cy, int64_t Cs, int64_t C,
    const double *__restrict__ cell_drift, int method, int m, int covar_flag,
    double min_vg_val, double *__restrict__ R, int64_t ldr,
    int32_t *__restrict__ nnb, const int *__restrict__ ccount) {
  if (sc->status) return;
  const int n = sc->n;
  // ccount (device, optional): the cells of this launch, at most Cs
  if (ccount) Cs = min(Cs, (int64_t)*ccount);
  if (blockIdx.x * (int64_t)256 >= Cs) return;
  __shared__ double stx[256], sty[256];
  __shared__ spk_vg vg;
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x * (int64_t)256 + tid;
  const bool valid = c < Cs;
  const double x = valid ? cx[c] : 0.0, y = valid ? cy[c] : 0.0;
  if (tid == 0) vg = *vgp;
  double best = INFINITY;
  int bi = 0;
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    if (base + tid < n) {
      stx[tid] = sx[idx[base + tid]];
      sty[tid] = sy[idx[base + tid]];
    }
    __syncthreads();
    const int cnt = min(256, n - base);
    if (valid) {
      for (int j = 0; j < cnt; ++j) {
        const double d = dist2(x, y, stx[j], sty[j]);
        // ties go to the lower station index (np.argmin), whatever order
        // the step's stations are compacted in
        if (d < best || (d == best && idx[base + j] < idx[bi])) {
          best = d;
          bi = base + j;
        }
        R[(int64_t)(base + j) * ldr + c] =
            vg_entry(vg, d, covar_flag, min_vg_val);
      }
    }
  }
  if (!valid) return;
  if (method != SPK_M_SK) R[(int64_t)n * ldr + c] = 1.0;
  if (method == SPK_M_EDK)
    for (int k = 0; k < m; ++k)
      R[(int64_t)(n + 1 + k) * ldr + c] = cell_drift[(int64_t)k * C + c];
  nnb[c] = bi;
}

// ---- MFMA contraction --------------------------------------------------------
// Block tile: 64 rows x 128 cells, k in chunks of BK, 256 threads = 4 waves;
// wave w owns cells [32 w, 32 w + 32) of the tile over all 64 rows as 4 x 2
// MFMA 16x16 tiles, so the per-cell row reductions of the epilogue need no
// cross-wave step.  The next k chunk is loaded into registers while the
// current one is multiplied out of LDS (two LDS buffers, one barrier per
// chunk).  A is column-major [P x P] (lda), B row-major [P x ldb cells];
// ldb is a multiple of 128 and every 128-cell tile lies inside it.
// MFMA 16x16x4 operand layout: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; the result row in slot r is Mfma<T>::row(l / 16, r).
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double2;  // 16-byte loads: two doubles
};

template <typename T, bool PLAIN, int BK>
__global__ __launch_bounds__(256, PLAIN ? 1 : 3) void k_gemm(
    const T *__restrict__ A, int64_t lda, int64_t strideA,
    const T *__restrict__ B, int64_t ldb, int64_t strideB,
    StepSc *__restrict__ sc, int p, int method, int want_var, int skm,
    const double *__restrict__ zact, int64_t ldz,
    const int32_t *__restrict__ nnb, int64_t ldn,
    const double *__restrict__ abuf, int64_t ldab, int64_t Cs,
    double *__restrict__ est, double *__restrict__ var, int64_t ldo,
    int sel, double *__restrict__ qpart, int nrb,
    const uint8_t *__restrict__ nzb, int64_t nzk,
    unsigned long long *__restrict__ nexec,
    const int *__restrict__ ccount, const int32_t *__restrict__ cmap) {
  using MF = Mfma<T>;
  // ccount (device, optional): the cells of this launch, at most Cs;
  // cmap (PLAIN, optional): output index of each cell
  if (ccount) Cs = min(Cs, (int64_t)*ccount);
  using V = typename MF::V;
  constexpr int BM = 64, BN = 128;
  // PLAIN: grid (cell tiles, steps), each block walks all row blocks.
  // TRI: grid (row blocks, cell tiles, steps), one row block per block;
  // blocks on the same XCD (linear id mod 8) take consecutive row blocks of
  // one cell tile so that the tile's K rows stay in that XCD's L2.
  int rbi = 0, cti = blockIdx.x;
  if (!PLAIN) {
    const int G = gridDim.x * gridDim.y;
    int lin = blockIdx.x + blockIdx.y * gridDim.x;
    if ((G & 7) == 0) lin = (lin & 7) * (G >> 3) + (lin >> 3);
    rbi = lin % gridDim.x;
    cti = lin / gridDim.x;
  }
  constexpr int VW = 16 / sizeof(T);         // elements per 16-byte load
  // row padding: FP32 rows 16 banks apart, so the four 16-lane groups of a
  // 32-bit LDS read (one row each) fall on distinct banks
  constexpr int PAD = sizeof(T) == 4 ? 16 : VW;
  constexpr int A_LD = BM * BK / VW / 256;   // vector loads per thread
  constexpr int B_LD = BN * BK / VW / 256;
  static_assert(A_LD >= 1 && B_LD >= 1, "tile too small");
  using VT = typename Vec4<T>::type;
  const int sl = PLAIN ? blockIdx.y : blockIdx.z;
  const StepSc st = sc[sl];
  if (st.status) return;
  // sel 1: only the FP32 steps (neither SPK_ST_FP64VAR nor SPK_ST_INT8VAR),
  // sel 2: only steps with SPK_ST_FP64VAR
  if (sel == 1 && (st.flags & (SPK_ST_FP64VAR | SPK_ST_INT8VAR))) return;
  if (sel == 2 && !(st.flags & SPK_ST_FP64VAR)) return;
  const int n = st.n;
  const int P = PLAIN ? n + p : n;
  const int64_t c0 = cti * (int64_t)BN;
  if (c0 >= Cs) return;
  const int r_lo = PLAIN ? 0 : rbi * BM;
  if (r_lo >= P) return;
  const int r_hi = PLAIN ? P : min(P, r_lo + BM);
  A += (int64_t)sl * strideA;
  B += (int64_t)sl * strideB;
  const double *z = zact + (int64_t)sl * ldz;
  __shared__ __align__(16) T As[2][BK][BM + PAD];
  __shared__ __align__(16) T Bs[2][BK][BN + PAD];
  __shared__ int s_list[256];  // nonzero chunks (TRI with a chunk map)
  __shared__ int s_cnt[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;

  double acc_e[2] = {0.0, 0.0}, acc_l[2] = {0.0, 0.0};
  double acc_v[2] = {0.0, 0.0}, acc_m[2] = {0.0, 0.0};

  // register stages of the next chunks' tiles; the FP32 contraction keeps
  // two chunks in flight, which covers an L2 miss behind the chunk being
  // multiplied
  constexpr bool PF2 = !PLAIN && sizeof(T) == 4;
  VT ra[A_LD], rb[B_LD], ra2[PF2 ? A_LD : 1], rb2[PF2 ? B_LD : 1];
  // A tile element group q: rows r..r+VW-1 of column k
  auto load_tiles = [&](auto &ra, auto &rb, int r0, int k0, int kend) {
#pragma unroll
    for (int q = 0; q < A_LD; ++q) {
      const int e = tid + q * 256;
      const int r = (e % (BM / VW)) * VW, k = e / (BM / VW);
      const int gk = k0 + k;
      if (gk < kend) {
        ra[q] = *reinterpret_cast<const VT *>(A + (r0 + r) +
                                              (int64_t)gk * lda);
        // rows beyond P are never used but may hold stale values
        T *v = reinterpret_cast<T *>(&ra[q]);
#pragma unroll
        for (int u = 0; u < VW; ++u)
          if (r0 + r + u >= P) v[u] = T(0);
      } else {
        ra[q] = VT{};
      }
    }
#pragma unroll
    for (int q = 0; q < B_LD; ++q) {
      const int e = tid + q * 256;
      const int c = (e % (BN / VW)) * VW, k = e / (BN / VW);
      const int gk = k0 + k;
      if (gk < kend)
        rb[q] = *reinterpret_cast<const VT *>(B + (int64_t)gk * ldb + c0 + c);
      else
        rb[q] = VT{};
    }
  };
  auto store_tiles = [&](const auto &ra, const auto &rb, int buf) {
#pragma unroll
    for (int q = 0; q < A_LD; ++q) {
      const int e = tid + q * 256;
      const int r = (e % (BM / VW)) * VW, k = e / (BM / VW);
      *reinterpret_cast<VT *>(&As[buf][k][r]) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < B_LD; ++q) {
      const int e = tid + q * 256;
      const int c = (e % (BN / VW)) * VW, k = e / (BN / VW);
      *reinterpret_cast<VT *>(&Bs[buf][k][c]) = rb[q];
    }
  };

  // FP32 MFMA sums are added in FP64 after every chunk (BK products), or
  // for the FP32 contraction after every second chunk (2 BK products)
  constexpr bool ACC64 = sizeof(T) == 4;
  for (int r0 = r_lo; r0 < r_hi; r0 += BM) {
    V acc[4][2];
    double acc64[ACC64 ? 4 : 1][2][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = V{0, 0, 0, 0};
    if (ACC64) {
#pragma unroll
      for (int a = 0; a < (ACC64 ? 4 : 1); ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
    }
    const int kend = PLAIN ? P : min(P, r0 + BM);
    int nch = (kend + BK - 1) / BK;
    // TRI with a map of the 16-row chunks of delta that hold a nonzero
    // value in this cell tile: only those chunks are multiplied (a zero
    // chunk adds exactly nothing)
    const bool skip = !PLAIN && BK == 16 && nzb != nullptr && nch <= 256;
    if (skip) {
      const uint8_t *nz = nzb + ((int64_t)sl * gridDim.y + cti) * nzk;
      const bool f = tid < nch && nz[tid] != 0;
      const unsigned long long bm = __ballot(f);
      if (lane == 0) s_cnt[wave] = __popcll(bm);
      __syncthreads();
      int off = 0;
      for (int w2 = 0; w2 < wave; ++w2) off += s_cnt[w2];
      if (f) s_list[off + __popcll(bm & ((1ull << lane) - 1ull))] = tid;
      const int tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
      __syncthreads();
      nch = tot;
    }
    // executed chunks, counted when profiling
    if (!PLAIN && nexec && tid == 0)
      atomicAdd(nexec, (unsigned long long)nch);
    auto kchunk = [&](int ch) { return skip ? s_list[ch] * BK : ch * BK; };
    // multiply chunk ch out of LDS buffer ch & 1 (FP32: then add to FP64)
    auto mma_chunk = [&](int buf, bool zs = true, bool fl = true) {
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        const int ka = kk * 4 + lg;
        T af[4], bf[2];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) af[ti] = As[buf][ka][ti * 16 + li];
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
          bf[tj] = Bs[buf][ka][wave * 32 + tj * 16 + li];
        // FP32 chunk sums start from zero (the previous chunk's went to FP64)
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj)
            acc[ti][tj] = MF::op(af[ti], bf[tj],
                                 ACC64 && zs && kk == 0 ? V{0, 0, 0, 0}
                                                        : acc[ti][tj]);
      }
      if (ACC64 && fl) {
#pragma unroll
        for (int ti = 0; ti < (ACC64 ? 4 : 1); ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              acc64[ti][tj][r] += (double)acc[ti][tj][r];
          }
      }
    };
    load_tiles(ra, rb, r0, nch > 0 ? kchunk(0) : 0, kend);
    __syncthreads();  // previous row block's readers are done with LDS
    store_tiles(ra, rb, 0);
    if constexpr (PF2) {
      // stage ra holds chunk ch + 1 on entry to chunk ch, even ch
      if (nch > 1) load_tiles(ra, rb, r0, kchunk(1), kend);
      __syncthreads();
      auto step = [&](int ch, auto &la, auto &lb, const auto &sa,
                      const auto &sb) {
        if (ch + 2 < nch) load_tiles(la, lb, r0, kchunk(ch + 2), kend);
        // FP32 sums of two chunks (32 products) start from zero and then
        // go to FP64
        mma_chunk(ch & 1, (ch & 1) == 0, (ch & 1) || ch + 1 >= nch);
        if (ch + 1 < nch) store_tiles(sa, sb, (ch & 1) ^ 1);
        __syncthreads();
      };
      for (int ch = 0; ch < nch; ch += 2) {
        step(ch, ra2, rb2, ra, rb);
        if (ch + 1 < nch) step(ch + 1, ra, rb, ra2, rb2);
      }
    } else {
      __syncthreads();
      for (int ch = 0; ch < nch; ++ch) {
        if (ch + 1 < nch) load_tiles(ra, rb, r0, kchunk(ch + 1), kend);
        mma_chunk(ch & 1);
        if (ch + 1 < nch) store_tiles(ra, rb, (ch & 1) ^ 1);
        __syncthreads();
      }
    }
    // epilogue for this row block
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const int64_t col = c0 + wave * 32 + tj * 16 + li;
      if (col >= Cs) continue;
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + ti * 16 + MF::row(lg, r);
          const double v = ACC64 ? acc64[ACC64 ? ti : 0][tj][r]
                                 : (double)acc[ti][tj][r];
          if (PLAIN) {
            if (row < n) {
              acc_e[tj] += v * z[row];
              acc_l[tj] += v;
            }
            if (want_var && row <= n)
              acc_v[tj] += v * (double)B[(int64_t)row * ldb + col];
            if (want_var && row == n) acc_m[tj] += v;
          } else {
            if (row < n) acc_v[tj] += v * v;
          }
        }
      }
    }
  }
  // the four lane groups of a wave share each column
#pragma unroll
  for (int tj = 0; tj < 2; ++tj) {
    acc_e[tj] += __shfl_xor(acc_e[tj], 16, 64);
    acc_e[tj] += __shfl_xor(acc_e[tj], 32, 64);
    acc_l[tj] += __shfl_xor(acc_l[tj], 16, 64);
    acc_l[tj] += __shfl_xor(acc_l[tj], 32, 64);
    acc_v[tj] += __shfl_xor(acc_v[tj], 16, 64);
    acc_v[tj] += __shfl_xor(acc_v[tj], 32, 64);
    acc_m[tj] += __shfl_xor(acc_m[tj], 16, 64);
    acc_m[tj] += __shfl_xor(acc_m[tj], 32, 64);
  }
  if (lg != 0) return;
  double *eo = est + (int64_t)sl * ldo;
  double *vo = var ? var + (int64_t)sl * ldo : nullptr;
  bool ill = false;
#pragma unroll
  for (int tj = 0; tj < 2; ++tj) {
    const int64_t col = c0 + wave * 32 + tj * 16 + li;
    if (col >= Cs) continue;
    if (PLAIN) {
      // residual check: the weights of an OK / EDK system sum to one
      // exactly, so their sum's distance from one is rounding of the
      // explicit inverse
      if (method != SPK_M_SK && fabs(acc_l[tj] - 1.0) > kSumTol) ill = true;
      const int64_t oc = cmap ? (int64_t)cmap[col] : col;
      if (skm) {
        // SK with the step mean m as known mean: m + lambda^T (z - m)
        eo[oc] = acc_e[tj] + st.mean * (1.0 - acc_l[tj]);
      } else if (isclose1(acc_l[tj])) {
        eo[oc] = acc_e[tj];
        if (want_var) vo[oc] = acc_v[tj] + acc_m[tj];
      } else {
        eo[oc] = z[nnb[(int64_t)sl * ldn + col]];
        if (want_var) vo[oc] = 0.0;
      }
    } else {
      // this row block's share of |L^-1 delta|^2
      qpart[((int64_t)sl * nrb + rbi) * ldab + col] = acc_v[tj];
    }
  }
  if (PLAIN && __ballot(ill) != 0ull && li == 0 &&
      !(st.flags & SPK_ST_ILLCOND))
    atomicOr(&sc[sl].flags, SPK_ST_ILLCOND);
}

// OK variance lambda.g + 2 mu (interp/steps.py:432-434): the FP64 base
// terms minus the deflated quadratic form, summed over row blocks in a
// fixed order.
__global__ __launch_bounds__(256) void k_var_reduce(
    const StepSc *__restrict__ sc, const double *__restrict__ abuf,
    const double *__restrict__ qpart, int64_t ldq, int nrb, int64_t Cs,
    const int32_t *__restrict__ perm, double *__restrict__ var, int64_t ldo) {
  const int sl = blockIdx.y;
  const StepSc *st = sc + sl;
  if (st->status) return;
  const int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (c >= Cs) return;
  const int nb = (st->n + 63) / 64;
  const double *q = qpart + (int64_t)sl * nrb * ldq + c;
  double acc = 0.0;
  for (int r = 0; r < nb; ++r) acc += q[(int64_t)r * ldq];
  var[(int64_t)sl * ldo + (perm ? (int64_t)perm[c] : c)] =
      abuf[(int64_t)sl * ldq + c] - acc;
}

// ---- FP64 triangular contraction, 128 x 128 tiles -----------------------------
// q_part = |rows r0..r0+127 of L^-1 delta|^2 per cell, L^-1 lower
// triangular (column-major, lda), delta row-major [n x ldb].  Block tile
// 128 rows x 128 cells, k in chunks of 16 (K runs to the tile's last row);
// 4 waves in 2 x 2, each 64 rows x 64 cells = 4 x 4 MFMA f64 16x16x4 tiles
// held in 64 accumulator registers.  The next chunk is loaded into
// registers while the current one is multiplied out of LDS.  Each wave
// writes the partial norm of its 64 rows to qpart[(step * nrb + rb) * ldq]
// with rb = 2 * (row tile) + (wave row), so the 64-row partials of the
// generic kernel and this one have the same layout.
// LOOP: grid (row tiles, cell tiles, 1), each block taking the tile of
// every one of the nsteps steps that the selection keeps in turn, for
// launches where few steps or none are selected (the mixed path's
// promoted steps): a block per (tile, step) that only exits still costs
// its dispatch.
template <bool LOOP>
__global__ __launch_bounds__(256, 2) void k_tri64(
    const double *__restrict__ A, int64_t lda, int64_t strideA,
    const double *__restrict__ B, int64_t ldb, int64_t strideB,
    const StepSc *__restrict__ sc, int64_t Cs, double *__restrict__ qpart,
    int64_t ldq, int nrb, int sel, int nsteps) {
  constexpr int BM = 128, BN = 128, BK = 16, PAD = 2;
  const int G = gridDim.x * gridDim.y;
  int lin = blockIdx.x + blockIdx.y * gridDim.x;
  if ((G & 7) == 0) lin = (lin & 7) * (G >> 3) + (lin >> 3);
  const int rti = lin % gridDim.x, cti = lin / gridDim.x;
  __shared__ __align__(16) double As[2][BK][BM + PAD];
  __shared__ __align__(16) double Bs[2][BK][BN + PAD];
  auto tile = [&](const int sl) {
  const StepSc st = sc[sl];
  if (st.status) return;
  if (sel == 1 && (st.flags & (SPK_ST_FP64VAR | SPK_ST_INT8VAR))) return;
  if (sel == 2 && !(st.flags & SPK_ST_FP64VAR)) return;
  const int n = st.n;
  const int r0 = rti * BM;
  if (r0 >= n) return;
  const int64_t c0 = cti * (int64_t)BN;
  if (c0 >= Cs) return;
  const int kend = min(n, r0 + BM);
  const double *Ap = A + (int64_t)sl * strideA;
  const double *Bp = B + (int64_t)sl * strideB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;
  // global -> register staging: A chunk = 16 columns x 128 rows, B chunk =
  // 16 rows x 128 cells; each thread moves 4 double2 of each
  double2 ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + q * 256;          // 0..1023 double2 slots
      const int r = (e & 63) * 2, k = e >> 6;  // A: row pair r, column k
      const int gk = k0 + k;
      // rows past n are zero and never read (lda is a multiple of 64 above
      // n, so a pair starting below n lies inside the column)
      if (gk < kend && r0 + r < n) {
        double2 v = *reinterpret_cast<const double2 *>(
            Ap + (r0 + r) + (int64_t)gk * lda);
        if (r0 + r + 1 >= n) v.y = 0.0;
        ra[q] = v;
      } else {
        ra[q] = double2{0.0, 0.0};
      }
      const int cc = (e & 63) * 2, kb = e >> 6;  // B: row kb, cell pair cc
      const int gkb = k0 + kb;
      rb[q] = gkb < kend ? *reinterpret_cast<const double2 *>(
                               Bp + (int64_t)gkb * ldb + c0 + cc)
                         : double2{0.0, 0.0};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + q * 256;
      const int r = (e & 63) * 2, k = e >> 6;
      *reinterpret_cast<double2 *>(&As[buf][k][r]) = ra[q];
      *reinterpret_cast<double2 *>(&Bs[buf][k][r]) = rb[q];
    }
  };
  d4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4_t{0, 0, 0, 0};
  const int nch = (kend + BK - 1) / BK;
  load(0);
  store(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) load((ch + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int ka = kk * 4 + lg;
      double af[4], bf[4];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) af[ti] = As[buf][ka][wr * 64 + ti * 16 + li];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) bf[tj] = Bs[buf][ka][wc * 64 + tj * 16 + li];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[ti], bf[tj],
                                                             acc[ti][tj], 0,
                                                             0, 0);
    }
    if (ch + 1 < nch) store(buf ^ 1);
    __syncthreads();
  }
  // squared norms of this wave's 64 rows, per cell
  double part[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = r0 + wr * 64 + ti * 16 + lg + 4 * r;
      if (row < n) {
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) part[tj] += acc[ti][tj][r] * acc[ti][tj][r];
      }
    }
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    part[tj] += __shfl_xor(part[tj], 16, 64);
    part[tj] += __shfl_xor(part[tj], 32, 64);
  }
  const int prow = 2 * rti + wr;
  if (lg != 0 || prow >= nrb) return;
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    const int64_t col = c0 + wc * 64 + tj * 16 + li;
    if (col < Cs) qpart[((int64_t)sl * nrb + prow) * ldq + col] = part[tj];
  }
  };
  if (!LOOP) {
    tile(blockIdx.z);
    return;
  }
  // the selected steps (at most kMaxStepsPerBatch), read at once by the
  // first wave; the block then takes their tiles in step order
  __shared__ int s_sel[kMaxStepsPerBatch];
  __shared__ int s_cnt;
  if (threadIdx.x < 64) {
    const int t = threadIdx.x;
    bool keep = false;
    if (t < nsteps && t < kMaxStepsPerBatch) {
      const StepSc *p = sc + t;
      const int fl = p->flags;
      keep = p->status == 0 && rti * 128 < p->n &&
             !(sel == 1 && (fl & (SPK_ST_FP64VAR | SPK_ST_INT8VAR))) &&
             !(sel == 2 && !(fl & SPK_ST_FP64VAR));
    }
    const unsigned long long m = __ballot(keep);
    if (keep) s_sel[__popcll(m & ((1ull << t) - 1ull))] = t;
    if (t == 0) s_cnt = __popcll(m);
  }
  __syncthreads();
  const int cnt = s_cnt;
  for (int i = 0; i < cnt; ++i) {
    // the previous step's tile is done with LDS
    if (i > 0) __syncthreads();
    tile(s_sel[i]);
  }
}
template __global__ void k_tri64<false>(const double *, int64_t, int64_t,
                                        const double *, int64_t, int64_t,
                                        const StepSc *, int64_t, double *,
                                        int64_t, int, int, int);
template __global__ void k_tri64<true>(const double *, int64_t, int64_t,
                                       const double *, int64_t, int64_t,
                                       const StepSc *, int64_t, double *,
                                       int64_t, int, int, int);

// ---- batched path: fused distance -> covariance -------------------------------
// For each active station i and cell c: k_ic = c0 - G(d_ic) (OK/EDK) or the
// covariance sum (SK).  The estimate k.alpha (+ constraint terms) is summed
// in FP64.  SK keeps the reference's sum-of-weights test (k.w == 1).
//
// OK variance.  With q = k^T C^-1 k and a = k.w the reference's
// lambda.g + 2 mu equals c0 - q + (a - 1)(a - 2) / b.  q is split with a
// local fit of k on the columns of C at the cell's nearest active stations
// J (up to kJ): for any beta and delta = k - C_J beta,
//     q = beta^T C_JJ beta + 2 beta^T delta_J + |L^-1 delta|^2,
// because C^-1 C_J = E_J.  beta solves C_JJ beta = k_J in FP64, so delta is
// the residual of local kriging, small near stations (where the variance is
// small) and wherever J screens the cell; that keeps the FP32 contraction
// term small next to the variance.  All other terms are FP64; abuf receives
// c0 - (beta terms) + (a-1)(a-2)/b and delta goes to K32 (the FP32
// contraction) or K64 (steps flagged SPK_ST_FP64VAR or SPK_ST_INT8VAR, and
// the FP64 batched path), with each cell's max |delta| in dmax for the int8
// planes.  Any beta gives the same variance in exact arithmetic, so J is
// cut where C_JJ's Cholesky pivot falls below 1e-8 of its diagonal.
// Cells come in a spatially compact order.  The nearest stations of a
// block's 256 cells fill a table of up to kSlots stations whose G columns
// are staged in LDS per station chunk; each cell's J is its kJ nearest
// stations in that table, so every cell gets a local fit even when the
// block's cells have more distinct nearest stations than slots.
// Cells are stripe-local indices c; with perm the outputs and the cell
// drifts are addressed at the original index perm[c].
// kChunk: stations staged per pass of k_kgen's main loop
constexpr int kJ = 4, kSlots = 16, kChunk = 128;

template <bool GEN>
__device__ inline double kg_vg(const VgFast &f, double h, int covar_flag,
                               double min_vg_val) {
  return GEN ? vg_fast_entry(f, h, covar_flag, min_vg_val)
             : vg_simple_entry(f, h, covar_flag, min_vg_val);
}

// 1 where a step's variogram has no power or hole-effect component
__global__ void k_vg_kinds(const spk_vg *__restrict__ vgs, int64_t T,
                           uint8_t *__restrict__ simple) {
  for (int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x; t < T;
       t += (int64_t)gridDim.x * 256)
    simple[t] = vg_simple(vgs[t]) ? 1 : 0;
}

// Squared cell-station distance of the nearest-station search, one
// rounding per operation (no fused multiply-add), so that every kernel
// that orders stations by it orders them alike.
__device__ inline double near_d2(double x, double y, double sx, double sy) {
#pragma clang fp contract(off)
  const double dx = x - sx, dy = y - sy;
  return (dx * dx) + (dy * dy);
}

// The kNear stations nearest to each cell over all N stations, in
// ascending (near_d2, rank) order, rank being the station's position in
// sorder (the order in which k_compact lists a step's active stations), as
// global station indices list[j * ldl + c] (-1 past N).  Once per call:
// for a step, the first listed station that is active is the cell's
// nearest active station with ties to the lower compacted index, unless
// none of the listed ones is active.  Stations are staged in rank order
// and inserted with a strict comparison, so equal distances keep rank
// order.  Grid ceil(C / 256), 256 threads.
constexpr int kNear = 32;
__global__ __launch_bounds__(256) void k_near_list(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ sorder, int64_t N,
    const double *__restrict__ cx, const double *__restrict__ cy, int64_t C,
    int32_t *__restrict__ list, int64_t ldl) {
  __shared__ double stx[256], sty[256];
  __shared__ int32_t sid[256];
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x * (int64_t)256 + tid;
  const bool valid = c < C;
  const double x = valid ? cx[c] : 0.0, y = valid ? cy[c] : 0.0;
  double d[kNear];
  int id[kNear];
#pragma unroll
  for (int k = 0; k < kNear; ++k) {
    d[k] = INFINITY;
    id[k] = -1;
  }
  for (int64_t base = 0; base < N; base += 256) {
    __syncthreads();
    if (base + tid < N) {
      const int g = sorder ? sorder[base + tid] : (int)(base + tid);
      stx[tid] = sx[g];
      sty[tid] = sy[g];
      sid[tid] = g;
    }
    __syncthreads();
    const int cnt = (int)min((int64_t)256, N - base);
    if (valid)
      for (int j = 0; j < cnt; ++j) {
        const double d2 = near_d2(x, y, stx[j], sty[j]);
        if (!(d2 < d[kNear - 1])) continue;
        const int g = sid[j];
        // sorted insert: slot k takes slot k - 1's entry when the new one
        // goes above it, else the new one when it goes above slot k
#pragma unroll
        for (int k = kNear - 1; k > 0; --k) {
          const double dk = d[k], dkm = d[k - 1];
          const int ik = id[k], ikm = id[k - 1];
          const bool up = d2 < dkm;
          d[k] = up ? dkm : (d2 < dk ? d2 : dk);
          id[k] = up ? ikm : (d2 < dk ? g : ik);
        }
        if (d2 < d[0]) {
          d[0] = d2;
          id[0] = g;
        }
      }
  }
  if (!valid) return;
#pragma unroll
  for (int k = 0; k < kNear; ++k) list[(int64_t)k * ldl + c] = id[k];
}

// pos[step][g]: the compacted index of station g in the step (idx), -1
// where it is not active.  Grid (steps), 256 threads.
__global__ __launch_bounds__(256) void k_stn_pos(
    const int32_t *__restrict__ idx, int64_t N,
    const StepSc *__restrict__ sc, int32_t *__restrict__ pos) {
  const int s = blockIdx.x;
  int32_t *ps = pos + (int64_t)s * N;
  const int32_t *my = idx + (int64_t)s * N;
  for (int64_t g = threadIdx.x; g < N; g += 256) ps[g] = -1;
  __syncthreads();
  const int n = sc[s].status == SPK_ST_NOSTNS ? 0 : sc[s].n;
  for (int q = threadIdx.x; q < n; q += 256) ps[my[q]] = q;
}

// Nearest active station of every cell of a stripe (ties to the lower
// index), for the local fit of k_kgen: the first active station of the
// cell's list from k_near_list (list, pos of k_stn_pos), or, for the
// blocks where a cell has no active station among its listed ones, a scan
// of the step's stations.  With promote set (mixed path) a step is moved
// to the FP64 contraction (SPK_ST_FP64VAR) when any block of 256 cells has
// more distinct nearest stations than k_kgen stages (kSlots): cells whose
// nearest station is not staged get a poorer local fit, a larger residual
// delta and so a larger FP32 rounding term.  The block partition is the
// one k_kgen uses.
__global__ __launch_bounds__(256) void k_nearest(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, int64_t N, StepSc *__restrict__ sc,
    const double *__restrict__ cx, const double *__restrict__ cy, int64_t Cs,
    int32_t *__restrict__ nrst, int64_t ldn, int promote,
    const int32_t *__restrict__ list, int64_t ldl,
    const int32_t *__restrict__ pos) {
  const int sl = blockIdx.y;
  StepSc *stp = sc + sl;
  if (stp->status) return;
  const int n = stp->n;
  const int32_t *my = idx + (int64_t)sl * N;
  __shared__ double stx[256], sty[256];
  __shared__ int slot_j[kSlots];
  __shared__ int ovf;
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x * (int64_t)256 + tid;
  const bool valid = c < Cs;
  if (tid < kSlots) slot_j[tid] = -1;
  if (tid == 0) ovf = 0;
  int bj = -1;
  if (valid && list) {
    const int32_t *ps = pos + (int64_t)sl * N;
    for (int k = 0; k < kNear; ++k) {
      const int g = list[(int64_t)k * ldl + c];
      if (g < 0) break;
      const int q = ps[g];
      if (q >= 0) {
        bj = q;
        break;
      }
    }
  }
  if (__syncthreads_or(valid && bj < 0 ? 1 : 0)) {
    const bool scan = valid && bj < 0;
    const double x = valid ? cx[c] : 0.0, y = valid ? cy[c] : 0.0;
    double bd = INFINITY;
    for (int base = 0; base < n; base += 256) {
      __syncthreads();
      if (base + tid < n) {
        const int g = my[base + tid];
        stx[tid] = sx[g];
        sty[tid] = sy[g];
      }
      __syncthreads();
      const int cnt = min(256, n - base);
      if (scan)
        for (int j = 0; j < cnt; ++j) {
          const double d2 = near_d2(x, y, stx[j], sty[j]);
          if (d2 < bd) {
            bd = d2;
            bj = base + j;
          }
        }
    }
  }
  if (valid) nrst[(int64_t)sl * ldn + c] = bj;
  if (!promote) return;
  __syncthreads();
  if (valid && bj >= 0) {
    bool placed = false;
    for (int k = 0; k < kSlots && !placed; ++k) {
      const int cur = atomicCAS(&slot_j[k], -1, bj);
      placed = cur == -1 || cur == bj;
    }
    if (!placed) ovf = 1;
  }
  __syncthreads();
  if (tid == 0 && ovf) atomicOr(&stp->flags, SPK_ST_FP64VAR);
}

template <bool GEN>
__global__ __launch_bounds__(256, 4) void k_kgen(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, int64_t N,
    const double *__restrict__ zact, int64_t ldz,
    StepSc *__restrict__ sc, const spk_vg *__restrict__ vgs,
    const double *__restrict__ cx, const double *__restrict__ cy, int64_t Cs,
    int64_t C, const double *__restrict__ cell_drift,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ nrst,
    int64_t ldn,
    const double *__restrict__ V, int64_t ldv, int64_t ld, int method, int m,
    double min_vg_val, int want_var, int all64, int skm,
    float *__restrict__ K32,
    double *__restrict__ K64, int64_t ldk, uint4 *__restrict__ Kb,
    const double *__restrict__ Gfull,
    double *__restrict__ abuf, double *__restrict__ est, int64_t ldo,
    double *__restrict__ dmax, uint8_t *__restrict__ nzb, int64_t nzt,
    int64_t nzk, int zskip) {
  const int sl = blockIdx.y;
  const StepSc *stp = sc + sl;
  if (stp->status) return;
  const int n = stp->n;
  const int32_t *my = idx + (int64_t)sl * N;
  const double *alpha = V + (int64_t)sl * ldv;
  const double *w = alpha + ld;
  const double *dal = w + ld;   // refinement correction of alpha (k_dual)
  __shared__ double stx[kChunk], sty[kChunk], sal[kChunk], sw[kChunk];
  __shared__ double sdal[kChunk];
  // one double of padding per slot row: lanes reading different slots at
  // the same station hit different banks
  __shared__ double sG[kSlots][kChunk + 1];
  __shared__ int slot_j[kSlots];
  __shared__ int n_slots;
  __shared__ VgFast vf;
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x * (int64_t)256 + tid;
  const bool valid = c < Cs;
  const double x = valid ? cx[c] : 0.0, y = valid ? cy[c] : 0.0;
  if (tid == 0) {
    vg_fast_init(vgs[sl], vf);
    n_slots = 0;
  }
  if (tid < kSlots) slot_j[tid] = -1;
  const int cflag = method == SPK_M_SK ? 1 : 0;
  const double c0 = stp->c0;
  const bool var_on = want_var && method == SPK_M_OK;
  // FP64 delta for the FP64 and the int8 contractions
  const bool use64 =
      all64 || (stp->flags & (SPK_ST_FP64VAR | SPK_ST_INT8VAR));
  // FP32-tier delta: FP32 (K32), or its three bfloat16 planes (Kb: octet
  // o of cell c of plane p at (p ld / 8 + o) ldk + c)
  float *K32s = (var_on && !use64 && !Kb) ? K32 + (int64_t)sl * ld * ldk
                                          : nullptr;
  uint4 *Kbs = (var_on && !use64 && Kb) ? Kb + (int64_t)sl * 3 * (ld / 8) * ldk + c
                                        : nullptr;
  const int64_t kbp = (ld / 8) * ldk;  // plane stride of Kbs
  double *K64s = (var_on && use64) ? K64 + (int64_t)sl * ld * ldk : nullptr;
  const double *Gs = var_on ? Gfull + (int64_t)sl * ld * ld : nullptr;
  double beta[kJ];
  int slt[kJ];
#pragma unroll
  for (int q = 0; q < kJ; ++q) {
    beta[q] = 0.0;
    slt[q] = -1;
  }
  double qloc = 0.0;
  if (var_on) {
    // the cell's nearest active station, from k_nearest
    const int bj = valid ? nrst[(int64_t)sl * ldn + c] : -1;
    __syncthreads();  // slot_j and n_slots initialised
    // the block's table of staged stations: the cells' nearest stations,
    // first come first served
    if (valid && bj >= 0) {
      for (int k = 0; k < kSlots; ++k) {
        const int cur = atomicCAS(&slot_j[k], -1, bj);
        if (cur == -1 || cur == bj) {
          if (cur == -1) atomicAdd(&n_slots, 1);
          break;
        }
      }
    }
    __syncthreads();
    const int ns = n_slots;
    if (tid < ns) {
      const int g = my[slot_j[tid]];
      stx[tid] = sx[g];
      sty[tid] = sy[g];
    }
    __syncthreads();
    // J: the cell's kJ nearest stations among the table (ascending)
    double jd[kJ];
#pragma unroll
    for (int q = 0; q < kJ; ++q) jd[q] = INFINITY;
    if (valid)
      for (int k = 0; k < ns; ++k) {
        const double dx = x - stx[k], dy = y - sty[k];
        double d2 = dx * dx + dy * dy;
        int id = k;
#pragma unroll
        for (int q = 0; q < kJ; ++q) {  // insertion, strict: stable
          if (d2 < jd[q]) {
            const double td = jd[q];
            const int ti = slt[q];
            jd[q] = d2;
            slt[q] = id;
            d2 = td;
            id = ti;
          }
        }
      }
    // local system on J in distance order: Cholesky of C_JJ = c0 - G_JJ,
    // cut at the first small pivot
    double Lj[kJ][kJ], kJv[kJ];
    int mm = 0;
#pragma unroll
    for (int q = 0; q < kJ; ++q) {
      kJv[q] = 0.0;
#pragma unroll
      for (int r = 0; r < kJ; ++r) Lj[q][r] = 0.0;
    }
    if (valid) {
      bool ok = true;
#pragma unroll
      for (int q = 0; q < kJ; ++q) {
        if (!ok || slt[q] < 0) {
          ok = false;
          slt[q] = -1;
          continue;
        }
        const int jq = slot_j[slt[q]];
        const double dq = c0 - Gs[(int64_t)jq * ld + jq];
#pragma unroll
        for (int r = 0; r < kJ; ++r) {
          if (r < q) {
            double v = c0 - Gs[(int64_t)slot_j[slt[r]] * ld + jq];
#pragma unroll
            for (int t2 = 0; t2 < kJ; ++t2)
              if (t2 < r) v -= Lj[q][t2] * Lj[r][t2];
            Lj[q][r] = v / Lj[r][r];
          }
        }
        double v = dq;
#pragma unroll
        for (int t2 = 0; t2 < kJ; ++t2)
          if (t2 < q) v -= Lj[q][t2] * Lj[q][t2];
        if (v > 1e-8 * dq && dq > 0.0) {
          Lj[q][q] = sqrt(v);
          kJv[q] = c0 - kg_vg<GEN>(vf, dist2(x, y, stx[slt[q]], sty[slt[q]]),
                                   0, min_vg_val);
          mm = q + 1;
        } else {
          ok = false;
          slt[q] = -1;
        }
      }
      // beta = C_JJ^-1 k_J; q_loc = 2 beta.k_J - beta^T C_JJ beta
      double t[kJ];
#pragma unroll
      for (int q = 0; q < kJ; ++q) {
        t[q] = 0.0;
        if (q < mm) {
          double v = kJv[q];
#pragma unroll
          for (int r = 0; r < kJ; ++r)
            if (r < q) v -= Lj[q][r] * t[r];
          t[q] = v / Lj[q][q];
        }
      }
      double tt = 0.0;
#pragma unroll
      for (int q = 0; q < kJ; ++q) tt += t[q] * t[q];
#pragma unroll
      for (int q = kJ - 1; q >= 0; --q) {
        if (q < mm) {
          double v = t[q];
#pragma unroll
          for (int r = 0; r < kJ; ++r)
            if (r > q && r < mm) v -= Lj[r][q] * beta[r];
          beta[q] = v / Lj[q][q];
        }
      }
      double bk = 0.0;
#pragma unroll
      for (int q = 0; q < kJ; ++q) bk += beta[q] * kJv[q];
      qloc = 2.0 * bk - tt;
    }
  }
  // slots of the local fit for the delta update; an unused one (slt -1)
  // has beta 0 and reads slot 0, which subtracts an exact zero
  int slq[kJ];
#pragma unroll
  for (int q = 0; q < kJ; ++q) slq[q] = slt[q] >= 0 ? slt[q] : 0;
  double dm = 0.0;  // max |delta| of the cell (int8 plane scaling)
  VgOne vo;
  const int one = vg_one_of(vf, vo);
  // Culling.  For a single bounded model whose far value nug + s equals c0
  // (and is above min_vg_val), k = c0 - gamma(h) is exactly 0 beyond a
  // distance: h >= r (spherical), h c1 <= -42 (exponential) or h^2 c1 <= -42
  // (Gaussian), where exp() < 2^-54 so that 1 - exp() rounds to 1.  A
  // sub-chunk of 16 stations whose bounding box is that far from every cell
  // of the block (the gap measured by dist2 itself, so that by monotone
  // rounding no cell-station distance the loop computes is below it), and
  // whose staged c0 - G entries against the block's slot stations are all
  // 0, gives k = 0 and delta = 0 exactly for every cell: it adds nothing to
  // the estimate, its delta chunk stays unmarked in the chunk map, and it is
  // not evaluated.  Its delta is not written either where every consumer
  // of this step's delta skips the unmarked chunks (zskip, and not an FP64
  // step); otherwise zeros are written.
  __shared__ double s_bb[4][4];  // per wave, then the block's in row 0
  __shared__ int s_cull[kChunk / 16];
  const bool cull_on = var_on && one >= 0 && c0 == vo.nug + vo.s &&
                       vo.nug + vo.s > min_vg_val;
  const bool zwrite = !(zskip && !all64 && !(stp->flags & SPK_ST_FP64VAR));
  if (cull_on) {
    double v0 = valid ? x : INFINITY, v1 = valid ? x : -INFINITY;
    double v2 = valid ? y : INFINITY, v3 = valid ? y : -INFINITY;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      v0 = fmin(v0, __shfl_xor(v0, off, 64));
      v1 = fmax(v1, __shfl_xor(v1, off, 64));
      v2 = fmin(v2, __shfl_xor(v2, off, 64));
      v3 = fmax(v3, __shfl_xor(v3, off, 64));
    }
    if ((tid & 63) == 0) {
      s_bb[tid >> 6][0] = v0;
      s_bb[tid >> 6][1] = v1;
      s_bb[tid >> 6][2] = v2;
      s_bb[tid >> 6][3] = v3;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
      for (int w2 = 1; w2 < 4; ++w2) {
        s_bb[0][0] = fmin(s_bb[0][0], s_bb[w2][0]);
        s_bb[0][1] = fmax(s_bb[0][1], s_bb[w2][1]);
        s_bb[0][2] = fmin(s_bb[0][2], s_bb[w2][2]);
        s_bb[0][3] = fmax(s_bb[0][3], s_bb[w2][3]);
      }
    }
    // read after the station loop's first barrier
  }
  double best = INFINITY;
  int bi = 0;
  double e = 0.0, a = 0.0;
  // the estimate's change by the last refinement step
  double ke = 0.0;
  // delta_i = k_i - sum_q (c0 - G_{i,J_q}) beta_q
  for (int base = 0; base < n; base += kChunk) {
    __syncthreads();
    if (tid < kChunk && base + tid < n) {
      const int g = my[base + tid];
      stx[tid] = sx[g];
      sty[tid] = sy[g];
      sal[tid] = alpha[base + tid];
      sw[tid] = w[base + tid];
      sdal[tid] = dal[base + tid];
    }
    const int cnt = min(kChunk, n - base);
    if (var_on) {
      const int ns = n_slots;
      for (int e2 = tid; e2 < ns * kChunk; e2 += 256) {
        const int k = e2 / kChunk, u = e2 % kChunk;
        // staged as covariances c0 - G: a station out of range of the cell
        // and of every J station then gives delta = 0 exactly
        if (u < cnt) sG[k][u] = c0 - Gs[(int64_t)slot_j[k] * ld + base + u];
      }
    }
    __syncthreads();
    if (cull_on) {
      if (tid < kChunk) {
        const bool have = tid < cnt;
        double u0 = have ? stx[tid] : INFINITY;
        double u1 = have ? stx[tid] : -INFINITY;
        double u2 = have ? sty[tid] : INFINITY;
        double u3 = have ? sty[tid] : -INFINITY;
        int z = 1;
        if (have)
          for (int k = 0; k < n_slots; ++k) z &= sG[k][tid] == 0.0 ? 1 : 0;
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) {
          u0 = fmin(u0, __shfl_xor(u0, off, 64));
          u1 = fmax(u1, __shfl_xor(u1, off, 64));
          u2 = fmin(u2, __shfl_xor(u2, off, 64));
          u3 = fmax(u3, __shfl_xor(u3, off, 64));
          z &= __shfl_xor(z, off, 64);
        }
        if ((tid & 15) == 0) {
          const double gx =
              fmax(0.0, fmax(u0 - s_bb[0][1], s_bb[0][0] - u1));
          const double gy =
              fmax(0.0, fmax(u2 - s_bb[0][3], s_bb[0][2] - u3));
          const double D = dist2(gx, gy, 0.0, 0.0);
          bool far;
          if (one == SPK_VG_SPH)
            far = D >= vo.r;
          else if (one == SPK_VG_EXP)
            far = D * vo.c1 <= -42.0;
          else
            far = (D * D) * vo.c1 <= -42.0;
          s_cull[tid >> 4] = (z && far) ? 1 : 0;
        }
      }
      __syncthreads();
    }
    unsigned nzf = 0;  // 16-station chunks of this pass with delta != 0
    if (valid) {
      if (var_on) {
        // the station loop, specialised on a single bounded model, over
        // sub-chunks of 16 stations
        auto var_loop = [&](auto tag) {
          constexpr int MT = decltype(tag)::value;
          // delta of station j of the pass (and its estimate terms)
          auto eval = [&](int j) {
            const double d = dist2(x, y, stx[j], sty[j]);
            double g;
            if constexpr (MT < 0)
              g = kg_vg<GEN>(vf, d, 0, min_vg_val);
            else
              g = vg_one<MT>(vo, d, min_vg_val);
            const double k = c0 - g;
            e += k * sal[j];
            a += k * sw[j];
            ke += k * sdal[j];
            double delta = k;
#pragma unroll
            for (int q = 0; q < kJ; ++q)
              delta -= sG[slq[q]][j] * beta[q];
            dm = fmax(dm, fabs(delta));
            if (delta != 0.0) nzf |= 1u << (j >> 4);
            return delta;
          };
          for (int j0 = 0; j0 < cnt; j0 += 16) {
            const int j1 = min(cnt, j0 + 16);
            if (MT >= 0 && cull_on && s_cull[j0 >> 4]) {
              if (zwrite) {
                if (Kbs) {
                  const int64_t o = (base + j0) >> 3;
#pragma unroll
                  for (int p = 0; p < 3; ++p) {
                    Kbs[p * kbp + o * ldk] = uint4{0u, 0u, 0u, 0u};
                    Kbs[p * kbp + (o + 1) * ldk] = uint4{0u, 0u, 0u, 0u};
                  }
                } else {
                  for (int j = j0; j < j1; ++j) {
                    if (K32s)
                      K32s[(int64_t)(base + j) * ldk + c] = 0.0f;
                    else
                      K64s[(int64_t)(base + j) * ldk + c] = 0.0;
                  }
                }
              }
              continue;
            }
            if (Kbs) {
              // both octets of the 16 stations, zero beyond the pass
              for (int h = 0; h < 2; ++h) {
                const int jb = j0 + 8 * h;
                uint32_t hw[4], mw[4], lw[4];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                  const int j = jb + u;
                  const float df = j < j1 ? (float)eval(j) : 0.0f;
                  uint16_t hh, mm2, ll;
                  split3f(df, hh, mm2, ll);
                  const int sh = (u & 1) * 16;
                  if (!(u & 1)) hw[u / 2] = mw[u / 2] = lw[u / 2] = 0u;
                  hw[u / 2] |= (uint32_t)hh << sh;
                  mw[u / 2] |= (uint32_t)mm2 << sh;
                  lw[u / 2] |= (uint32_t)ll << sh;
                }
                const int64_t o = (base + jb) >> 3;
                Kbs[o * ldk] = uint4{hw[0], hw[1], hw[2], hw[3]};
                Kbs[kbp + o * ldk] = uint4{mw[0], mw[1], mw[2], mw[3]};
                Kbs[2 * kbp + o * ldk] = uint4{lw[0], lw[1], lw[2], lw[3]};
              }
            } else {
              for (int j = j0; j < j1; ++j) {
                const double delta = eval(j);
                if (K32s)
                  K32s[(int64_t)(base + j) * ldk + c] = (float)delta;
                else
                  K64s[(int64_t)(base + j) * ldk + c] = delta;
              }
            }
          }
        };
        switch (one) {
          case SPK_VG_SPH:
            var_loop(std::integral_constant<int, SPK_VG_SPH>{});
            break;
          case SPK_VG_EXP:
            var_loop(std::integral_constant<int, SPK_VG_EXP>{});
            break;
          case SPK_VG_GAU:
            var_loop(std::integral_constant<int, SPK_VG_GAU>{});
            break;
          default:
            var_loop(std::integral_constant<int, -1>{});
        }
      } else {
        for (int j = 0; j < cnt; ++j) {
          const double d = dist2(x, y, stx[j], sty[j]);
          if (d < best || (d == best && my[base + j] < my[bi])) {
            best = d;
            bi = base + j;
          }
          const double g = kg_vg<GEN>(vf, d, cflag, min_vg_val);
          const double k = cflag ? g : c0 - g;
          e += k * sal[j];
          a += k * sw[j];
          ke += k * sdal[j];
        }
      }
    }
    // the chunk map of this wave's 64 cells (one 128-cell tile of the
    // contraction): any lane's nonzero delta marks the chunk
    if (var_on && nzb) {
      unsigned m = nzf;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) m |= __shfl_xor(m, off, 64);
      if ((tid & 63) == 0 && m) {
        const int64_t cw = blockIdx.x * (int64_t)256 + (tid & ~63);
        uint8_t *nz = nzb + ((int64_t)sl * nzt + (cw >> 7)) * nzk +
                      (base >> 4);
        for (int b2 = 0; b2 < kChunk / 16; ++b2)
          if ((m >> b2) & 1u) nz[b2] = 1;
      }
    }
  }
  if (!valid) return;
  const int64_t oc = perm ? (int64_t)perm[c] : c;
  double *eo = est + (int64_t)sl * ldo;
  const double *z = zact + (int64_t)sl * ldz;
  // Residual check: the second refinement step changed this cell's
  // estimate by de, of the size of the error the first one left, and the
  // result is at least as good; a change above kRefineTol of the cell's
  // scale (its estimate, at least 1e-3 of the step's mean |value|) marks
  // the step ill-conditioned: no FP64 solve of it is reproducible to 1e-6.
  // (The variance is far less sensitive: in the dual form its error stays
  // near 1e-9 where the estimate's reaches 1e-6.)
  double de = ke;
  double ev;
  if (method == SPK_M_OK) {
    ev = e + stp->a0;
    de += stp->dh[0];
  } else if (method == SPK_M_EDK) {
    ev = e + stp->h[0];
    de += stp->dh[0];
    for (int k = 0; k < m; ++k) {
      const double f = cell_drift[(int64_t)k * C + oc];
      ev += stp->h[1 + k] * f;
      de += stp->dh[1 + k] * f;
    }
  } else if (skm) {
    // SK with the step mean m as known mean: m + k^T C^-1 (z - m 1)
    // = k.alpha + m (1 - k.w)
    ev = e + stp->mean * (1.0 - a);
  } else {  // the reference's SK
    ev = isclose1(a) ? e : z[bi];
  }
  const bool ill = fabs(de) > kRefineTol * fmax(fabs(ev), 1e-3 * stp->zabs);
  if (ill && !(stp->flags & SPK_ST_ILLCOND))
    atomicOr(&sc[sl].flags, SPK_ST_ILLCOND);
  eo[oc] = ev;
  if (method == SPK_M_OK) {
    if (var_on)
      abuf[(int64_t)sl * ldk + c] =
          c0 - qloc + (a - 1.0) * (a - 2.0) / stp->b;
    if (var_on && dmax) dmax[(int64_t)sl * ldk + c] = dm;
  }
}

// ---- IDW and nearest neighbour -------------------------------------------------
__global__ __launch_bounds__(256) void k_idw_nnb(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, int64_t N,
    const double *__restrict__ zact, int64_t ldz,
    const StepSc *__restrict__ sc, const double *__restrict__ cx,
    const double *__restrict__ cy, int64_t C, int method, double p,
    double *__restrict__ est, int *__restrict__ bad) {
#pragma clang fp contract(off)
  const int sl = blockIdx.y;
  const StepSc st = sc[sl];
  if (st.status) return;
  const int n = st.n;
  const int32_t *my = idx + (int64_t)sl * N;
  const double *z = zact + (int64_t)sl * ldz;
  __shared__ double stx[256], sty[256], sz[256];
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x * (int64_t)256 + tid;
  const bool valid = c < C;
  const double x = valid ? cx[c] : 0.0, y = valid ? cy[c] : 0.0;
  double *eo = est + (int64_t)sl * C;
  if (method != SPK_M_NNB && p == 2.0) {
    // p = 2: the reference's weights 1 / (d / dmax)^2 are dmax^2 / s with
    // s = dx^2 + dy^2, and the per-cell factor dmax^2 cancels in
    // sum(w z) / sum(w); one pass with w = 1 / s (no square root, one
    // division).  The weight-sum check is made on the normalised sum
    // max(s) sum(1 / s).
    double ws = 0.0, wz = 0.0, smax = 0.0;
    for (int base = 0; base < n; base += 256) {
      __syncthreads();
      if (base + tid < n) {
        const int g = my[base + tid];
        stx[tid] = sx[g];
        sty[tid] = sy[g];
        sz[tid] = z[base + tid];
      }
      __syncthreads();
      const int cnt = min(256, n - base);
      if (valid)
        for (int j = 0; j < cnt; ++j) {
          const double dx = x - stx[j], dy = y - sty[j];
          const double s2 = (dx * dx) + (dy * dy);
          smax = fmax(smax, s2);
          const double wj = 1.0 / s2;
          ws = ws + wj;
          wz = wz + wj * sz[j];
        }
    }
    if (!valid) return;
    const double wn = smax > 0.0 ? ws * smax : ws;
    if (!(wn >= 1e-14)) atomicOr(bad, 1);  // interp/steps.py:305
    eo[c] = wz / ws;
    return;
  }
  double dmax = 0.0, best = INFINITY;
  int bi = 0;
  // pass 1: max distance (IDW) / nearest station (NNB)
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    if (base + tid < n) {
      const int g = my[base + tid];
      stx[tid] = sx[g];
      sty[tid] = sy[g];
      sz[tid] = z[base + tid];
    }
    __syncthreads();
    const int cnt = min(256, n - base);
    if (valid)
      for (int j = 0; j < cnt; ++j) {
        const double d = dist2(x, y, stx[j], sty[j]);
        dmax = fmax(dmax, d);
        if (d < best) {
          best = d;
          bi = base + j;
        }
      }
  }
  if (method == SPK_M_NNB) {
    if (valid) eo[c] = z[bi];
    return;
  }
  // pass 2: w = 1 / (d / dmax)^p, sums in station order (pyx:784-808)
  double ws = 0.0, wz = 0.0;
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    if (base + tid < n) {
      const int g = my[base + tid];
      stx[tid] = sx[g];
      sty[tid] = sy[g];
      sz[tid] = z[base + tid];
    }
    __syncthreads();
    const int cnt = min(256, n - base);
    if (valid)
      for (int j = 0; j < cnt; ++j) {
        double d = dist2(x, y, stx[j], sty[j]);
        if (dmax > 0) d = d / dmax;
        const double wj = 1.0 / (p == 2.0 ? d * d : pow(d, p));
        ws = ws + wj;
        wz = wz + wj * sz[j];
      }
  }
  if (!valid) return;
  if (!(ws >= 1e-14)) atomicOr(bad, 1);  // interp/steps.py:305
  eo[c] = wz / ws;
}

// ---- special steps ---------------------------------------------------------------
__global__ __launch_bounds__(256) void k_special(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, int64_t N,
    const double *__restrict__ zact, int64_t ldz,
    const StepSc *__restrict__ sc, const double *__restrict__ cx,
    const double *__restrict__ cy, int64_t C, double *__restrict__ est,
    double *__restrict__ var) {
  const int sl = blockIdx.y;
  const StepSc st = sc[sl];
  if (st.status == 0 || st.status == SPK_ST_NOSTNS) return;
  const int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x;
  const double *z = zact + (int64_t)sl * ldz;
  double *eo = est + (int64_t)sl * C;
  double *vo = var ? var + (int64_t)sl * C : nullptr;
  if (st.status == SPK_ST_ONESTN) {
    if (c < C) {
      eo[c] = z[0];
      if (vo) vo[c] = NAN;  // est_vars stays NaN (interp/steps.py:282-283)
    }
    return;
  }
  if (st.status == SPK_ST_MEANFILL) {
    if (c < C) {
      eo[c] = st.mean;
      if (vo) vo[c] = 0.0;
    }
    return;
  }
  // failed system: nearest active station (interp/steps.py:372-384)
  const int32_t *my = idx + (int64_t)sl * N;
  __shared__ double stx[256], sty[256];
  const int n = st.n;
  const bool valid = c < C;
  const double x = valid ? cx[c] : 0.0, y = valid ? cy[c] : 0.0;
  double best = INFINITY;
  int bi = 0;
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    if (base + (int)threadIdx.x < n) {
      stx[threadIdx.x] = sx[my[base + threadIdx.x]];
      sty[threadIdx.x] = sy[my[base + threadIdx.x]];
    }
    __syncthreads();
    const int cnt = min(256, n - base);
    // ties go to the lower station index, whatever order the step's
    // stations are compacted in
    if (valid)
      for (int j = 0; j < cnt; ++j) {
        const double d = dist2(x, y, stx[j], sty[j]);
        if (d < best || (d == best && my[base + j] < my[bi])) {
          best = d;
          bi = base + j;
        }
      }
  }
  if (!valid) return;
  eo[c] = z[bi];
  if (vo) vo[c] = 0.0;
}

// n ints set to v (a kernel, not hipMemsetAsync, so that a caller may
// capture it in a graph)
__global__ void k_fill_int(int *__restrict__ p, int n, int v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

__global__ void k_status_out(const StepSc *__restrict__ sc, int S,
                             int32_t *__restrict__ info,
                             double *__restrict__ kappa) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) {
    info[s] = sc[s].status | sc[s].flags;
    if (kappa) kappa[s] = sc[s].kappa;
  }
}

__global__ void k_clip_cast(const double *__restrict__ in,
                            float *__restrict__ out, int64_t n, int has_min,
                            double mn, int has_max, double mx) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = in[i];
    if (has_min && v < mn) v = mn;
    if (has_max && v > mx) v = mx;
    out[i] = (float)v;
  }
}

// ---- kernel profiling (hipEvents around the heavy launches) ----------------
Prof &prof_of(spk_ctx *ctx) {
  static std::vector<std::pair<spk_ctx *, Prof *>> all;
  for (auto &kv : all)
    if (kv.first == ctx) return *kv.second;
  all.push_back({ctx, new Prof()});
  return *all.back().second;
}

hipEvent_t prof_event(Prof &pf) {
  if (!pf.pool.empty()) {
    hipEvent_t e = pf.pool.back();
    pf.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

static int prof_collect(spk_ctx *ctx) {
  Prof &pf = prof_of(ctx);
  if (pf.pending.empty()) return SPK_OK;
  SPK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->aux_stream) SPK_HIP(hipStreamSynchronize(ctx->aux_stream));
  for (auto &r : pf.pending) {
    float ms = 0.f;
    SPK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    bool found = false;
    for (auto &kv : pf.stats)
      if (kv.first == r.name) {
        kv.second.first += ms;
        kv.second.second += 1;
        found = true;
      }
    if (!found) pf.stats.push_back({r.name, {(double)ms, (int64_t)1}});
    pf.pool.push_back(r.a);
    pf.pool.push_back(r.b);
  }
  pf.pending.clear();
  return SPK_OK;
}

static int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

}  // namespace spk

using namespace spk;

// IDW / nearest neighbour over T steps, S steps per launch
// (interp/steps.py:285-313, and its special steps :280-283): the kernels
// of spk_interp_steps_dev and spk_idw_steps_noalloc_dev.  sx.. cy are the
// re-centred coordinates; idx [S x N], zact [S x ldz], sc [S] and bad [1]
// are the caller's scratch.  The weight-sum flag is cleared by a fill
// kernel, and read back at the end (the reference asserts, steps.py:305).
static int idw_nnb_steps(spk_ctx *ctx, const spk_opts *opts, int64_t T,
                         int64_t N, int64_t C, const double *vals,
                         const double *sx, const double *sy,
                         const double *cx, const double *cy, int32_t *idx,
                         double *zact, int64_t ldz, StepSc *sc, int64_t S,
                         int *bad, double *est, int32_t *info,
                         double *kappa) {
  hipStream_t sm = ctx->stream;
  const int method = opts->method;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, sm, bad, 1, 0);
  for (int64_t t0 = 0; t0 < T; t0 += S) {
    const int Sb = (int)std::min<int64_t>(S, T - t0);
    double *eb = est + t0 * C;
    hipLaunchKernelGGL(k_compact, dim3(Sb), dim3(256), 0, sm, vals + t0 * N,
                       N, idx, zact, ldz, sc, (const spk_vg *)nullptr,
                       method, opts->min_var_thr, opts->min_vg_val,
                       (const int32_t *)nullptr);
    {
      ProfScope ps(ctx, "idw_nnb");
      if (method == SPK_M_IDW)
        // num = (Z . M) W^T and den = M W^T on FP64 MFMA, W = d^-p
        hipLaunchKernelGGL(k_idw_mfma,
                           dim3((unsigned)ceil_div(C, 128),
                                (unsigned)ceil_div(Sb, 32)),
                           dim3(256), 0, sm, sx, sy, N, vals + t0 * N, Sb,
                           (const StepSc *)sc, (const int32_t *)idx,
                           (const double *)zact, ldz, cx, cy, C,
                           opts->idw_exp, eb, bad);
      else
        hipLaunchKernelGGL(k_idw_nnb, dim3((unsigned)ceil_div(C, 256), Sb),
                           dim3(256), 0, sm, sx, sy, idx, N, zact, ldz, sc,
                           cx, cy, C, method, opts->idw_exp, eb, bad);
    }
    hipLaunchKernelGGL(k_special, dim3((unsigned)ceil_div(C, 256), Sb),
                       dim3(256), 0, sm, sx, sy, idx, N, zact, ldz, sc, cx,
                       cy, C, eb, (double *)nullptr);
    hipLaunchKernelGGL(k_status_out, dim3((unsigned)ceil_div(Sb, 256)),
                       dim3(256), 0, sm, sc, Sb, info + t0,
                       kappa ? kappa + t0 : nullptr);
    SPK_TRY(check_launch("idw/nnb"));
  }
  int hbad = 0;
  SPK_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, sm));
  SPK_HIP(hipStreamSynchronize(sm));
  if (hbad)
    return set_error(SPK_EARG, "IDW weight sum below 1e-14 "
                     "(interp/steps.py:305)");
  return SPK_OK;
}

extern "C" {

int spk_prof_enable(spk_ctx *ctx, int on) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  Prof &pf = prof_of(ctx);
  pf.on = on != 0;
  if (pf.on && !pf.work) {
    SPK_HIP(hipSetDevice(ctx->device));
    SPK_TRY(ws_get(ctx, "prof_work", 64, (void **)&pf.work));
    SPK_HIP(hipMemsetAsync(pf.work, 0, 64, ctx->stream));
    SPK_TRY(sync(ctx));
  }
  return SPK_OK;
}

int spk_prof_reset(spk_ctx *ctx) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  SPK_TRY(prof_collect(ctx));
  Prof &pf = prof_of(ctx);
  pf.stats.clear();
  pf.counts.clear();
  if (pf.work) {
    SPK_HIP(hipMemsetAsync(pf.work, 0, 64, ctx->stream));
    SPK_TRY(sync(ctx));
  }
  return SPK_OK;
}

int spk_prof_work(spk_ctx *ctx, const char *name, double *value) {
  if (!ctx || !name || !value) return set_error(SPK_EARG, "null argument");
  Prof &pf = prof_of(ctx);
  *value = 0.0;
  const std::string nm(name);
  int slot;
  double per;
  if (nm == "gemm_tri_f32_macs") {
    slot = 0;  // 16 k x 64 rows x 128 cells multiply-adds per chunk
    per = 16.0 * 64.0 * 128.0;
  } else if (nm == "gemm_tri_bf16_macs") {
    slot = 2;  // 32 k x 64 rows x 128 cells x 6 plane products per chunk
    per = 32.0 * 64.0 * 128.0 * 6.0;
  } else if (nm == "gemm_tri_i8_ops") {
    slot = 1;  // 2 x 64 k x 128 rows x 64 cells ops per chunk and product
    per = 2.0 * 64.0 * 128.0 * 64.0;
  } else {
    for (auto &kv : pf.counts)
      if (kv.first == nm) {
        *value = kv.second;
        return SPK_OK;
      }
    return set_error(SPK_EARG, "unknown work counter %s", name);
  }
  if (!pf.work) return SPK_OK;
  SPK_TRY(prof_collect(ctx));
  unsigned long long cnt = 0;
  SPK_TRY(d2h(ctx, &cnt, pf.work + slot, sizeof(cnt)));
  SPK_TRY(sync(ctx));
  *value = (double)cnt * per;
  return SPK_OK;
}

int spk_prof_get(spk_ctx *ctx, const char *name, double *total_ms,
                 int64_t *count) {
  if (!ctx || !name || !total_ms || !count)
    return set_error(SPK_EARG, "null argument");
  SPK_TRY(prof_collect(ctx));
  *total_ms = 0.0;
  *count = 0;
  for (auto &kv : prof_of(ctx).stats)
    if (kv.first == name) {
      *total_ms = kv.second.first;
      *count = kv.second.second;
    }
  return SPK_OK;
}

int spk_clip_cast_dev(spk_ctx *ctx, const double *in64, float *out32,
                      int64_t n, int has_min, double min_cut, int has_max,
                      double max_cut) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (n <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  int64_t g = std::min<int64_t>(ceil_div(n, 256), 65536);
  hipLaunchKernelGGL(k_clip_cast, dim3((unsigned)g), dim3(256), 0, ctx->stream,
                     in64, out32, n, has_min, min_cut, has_max, max_cut);
  return check_launch("k_clip_cast");
}

// Float32 host fields of rows t0 .. t0 + rows - 1 (est clipped as
// interp/steps.py:466-476, var cast): cast on stream st into staging
// (2 x rows x C floats) and copied to the host arrays on st.
static int f32_out_rows(spk_ctx *ctx, hipStream_t st, const spk_f32_out *o,
                        const double *est, const double *var, int64_t C,
                        int64_t t0, int64_t rows, float *stage) {
  const int64_t n = rows * C;
  const unsigned g = (unsigned)std::min<int64_t>(ceil_div(n, 256), 65536);
  hipLaunchKernelGGL(k_clip_cast, dim3(g), dim3(256), 0, st, est + t0 * C,
                     stage, n, o->has_min, o->min_cut, o->has_max,
                     o->max_cut);
  if (o->var32 && var)
    hipLaunchKernelGGL(k_clip_cast, dim3(g), dim3(256), 0, st, var + t0 * C,
                       stage + n, n, 0, 0.0, 0, 0.0);
  SPK_TRY(check_launch("k_clip_cast"));
  SPK_HIP(hipMemcpyAsync(o->est32 + t0 * C, stage, (size_t)n * sizeof(float),
                         hipMemcpyDeviceToHost, st));
  if (o->var32 && var)
    SPK_HIP(hipMemcpyAsync(o->var32 + t0 * C, stage + n,
                           (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                           st));
  return SPK_OK;
}

static int interp_steps_impl(spk_ctx *ctx, const spk_problem *prob,
                             const spk_opts *opts, int64_t T,
                             const double *vals, const spk_vg *vgs,
                             double *est, double *var, int32_t *info,
                             double *kappa, const spk_f32_out *o32) {
  if (!ctx || !prob || !opts) return set_error(SPK_EARG, "null argument");
  if (T < 0) return set_error(SPK_ESHAPE, "negative step count");
  if (T == 0) return SPK_OK;
  const int64_t N = prob->n_stns, C = prob->n_cells;
  if (N < 1 || C < 1) return set_error(SPK_ESHAPE, "need stations and cells");
  if (N > (int64_t)1 << 20)
    return set_error(SPK_ESHAPE, "more than 2^20 stations");
  if (!vals || !est || !info || !prob->stn_x || !prob->stn_y ||
      !prob->cell_x || !prob->cell_y)
    return set_error(SPK_EARG, "null device pointer");
  const int method = opts->method;
  if (method < SPK_M_OK || method > SPK_M_NNB)
    return set_error(SPK_EARG, "unknown method %d", method);
  const bool krig = method <= SPK_M_EDK;
  if (krig && !vgs) return set_error(SPK_EARG, "kriging needs variograms");
  const int m = method == SPK_M_EDK ? prob->n_drifts : 0;
  if (method == SPK_M_EDK &&
      (m < 1 || m > kMaxP - 1 || !prob->stn_drift || !prob->cell_drift))
    return set_error(SPK_EARG, "EDK needs 1..%d drifts", kMaxP - 1);
  const int prec = opts->precision;
  if (prec < SPK_P_FP64 || prec > SPK_P_MIXED_F32)
    return set_error(SPK_EARG, "unknown precision %d", prec);
  const bool want_var = opts->want_var && method == SPK_M_OK && var;
  const int skm = method == SPK_M_SK && opts->sk_mean ? 1 : 0;
  SPK_HIP(hipSetDevice(ctx->device));
  hipStream_t sm = ctx->stream;

  // re-centred coordinates (exact for origins on the coordinates' grid)
  double *sx, *sy, *cx, *cy;
  SPK_TRY(ws_get(ctx, "stn_xy", 2 * N * sizeof(double), (void **)&sx));
  sy = sx + N;
  SPK_TRY(ws_get(ctx, "cell_xy", 2 * C * sizeof(double), (void **)&cx));
  cy = cx + C;
  {
    const unsigned gs = (unsigned)std::min<int64_t>(ceil_div(N, 256), 4096);
    const unsigned gc = (unsigned)std::min<int64_t>(ceil_div(C, 256), 65536);
    hipLaunchKernelGGL(k_recenter, dim3(gs), dim3(256), 0, sm, prob->stn_x, sx,
                       N, prob->origin_x);
    hipLaunchKernelGGL(k_recenter, dim3(gs), dim3(256), 0, sm, prob->stn_y, sy,
                       N, prob->origin_y);
    hipLaunchKernelGGL(k_recenter, dim3(gc), dim3(256), 0, sm, prob->cell_x, cx,
                       C, prob->origin_x);
    hipLaunchKernelGGL(k_recenter, dim3(gc), dim3(256), 0, sm, prob->cell_y, cy,
                       C, prob->origin_y);
    SPK_TRY(check_launch("k_recenter"));
  }
  const int64_t ldz = round_up(N, 64);
  const int p = method == SPK_M_OK ? 1 : method == SPK_M_EDK ? 1 + m : 0;
  const int64_t ld = round_up(N + (p > 0 ? p : 1), 64);
  const int64_t strideM = ld * ld;
  // the mixed path; its FP32-tier steps contract split-BF16 planes
  // (SPK_P_MIXED, systems up to kBfMaxLd) or FP32 operands
  const bool mixed = prec == SPK_P_MIXED || prec == SPK_P_MIXED_F32;
  const bool bfop = prec == SPK_P_MIXED && ld <= kBfMaxLd;
  const double thr = opts->min_var_thr;
  const double mvg = opts->min_vg_val;
  // ---- IDW / nearest neighbour: no systems ---------------------------------
  if (!krig) {
    const int64_t S = std::min<int64_t>(T, 512);
    int32_t *idx;
    double *zact;
    StepSc *sc;
    int *bad;
    SPK_TRY(ws_get(ctx, "bad", 64, (void **)&bad));
    SPK_TRY(ws_get(ctx, "idx0", S * N * sizeof(int32_t), (void **)&idx));
    SPK_TRY(ws_get(ctx, "zact0", S * ldz * sizeof(double), (void **)&zact));
    SPK_TRY(ws_get(ctx, "sc0", S * sizeof(StepSc), (void **)&sc));
    SPK_TRY(idw_nnb_steps(ctx, opts, T, N, C, vals, sx, sy, cx, cy, idx,
                          zact, ldz, sc, S, bad, est, info, kappa));
    if (!o32) return SPK_OK;
    // the float32 fields after the whole call, in slices of the staging
    const int64_t rows =
        std::max<int64_t>(1, std::min<int64_t>(T, ((int64_t)256 << 20) / C));
    float *stage;
    SPK_TRY(ws_get(ctx, "f32_stage", (size_t)2 * rows * C * sizeof(float),
                   (void **)&stage));
    for (int64_t t0 = 0; t0 < T; t0 += rows)
      SPK_TRY(f32_out_rows(ctx, sm, o32, est, nullptr, C, t0,
                           std::min<int64_t>(rows, T - t0), stage));
    return SPK_OK;
  }

  // ---- kriging --------------------------------------------------------------
  // Batches of S steps.  The systems of batch b+1 are factorised on a second
  // stream while the cells of batch b are contracted on the caller's stream;
  // the per-batch system workspaces are double-buffered.
  const bool batched = prec != SPK_P_FP64;
  const int64_t per_step =
      strideM * 8 * 2 +
      (prec == SPK_P_FP64 ? strideM * 8
                          : (want_var ? strideM * 8 : 0) +
                                (mixed && want_var
                                     ? strideM * (bfop ? 6 : 4)
                                     : 0));
  const int64_t budget = (int64_t)20 << 30;
  int64_t S = std::max<int64_t>(
      1, std::min<int64_t>(budget / per_step, kMaxStepsPerBatch));
  S = std::min<int64_t>(S, T);
  const int nsets = T > S ? 2 : 1;
  struct Set {
    int32_t *idx;
    double *zact, *A, *Li, *M, *V, *T0, *Gm;
    float *L32;
    uint4 *Lb;   // split-BF16 planes of L^-1 (bfop)
    int8_t *L8;  // int8 planes of L^-1 rows (mixed path)
    int *ea;     // their row exponents
    StepSc *sc;
    // the batch's variogram records, copied so that the captured
    // factorisation sequences read a buffer whose address is the same
    // for every batch of the set
    spk_vg *vg;
    // compacted index of every station in each step, -1 where inactive
    // (k_stn_pos; read by k_nearest)
    int32_t *spos;
    // steps of the batch whose Cholesky failed: solved by the
    // pseudo-inverse fallback (spk_pinv.hip) and the plain contraction
    std::vector<int> pinv;
    // the steps each variance contraction takes (split-BF16 or FP32
    // operands, int8 with 5 and with 6 planes) and the largest n of each
    StepList lst32, lst8a, lst8b;
    int n32, n8a, n8b;
  } set[2] = {};
  // pinned host copy of a batch's step records, read after its Cholesky
  {
    const size_t need = (size_t)S * sizeof(StepSc);
    if (ctx->h_sc_bytes < need) {
      if (ctx->h_sc) SPK_HIP(hipHostFree(ctx->h_sc));
      ctx->h_sc = nullptr;
      ctx->h_sc_bytes = 0;
      SPK_HIP(hipHostMalloc(&ctx->h_sc, need));
      ctx->h_sc_bytes = need;
    }
  }
  StepSc *h_sc = static_cast<StepSc *>(ctx->h_sc);
  // int8 plane layout: rows padded to R (a multiple of the 128-row tile),
  // k to KS (a multiple of the 64-deep MFMA)
  const int64_t R8 = round_up(N, 128), KS8 = round_up(N, 64);
  for (int k = 0; k < nsets; ++k) {
    char nm[32];
    Set &q = set[k];
    auto get = [&](const char *base, size_t bytes, void **out) {
      snprintf(nm, sizeof(nm), "%s%d", base, k);
      return ws_get(ctx, nm, bytes, out);
    };
    SPK_TRY(get("idx", S * N * sizeof(int32_t), (void **)&q.idx));
    SPK_TRY(get("zact", S * ldz * sizeof(double), (void **)&q.zact));
    SPK_TRY(get("sc", S * sizeof(StepSc), (void **)&q.sc));
    SPK_TRY(get("vgb", S * sizeof(spk_vg), (void **)&q.vg));
    q.spos = nullptr;
    if (batched && want_var && method == SPK_M_OK)
      SPK_TRY(get("spos", S * N * sizeof(int32_t), (void **)&q.spos));
    SPK_TRY(get("A", S * strideM * 8, (void **)&q.A));
    SPK_TRY(get("Li", S * strideM * 8, (void **)&q.Li));
    SPK_TRY(get("T0", S * ld * kT0Cols * 8, (void **)&q.T0));
    if (!batched) {
      SPK_TRY(get("M", S * strideM * 8, (void **)&q.M));
    } else {
      SPK_TRY(get("V", S * 4 * ld * 8, (void **)&q.V));
      if (want_var) SPK_TRY(get("Gm", S * strideM * 8, (void **)&q.Gm));
      if (mixed && want_var) {
        if (bfop)
          SPK_TRY(get("Lb", S * strideM * 6, (void **)&q.Lb));
        else
          SPK_TRY(get("L32", S * strideM * 4, (void **)&q.L32));
        SPK_TRY(get("L8", S * kInt8Planes * R8 * KS8, (void **)&q.L8));
        SPK_TRY(get("ea", S * R8 * sizeof(int), (void **)&q.ea));
      }
    }
  }
  // cell stripes for the materialised right-hand sides / deflated covariances:
  // the budget is an eighth of the device's memory, at least 12 GiB and at
  // most 48 GiB (36 GB of a 288 GB MI355X).  Longer stripes mean fewer
  // launches and fewer partly idle launch tails; a cell's results do not
  // depend on the stripe length (stripes are whole 128-cell tiles).
  int64_t Cs = C;
  size_t rhs_budget = (size_t)12 << 30;
  {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess)
      rhs_budget = std::min<size_t>(std::max<size_t>(rhs_budget, tot / 8),
                                    (size_t)48 << 30);
  }
  if (!batched) {
    Cs = std::min<int64_t>(C, (int64_t)(rhs_budget / ((N + p + 64) * 8)));
  } else if (want_var) {
    // bytes per (step, station, cell): FP64 delta, plus the FP32 delta
    // or its three bfloat16 planes and the int8 planes in the mixed path
    const int64_t eb = mixed ? (bfop ? 14 : 12) + kInt8Planes : 8;
    Cs = std::min<int64_t>(C, (int64_t)(2 * rhs_budget / (S * ld * eb)));
  }
  // whole 256-cell blocks of k_kgen: a block's cells share a table of
  // nearest stations for their local fits, so with stripes of whole blocks
  // a cell's block, and its results, do not depend on the stripe length
  Cs = std::max<int64_t>(256, Cs / 256 * 256);
  if (Cs > C) Cs = C;
  const int64_t ldk = round_up(Cs, 128);
  const int nrb = (int)ceil_div(N, 64);
  double *R = nullptr, *abuf = nullptr, *K64 = nullptr, *qpart = nullptr;
  double *kap_lam = nullptr;  // the condition estimate's two eigenvalues
  double *dual_ds = nullptr;  // per-step scratch of the dual-vector kernels
  float *K32 = nullptr;
  uint4 *Kb = nullptr;
  int32_t *nnb = nullptr, *nrst = nullptr;
  int8_t *B8 = nullptr;
  int *eb8 = nullptr;
  double *dmax = nullptr;
  // map of the 16-row chunks of delta with a nonzero value, per step and
  // 128-cell tile (mixed path): the FP32 contraction skips the others
  uint8_t *nzb = nullptr;
  const int64_t nzk = ld / 16;
  // map of the 64 x 64 (cells x stations) blocks whose int8 planes hold a
  // nonzero digit, written by k_slice_cols and read by k_tri8
  uint8_t *nz8 = nullptr;
  // every consumer of the mixed path's delta (FP32 and int8 contractions,
  // int8 slicing) skips the unmarked chunks when the FP32 contraction's
  // chunk list (256 chunks of 16 rows) covers the triangle: kgen then
  // leaves culled chunks unwritten
  const int zskip = mixed && ceil_div(N, 16) <= 256 ? 1 : 0;
  // every batched call estimates each step's condition (k_kap_*)
  if (batched) {
    SPK_TRY(ws_get(ctx, "kap_lam", S * 2 * 8, (void **)&kap_lam));
    SPK_TRY(ws_get(ctx, "dual_ds", S * kDualSc * 8, (void **)&dual_ds));
  }
  if (!batched) {
    SPK_TRY(ws_get(ctx, "R", (N + p + 64) * ldk * 8, (void **)&R));
    SPK_TRY(ws_get(ctx, "nnb", ldk * 4, (void **)&nnb));
  } else if (want_var) {
    // FP64 buffer for the batched FP64 path and for the steps the mixed
    // path promotes to FP64 (SPK_ST_FP64VAR)
    SPK_TRY(ws_get(ctx, "K64", S * ld * ldk * 8, (void **)&K64));
    if (mixed) {
      if (bfop)
        SPK_TRY(ws_get(ctx, "Kb", S * ld * ldk * 6, (void **)&Kb));
      else
        SPK_TRY(ws_get(ctx, "K32", S * ld * ldk * 4, (void **)&K32));
      SPK_TRY(ws_get(ctx, "B8", S * kInt8Planes * ldk * KS8, (void **)&B8));
      SPK_TRY(ws_get(ctx, "eb8", S * ldk * sizeof(int), (void **)&eb8));
      SPK_TRY(ws_get(ctx, "dmax", S * ldk * 8, (void **)&dmax));
      SPK_TRY(ws_get(ctx, "nzb", S * (ldk / 128) * nzk, (void **)&nzb));
      SPK_TRY(ws_get(ctx, "nz8", S * (ldk / 64) * (KS8 / 64), (void **)&nz8));
    }
    SPK_TRY(ws_get(ctx, "abuf", S * ldk * 8, (void **)&abuf));
    SPK_TRY(ws_get(ctx, "qpart", S * nrb * ldk * 8, (void **)&qpart));
    if (method == SPK_M_OK)
      SPK_TRY(ws_get(ctx, "nrst", S * ldk * 4, (void **)&nrst));
  }
  // SPK_P_MIXED: condition estimate above which a step's variance leaves
  // FP32.  The FP32 contraction's error is rounding in L^-1 delta amplified
  // by cancellation; it grows with the condition estimate and with the
  // size of delta against the variance.  Measured on 1 024 hourly steps
  // (2 000 stations, 500 m cells; error relative to the variance with a
  // floor of 1e-3 x the step's mean variance): spherical and exponential
  // models stay below 2.1e-7 up to the largest condition estimate seen
  // (7e4); Gaussian models stay below 5.5e-7 up to 2e3 and exceed 1e-6
  // from 2.4e3.  A step whose cells have more distinct nearest stations
  // per block than the local fit stages is promoted to FP64 by k_nearest.
  // The batched FP64 path runs every step in FP64; a caller's limit
  // applies to all models.
  const double kap = opts->fp64var_kappa;
  const double fp64var_thr =
      mixed ? (kap > 0.0 ? kap : 1e5) : INFINITY;
  const double fp64var_thr_gau =
      mixed ? (kap > 0.0 ? kap : 1e3) : INFINITY;
  // Above the FP32 limit and up to this one the variance uses exact int8
  // plane products (five 7-bit planes per operand and the plane products
  // up to the fourth diagonal up to condition estimate 1e7, six planes and
  // the fifth diagonal above).  Measured against the batched FP64
  // contraction on 768 hourly steps (500 m cells; relative, floor 1e-3 x
  // the step's mean variance): 4e-9 below condition estimate 1e3, 1.2e-8 to
  // 1e5, 7.8e-8 to 1e6, 3e-7 to 1e7 (five planes), 1e-7 to 1e10, 4.5e-7 to
  // 1e12 and 2.2e-6 above (six planes; above 1e12 the step is flagged
  // ill-conditioned).  By default every step above the FP32 limit uses the
  // planes; FP64 remains for steps whose local fit is short of stations
  // (k_nearest) and for a caller's limit.
  // The int32 diagonal sums stay exact while 6 * n * 127^2 < 2^31, i.e.
  // for up to 22 000 active stations; larger systems use FP64.
  const double k8 = opts->int8var_kappa;
  const double int8var_thr =
      !mixed || N > 20000
          ? -INFINITY
          : (k8 > 0.0 ? k8 : (k8 < 0.0 ? -INFINITY : INFINITY));

  // which steps have only simple variogram components (the cell kernels
  // are specialised for them)
  std::vector<uint8_t> vg_simple_h;
  if (batched) {
    uint8_t *vk;
    SPK_TRY(ws_get(ctx, "vg_kinds", T, (void **)&vk));
    hipLaunchKernelGGL(k_vg_kinds,
                       dim3((unsigned)std::min<int64_t>(ceil_div(T, 256), 4096)),
                       dim3(256), 0, sm, vgs, T, vk);
    SPK_TRY(check_launch("k_vg_kinds"));
    vg_simple_h.resize(T);
    SPK_TRY(d2h(ctx, vg_simple_h.data(), vk, T));
    SPK_TRY(sync(ctx));
  }

  // spatially compact cell order for the batched cell kernels
  int32_t *perm = nullptr;
  double *pcx = nullptr, *pcy = nullptr;
  if (batched && C >= 1024) {
    SPK_TRY(ws_get(ctx, "cell_perm", C * sizeof(int32_t), (void **)&perm));
    SPK_TRY(ws_get(ctx, "cell_pxy", 2 * C * sizeof(double), (void **)&pcx));
    pcy = pcx + C;
    SPK_TRY(cell_order(ctx, cx, cy, C, perm, pcx, pcy));
  }
  // the batched path compacts each step's stations in a spatially compact
  // order too, so that the stations near a block of cells are few runs of
  // the triangle's rows
  int32_t *sorder = nullptr;
  if (batched && N >= 2) {
    double *psx;
    SPK_TRY(ws_get(ctx, "stn_order", N * sizeof(int32_t), (void **)&sorder));
    SPK_TRY(ws_get(ctx, "stn_pxy", 2 * N * sizeof(double), (void **)&psx));
    SPK_TRY(cell_order(ctx, sx, sy, N, sorder, psx, psx + N));
  }
  // every cell's nearest stations over all of them, for k_nearest
  int32_t *near_list = nullptr;
  if (nrst) {
    SPK_TRY(ws_get(ctx, "near_list", (size_t)kNear * C * sizeof(int32_t),
                   (void **)&near_list));
    hipLaunchKernelGGL(k_near_list, dim3((unsigned)ceil_div(C, 256)),
                       dim3(256), 0, sm, sx, sy, (const int32_t *)sorder, N,
                       (const double *)(perm ? pcx : cx),
                       (const double *)(perm ? pcy : cy), C, near_list, C);
    SPK_TRY(check_launch("k_near_list"));
  }

  // second stream + events for the overlap
  if (!ctx->aux_stream)
    SPK_HIP(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
  for (int k = 0; k < 6; ++k)
    if (!ctx->ev[k])
      SPK_HIP(hipEventCreateWithFlags(&ctx->ev[k], hipEventDisableTiming));
  // the factorisations run on the second stream also for a single batch:
  // their graphs are captured there, never on the caller's stream (which
  // may be the null stream)
  hipStream_t sf = ctx->aux_stream;
  hipEvent_t ev_fact[2] = {ctx->ev[0], ctx->ev[1]};
  hipEvent_t ev_done[2] = {ctx->ev[2], ctx->ev[3]};
  // batched path: a batch's factorisation up to the pseudo-inverse
  // fallback, which only the contraction of the failed steps waits for
  hipEvent_t ev_pre[2] = {ctx->ev[4], ctx->ev[5]};
  // the factor stream must see the re-centred coordinates
  if (sf != sm) {
    SPK_HIP(hipEventRecord(ev_done[0], sm));
    SPK_HIP(hipStreamWaitEvent(sf, ev_done[0], 0));
  }

  auto factorise = [&](int64_t t0, int Sb, Set &q, hipEvent_t pre) -> int {
    const spk_vg *gb = vgs + t0;
    SPK_HIP(hipMemcpyAsync(q.vg, gb, (size_t)Sb * sizeof(spk_vg),
                           hipMemcpyDeviceToDevice, sf));
    hipLaunchKernelGGL(k_compact, dim3(Sb), dim3(256), 0, sf, vals + t0 * N,
                       N, q.idx, q.zact, ldz, q.sc, gb, method, thr, mvg,
                       (const int32_t *)sorder);
    SPK_TRY(check_launch("k_compact"));
    if (q.spos) {
      hipLaunchKernelGGL(k_stn_pos, dim3(Sb), dim3(256), 0, sf, q.idx, N,
                         q.sc, q.spos);
      SPK_TRY(check_launch("k_stn_pos"));
    }
    const int tiles = (int)ceil_div(N, 64);
    {
      ProfScope ps(ctx, "assemble", sf);
      hipLaunchKernelGGL(k_assemble, dim3(tiles * tiles, Sb), dim3(256), 0, sf,
                         sx, sy, q.idx, N, q.sc, (const spk_vg *)q.vg,
                         method == SPK_M_SK ? 1 : 0, mvg, q.A, ld, strideM,
                         q.Gm);
    }
    SPK_TRY(check_launch("k_assemble"));
    // The factorisation's launch sequences (about 190 launches for the
    // Cholesky panels of 2 000 rows, 100 for the condition estimate) run
    // as captured graphs, one per sequence and set of buffers.
    GraphKey gk;
    gk << q.A << q.Li << q.sc << q.vg << q.T0 << ld << strideM << N << Sb
       << method;
    {
      ProfScope ps(ctx, "potrf", sf);
      SPK_TRY(graph_run(ctx, "potrf", gk.v, sf, [&]() -> int {
        // the systems have at most N rows: the panels of the largest
        if (method != SPK_M_SK)
          hipLaunchKernelGGL(k_chol_shift, dim3(Sb), dim3(256), 0, sf, q.A,
                             ld, strideM, q.sc, (const spk_vg *)q.vg);
        for (int k0 = 0; k0 < N; k0 += 32) {
          hipLaunchKernelGGL(k_chol_diag, dim3(1, Sb), dim3(256), 0, sf, q.A,
                             ld, strideM, q.sc, k0);
          const int64_t below = N - (k0 + 32);
          if (below <= 0) break;
          hipLaunchKernelGGL(k_chol_solve,
                             dim3((unsigned)ceil_div(below, 256), Sb),
                             dim3(256), 0, sf, q.A, ld, strideM, q.sc, k0);
          const int64_t nt = ceil_div(below, 64);
          hipLaunchKernelGGL(k_chol_trail,
                             dim3((unsigned)(nt * (nt + 1) / 2), Sb),
                             dim3(256), 0, sf, q.A, ld, strideM, q.sc, k0);
        }
        return check_launch("k_potrf");
      }));
    }
    // which steps failed the Cholesky (the host waits for this batch's
    // factorisation while the previous batch's cells contract)
    SPK_HIP(hipMemcpyAsync(h_sc, q.sc, (size_t)Sb * sizeof(StepSc),
                           hipMemcpyDeviceToHost, sf));
    SPK_HIP(hipStreamSynchronize(sf));
    q.pinv.clear();
    for (int s2 = 0; s2 < Sb; ++s2)
      if (h_sc[s2].status == SPK_ST_FAILED) q.pinv.push_back(s2);
    {
      ProfScope ps(ctx, "trtri", sf);
      const dim3 gt((unsigned)ceil_div(N, 64), Sb);
      hipLaunchKernelGGL(k_trtri_diag, gt, dim3(256), 0, sf, q.A, q.Li, ld,
                         strideM, q.sc);
      hipLaunchKernelGGL(k_trtri_cols, gt, dim3(256), 0, sf, q.A, q.Li, ld,
                         strideM, q.sc);
    }
    SPK_TRY(check_launch("k_trtri"));
    if (!batched) {
      ProfScope ps(ctx, "minv", sf);
      hipLaunchKernelGGL(k_minv, dim3(Sb), dim3(256), 0, sf, q.Li, ld, strideM,
                         q.sc, q.idx, N, prob->stn_drift, m, method, q.M, q.T0,
                         gb, mvg);
    } else {
      {
        // condition estimate (power iteration over (row blocks, steps))
        ProfScope ps(ctx, "kappa", sf);
        const unsigned gl = (unsigned)ceil_div(N, 256);
        const unsigned gu = (unsigned)ceil_div(N, 4);
        GraphKey kk = gk;
        kk << kap_lam << fp64var_thr << fp64var_thr_gau << int8var_thr;
        SPK_TRY(graph_run(ctx, "kappa", kk.v, sf, [&]() -> int {
          for (int which = 0; which < 2; ++which) {
            const double *Tm = which == 0 ? q.A : q.Li;
            hipLaunchKernelGGL(k_kap_start, dim3(Sb), dim3(256), 0, sf, q.sc,
                               q.T0, ld, which);
            for (int it = 0; it < 12; ++it) {
              hipLaunchKernelGGL(k_kap_norm, dim3(Sb), dim3(256), 0, sf, q.sc,
                                 q.T0, ld);
              if (which == 0)
                hipLaunchKernelGGL(k_kap_umv, dim3(gu, Sb), dim3(256), 0, sf, Tm,
                                   ld, strideM, q.sc, q.T0, ld, 1);
              else
                hipLaunchKernelGGL(k_kap_lmv, dim3(gl, Sb), dim3(256), 0, sf, Tm,
                                   ld, strideM, q.sc, q.T0, ld, 1);
              hipLaunchKernelGGL(k_kap_rq, dim3(Sb), dim3(256), 0, sf, q.sc,
                                 q.T0, ld, kap_lam, which);
              if (which == 0)
                hipLaunchKernelGGL(k_kap_lmv, dim3(gl, Sb), dim3(256), 0, sf, Tm,
                                   ld, strideM, q.sc, q.T0, ld, 0);
              else
                hipLaunchKernelGGL(k_kap_umv, dim3(gu, Sb), dim3(256), 0, sf, Tm,
                                   ld, strideM, q.sc, q.T0, ld, 0);
            }
          }
          hipLaunchKernelGGL(k_kap_flags, dim3(Sb), dim3(64), 0, sf, q.sc,
                             (const double *)kap_lam, (const spk_vg *)q.vg,
                             fp64var_thr, fp64var_thr_gau, int8var_thr);
          return check_launch("k_kap");
        }));
      }
      {
        ProfScope ps(ctx, "dual", sf);
        const unsigned gl = (unsigned)ceil_div(N, 256);
        const unsigned gu = (unsigned)ceil_div(N, 4);
        const int pq = method == SPK_M_EDK ? 1 + m : 1;
        GraphKey dk = gk;
        dk << q.zact << ldz << q.idx << prob->stn_drift << m << q.V
           << dual_ds << mvg << sx << sy;
        SPK_TRY(graph_run(ctx, "dual", dk.v, sf, [&]() -> int {
          hipLaunchKernelGGL(k_dual_rhs, dim3(Sb), dim3(256), 0, sf, q.sc,
                             q.zact, ldz, q.idx, N, prob->stn_drift, m, method,
                             q.T0, ld);
          // U = Li [F z], X = C^-1 [F z]
          hipLaunchKernelGGL(k_dual_lmv, dim3(gl, Sb), dim3(256), 0, sf, q.Li, ld,
                             strideM, q.sc, q.T0, ld, 0, 1, pq + 1);
          hipLaunchKernelGGL(k_dual_umv, dim3(gu, Sb), dim3(256), 0, sf, q.Li, ld,
                             strideM, q.sc, q.T0, ld, 1, 0, pq + 1);
          hipLaunchKernelGGL(k_dual_mid, dim3(Sb), dim3(256), 0, sf, q.sc, m,
                             method, q.T0, ld, q.V, 4 * ld, dual_ds);
          for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(k_dual_resid, dim3(gl, Sb), dim3(256), 0, sf,
                               q.sc, q.zact, ldz, q.idx, N, prob->stn_drift, m,
                               method, q.T0, ld, q.V, 4 * ld, dual_ds,
                               (const spk_vg *)q.vg, mvg, (const double *)sx,
                               (const double *)sy);
            // Y2 = Li [r1 rw], Z2 = C^-1 [r1 rw]
            hipLaunchKernelGGL(k_dual_lmv, dim3(gl, Sb), dim3(256), 0, sf, q.Li,
                               ld, strideM, q.sc, q.T0, ld, 2, 3, 2);
            hipLaunchKernelGGL(k_dual_umv, dim3(gu, Sb), dim3(256), 0, sf, q.Li,
                               ld, strideM, q.sc, q.T0, ld, 3, 4, 2);
            hipLaunchKernelGGL(k_dual_upd, dim3(Sb), dim3(256), 0, sf, q.sc,
                               q.idx, N, prob->stn_drift, m, method, q.T0, ld,
                               q.V, 4 * ld, dual_ds);
          }
          hipLaunchKernelGGL(k_dual_fin, dim3(Sb), dim3(64), 0, sf, q.sc, m,
                             method, (const double *)dual_ds);
          return check_launch("k_dual");
        }));
      }
      // Steps whose Cholesky succeeded with a condition estimate above
      // kPinvKappa take the pseudo-inverse route too: their solve is
      // rounding noise, where the reference's pinv truncates the
      // eigenvalues under its cutoff (and sends most cells to the nearest
      // station).  Marked failed before the slicing kernels, which skip them.
      SPK_TRY(check_launch("k_dual"));
      SPK_HIP(hipMemcpyAsync(h_sc, q.sc, (size_t)Sb * sizeof(StepSc),
                             hipMemcpyDeviceToHost, sf));
      SPK_HIP(hipStreamSynchronize(sf));
      for (int s2 = 0; s2 < Sb; ++s2)
        if (h_sc[s2].status == 0 && !(h_sc[s2].kappa <= kPinvKappa)) {
          q.pinv.push_back(s2);
          hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, sf,
                             &q.sc[s2].status, 1, SPK_ST_FAILED);
        }
      // the contraction each step takes, from its flags (k_nearest may
      // still promote a step to FP64 during the contraction; the kernels
      // check the flags again)
      q.lst32.cnt = q.lst8a.cnt = q.lst8b.cnt = 0;
      q.n32 = q.n8a = q.n8b = 0;
      for (int s2 = 0; s2 < Sb; ++s2) {
        const StepSc &h = h_sc[s2];
        if (h.status != 0 || !(h.kappa <= kPinvKappa) ||
            (h.flags & SPK_ST_FP64VAR))
          continue;
        if (h.flags & SPK_ST_INT8VAR) {
          StepList &l = h.planes == 5 ? q.lst8a : q.lst8b;
          int &nm = h.planes == 5 ? q.n8a : q.n8b;
          l.s[l.cnt++] = s2;
          nm = std::max(nm, h.n);
        } else {
          q.lst32.s[q.lst32.cnt++] = s2;
          q.n32 = std::max(q.n32, h.n);
        }
      }
      if (q.L8) {
        ProfScope ps(ctx, "slice_rows", sf);
        hipLaunchKernelGGL(k_slice_rows, dim3((unsigned)(R8 / 64), Sb),
                           dim3(256), 0, sf, q.Li, ld, strideM, q.sc, q.L8,
                           KS8, R8, q.ea);
      }
      if (q.L32) {
        ProfScope ps(ctx, "cast_tri", sf);
        hipLaunchKernelGGL(k_cast_tri, dim3(1024, Sb), dim3(256), 0, sf, q.Li,
                           q.L32, ld, strideM, q.sc);
      }
      if (q.Lb) {
        ProfScope ps(ctx, "split_tri", sf);
        hipLaunchKernelGGL(k_split_tri<kBfA32>, dim3(1024, Sb), dim3(256), 0, sf,
                           q.Li,
                           ld, strideM, q.sc, q.Lb);
      }
    }
    SPK_TRY(check_launch("factorise"));
    if (sf != sm) SPK_HIP(hipEventRecord(pre, sf));
    // pseudo-inverses of the failed steps' bordered systems, written where
    // the plain contraction reads K^-1: the M slab (plain path) or the
    // step's L^-1 slab, which a failed step leaves unused (batched path)
    if (!q.pinv.empty()) {
      ProfScope ps(ctx, "pinv", sf);
      SPK_TRY(pinv_steps(ctx, sf, sx, sy, q.idx, N, q.sc, h_sc, q.pinv, gb,
                         method, m, prob->stn_drift, mvg,
                         batched ? q.Li : q.M, strideM, ld));
    }
    return SPK_OK;
  };

  auto contract = [&](int64_t t0, int Sb, Set &q, hipEvent_t fact) -> int {
    const spk_vg *gb = vgs + t0;
    bool simple = batched;
    for (int s = 0; batched && s < Sb; ++s) simple = simple && vg_simple_h[t0 + s];
    double *eb = est + t0 * C;
    double *vbo = want_var ? var + t0 * C : nullptr;
    if (!batched) {
      for (int s = 0; s < Sb; ++s) {
        if (std::find(q.pinv.begin(), q.pinv.end(), s) != q.pinv.end())
          hipLaunchKernelGGL(k_mark_pinv, dim3(1), dim3(64), 0, sm, q.sc + s);
        for (int64_t cb = 0; cb < C; cb += Cs) {
          const int64_t cn = std::min<int64_t>(Cs, C - cb);
          {
            ProfScope ps(ctx, "rhs_plain");
            hipLaunchKernelGGL(
                k_rhs_plain, dim3((unsigned)ceil_div(cn, 256)), dim3(256), 0,
                sm, sx, sy, q.idx + s * N, q.sc + s, gb + s, cx + cb, cy + cb,
                cn, C, prob->cell_drift ? prob->cell_drift + cb : nullptr,
                method, m, method == SPK_M_SK ? 1 : 0, mvg, R, ldk, nnb,
                (const int *)nullptr);
          }
          SPK_TRY(check_launch("k_rhs_plain"));
          {
            ProfScope ps(ctx, "gemm_plain_f64");
            // k depth per LDS stage 8, 16 or 32 (the same products in the
            // same order); a tuning run takes the first 1 024 cell tiles,
            // whole waves of blocks at one, two or four blocks per CU
            auto run = [&](int v, bool timing) -> int {
              const int64_t cr =
                  timing ? std::min<int64_t>(cn, 1024 * 128) : cn;
              auto k = v == 0   ? k_gemm<double, true, 8>
                       : v == 2 ? k_gemm<double, true, 32>
                                : k_gemm<double, true, 16>;
              hipLaunchKernelGGL(
                  k, dim3((unsigned)ceil_div(cr, 128), 1), dim3(256), 0, sm,
                  q.M + s * strideM, ld, (int64_t)0, R, ldk, (int64_t)0,
                  q.sc + s, p, method, want_var ? 1 : 0, skm,
                  q.zact + s * ldz, ldz, nnb, ldk, (const double *)nullptr,
                  (int64_t)0, cr, eb + s * C + cb,
                  vbo ? vbo + s * C + cb : nullptr, C, 0, (double *)nullptr,
                  0, (const uint8_t *)nullptr, (int64_t)0,
                  (unsigned long long *)nullptr, (const int *)nullptr,
                  (const int32_t *)nullptr);
              return check_launch("k_gemm<double,plain>");
            };
            // steps the kernel skips (no system) are not timed
            const int hs = h_sc[s].status;
            if (hs != 0 && hs != SPK_ST_FAILED)
              SPK_TRY(run(1, false));
            else
              SPK_TRY(tune_launch(ctx, "gemm_plain_f64",
                                  (int)ceil_div(h_sc[s].n + p, 1024), sm, 3,
                                  1, run));
          }
        }
      }
    } else {
      for (int64_t cb = 0; cb < C; cb += Cs) {
        const int64_t cn = std::min<int64_t>(Cs, C - cb);
        // with a cell order, stripes are ranges of the permuted cells and
        // outputs / drifts are addressed through perm
        const double *cd = prob->cell_drift
                               ? (perm ? prob->cell_drift
                                       : prob->cell_drift + cb)
                               : nullptr;
        if (nrst) {
          ProfScope ps(ctx, "nearest");
          hipLaunchKernelGGL(k_nearest, dim3((unsigned)ceil_div(cn, 256), Sb),
                             dim3(256), 0, sm, sx, sy, q.idx, N, q.sc,
                             (perm ? pcx : cx) + cb, (perm ? pcy : cy) + cb,
                             cn, nrst, ldk, mixed ? 1 : 0,
                             q.spos ? (const int32_t *)near_list + cb
                                    : (const int32_t *)nullptr,
                             C, (const int32_t *)q.spos);
          SPK_TRY(check_launch("k_nearest"));
        }
        const int64_t nzt = ceil_div(cn, 128);
        if (nzb && want_var)
          SPK_HIP(hipMemsetAsync(nzb, 0, (size_t)Sb * nzt * nzk, sm));
        {
          ProfScope ps(ctx, "kgen");
          hipLaunchKernelGGL(simple ? k_kgen<false> : k_kgen<true>,
                             dim3((unsigned)ceil_div(cn, 256), Sb),
                             dim3(256), 0, sm, sx, sy, q.idx, N, q.zact, ldz,
                             q.sc, gb, (perm ? pcx : cx) + cb,
                             (perm ? pcy : cy) + cb, cn, C, cd,
                             perm ? perm + cb : (const int32_t *)nullptr,
                             (const int32_t *)nrst, ldk, q.V,
                             4 * ld, ld, method, m, mvg, want_var ? 1 : 0,
                             mixed ? 0 : 1, skm, K32, K64, ldk, Kb,
                             q.Gm,
                             abuf, perm ? eb : eb + cb, C, dmax, nzb,
                             nzt, nzk, zskip);
        }
        SPK_TRY(check_launch("k_kgen"));
        if (!want_var) continue;
        if (B8) {
          {
            ProfScope ps(ctx, "slice_cols");
            const dim3 gs((unsigned)(ldk / 64), Sb);
            hipLaunchKernelGGL((k_slice_cols<double>), gs, dim3(256), 0, sm,
                               (const double *)K64, ld, ldk, cn, q.sc,
                               (const double *)dmax, B8, KS8, eb8,
                               (const uint8_t *)nzb, nzk, nz8);
          }
          {
            ProfScope ps(ctx, "gemm_tri_i8");
            unsigned long long *w8 =
                prof_of(ctx).on ? prof_of(ctx).work + 1 : nullptr;
            if (q.lst8a.cnt > 0)
              hipLaunchKernelGGL(
                  (k_tri8<5, 4>),
                  dim3((unsigned)ceil_div(q.n8a, 64), (unsigned)(ldk / 128),
                       (unsigned)q.lst8a.cnt),
                  dim3(512), 0, sm, (const int8_t *)q.L8, R8,
                  (const int *)q.ea, (const int8_t *)B8, ldk,
                  (const int *)eb8, KS8, q.sc, q.lst8a, cn, qpart, nrb,
                  (const uint8_t *)nz8, w8);
            if (q.lst8b.cnt > 0)
              hipLaunchKernelGGL(
                  (k_tri8<6, 5>),
                  dim3((unsigned)ceil_div(q.n8b, 64), (unsigned)(ldk / 128),
                       (unsigned)q.lst8b.cnt),
                  dim3(512), 0, sm, (const int8_t *)q.L8, R8,
                  (const int *)q.ea, (const int8_t *)B8, ldk,
                  (const int *)eb8, KS8, q.sc, q.lst8b, cn, qpart, nrb,
                  (const uint8_t *)nz8, w8);
          }
          SPK_TRY(check_launch("k_tri8"));
        }
        const dim3 grid((unsigned)nrb, (unsigned)ceil_div(cn, 128), Sb);
        if (bfop) {
          ProfScope ps(ctx, "gemm_tri_bf16");
          // delta of the next chunk loaded after (0) or while (1) a chunk
          // is multiplied: the same products in the same order
          auto run = [&](int v, bool timing) -> int {
            unsigned long long *w =
                prof_of(ctx).on && !timing ? prof_of(ctx).work + 2 : nullptr;
            auto k = v == 1 ? k_tri_bf<1, kBfA32> : k_tri_bf<0, kBfA32>;
            hipLaunchKernelGGL(
                k, dim3((unsigned)ceil_div(q.n32, 64), (unsigned)ceil_div(cn, 128),
                     (unsigned)q.lst32.cnt),
                dim3(256), 0, sm, (const uint4 *)q.Lb, ld, (const uint4 *)Kb,
                ldk, q.sc, q.lst32, cn, qpart, nrb, (const uint8_t *)nzb, nzk,
                w);
            return check_launch("k_tri_bf");
          };
          if (q.lst32.cnt > 0)
            SPK_TRY(tune_launch(ctx, "gemm_tri_bf16",
                                (int)ceil_div(q.n32, 1024), sm, 2, 0, run));
        } else if (mixed) {
          ProfScope ps(ctx, "gemm_tri_f32");
          // k chunks of 16: FP32 sums of 32 products, 3 waves per SIMD
          hipLaunchKernelGGL((k_gemm<float, false, 16>), grid, dim3(256), 0, sm,
                             q.L32, ld, strideM, (const float *)K32, ldk,
                             ld * ldk, q.sc, p, method, 1, 0, q.zact, ldz,
                             (const int32_t *)nullptr, (int64_t)0, abuf, ldk,
                             cn, eb + cb, vbo + cb, C, 1, qpart, nrb,
                             (const uint8_t *)nzb, nzk,
                             prof_of(ctx).on ? prof_of(ctx).work : nullptr,
                             (const int *)nullptr, (const int32_t *)nullptr);
        }
        {
          ProfScope ps(ctx, "gemm_tri_f64");
          // the mixed path's promoted steps are few (and known only on
          // the device): one block per tile walks the steps
          const dim3 g64((unsigned)ceil_div(N, 128),
                         (unsigned)ceil_div(cn, 128), mixed ? 1 : Sb);
          hipLaunchKernelGGL(mixed ? k_tri64<true> : k_tri64<false>, g64,
                             dim3(256), 0, sm, q.Li, ld, strideM,
                             (const double *)K64, ldk, ld * ldk, q.sc, cn,
                             qpart, ldk, nrb, mixed ? 2 : 0, Sb);
        }
        SPK_TRY(check_launch("k_gemm<tri>"));
        {
          ProfScope ps(ctx, "var_reduce");
          hipLaunchKernelGGL(k_var_reduce, dim3((unsigned)ceil_div(cn, 256), Sb),
                             dim3(256), 0, sm, q.sc, abuf, qpart, ldk, nrb, cn,
                             perm ? perm + cb : (const int32_t *)nullptr,
                             perm ? vbo : vbo + cb, C);
        }
        SPK_TRY(check_launch("k_var_reduce"));
      }
      // the cell kernels above skip the failed steps and started once the
      // factorisation was done; the rest waits for the pseudo-inverses
      if (sf != sm) SPK_HIP(hipStreamWaitEvent(sm, fact, 0));
      // steps solved by the pseudo-inverse fallback: estimates and the
      // reference's per-cell sum test from the dual vectors of pinv(K)
      // (O(n) per cell); the plain contraction then gives the OK variance
      // (and the reference's estimate arithmetic) of the cells that pass
      if (!q.pinv.empty()) {
        const int64_t Csp = std::min<int64_t>(
            C, std::max<int64_t>(
                   256, (int64_t)(rhs_budget / ((N + p + 64) * 8))));
        const int64_t ldkp = round_up(Csp, 128);
        const int nstr = (int)ceil_div(C, Csp);
        const int want_list = (method == SPK_M_OK && want_var) ? 1 : 0;
        double *bwp, *gxy = nullptr, *Rp = nullptr;
        int32_t *plist, *nnbp = nullptr;
        int *pcnt;
        SPK_TRY(ws_get(ctx, "pinv_bw", 2 * ld * 8, (void **)&bwp));
        SPK_TRY(ws_get(ctx, "pinv_list", C * 4, (void **)&plist));
        SPK_TRY(ws_get(ctx, "pinv_cnt", (1 + nstr) * 4, (void **)&pcnt));
        if (want_list) {
          SPK_TRY(ws_get(ctx, "pinv_gxy", 2 * C * 8, (void **)&gxy));
          SPK_TRY(ws_get(ctx, "R", (N + p + 64) * ldkp * 8, (void **)&Rp));
          SPK_TRY(ws_get(ctx, "nnb", ldkp * 4, (void **)&nnbp));
        }
        for (int s : q.pinv) {
          ProfScope ps(ctx, "pinv_contract");
          const double *Ms = q.Li + s * strideM;
          hipLaunchKernelGGL(k_mark_pinv, dim3(1), dim3(64), 0, sm, q.sc + s);
          hipLaunchKernelGGL(k_pinv_vec, dim3((unsigned)ceil_div(N + p, 256)),
                             dim3(256), 0, sm, Ms, ld, p, q.sc + s,
                             q.zact + s * ldz, bwp);
          hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, sm, pcnt, 1, 0);
          hipLaunchKernelGGL(k_pinv_cells, dim3((unsigned)ceil_div(C, 256)),
                             dim3(256), 0, sm, sx, sy, q.idx + s * N, q.sc + s,
                             gb + s, cx, cy, C, C, (int64_t)0,
                             prob->cell_drift, method, m, skm, mvg, bwp, ld,
                             q.zact + s * ldz, eb + s * C,
                             vbo ? vbo + s * C : nullptr, want_list, plist,
                             pcnt);
          SPK_TRY(check_launch("k_pinv_cells"));
          if (!want_list) continue;
          hipLaunchKernelGGL(
              k_pinv_gather,
              dim3((unsigned)std::min<int64_t>(ceil_div(C, 256), 4096)),
              dim3(256), 0, sm, plist, pcnt, cx, cy, gxy, gxy + C, Csp, nstr,
              pcnt + 1);
          for (int k2 = 0; k2 < nstr; ++k2) {
            const int64_t off = (int64_t)k2 * Csp;
            const int64_t cn = std::min<int64_t>(Csp, C - off);
            hipLaunchKernelGGL(
                k_rhs_plain, dim3((unsigned)ceil_div(cn, 256)), dim3(256), 0,
                sm, sx, sy, q.idx + s * N, q.sc + s, gb + s, gxy + off,
                gxy + C + off, cn, C, (const double *)nullptr, method, m, 0,
                mvg, Rp, ldkp, nnbp, (const int *)(pcnt + 1 + k2));
            hipLaunchKernelGGL((k_gemm<double, true, 16>),
                               dim3((unsigned)ceil_div(cn, 128), 1), dim3(256),
                               0, sm, Ms, ld, (int64_t)0, Rp, ldkp, (int64_t)0,
                               q.sc + s, p, method, 1, skm, q.zact + s * ldz,
                               ldz, nnbp, ldkp, (const double *)nullptr,
                               (int64_t)0, cn, eb + s * C, vbo + s * C, C, 0,
                               (double *)nullptr, 0, (const uint8_t *)nullptr,
                               (int64_t)0, (unsigned long long *)nullptr,
                               (const int *)(pcnt + 1 + k2),
                               (const int32_t *)(plist + off));
          }
          SPK_TRY(check_launch("pinv contraction"));
        }
      }
    }
    {
      ProfScope ps(ctx, "special");
      hipLaunchKernelGGL(k_special, dim3((unsigned)ceil_div(C, 256), Sb),
                         dim3(256), 0, sm, sx, sy, q.idx, N, q.zact, ldz, q.sc,
                         cx, cy, C, eb, vbo);
    }
    SPK_TRY(check_launch("k_special"));
    hipLaunchKernelGGL(k_status_out, dim3((unsigned)ceil_div(Sb, 256)),
                       dim3(256), 0, sm, q.sc, Sb, info + t0,
                       kappa ? kappa + t0 : nullptr);
    SPK_TRY(check_launch("k_status_out"));
    return SPK_OK;
  };

  const int64_t nb = ceil_div(T, S);
  auto bsize = [&](int64_t b) { return (int)std::min<int64_t>(S, T - b * S); };
  // Batch b's contraction is queued before batch b + 1's factorisation:
  // the host reads each batch's Cholesky outcome (for the pseudo-inverse
  // fallback) while the previous batch's cells contract.
  // The batched path's cell kernels start after the factorisation proper
  // (ev_pre); the pseudo-inverses of the failed steps, which those kernels
  // skip, run beside them on the factor stream until the contraction of
  // the failed steps waits for them (ev_fact).  The plain path contracts
  // every step with its K^-1 and waits for all of it.
  // float32 host fields: batch b's are cast and copied on a third stream
  // while later batches compute (staging of at most 512 MB, in slices)
  hipStream_t scp = nullptr;
  float *stage = nullptr;
  int64_t srows = 0;
  if (o32) {
    if (!ctx->copy_stream)
      SPK_HIP(hipStreamCreateWithFlags(&ctx->copy_stream,
                                       hipStreamNonBlocking));
    scp = ctx->copy_stream;
    srows = std::max<int64_t>(
        1, std::min<int64_t>(S, ((int64_t)256 << 20) / C));
    SPK_TRY(ws_get(ctx, "f32_stage", (size_t)2 * srows * C * sizeof(float),
                   (void **)&stage));
  }
  SPK_TRY(factorise(0, bsize(0), set[0], ev_pre[0]));
  if (sf != sm) SPK_HIP(hipEventRecord(ev_fact[0], sf));
  for (int64_t b = 0; b < nb; ++b) {
    const int k = (int)(b & 1);
    if (sf != sm)
      SPK_HIP(hipStreamWaitEvent(sm, batched ? ev_pre[k] : ev_fact[k], 0));
    SPK_TRY(contract(b * S, bsize(b), set[nsets > 1 ? k : 0], ev_fact[k]));
    SPK_HIP(hipEventRecord(ev_done[k], sm));
    if (o32) {
      SPK_HIP(hipStreamWaitEvent(scp, ev_done[k], 0));
      for (int64_t r0 = 0; r0 < bsize(b); r0 += srows)
        SPK_TRY(f32_out_rows(ctx, scp, o32, est, want_var ? var : nullptr, C,
                             b * S + r0,
                             std::min<int64_t>(srows, bsize(b) - r0), stage));
    }
    if (b + 1 < nb) {
      const int k2 = k ^ 1;
      // set k2 was last read by the contraction of batch b - 1
      if (b >= 1 && sf != sm) SPK_HIP(hipStreamWaitEvent(sf, ev_done[k2], 0));
      SPK_TRY(factorise((b + 1) * S, bsize(b + 1), set[nsets > 1 ? k2 : 0],
                        ev_pre[k2]));
      if (sf != sm) SPK_HIP(hipEventRecord(ev_fact[k2], sf));
    }
  }
  // nothing may remain queued on the factor or copy stream after return
  if (sf != sm) {
    SPK_HIP(hipEventRecord(ev_fact[0], sf));
    SPK_HIP(hipStreamWaitEvent(sm, ev_fact[0], 0));
  }
  if (o32) {
    SPK_HIP(hipEventRecord(ev_done[0], scp));
    SPK_HIP(hipStreamWaitEvent(sm, ev_done[0], 0));
  }
  return SPK_OK;
}

int spk_interp_steps_dev(spk_ctx *ctx, const spk_problem *prob,
                         const spk_opts *opts, int64_t T, const double *vals,
                         const spk_vg *vgs, double *est, double *var,
                         int32_t *info, double *kappa) {
  return interp_steps_impl(ctx, prob, opts, T, vals, vgs, est, var, info,
                           kappa, nullptr);
}

int spk_interp_steps_f32_dev(spk_ctx *ctx, const spk_problem *prob,
                             const spk_opts *opts, int64_t T,
                             const double *vals, const spk_vg *vgs,
                             double *est, double *var, int32_t *info,
                             double *kappa, const spk_f32_out *out32) {
  if (!out32 || !out32->est32)
    return set_error(SPK_EARG, "spk_interp_steps_f32_dev needs est32");
  return interp_steps_impl(ctx, prob, opts, T, vals, vgs, est, var, info,
                           kappa, out32);
}

// spk_interp_steps_dev for IDW / nearest neighbour with no allocation:
// every buffer beyond the inputs and outputs comes from the library's
// file-scope scratch array, whose address on this device was resolved by
// spk_ctx_create.  Same kernels, same results; at most kNaMaxStns
// stations and kNaMaxCells cells.  Calls that overlap on one device share
// that array and race.
int spk_idw_steps_noalloc_dev(spk_ctx *ctx, const spk_problem *prob,
                              const spk_opts *opts, int64_t T,
                              const double *vals, double *est, int32_t *info,
                              double *kappa) {
  if (!ctx || !prob || !opts) return set_error(SPK_EARG, "null argument");
  if (T < 0) return set_error(SPK_ESHAPE, "negative step count");
  const int64_t N = prob->n_stns, C = prob->n_cells;
  if (N < 1 || C < 1) return set_error(SPK_ESHAPE, "need stations and cells");
  if (N > kNaMaxStns || C > kNaMaxCells)
    return set_error(SPK_ESHAPE,
                     "%lld stations and %lld cells exceed the scratch of the "
                     "allocation-free entry point (at most %lld stations and "
                     "%lld cells): use spk_interp_steps_dev",
                     (long long)N, (long long)C, (long long)kNaMaxStns,
                     (long long)kNaMaxCells);
  if (opts->method != SPK_M_IDW && opts->method != SPK_M_NNB)
    return set_error(SPK_EARG, "the allocation-free entry point runs IDW "
                     "and nearest neighbour only (method %d)", opts->method);
  if (!vals || !est || !info || !prob->stn_x || !prob->stn_y ||
      !prob->cell_x || !prob->cell_y)
    return set_error(SPK_EARG, "null device pointer");
  if (T == 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  unsigned char *base = nullptr;
  SPK_TRY(noalloc_scratch(ctx->device, &base));
  double *sx = reinterpret_cast<double *>(base + kNaOffStn);
  double *sy = sx + N;
  double *cx = reinterpret_cast<double *>(base + kNaOffCell);
  double *cy = cx + C;
  int *bad = reinterpret_cast<int *>(base + kNaOffBad);
  int32_t *idx = reinterpret_cast<int32_t *>(base + kNaOffIdx);
  double *zact = reinterpret_cast<double *>(base + kNaOffZ);
  StepSc *sc = reinterpret_cast<StepSc *>(base + kNaOffSc);
  const int64_t ldz = round_up(N, 64);  // <= kNaMaxStns
  hipStream_t sm = ctx->stream;
  const unsigned gs = (unsigned)std::min<int64_t>(ceil_div(N, 256), 4096);
  const unsigned gc = (unsigned)std::min<int64_t>(ceil_div(C, 256), 65536);
  hipLaunchKernelGGL(k_recenter, dim3(gs), dim3(256), 0, sm, prob->stn_x, sx,
                     N, prob->origin_x);
  hipLaunchKernelGGL(k_recenter, dim3(gs), dim3(256), 0, sm, prob->stn_y, sy,
                     N, prob->origin_y);
  hipLaunchKernelGGL(k_recenter, dim3(gc), dim3(256), 0, sm, prob->cell_x, cx,
                     C, prob->origin_x);
  hipLaunchKernelGGL(k_recenter, dim3(gc), dim3(256), 0, sm, prob->cell_y, cy,
                     C, prob->origin_y);
  SPK_TRY(check_launch("k_recenter"));
  return idw_nnb_steps(ctx, opts, T, N, C, vals, sx, sy, cx, cy, idx, zact,
                       ldz, sc, std::min<int64_t>(T, kNaSteps), bad, est,
                       info, kappa);
}

}  // extern "C"
