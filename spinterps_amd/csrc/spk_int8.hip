// Variance contraction by exact int8 plane products (mixed path).
//
// For steps above the FP32 limit, |L^-1 delta|^2 is formed from integer
// products (the kernels are written for S planes and diagonals up to P;
// the library uses S = 5, P = 4 up to condition estimate 1e7 and S = 6,
// P = 5 above):
//   * each row i of L^-1 is scaled by 2^-e_i (max_k |L^-1_ik| < 2^e_i) and
//     each cell column of delta by 2^-f_c, so every scaled entry u lies in
//     (-1, 1);
//   * u is cut into S signed 7-bit planes by repeated truncation,
//     u = sum_s a_s 2^-7(s+1) + r with |a_s| <= 127 and |r| < 2^-7 S (the
//     planes are the leading digits, so the first 4 of 5 planes are the
//     4-plane cut);
//   * Z_d = sum_{s+t=d} A_s B_t is summed exactly in int32 by
//     mfma_i32_16x16x64_i8 (|Z_d| <= 6 * 2048 * 127^2 < 2^31), for the
//     diagonals d <= P;
//   * y_ic = 2^(e_i + f_c) sum_d 2^-7(d+2) Z_d,ic in FP64 and the squared
//     norms go to the same 64-row partials as the FP64 contraction.
// The only rounding is the truncation of the planes and of the dropped
// diagonals (relative 2^-7S of the row and column maxima) and the final
// FP64 combination; nothing is lost to cancellation in the sums.
#include "spk_krige.h"

namespace spk {

typedef int i4_t __attribute__((ext_vector_type(4)));

// planes stored per operand (the 5-plane cut)
constexpr int kPlanes = kInt8Planes;
// 64-k chunks up to which the int8 kernels skip the zero chunks of delta
// (the int8 steps have at most 20 000 stations, 313 chunks)
constexpr int kMaxChunks8 = 512;

// Cut x / 2^e (|x| < 2^e) into kPlanes 7-bit planes: the base-128 digits
// of trunc(|x| 2^(7 kPlanes - e)), most significant first, with the sign of
// x (the same digits as truncating 128 x / 2^e plane by plane).  The
// scaling is exact; the digits above 2^28 and below are split in FP64 so
// that each part converts exactly to int32.
__device__ inline void planes_of(double x, int e, int8_t *out) {
  static_assert(7 * kPlanes - 28 <= 30, "high digits fit in int32");
  const double w = ldexp(fabs(x), 7 * kPlanes - e);  // < 2^(7 kPlanes)
  const double hi = trunc(ldexp(w, -28));
  const int lo = (int)(w - ldexp(hi, 28));           // < 2^28
  const long long t = ((long long)(int)hi << 28) | (long long)lo;
  const int sg = x < 0.0 ? -1 : 1;
#pragma unroll
  for (int s = 0; s < kPlanes; ++s)
    out[s] = (int8_t)(sg * (int)((t >> (7 * (kPlanes - 1 - s))) & 127));
}

// exponent e with |x| < 2^e (0 for x == 0)
__device__ inline int exp_above(double x) {
  int e = 0;
  if (x != 0.0) frexp(x, &e);
  return e;
}

__device__ inline bool int8_step(const StepSc &st) {
  return !st.status && (st.flags & SPK_ST_INT8VAR) &&
         !(st.flags & SPK_ST_FP64VAR);
}

// L^-1 of each int8 step (FP64, column-major lower triangle, ld) -> planes
// A8[step][s][R][KS] (row i, k contiguous; zero above the diagonal and
// beyond n) and row exponents ea[step][R].  Grid (R / 64, steps), 256
// threads: 64 rows per block, transposed through LDS in 64 x 64 tiles.
__global__ __launch_bounds__(256) void k_slice_rows(
    const double *__restrict__ Li, int64_t ld, int64_t stride,
    const StepSc *__restrict__ sc, int8_t *__restrict__ A8, int64_t KS,
    int64_t R, int *__restrict__ ea) {
  const int sl = blockIdx.y;
  const StepSc st = sc[sl];
  if (!int8_step(st)) return;
  const int n = st.n;
  const int r0 = blockIdx.x * 64;
  const double *L = Li + (int64_t)sl * stride;
  int8_t *Ap = A8 + (int64_t)sl * kPlanes * R * KS;
  __shared__ double tile[64][65];
  __shared__ double red[4][64];
  __shared__ int rexp[64];
  const int tid = threadIdx.x, rr = tid & 63, kq = tid >> 6;
  const int row = r0 + rr;
  double mx = 0.0;
  if (row < n)
    for (int k = kq; k <= row; k += 4)
      mx = fmax(mx, fabs(L[row + (int64_t)k * ld]));
  red[kq][rr] = mx;
  __syncthreads();
  if (tid < 64) {
    const double m = fmax(fmax(red[0][tid], red[1][tid]),
                          fmax(red[2][tid], red[3][tid]));
    const int e = exp_above(m);
    rexp[tid] = e;
    ea[(int64_t)sl * R + r0 + tid] = e;
  }
  __syncthreads();
  // writer lane: row wrow, 16 k's starting at 16 wseg
  const int wrow = tid >> 2, wseg = tid & 3;
  for (int k0 = 0; k0 < KS; k0 += 64) {
    for (int e = tid; e < 64 * 64; e += 256) {
      const int kk = e >> 6, r2 = e & 63;
      const int k = k0 + kk, i = r0 + r2;
      tile[kk][r2] = (i < n && k <= i) ? L[i + (int64_t)k * ld] : 0.0;
    }
    __syncthreads();
    int8_t p[kPlanes][16];
    const int e = rexp[wrow];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      int8_t q[kPlanes];
      planes_of(tile[wseg * 16 + j][wrow], e, q);
#pragma unroll
      for (int s = 0; s < kPlanes; ++s) p[s][j] = q[s];
    }
#pragma unroll
    for (int s = 0; s < kPlanes; ++s)
      *reinterpret_cast<i4_t *>(Ap + (int64_t)s * R * KS +
                                (int64_t)(r0 + wrow) * KS + k0 + wseg * 16) =
          *reinterpret_cast<const i4_t *>(p[s]);
    __syncthreads();
  }
}

// delta of each int8 step (T = double, [step][k][c] with ldk, rows < n
// valid, cells < Cs valid) -> its first S planes (S = the step's plane
// count) B8[step][t][ldk][KS] (cell c, k contiguous; zero beyond n and beyond
// Cs) and cell exponents eb[step][ldk] from the per-cell maxima dmax[step][ldk]
// that k_kgen leaves (for FP32 input raised by 2^-22 so that the rounded
// values stay below 2^e).  Grid (ldk / 64, steps), 256 threads: each block
// takes its 64 cells through the k range in chunks of 64, transposed through
// LDS (one launch for both plane counts, and few blocks that only exit).
// With a chunk map (nzb[step][128-cell tile][16-row chunk], a nonzero byte
// where k_kgen found a nonzero delta) and at most kMaxChunks8 chunks of 64 k,
// a chunk that holds no nonzero delta in the block's 128-cell tile gets no
// planes: k_tri8 skips the same chunks.  Within a chunk, the
// 16-row chunks the map leaves unmarked are taken as zeros without reading
// delta.  Every chunk's entry of nz8[step][cell tile][chunk] (ldk / 64 tiles
// of 64 cells, KS / 64 chunks) is written: 1 where one of the S planes holds
// a nonzero digit, which is what k_tri8 multiplies.
template <typename T>
__global__ __launch_bounds__(256) void k_slice_cols(
    const T *__restrict__ K, int64_t ld, int64_t ldk, int64_t Cs,
    const StepSc *__restrict__ sc, const double *__restrict__ dmax,
    int8_t *__restrict__ B8, int64_t KS, int *__restrict__ eb,
    const uint8_t *__restrict__ nzb, int64_t nzk, uint8_t *__restrict__ nz8) {
  const int sl = blockIdx.y;
  const StepSc st = sc[sl];
  if (!int8_step(st)) return;
  const int S = st.planes;
  const int n = st.n;
  const int64_t c0 = blockIdx.x * (int64_t)64;
  const T *D = K + (int64_t)sl * ld * ldk;
  int8_t *Bp = B8 + (int64_t)sl * kPlanes * ldk * KS;
  __shared__ double tile[64][65];
  __shared__ int cexp[64];
  const int tid = threadIdx.x;
  if (tid < 64) {
    const int64_t c = c0 + tid;
    const double m = c < Cs ? dmax[(int64_t)sl * ldk + c] : 0.0;
    const int e = exp_above(sizeof(T) == 4 ? m * (1.0 + 0x1p-22) : m);
    cexp[tid] = e;
    eb[(int64_t)sl * ldk + c] = e;
  }
  const int nkc = (int)(KS >> 6);
  uint8_t *nz8r = nz8 + ((int64_t)sl * (ldk >> 6) + blockIdx.x) * nkc;
  const bool mapped = nzb != nullptr;
  const bool skip = mapped && nkc <= kMaxChunks8;
  const uint8_t *nz =
      mapped ? nzb + ((int64_t)sl * ((Cs + 127) >> 7) + (c0 >> 7)) * nzk
             : nullptr;
  const int wc = tid >> 2, wseg = tid & 3;
  for (int kc = 0; kc < nkc; ++kc) {
    const int k0 = kc * 64;
    // the 16-row chunks the map leaves unmarked hold delta = 0 for every
    // cell of the tile (k_kgen may leave them unwritten): zeros, not reads
    unsigned live = 15u;
    if (mapped) {
      live = 0u;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int64_t i = (k0 >> 4) + b;
        if (i < nzk && nz[i] != 0) live |= 1u << b;
      }
    }
    if (skip && live == 0u) {
      if (tid == 0) nz8r[kc] = 0;
      continue;
    }
    // the previous chunk's readers finished at its closing barrier
    for (int e = tid; e < 64 * 64; e += 256) {
      const int kk = e >> 6, cc = e & 63;
      const int k = k0 + kk;
      const int64_t c = c0 + cc;
      tile[kk][cc] = (k < n && c < Cs && ((live >> (kk >> 4)) & 1u))
                         ? (double)D[(int64_t)k * ldk + c]
                         : 0.0;
    }
    __syncthreads();
    int8_t p[kPlanes][16];
    const int e = cexp[wc];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      int8_t q[kPlanes];
      planes_of(tile[wseg * 16 + j][wc], e, q);
#pragma unroll
      for (int s2 = 0; s2 < kPlanes; ++s2) p[s2][j] = q[s2];
    }
    int any = 0;
#pragma unroll
    for (int s2 = 0; s2 < kPlanes; ++s2) {
      if (s2 >= S) break;
      const i4_t v = *reinterpret_cast<const i4_t *>(p[s2]);
      any |= v[0] | v[1] | v[2] | v[3];
      *reinterpret_cast<i4_t *>(Bp + (int64_t)s2 * ldk * KS +
                                (c0 + wc) * KS + k0 + wseg * 16) = v;
    }
    any = __syncthreads_or(any);
    if (tid == 0) nz8r[kc] = any ? 1 : 0;
  }
}
template __global__ void k_slice_cols<double>(
    const double *, int64_t, int64_t, int64_t, const StepSc *,
    const double *, int8_t *, int64_t, int *, const uint8_t *, int64_t,
    uint8_t *);

// q_part = |rows r0..r0+63 of L^-1 delta|^2 per cell from the first S
// planes, diagonals s + t <= P.  Block tile 64 rows x 128 cells, k in
// chunks of 64 (one MFMA depth) up to the tile's last row; 8 waves in
// 2 x 4, each 32 rows x 32 cells = 2 x 2 MFMA tiles per diagonal, the
// P + 1 diagonal sums held in int32 for the whole k range.  The wave tile
// keeps a wave near 200 registers, so two waves share each SIMD and one
// wave's loads, barriers and epilogue overlap the other's products.  A
// step's L^-1 planes do not fit in an XCD's L2 and are read again for
// every cell tile, so the tile is wide in cells (128: half the reads of
// a 64-cell tile) and short in rows; the cell tile's delta planes stay in
// L2 while its row blocks run on that XCD (the XCD remap of k_tri64).
// Each chunk's planes are loaded into registers while the previous chunk is
// multiplied out of one of two LDS buffers and then stored into the other,
// one barrier per chunk (64-byte rows with their 16-byte segments swizzled
// by row, so that both the 16-row reads and the 4-row stores of a 16-lane
// group hit distinct banks).  The two 32-row halves of the 64-row partial
// are added through LDS in a fixed order.  With the digit map nz8 of
// k_slice_cols (per 64 cells) only the 64-k chunks whose delta planes hold
// a nonzero digit in one of the block's two 64-cell halves are multiplied,
// and a half without one is taken as zero, not read (its planes may be
// unwritten); the skipped products are exactly zero.  nexec, when given,
// counts the plane products executed.  Grid (row blocks of 64, cell tiles
// of 128, the steps of lst), 512 threads.
template <int S, int P>
__global__ __launch_bounds__(512) void k_tri8(
    const int8_t *__restrict__ A8, int64_t R, const int *__restrict__ ea,
    const int8_t *__restrict__ B8, int64_t ldk, const int *__restrict__ eb,
    int64_t KS, const StepSc *__restrict__ sc, StepList lst, int64_t Cs,
    double *__restrict__ qpart, int nrb, const uint8_t *__restrict__ nz8,
    unsigned long long *nexec) {
  constexpr int BM = 64, BN = 128, BK = 64, RS = 64, NT = 512;
  // 16-byte segment seg of LDS row r is kept at seg ^ ((r >> 1) & 3): the
  // 16-byte reads of 16 rows by the four lane groups (one segment each)
  // then run at the conflict-free rate (measured on gfx950: 1.7x faster
  // than with seg ^ ((r >> 2) & 3) or no swizzle)
  auto swz = [](int r, int seg) {
    return r * RS + ((seg ^ ((r >> 1) & 3)) << 4);
  };
  // A chunk: S planes x 64 rows x 4 x 16 B; B chunk: S planes x 128 cells
  // x 4 x 16 B; NAQ / NBQ 16-byte loads per thread
  constexpr int NAQ = (S * 256 + NT - 1) / NT, NBQ = (S * 512) / NT;
  static_assert(S <= kPlanes && P <= 2 * (S - 1), "shape");
  const int G = gridDim.x * gridDim.y;
  int lin = blockIdx.x + blockIdx.y * gridDim.x;
  lin = xcd_remap(lin, G);
  const int rti = lin % gridDim.x, cti = lin / gridDim.x;
  const int sl = lst.s[blockIdx.z];
  const StepSc st = sc[sl];
  if (!int8_step(st) || st.planes != S) return;
  const int n = st.n;
  const int r0 = rti * BM;
  if (r0 >= n) return;
  const int64_t c0 = cti * (int64_t)BN;
  if (c0 >= Cs) return;
  const int kend = min(n, r0 + BM);
  const int8_t *A = A8 + (int64_t)sl * kPlanes * R * KS;
  const int8_t *B = B8 + (int64_t)sl * kPlanes * ldk * KS;
  // two LDS buffers: chunk ch + 1 is stored while chunk ch is multiplied
  __shared__ __align__(16) int8_t As[2][S][BM * RS];
  __shared__ __align__(16) int8_t Bs[2][S][BN * RS];
  __shared__ int s_list[kMaxChunks8];  // chunk << 2 | halves with a digit
  __shared__ int s_cnt[NT / 64];
  __shared__ double s_half[4][2][16];  // [wc][tj][column]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int wr = wave >> 2, wc = wave & 3;
  i4_t ra[NAQ], rb[NBQ];
  auto load = [&](int e) {
    const int k0 = (e >> 2) * BK;
#pragma unroll
    for (int q = 0; q < NAQ; ++q) {
      const int i = tid + q * NT, s = i >> 8, row = (i & 255) >> 2,
                seg = i & 3;
      if (s < S)
        ra[q] = *reinterpret_cast<const i4_t *>(
            A + (int64_t)s * R * KS + (int64_t)(r0 + row) * KS + k0 +
            seg * 16);
    }
#pragma unroll
    for (int q = 0; q < NBQ; ++q) {
      const int i = tid + q * NT, t = i >> 9, cell = (i & 511) >> 2,
                seg = i & 3;
      if ((e >> (cell >> 6)) & 1)
        rb[q] = *reinterpret_cast<const i4_t *>(
            B + (int64_t)t * ldk * KS + (c0 + cell) * KS + k0 + seg * 16);
      else
        rb[q] = i4_t{0, 0, 0, 0};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < NAQ; ++q) {
      const int i = tid + q * NT, s = i >> 8, row = (i & 255) >> 2,
                seg = i & 3;
      if (s < S) *reinterpret_cast<i4_t *>(&As[buf][s][swz(row, seg)]) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NBQ; ++q) {
      const int i = tid + q * NT, t = i >> 9, cell = (i & 511) >> 2,
                seg = i & 3;
      *reinterpret_cast<i4_t *>(&Bs[buf][t][swz(cell, seg)]) = rb[q];
    }
  };
  i4_t acc[P + 1][2][2];
#pragma unroll
  for (int d = 0; d <= P; ++d)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[d][a][b] = i4_t{0, 0, 0, 0};
  // the list holds up to kMaxChunks8 chunks: the library sends only steps
  // with at most 20 000 stations (313 chunks) to the int8 planes
  const int nkc = min((kend + BK - 1) / BK, kMaxChunks8);
  const bool skip = nz8 != nullptr && (KS >> 6) <= kMaxChunks8;
  const uint8_t *nzd =
      skip ? nz8 + ((int64_t)sl * (ldk >> 6) + 2 * cti) * (KS >> 6) : nullptr;
  int nch = 0;
  for (int b0 = 0; b0 < nkc; b0 += NT) {
    const int ch = b0 + tid;
    int mk = 0;
    if (ch < nkc)
      mk = skip ? (nzd[ch] ? 1 : 0) | (nzd[(KS >> 6) + ch] ? 2 : 0) : 3;
    const unsigned long long bm = __ballot(mk != 0);
    if (lane == 0) s_cnt[wave] = __popcll(bm);
    __syncthreads();
    int off = nch;
    for (int w2 = 0; w2 < wave; ++w2) off += s_cnt[w2];
    if (mk) s_list[off + __popcll(bm & ((1ull << lane) - 1ull))] =
        (ch << 2) | mk;
#pragma unroll
    for (int w2 = 0; w2 < NT / 64; ++w2) nch += s_cnt[w2];
    __syncthreads();
  }
  if (nexec && tid == 0) {
    int prods = 0;
    for (int s = 0; s < S; ++s)
      for (int t = 0; t < S; ++t) prods += s + t <= P;
    atomicAdd(nexec, (unsigned long long)nch * prods);
  }
  if (nch > 0) {
    load(s_list[0]);
    store(0);
  }
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) load(s_list[ch + 1]);
    i4_t af[S][2];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
        af[s][ti] = *reinterpret_cast<const i4_t *>(
            &As[buf][s][swz(wr * 32 + ti * 16 + li, lg)]);
#pragma unroll
    for (int t = 0; t < S; ++t) {
      i4_t bf[2];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
        bf[tj] = *reinterpret_cast<const i4_t *>(
            &Bs[buf][t][swz(wc * 32 + tj * 16 + li, lg)]);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (s + t > P) continue;
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj)
            acc[s + t][ti][tj] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                af[s][ti], bf[tj], acc[s + t][ti][tj], 0, 0, 0);
      }
    }
    // the other buffer was last read in chunk ch - 1, before the barrier
    // that ended it
    if (ch + 1 < nch) store(buf ^ 1);
    __syncthreads();
  }
  // y = 2^(e_i + f_c) sum_d 2^-7(d+2) Z_d; result slot r of lane group lg
  // is row 4 lg + r of the 16 x 16 tile
  const int *er = ea + (int64_t)sl * R;
  double part[2] = {0.0, 0.0};
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = r0 + wr * 32 + ti * 16 + 4 * lg + r;
      if (row >= n) continue;
      const int e2 = 2 * er[row];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        double y = 0.0;
#pragma unroll
        for (int d = P; d >= 0; --d)
          y += ldexp((double)acc[d][ti][tj][r], -7 * (d + 2));
        part[tj] += ldexp(y * y, e2);
      }
    }
#pragma unroll
  for (int tj = 0; tj < 2; ++tj) {
    part[tj] += __shfl_xor(part[tj], 16, 64);
    part[tj] += __shfl_xor(part[tj], 32, 64);
  }
  // rows 32-63 of the 64-row partial go through LDS to the wave holding
  // rows 0-31, which adds them to its own
  if (wr == 1 && lg == 0) {
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) s_half[wc][tj][li] = part[tj];
  }
  __syncthreads();
  if (wr != 0 || lg != 0 || rti >= nrb) return;
#pragma unroll
  for (int tj = 0; tj < 2; ++tj) {
    const int64_t col = c0 + wc * 32 + tj * 16 + li;
    const double v = part[tj] + s_half[wc][tj][li];
    if (col < Cs)
      qpart[((int64_t)sl * nrb + rti) * ldk + col] =
          ldexp(v, 2 * eb[(int64_t)sl * ldk + col]);
  }
}
template __global__ void k_tri8<5, 4>(const int8_t *, int64_t, const int *,
                                      const int8_t *, int64_t, const int *,
                                      int64_t, const StepSc *, StepList,
                                      int64_t, double *, int,
                                      const uint8_t *,
                                      unsigned long long *);
template __global__ void k_tri8<6, 5>(const int8_t *, int64_t, const int *,
                                      const int8_t *, int64_t, const int *,
                                      int64_t, const StepSc *, StepList,
                                      int64_t, double *, int,
                                      const uint8_t *,
                                      unsigned long long *);

}  // namespace spk
