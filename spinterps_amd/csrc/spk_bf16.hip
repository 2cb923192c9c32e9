// The variance contraction of the mixed path's FP32-tier steps on split
// BF16 operands.
//
// q(c) = |L^-1 delta_c|^2 (the deflated quadratic form of the OK variance,
// interp/steps.py:432-434 in the dual form) needs about FP32 accuracy in
// the operands and the partial sums.  Each operand x is split into three
// bfloat16 terms, x = hi + mid + lo (each the nearest bfloat16 to what the
// previous terms leave, so together 24 or more significant bits), and
// L^-1 delta is formed from the six plane products whose magnitude reaches
// 2^-24 of hi.hi (hi.hi, hi.mid, mid.hi, hi.lo, mid.mid, lo.hi) on
// v_mfma_f32_16x16x32_bf16: bfloat16 products are exact in FP32, each
// 32-deep chunk is summed in FP32 by the MFMAs and the chunks are added
// in FP64.  Six such MFMAs (16 cycles each) replace the eight
// v_mfma_f32_16x16x4_f32 (32 cycles each) of the FP32 contraction per
// 16 x 16 tile and 32 rows.
//
//  k_split_tri  L^-1 (FP64, lower triangle) -> three planes, once per step
//               on the factor stream (k_kgen writes delta's planes)
//  k_tri_bf     the contraction, one 64-row block of L^-1 against a
//               128-cell tile of delta per workgroup
#include "spk_krige.h"

namespace spk {

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// one bfloat16 octet of each plane from eight values (hi, mid, lo planes)
__device__ inline void split8d(const double (&x)[8], uint4 &h, uint4 &m,
                               uint4 &l) {
  uint32_t hw[4], mw[4], lw[4];
#pragma unroll
  for (int u = 0; u < 8; u += 2) {
    uint16_t h0, m0, l0, h1, m1, l1;
    split3d(x[u], h0, m0, l0);
    split3d(x[u + 1], h1, m1, l1);
    hw[u / 2] = (uint32_t)h0 | ((uint32_t)h1 << 16);
    mw[u / 2] = (uint32_t)m0 | ((uint32_t)m1 << 16);
    lw[u / 2] = (uint32_t)l0 | ((uint32_t)l1 << 16);
  }
  h = uint4{hw[0], hw[1], hw[2], hw[3]};
  m = uint4{mw[0], mw[1], mw[2], mw[3]};
  l = uint4{lw[0], lw[1], lw[2], lw[3]};
}

// L^-1 of the FP32-tier steps in octets of 8 consecutive k of one row:
// AF 0: the three bfloat16 planes, octet o of row r of plane p at
// Lb[s][(p KO + o) ld + r] (KO = ld / 8); AF 1: FP32 values (split into
// the planes by k_tri_bf as it stages them), the octet's two 16-byte
// halves at Lb[s][2 (o ld + r)] and the next.  Zero above the diagonal and
// beyond n, for rows up to round_up(n, 64) and octets up to
// round_up(n, 32) / 8 (what k_tri_bf reads).  Grid (any, S).
template <int AF>
__global__ __launch_bounds__(256) void k_split_tri(
    const double *__restrict__ Li, int64_t ld, int64_t stride,
    const StepSc *__restrict__ sc, uint4 *__restrict__ Lb) {
  const int s = blockIdx.y;
  if (sc[s].status || (sc[s].flags & (SPK_ST_FP64VAR | SPK_ST_INT8VAR)))
    return;
  const int n = sc[s].n;
  const int64_t Rn = (n + 63) / 64 * 64, KOn = (n + 31) / 32 * 4;
  const int64_t KO = ld / 8;
  const double *src = Li + (int64_t)s * stride;
  uint4 *dst = Lb + (int64_t)s * 3 * KO * ld;
  for (int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x; e < KOn * Rn;
       e += (int64_t)gridDim.x * 256) {
    const int r = (int)(e % Rn), o = (int)(e / Rn);
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = 8 * o + u;
      x[u] = (r < n && k <= r) ? src[r + (int64_t)k * ld] : 0.0;
    }
    if constexpr (AF) {
      uint32_t f[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f[u] = __float_as_uint((float)x[u]);
      dst[2 * (o * ld + r)] = uint4{f[0], f[1], f[2], f[3]};
      dst[2 * (o * ld + r) + 1] = uint4{f[4], f[5], f[6], f[7]};
    } else {
      uint4 h, m, l;
      split8d(x, h, m, l);
      dst[(0 * KO + o) * ld + r] = h;
      dst[(1 * KO + o) * ld + r] = m;
      dst[(2 * KO + o) * ld + r] = l;
    }
  }
}

// q partials of the FP32-tier steps: qpart[(s nrb + rb) ldk + c] =
// sum over the rows of row block rb of (L^-1 delta)[r, c]^2.
// Grid (row blocks of 64, cell tiles of 128, the steps of lst), 256
// threads; wave w
// takes cells 32 w .. 32 w + 31 of the tile (two 16-column MFMA tiles) and
// all 64 rows (four).  Lb: planes of L^-1 (k_split_tri); Db: planes of
// delta, octet o of cell c of plane p at Db[s][(p KO + o) ldk + c].
// The 32-row chunks multiplied are those with a nonzero 16-row half in the
// tile's chunk map (k_kgen); an unmarked half is taken as zero, not read.
// Rows of L^-1 beyond n are zero, so they add nothing.
// PF 1: delta of the next chunk is loaded while a chunk is multiplied (two
// register sets, 2 waves per SIMD); PF 0: after it (one set, 3 waves).
// AF: the format of L^-1 (k_split_tri): 0 planes, 1 FP32 split here.
template <int PF, int AF>
__global__ __launch_bounds__(256, PF ? 2 : 3) void k_tri_bf(
    const uint4 *__restrict__ Lb, int64_t ld, const uint4 *__restrict__ Db,
    int64_t ldk, const StepSc *__restrict__ sc, StepList lst, int64_t Cs,
    double *__restrict__ qpart, int nrb, const uint8_t *__restrict__ nzb,
    int64_t nzk, unsigned long long *__restrict__ nexec) {
  // blocks on the same XCD (linear id mod 8) take consecutive row blocks
  // of one cell tile, so that the tile's delta stays in that XCD's L2
  const int G = gridDim.x * gridDim.y;
  int lin = blockIdx.x + blockIdx.y * gridDim.x;
  lin = xcd_remap(lin, G);
  const int rbi = lin % gridDim.x, cti = lin / gridDim.x;
  const int sl = lst.s[blockIdx.z];
  const int status = sc[sl].status, flags = sc[sl].flags, n = sc[sl].n;
  if (status || (flags & (SPK_ST_FP64VAR | SPK_ST_INT8VAR))) return;
  const int64_t c0 = cti * (int64_t)128;
  if (c0 >= Cs) return;
  const int r0 = rbi * 64;
  if (r0 >= n) return;
  const int kend = min(n, r0 + 64);
  const int64_t KO = ld / 8;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  // thread tid stages octet (tid / 16) % 4 of row tid % 16 + 16 (tid / 64)
  // of each A plane: its LDS store has the pattern of the reads below
  const int so = (tid >> 4) & 3, sr = (tid & 15) + 16 * (tid >> 6);
  // per-thread bases of the operand loads; chunk q adds q times the
  // chunk stride, plane p p times the plane stride
  const uint4 *La = Lb + (int64_t)sl * 3 * KO * ld +
                    (AF ? 2 * ((int64_t)so * ld + r0 + sr)
                        : (int64_t)so * ld + r0 + sr);
  const int64_t a_q = AF ? 8 * ld : 4 * ld, a_p = KO * ld;
  const uint4 *Dl = Db + (int64_t)sl * 3 * KO * ldk + (int64_t)lg * ldk + c0 +
                    wave * 32 + li;
  const int64_t b_q = 4 * ldk, b_p = KO * ldk;

  // A planes of one chunk: 64 rows x 4 octets, rows 96 bytes apart: the
  // 16-byte accesses of 16 rows by the four lane groups (one octet each)
  // then run at the conflict-free rate (measured on gfx950: rows 64 or
  // 80 bytes apart are 1.7x slower)
  __shared__ uint4 As[2][3][64][6];
  __shared__ int s_list[256];
  __shared__ int s_cnt[4];
  // chunk list: entry (q << 2) | (half 0 marked) | (half 1 marked) << 1
  const int nq = (kend + 31) / 32;  // at most ld / 32 <= 256 (host)
  {
    const uint8_t *nz = nzb + ((int64_t)sl * gridDim.y + cti) * nzk;
    int mk = 0;
    if (tid < nq) mk = (nz[2 * tid] ? 1 : 0) | (nz[2 * tid + 1] ? 2 : 0);
    const unsigned long long bm = __ballot(mk != 0);
    if (lane == 0) s_cnt[wave] = __popcll(bm);
    __syncthreads();
    int off = 0;
    for (int w2 = 0; w2 < wave; ++w2) off += s_cnt[w2];
    if (mk) s_list[off + __popcll(bm & ((1ull << lane) - 1ull))] =
        (tid << 2) | mk;
  }
  __syncthreads();
  const int cnt = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
  if (nexec && tid == 0 && cnt) atomicAdd(nexec, (unsigned long long)cnt);

  double acc[4][2][4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[ti][tj][r] = 0.0;

  constexpr int NA = AF ? 2 : 3;  // 16-byte loads per thread and chunk
  auto load_a = [&](uint4 (&ra)[NA], int e) {
    const uint4 *src = La + (e >> 2) * a_q;
    if constexpr (AF) {
      ra[0] = src[0];
      ra[1] = src[1];
    } else {
#pragma unroll
      for (int p = 0; p < 3; ++p) ra[p] = src[p * a_p];
    }
  };
  auto store_a = [&](const uint4 (&ra)[NA], int buf) {
    if constexpr (AF) {
      // two values per conversion: the packed words are the planes'
      const uint32_t f[8] = {ra[0].x, ra[0].y, ra[0].z, ra[0].w,
                             ra[1].x, ra[1].y, ra[1].z, ra[1].w};
      uint32_t hw[4], mw[4], lw[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        split3f2(__uint_as_float(f[2 * u]), __uint_as_float(f[2 * u + 1]),
                 hw[u], mw[u], lw[u]);
      As[buf][0][sr][so] = uint4{hw[0], hw[1], hw[2], hw[3]};
      As[buf][1][sr][so] = uint4{mw[0], mw[1], mw[2], mw[3]};
      As[buf][2][sr][so] = uint4{lw[0], lw[1], lw[2], lw[3]};
    } else {
#pragma unroll
      for (int p = 0; p < 3; ++p) As[buf][p][sr][so] = ra[p];
    }
  };
  // lane group lg holds octet lg of the chunk: half lg / 2, taken as zero
  // (not read) when unmarked
  auto load_b = [&](uint4 (&rb)[3][2], int e) {
    const uint4 *src = Dl + (e >> 2) * b_q;
    const bool on = (e >> (lg >> 1)) & 1;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) rb[p][tj] = uint4{0u, 0u, 0u, 0u};
    if (on) {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) rb[p][tj] = src[p * b_p + tj * 16];
    }
  };
  auto mma = [&](int buf, const uint4 (&rb)[3][2]) {
    bf8_t b[3][2];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
        b[p][tj] = __builtin_bit_cast(bf8_t, rb[p][tj]);
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
      bf8_t a[3];
#pragma unroll
      for (int p = 0; p < 3; ++p)
        a[p] = __builtin_bit_cast(bf8_t, As[buf][p][ti * 16 + li][lg]);
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        // the small products first; the chunk's FP32 sum starts at zero
        f4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0][tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1][tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2][tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0][tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1][tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0][tj], c, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[ti][tj][r] += (double)c[r];
      }
    }
  };

  if (cnt > 0) {
    uint4 ra[NA], rb0[3][2], rb1[PF ? 3 : 1][2];
    load_a(ra, s_list[0]);
    load_b(rb0, s_list[0]);
    store_a(ra, 0);
    if (cnt > 1) load_a(ra, s_list[1]);
    __syncthreads();
    // chunk i: A of chunk i + 1 goes to the other LDS buffer and A of
    // i + 2 is loaded; delta of chunk i + 1 is loaded before (PF) or after
    // chunk i's products
    if constexpr (PF) {
      auto step = [&](int i, const uint4 (&cur)[3][2], uint4 (&nxt)[3][2]) {
        if (i + 1 < cnt) load_b(nxt, s_list[i + 1]);
        mma(i & 1, cur);
        if (i + 1 < cnt) {
          store_a(ra, (i + 1) & 1);
          if (i + 2 < cnt) load_a(ra, s_list[i + 2]);
        }
        __syncthreads();
      };
      for (int i = 0; i < cnt; i += 2) {
        step(i, rb0, rb1);
        if (i + 1 < cnt) step(i + 1, rb1, rb0);
      }
    } else {
      for (int i = 0; i < cnt; ++i) {
        mma(i & 1, rb0);
        if (i + 1 < cnt) {
          load_b(rb0, s_list[i + 1]);
          store_a(ra, (i + 1) & 1);
          if (i + 2 < cnt) load_a(ra, s_list[i + 2]);
        }
        __syncthreads();
      }
    }
  }
  double qv[2] = {0.0, 0.0};
#pragma unroll
  for (int tj = 0; tj < 2; ++tj)
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int r = 0; r < 4; ++r) qv[tj] += acc[ti][tj][r] * acc[ti][tj][r];
#pragma unroll
  for (int tj = 0; tj < 2; ++tj) {
    qv[tj] += __shfl_xor(qv[tj], 16, 64);
    qv[tj] += __shfl_xor(qv[tj], 32, 64);
  }
  if (lg != 0) return;
#pragma unroll
  for (int tj = 0; tj < 2; ++tj) {
    const int64_t col = c0 + wave * 32 + tj * 16 + li;
    if (col < Cs) qpart[((int64_t)sl * nrb + rbi) * ldk + col] = qv[tj];
  }
}

template __global__ void k_tri_bf<0, 0>(
    const uint4 *, int64_t, const uint4 *, int64_t, const StepSc *, StepList,
    int64_t, double *, int, const uint8_t *, int64_t, unsigned long long *);
template __global__ void k_tri_bf<0, 1>(
    const uint4 *, int64_t, const uint4 *, int64_t, const StepSc *, StepList,
    int64_t, double *, int, const uint8_t *, int64_t, unsigned long long *);
template __global__ void k_tri_bf<1, 0>(
    const uint4 *, int64_t, const uint4 *, int64_t, const StepSc *, StepList,
    int64_t, double *, int, const uint8_t *, int64_t, unsigned long long *);
template __global__ void k_tri_bf<1, 1>(
    const uint4 *, int64_t, const uint4 *, int64_t, const StepSc *, StepList,
    int64_t, double *, int, const uint8_t *, int64_t, unsigned long long *);
template __global__ void k_split_tri<0>(const double *, int64_t, int64_t,
                                        const StepSc *, uint4 *);
template __global__ void k_split_tri<1>(const double *, int64_t, int64_t,
                                        const StepSc *, uint4 *);

}  // namespace spk
