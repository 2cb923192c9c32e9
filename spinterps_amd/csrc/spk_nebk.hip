// Neighbourhood interpolation: each cell uses only n_nebs of the step's
// active stations, its nearest (nrst) or n_nebs taken sector by sector
// (pie), as interp/steps.py:592-608 and :732-831 do with the neighbour
// groups of interp/grps.py:103-288.
//
// Per batch of steps, all on the context's stream:
//   k_neb_mask     stations that no cell selects with every station present
//                  count as missing (interp/steps.py:592-608)
//   k_compact      each step's active stations (shared with spk_factor.hip)
//   k_neb_nrst     each cell's k nearest active stations: the block culls
//                  candidates for a tile of 64 cells with a distance bound,
//                  then each cell finds its k-th smallest distance by
//                  bisection over ballot counts
//   k_neb_pie      (pie) sector-balanced selection, one wavefront per cell
//   radix sort     cells by (step, hash of the station set), so cells that
//                  share a set become adjacent
//   k_neb_mark ... groups of identical sets (interp/grps.py:103-139), cut
//                  into chunks of at most 64 cells
//   k_neb_krige    one wavefront per chunk: the set's covariance system,
//                  its Cholesky factor and triangular inverse in LDS, the
//                  dual-form vectors, then every cell of the chunk
//   k_neb_pinv_*   chunks whose Cholesky fails (or whose condition estimate
//                  exceeds 1e14) are solved as the reference solves every
//                  group: pinv of the bordered system (rocSOLVER eigen-
//                  decomposition, numpy's cutoff), lambda = pinv(K) rhs
//                  per cell and the sum test
//
// Group results depend only on the group's station set, so grouping is a
// matter of speed: each cell gets what its own set gives.
#include <hipcub/hipcub.hpp>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "spk_krige.h"
#include "spk_neb.h"
#include "spk_prof.h"

namespace spk {

constexpr int kNebMax = 64;     // stations per neighbourhood
constexpr int kNebChunk = 64;   // cells per chunk of a group
constexpr int kNebTile = 64;    // cells per selection tile
constexpr int kNebSlots = 8;    // candidates per lane held in registers
constexpr int kNebCand = kNebSlots * 64;  // candidates staged per tile
constexpr int kStepShift = 56;  // step index above the 56-bit set hash

__device__ inline bool neb_close1(double x) {
  return fabs(x - 1.0) <= 1e-8 + 1e-5;  // np.isclose(x, 1.0)
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)v, off, 64);
    const unsigned hi = __shfl_xor((unsigned)(v >> 32), off, 64);
    v += ((unsigned long long)hi << 32) | lo;
  }
  return v;
}

__device__ inline unsigned long long uniform_u64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

__device__ inline unsigned long long lanes_below(int lane) {
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// splitmix64 finaliser: the contribution of station position v at rank r
__device__ inline unsigned long long set_mix(unsigned v, unsigned r) {
  unsigned long long z = ((unsigned long long)r << 32) + v +
                         0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ inline unsigned long long dbits(double d) {
  return (unsigned long long)__double_as_longlong(d);
}

static inline int64_t round_up(int64_t a, int64_t b) {
  return ceil_div(a, b) * b;
}

// IDW of one cell over a station set in the reference's arithmetic
// (interp/steps.py:293-313, pyx:784-808): distances normalised by their
// maximum, w = 1 / d^p, sum(w z) / sum(w); bit 4 of bad where the weight
// sum is below 1e-14 (interp/steps.py:305).
__device__ inline double idw_set_cell(double x, double y, const double *xs,
                                      const double *ys, const double *zs,
                                      int k, double pw, int *bad) {
#pragma clang fp contract(off)
  double dmax = 0.0;
  for (int j = 0; j < k; ++j) dmax = fmax(dmax, dist2(x, y, xs[j], ys[j]));
  double ws = 0.0, wz = 0.0;
  for (int j = 0; j < k; ++j) {
    double d = dist2(x, y, xs[j], ys[j]);
    if (dmax > 0) d = d / dmax;
    const double wj = 1.0 / (pw == 2.0 ? d * d : pow(d, pw));
    ws = ws + wj;
    wz = wz + wj * zs[j];
  }
  if (!(ws >= 1e-14)) atomicOr(bad, 4);
  return wz / ws;
}

// Start of column s of a k x k lower triangle packed by columns.
__device__ inline int cofs(int s, int k) { return s * k - (s * (s - 1)) / 2; }

// Lane j's value of v in every lane (j uniform across the wavefront).
__device__ inline double bcast(double v, int j) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline int bcast_i(int v, int j) {
  return __builtin_amdgcn_readlane(v, j);
}

// sum_{l0 <= l < l1} a[l] b[l] with four partial sums, so that the LDS
// loads of four terms are in flight together
__device__ inline double dot_lds(const double *a, const double *b, int l0,
                                 int l1) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int l = l0;
  for (; l + 4 <= l1; l += 4) {
    const double a0 = a[l], a1 = a[l + 1], a2 = a[l + 2], a3 = a[l + 3];
    const double b0 = b[l], b1 = b[l + 1], b2 = b[l + 2], b3 = b[l + 3];
    s0 += a0 * b0;
    s1 += a1 * b1;
    s2 += a2 * b2;
    s3 += a3 * b3;
  }
  for (; l < l1; ++l) s0 += a[l] * b[l];
  return (s0 + s1) + (s2 + s3);
}

// ---- set-up ----------------------------------------------------------------

// A pseudo-step with every station active, for the selection that decides
// which stations are ever used.
__global__ void k_neb_all_init(int32_t *__restrict__ idx, int64_t N,
                               StepSc *__restrict__ sc) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    idx[i] = (int32_t)i;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->n = (int)N;
    sc->status = 0;
  }
}

__global__ void k_neb_zero(unsigned char *__restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// keep[s] = 1 for every station some cell selected (positions of the
// pseudo-step are station indices).
__global__ void k_neb_keep(const int32_t *__restrict__ nb, int64_t C, int K,
                           int k, unsigned char *__restrict__ keep) {
  const int64_t n = C * (int64_t)k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    keep[nb[(e / k) * K + e % k]] = 1;
}

// Values of stations that are never selected count as missing.
__global__ void k_neb_mask(const double *__restrict__ vals, int64_t n,
                           int64_t N, const unsigned char *__restrict__ keep,
                           double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = keep[i % N] ? vals[i] : NAN;
}

// ---- Z order of the cells ----------------------------------------------------
// One workgroup: the bounding box of the cells, bb = (x0, x1, y0, y1).
__global__ __launch_bounds__(1024) void k_neb_bbox(
    const double *__restrict__ cx, const double *__restrict__ cy, int64_t C,
    double *__restrict__ bb) {
  __shared__ double red[4][16];
  double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int64_t c = threadIdx.x; c < C; c += 1024) {
    v[0] = fmin(v[0], cx[c]);
    v[1] = fmax(v[1], cx[c]);
    v[2] = fmin(v[2], cy[c]);
    v[3] = fmax(v[3], cy[c]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v[0] = wave_min(v[0]);
  v[1] = wave_max(v[1]);
  v[2] = wave_min(v[2]);
  v[3] = wave_max(v[3]);
  if (lane == 0)
    for (int q = 0; q < 4; ++q) red[q][wave] = v[q];
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    double r = red[q][0];
    for (int w = 1; w < 16; ++w)
      r = (q & 1) ? fmax(r, red[q][w]) : fmin(r, red[q][w]);
    bb[q] = r;
  }
}

__device__ inline unsigned spread16(unsigned v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// code[c]: coordinates quantised to 16 bits over the bounding box and
// interleaved (Z order); id[c] = c.
__global__ void k_neb_morton(const double *__restrict__ cx,
                             const double *__restrict__ cy, int64_t C,
                             const double *__restrict__ bb,
                             unsigned *__restrict__ code,
                             int32_t *__restrict__ id) {
  const double x0 = bb[0], y0 = bb[2];
  const double wx = bb[1] - x0, wy = bb[3] - y0;
  const double sx = wx > 0.0 ? 65535.0 / wx : 0.0;
  const double sy = wy > 0.0 ? 65535.0 / wy : 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < C;
       c += (int64_t)gridDim.x * blockDim.x) {
    const double fx = fmin(fmax((cx[c] - x0) * sx, 0.0), 65535.0);
    const double fy = fmin(fmax((cy[c] - y0) * sy, 0.0), 65535.0);
    code[c] = spread16((unsigned)fx) | (spread16((unsigned)fy) << 1);
    id[c] = (int32_t)c;
  }
}

// ---- nrst selection ----------------------------------------------------------
// grid (ceil(C / 64), S), 256 threads.  Tile i holds cells perm[64 i ..
// 64 i + 63], consecutive in Z order, so a tile is a compact patch.  For
// the tile's bounding box
// (centre c, half diagonal h) the block finds a T with at least k active
// stations within T of c; the k nearest of any cell x of the tile then lie
// within T + 2h of c (r_k(x) <= r_k(c) + h, and |s - c| <= |s - x| + h),
// so only those candidates are staged.  Each cell's selection is exact:
// the k smallest (distance, position) pairs, distances as in pyx:141.
// Output: nb[(t C + c) K + r] the r-th selected position into the step's
// active list, ascending; key = step << 56 | hash of the set; cell = c.
__global__ __launch_bounds__(256) void k_neb_nrst(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, int64_t N, const StepSc *__restrict__ sc,
    const double *__restrict__ cx, const double *__restrict__ cy, int64_t C,
    const int32_t *__restrict__ perm, int K, int32_t *__restrict__ nb,
    unsigned long long *__restrict__ key, int32_t *__restrict__ cell) {
  const int t = blockIdx.y;
  const int64_t c0 = (int64_t)blockIdx.x * kNebTile;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = sc[t].n;
  const int k = min(K, n);
  const int32_t *my = idx + (int64_t)t * N;
  const unsigned long long tkey = (unsigned long long)t << kStepShift;
  const int64_t rowbase = (int64_t)t * C;
  __shared__ double cand_x[kNebCand], cand_y[kNebCand];
  __shared__ int cand_j[kNebCand];
  __shared__ double bb[4];
  __shared__ int wcnt[4];
  __shared__ double wred[4];
  if (n <= 0) {
    if (tid < kNebTile && c0 + tid < C) {
      const int64_t c = perm[c0 + tid];
      key[rowbase + c] = tkey;
      cell[rowbase + c] = (int32_t)c;
    }
    return;
  }
  // bounding box of the tile's cells
  if (wave == 0) {
    const int64_t c = perm[c0 + lane < C ? c0 + lane : c0];
    const double x = cx[c], y = cy[c];
    const double x0 = wave_min(x), x1 = wave_max(x);
    const double y0 = wave_min(y), y1 = wave_max(y);
    if (lane == 0) {
      bb[0] = x0;
      bb[1] = x1;
      bb[2] = y0;
      bb[3] = y1;
    }
  }
  __syncthreads();
  const double xc = 0.5 * (bb[0] + bb[1]), yc = 0.5 * (bb[2] + bb[3]);
  const double hx = bb[1] - bb[0], hy = bb[3] - bb[2];
  const double h = 0.5 * sqrt(hx * hx + hy * hy);
  // bisection for T^2: count(|s - c|^2 <= T^2) >= k
  double mx = 0.0;
  for (int j = tid; j < n; j += 256) {
    const double dx = sx[my[j]] - xc, dy = sy[my[j]] - yc;
    mx = fmax(mx, dx * dx + dy * dy);
  }
  mx = wave_max(mx);
  if (lane == 0) wred[wave] = mx;
  __syncthreads();
  double hi = fmax(fmax(wred[0], wred[1]), fmax(wred[2], wred[3]));
  double lo = 0.0;
  for (int it = 0; it < 14; ++it) {
    const double mid = 0.5 * (lo + hi);
    int cnt = 0;
    for (int j = tid; j < n; j += 256) {
      const double dx = sx[my[j]] - xc, dy = sy[my[j]] - yc;
      cnt += (dx * dx + dy * dy) <= mid;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    __syncthreads();  // wcnt of the previous round is read
    if (lane == 0) wcnt[wave] = cnt;
    __syncthreads();
    const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    if (tot >= k)
      hi = mid;
    else
      lo = mid;
  }
  const double R = (sqrt(hi) + 2.0 * h) * (1.0 + 1e-12) + 1e-6;
  const double R2 = R * R;
  // stage the candidates in position order
  int base = 0;
  for (int j0 = 0; j0 < n; j0 += 256) {
    const int j = j0 + tid;
    double x = 0.0, y = 0.0;
    bool f = false;
    if (j < n) {
      x = sx[my[j]];
      y = sy[my[j]];
      const double dx = x - xc, dy = y - yc;
      f = (dx * dx + dy * dy) <= R2;
    }
    const unsigned long long bal = __ballot(f);
    __syncthreads();  // wcnt of the previous round is read
    if (lane == 0) wcnt[wave] = __popcll(bal);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; ++w) off += wcnt[w];
    const int at = off + __popcll(bal & lanes_below(lane));
    if (f && at < kNebCand) {
      cand_x[at] = x;
      cand_y[at] = y;
      cand_j[at] = j;
    }
    base += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
  }
  __syncthreads();
  const int m = base;
  const bool staged = m <= kNebCand;
  // each wave: 16 cells of the tile
  for (int u = 0; u < kNebTile / 4; ++u) {
    const int64_t cs = c0 + wave * (kNebTile / 4) + u;
    if (cs >= C) break;
    const int64_t c = perm[cs];
    const double x = cx[c], y = cy[c];
    int32_t *out = nb + (rowbase + c) * K;
    unsigned long long hsum = 0;
    if (staged) {
      // distances to the staged candidates, kNebSlots per lane
      const int ns = (m + 63) >> 6;
      double dv[kNebSlots];
      double dmax = 0.0;
#pragma unroll
      for (int s = 0; s < kNebSlots; ++s) {
        const int q = s * 64 + lane;
        dv[s] = INFINITY;
        if (s < ns && q < m) {
          dv[s] = dist2(x, y, cand_x[q], cand_y[q]);
          dmax = fmax(dmax, dv[s]);
        }
      }
      dmax = wave_max(dmax);
      // bisection on the distance until exactly k lie at or below it
      // (count(<= lo) < k <= count(<= hi)); a near tie that stops it falls
      // back to the exact order on (distance, position) below
      double lo = -1.0, hi = dmax, thr = -1.0;
      for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break;
        int cnt = 0;
#pragma unroll
        for (int s = 0; s < kNebSlots; ++s)
          if (s < ns) cnt += __popcll(__ballot(dv[s] <= mid));
        if (cnt == k) {
          thr = mid;
          break;
        }
        if (cnt > k)
          hi = mid;
        else
          lo = mid;
      }
      int nsel = 0;
      if (thr >= 0.0) {
#pragma unroll
        for (int s = 0; s < kNebSlots; ++s) {
          if (s >= ns) break;
          const bool take = dv[s] <= thr;
          const unsigned long long bt = __ballot(take);
          if (take) {
            const int r = nsel + __popcll(bt & lanes_below(lane));
            const int v = cand_j[s * 64 + lane];
            out[r] = v;
            hsum += set_mix((unsigned)v, (unsigned)r);
          }
          nsel += __popcll(bt);
        }
      } else {
        // exact: the k-th smallest key (distance bits, monotone for
        // non-negative doubles), ties at it taken by position
        unsigned long long kv[kNebSlots];
#pragma unroll
        for (int s = 0; s < kNebSlots; ++s)
          kv[s] = dv[s] < INFINITY ? dbits(dv[s]) : ~0ull;
        unsigned long long vlo = 0, vhi = uniform_u64(dbits(dmax));
        while (vlo < vhi) {
          const unsigned long long mid = vlo + ((vhi - vlo) >> 1);
          int cnt = 0;
#pragma unroll
          for (int s = 0; s < kNebSlots; ++s)
            cnt += __popcll(__ballot(kv[s] <= mid));
          if (cnt >= k)
            vhi = mid;
          else
            vlo = mid + 1;
        }
        const unsigned long long tk = vhi;
        int nlt = 0;
#pragma unroll
        for (int s = 0; s < kNebSlots; ++s)
          nlt += __popcll(__ballot(kv[s] < tk));
        const int need = k - nlt;  // ties at tk taken by position
        int neq = 0;
#pragma unroll
        for (int s = 0; s < kNebSlots; ++s) {
          const int q = s * 64 + lane;
          const bool lt = kv[s] < tk, eq = kv[s] == tk;
          const unsigned long long beq = __ballot(eq);
          const bool take =
              lt || (eq && neq + __popcll(beq & lanes_below(lane)) < need);
          const unsigned long long bt = __ballot(take);
          if (take) {
            const int r = nsel + __popcll(bt & lanes_below(lane));
            const int v = cand_j[q];
            out[r] = v;
            hsum += set_mix((unsigned)v, (unsigned)r);
          }
          nsel += __popcll(bt);
          neq += __popcll(beq);
        }
      }
    } else {
      // many candidates: repeated wave arg-min over the active stations
      __shared__ int selw[4][kNebMax];
      double cd = -INFINITY;
      int ci = -1;
      for (int r = 0; r < k; ++r) {
        double bd = INFINITY;
        int bi = 0x7fffffff;
        for (int j = lane; j < n; j += 64) {
          const double d = dist2(x, y, sx[my[j]], sy[my[j]]);
          if (lex_less(cd, ci, d, j) && lex_less(d, j, bd, bi)) {
            bd = d;
            bi = j;
          }
        }
        wave_argmin(bd, bi);
        cd = bd;
        ci = bi;
        if (lane == 0) selw[wave][r] = bi;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      for (int r = lane; r < k; r += 64) {
        const int v = selw[wave][r];
        int rank = 0;
        for (int q = 0; q < k; ++q) rank += selw[wave][q] < v;
        out[rank] = v;
        hsum += set_mix((unsigned)v, (unsigned)rank);
      }
      __builtin_amdgcn_wave_barrier();
    }
    hsum = wave_sum_u64(hsum);
    if (lane == 0) {
      key[rowbase + c] = tkey | (hsum >> (64 - kStepShift));
      cell[rowbase + c] = (int32_t)c;
    }
  }
}

// ---- pie selection -----------------------------------------------------------
// grid (C, S), 64 threads, pie_lds_bytes(N, n_pies, K) of LDS.  Every set
// has K entries.  A step with one station selects it K times; one with
// 1 < n < n_nebs sets bit 2 of bad (the reference's row assignment fails
// there, interp/grps.py:245) and writes positions below n so that later
// kernels stay in bounds.
__global__ __launch_bounds__(64) void k_neb_pie(
    const double *__restrict__ sx, const double *__restrict__ sy,
    const int32_t *__restrict__ idx, int64_t N, const StepSc *__restrict__ sc,
    const double *__restrict__ cx, const double *__restrict__ cy, int64_t C,
    int K, int n_pies, int32_t *__restrict__ nb,
    unsigned long long *__restrict__ key, int32_t *__restrict__ cell,
    int *__restrict__ bad) {
  const int t = blockIdx.y;
  const int64_t c = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = sc[t].n;
  const int32_t *my = idx + (int64_t)t * N;
  const int64_t row = (int64_t)t * C + c;
  const unsigned long long tkey = (unsigned long long)t << kStepShift;
  int32_t *out = nb + row * K;
  extern __shared__ double smem[];
  unsigned long long hsum = 0;
  if (n <= 0) {
    if (lane == 0) {
      key[row] = tkey;
      cell[row] = (int32_t)c;
    }
    return;
  }
  if (n == 1) {
    // every row of the reference's selection is the one station, repeated
    // n_nebs times by numpy's broadcast (interp/grps.py:245): a set of K
    // copies, whose singular system the pseudo-inverse solves
    for (int r = lane; r < K; r += 64) {
      out[r] = 0;
      hsum += set_mix(0u, (unsigned)r);
    }
  } else if (n < K) {
    if (lane == 0) atomicOr(bad, 2);
    for (int r = lane; r < K; r += 64) {
      out[r] = r < n ? r : 0;
      hsum += set_mix((unsigned)(r < n ? r : 0), (unsigned)r);
    }
  } else {
    int *sel = nullptr;
    const int nsel = pie_select(
        lane, cx[c], cy[c], n, K, n_pies,
        [&](int j) { return make_double2(sx[my[j]], sy[my[j]]); },
        (char *)smem, &sel, bad);
    if (nsel < K) {
      if (lane == 0) atomicOr(bad, 2);
      for (int r = lane; r < K; r += 64) out[r] = r;
    } else {
      for (int r = lane; r < K; r += 64) {
        const int v = sel[r];
        int rank = 0;
        for (int q = 0; q < K; ++q) rank += sel[q] < v;
        out[rank] = v;
        hsum += set_mix((unsigned)v, (unsigned)rank);
      }
    }
  }
  hsum = wave_sum_u64(hsum);
  if (lane == 0) {
    key[row] = tkey | (hsum >> (64 - kStepShift));
    cell[row] = (int32_t)c;
  }
}

// ---- groups and chunks -------------------------------------------------------
// Entries of a step's sets: min(K, n) for nrst, K for pie (n > 0).
__host__ __device__ inline int neb_set_size(int K, int kfix, int n) {
  return n <= 0 ? 0 : kfix ? K : min(K, n);
}

// After the sort, position p holds a cell of step p / C (the step index
// leads the key).  gs[p] = p where a new group starts (new step, new key,
// or a different set than the predecessor's), else 0.
__global__ void k_neb_mark(const unsigned long long *__restrict__ key,
                           const int32_t *__restrict__ cell,
                           const int32_t *__restrict__ nb,
                           const StepSc *__restrict__ sc, int64_t C, int K,
                           int kfix, int64_t total, int32_t *__restrict__ gs) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * blockDim.x) {
    bool nw = (p % C) == 0 || key[p] != key[p - 1];
    if (!nw) {
      const int64_t t = p / C;
      const int k = neb_set_size(K, kfix, sc[t].n);
      const int32_t *a = nb + (t * C + cell[p]) * K;
      const int32_t *b = nb + (t * C + cell[p - 1]) * K;
      for (int r = 0; r < k && !nw; ++r) nw = a[r] != b[r];
    }
    gs[p] = nw ? (int32_t)p : 0;
  }
}

// cf[p] = 1 at the first cell of each chunk (every kNebChunk cells of a
// group), from the group start gstart[p].
__global__ void k_neb_cflag(const int32_t *__restrict__ gstart, int64_t total,
                            int32_t *__restrict__ cf) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * blockDim.x)
    cf[p] = ((p - gstart[p]) % kNebChunk) == 0;
}

__global__ void k_neb_cwrite(const int32_t *__restrict__ cf,
                             const int32_t *__restrict__ cid, int64_t total,
                             int32_t *__restrict__ cstart,
                             int *__restrict__ nch) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * blockDim.x) {
    if (cf[p]) cstart[cid[p]] = (int32_t)p;
    if (p == total - 1) *nch = cid[p] + cf[p];
  }
}

// ---- per-chunk interpolation ---------------------------------------------------
struct NebArgs {
  const double *sx, *sy;
  int64_t N;
  const int32_t *idx;
  const double *zact;
  int64_t ldz;
  const StepSc *sc;
  const spk_vg *vgs;
  const double *cx, *cy;
  int64_t C;
  const double *stn_drift;   // [N x m]
  const double *cell_drift;  // [m x C]
  int m;
  const int32_t *nb;
  int K;
  int kfix;  // 1: every set has K entries (pie)
  const int32_t *cells;   // sorted cells
  const int32_t *cstart;  // chunk starts
  const int *nch;
  int64_t total;
  int method, want_var, skm;
  double thr, mvg, idw_exp;
  double *est, *var;  // [S x C] of the batch
  int *sflags;
  unsigned long long *skappa;
  int *fail_list, *fail_cnt, *bad;
};

__host__ __device__ inline int neb_p(int method, int m) {
  return method == SPK_M_OK ? 1 : method == SPK_M_EDK ? 1 + m : 0;
}

// Dynamic LDS of k_neb_krige in doubles for K stations and p border
// columns: packed lower triangle, 4 vectors, the border F, the chunk's
// cell coordinates and indices.
__host__ __device__ inline int neb_lds_doubles(int K, int p) {
  return K * (K + 1) / 2 + 4 * K + K * (p > 0 ? p : 1) + 2 * kNebChunk +
         kNebChunk / 2;
}

// Chunk ch of the batch: the step t, the cells [p0, p1) and the set.
struct NebChunk {
  int64_t p0, p1;
  int t, n, k;
};

__device__ inline NebChunk neb_chunk(const NebArgs &a, int ch, int nch) {
  NebChunk q;
  q.p0 = a.cstart[ch];
  q.p1 = ch + 1 < nch ? a.cstart[ch + 1] : a.total;
  q.t = (int)(q.p0 / a.C);
  q.n = a.sc[q.t].n;
  q.k = neb_set_size(a.K, a.kfix, q.n);
  return q;
}

// One 64-thread workgroup per chunk (grid-stride over the chunks).  Lane j
// holds station j of the set.  Kriging in covariance form: C = c0 - G
// (OK/EDK; the constant c0 cancels because the weights sum to one, which
// is also why the reference's nugget has no effect on them) or the
// covariance sum (SK); C = L L^T and M = L^-1 in place; with F the border
// columns (ones, then drifts) and z the set's values
//   beta = (F^T C^-1 F)^-1 F^T C^-1 z,  alpha = C^-1 (z - F beta),
//   estimate = c . alpha + f . beta  (c, f the cell's covariances and
//   border row), the reference's lambda^T z for lambda = K^-1 rhs;
//   OK variance = c0 - |M c|^2 + (a - 1)(a - 2) / b with a = 1^T C^-1 c,
//   b = 1^T C^-1 1, the reference's lambda^T gamma + 2 mu
//   (interp/steps.py:416-434).
template <int PT, bool PH, bool KRIG>
__global__ __launch_bounds__(64) void k_neb_krige(NebArgs a) {
  constexpr int PA = PT > 0 ? PT : 1;
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const int nch = *a.nch;
  const int p = min(neb_p(a.method, a.m), PT);
  const int P = p > 0 ? p : 1;
  const int K = a.K;
  double *A = lds;                    // packed lower triangle
  double *xs = A + K * (K + 1) / 2;   // the set's stations
  double *ys = xs + K;
  double *zs = ys + K;
  double *tv = zs + K;                // 1 / L_jj
  double *F = tv + K;                 // [K x P] border columns
  double *cxl = F + K * P;            // the chunk's cells
  double *cyl = cxl + kNebChunk;
  int *cil = (int *)(cyl + kNebChunk);
  const bool sk = a.method == SPK_M_SK;
  for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
    const NebChunk q = neb_chunk(a, ch, nch);
    if (q.n <= 0) continue;
    const int k = q.k, t = q.t;
    const int np = (int)(q.p1 - q.p0);  // at most kNebChunk
    double *eo = a.est + (int64_t)t * a.C;
    double *vo = a.var ? a.var + (int64_t)t * a.C : nullptr;
    __syncthreads();  // the previous chunk's readers are done
    // the chunk's cells, all loads in flight together
    int cme = 0;
    if (lane < np) {
      cme = a.cells[q.p0 + lane];
      cil[lane] = cme;
      cxl[lane] = a.cx[cme];
      cyl[lane] = a.cy[cme];
    }
    const int32_t *list =
        a.nb + ((int64_t)t * a.C + bcast_i(cme, 0)) * K;
    double zl = 0.0;
    if (lane < k) {
      // positions are below n by construction; the clamp keeps a
      // selection that met NaN coordinates inside the step's list
      const int pos = min(max(list[lane], 0), q.n - 1);
      const int s = a.idx[(int64_t)t * a.N + pos];
      xs[lane] = a.sx[s];
      ys[lane] = a.sy[s];
      zl = a.zact[(int64_t)t * a.ldz + pos];
      zs[lane] = zl;
      if (p > 0) {
        F[lane * P] = 1.0;
        for (int qd = 1; qd < p; ++qd)
          F[lane * P + qd] = a.stn_drift[(int64_t)s * a.m + (qd - 1)];
      }
    }
    const double mean = wave_sum(lane < k ? zl : 0.0) / k;
    const bool anyhi = __ballot(lane < k && zl >= a.thr) != 0;
    __syncthreads();
    // one station: its value everywhere, variance untouched (NaN)
    // (interp/steps.py:280-283)
    if (k == 1) {
      if (lane < np) {
        eo[cme] = zs[0];
        if (vo && a.want_var) vo[cme] = NAN;
      }
      if (lane == 0) atomicOr(&a.sflags[t], SPK_ST_ONESTN);
      continue;
    }
    if constexpr (!KRIG) {
      if (a.method == SPK_M_NNB) {
        if (lane < np) {
          const double x = cxl[lane], y = cyl[lane];
          double bd = INFINITY;
          int bj = 0;
          for (int j = 0; j < k; ++j) {
            const double d = dist2(x, y, xs[j], ys[j]);
            if (d < bd) {
              bd = d;
              bj = j;
            }
          }
          eo[cme] = zs[bj];
        }
        continue;
      }
    }
    bool fill = !anyhi;
    if constexpr (KRIG) {
      const spk_vg &vg = a.vgs[t];
      fill = fill || (vg.sill_sum <= a.mvg) || (vg.range_max_raw <= a.mvg);
    }
    if (fill) {
      // interp/steps.py:760-765, 325-331: the group's mean, variance 0
      if (lane < np) {
        eo[cme] = mean;
        if (vo && a.want_var) vo[cme] = 0.0;
      }
      if (lane == 0) atomicOr(&a.sflags[t], SPK_ST_MEANFILL);
      continue;
    }
    if constexpr (!KRIG) {
      // IDW, interp/steps.py:293-313 with pyx:784-808 over the set
      if (lane < np)
        eo[cme] = idw_set_cell(cxl[lane], cyl[lane], xs, ys, zs, k,
                               a.idw_exp, a.bad);
      continue;
    } else {
    // ---- kriging ----
    // C in LDS as a packed lower triangle by columns: entry (r, s), r >= s,
    // at cofs(s) + r - s, so that a fixed column is contiguous over the
    // lanes that own its rows
    const spk_vg &vg = a.vgs[t];
    const int ntri = k * (k + 1) / 2;
    bool bounded = true;
    for (int qd = 0; qd < vg.n; ++qd) {
      const int ty = vg.type[qd];
      if (ty == SPK_VG_RNG || ty == SPK_VG_POW || ty == SPK_VG_HOL)
        bounded = false;
    }
    for (int e = lane; e < ntri; e += 64) {
      // column s of entry e: the largest s with cofs(s) <= e
      const float b2 = (float)(2 * k + 1);
      int s = (int)((b2 - sqrtf(b2 * b2 - 8.0f * (float)e)) * 0.5f);
      s = max(0, min(s, k - 1));
      while (s > 0 && cofs(s, k) > e) --s;
      while (s + 1 < k && cofs(s + 1, k) <= e) ++s;
      const int r = s + (e - cofs(s, k));
      const double g = vg_entry_t<PH>(vg, dist2(xs[r], ys[r], xs[s], ys[s]),
                                      sk ? 1 : 0, a.mvg);
      A[e] = (sk || !bounded) ? g : vg.sill_sum - g;
    }
    double c0 = sk ? 0.0 : vg.sill_sum;
    if (!sk && !bounded) {
      __syncthreads();
      double mxa = 0.0;
      for (int e = lane; e < ntri; e += 64) mxa = fmax(mxa, fabs(A[e]));
      c0 = 2.0 * wave_max(mxa) + fabs(vg.sill_sum);
      for (int e = lane; e < ntri; e += 64) A[e] = c0 - A[e];
    }
    __syncthreads();
    const double dmaxc = wave_max(lane < k ? A[cofs(lane, k)] : 0.0);
    // Cholesky by columns (left-looking): lane r forms
    // L_rj = (C_rj - sum_{t<j} L_rt L_jt) / L_jj as one dot product over
    // its row, whose loads are independent of each other
    bool fail = false;
    double pmin = INFINITY;
    double *dinv = tv;  // 1 / L_jj
    for (int j = 0; j < k; ++j) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      const bool act = lane >= j && lane < k;
      if (act) {
        int o = lane;       // entry (lane, t) at cofs(t) + lane - t
        int oj = j;         // entry (j, t)
        int t = 0;
        for (; t + 4 <= j; t += 4) {
          const int d1 = k - t - 1, d2 = k - t - 2, d3 = k - t - 3;
          const double r0 = A[o], q0 = A[oj];
          const double r1 = A[o + d1], q1 = A[oj + d1];
          const double r2 = A[o + d1 + d2], q2 = A[oj + d1 + d2];
          const double r3 = A[o + d1 + d2 + d3], q3 = A[oj + d1 + d2 + d3];
          s0 += r0 * q0;
          s1 += r1 * q1;
          s2 += r2 * q2;
          s3 += r3 * q3;
          const int step = d1 + d2 + d3 + (k - t - 4);
          o += step;
          oj += step;
        }
        for (; t < j; ++t) {
          s0 += A[o] * A[oj];
          o += k - t - 1;
          oj += k - t - 1;
        }
      }
      const double sr = (s0 + s1) + (s2 + s3);
      const double arj = act ? A[cofs(j, k) + lane - j] - sr : 0.0;
      const double dj = bcast(arj, j);  // the pivot C_jj - |L_j,<j|^2
      if (!(dj > 0.0) || !isfinite(dj)) {
        fail = true;
        break;
      }
      pmin = fmin(pmin, dj);
      const double ljj = sqrt(dj);
      const double rl = 1.0 / ljj;
      if (act) A[cofs(j, k) + lane - j] = lane == j ? ljj : arj * rl;
      if (lane == 0) dinv[j] = rl;
      __syncthreads();
    }
    if (fail) {
      if (lane == 0) a.fail_list[atomicAdd(a.fail_cnt, 1)] = ch;
      continue;
    }
    // condition estimate max diag(C) / min pivot (a lower bound of the
    // condition number); sets beyond 1e10 take the pseudo-inverse as the
    // reference computes it
    const double kap = dmaxc / pmin;
    if (!(kap <= 1e10)) {
      if (lane == 0) a.fail_list[atomicAdd(a.fail_cnt, 1)] = ch;
      continue;
    }
    // forward solves by columns, lane r holding element r:
    // w = L^-1 1, u = L^-1 (z - shift), G = L^-1 F (column 0 of F is 1)
    const double zsh = a.skm && sk ? mean : 0.0;
    double wl = lane < k ? 1.0 : 0.0;
    double ul = lane < k ? zs[lane] - zsh : 0.0;
    double gl[PA];
    for (int qd = 0; qd < PA; ++qd)
      gl[qd] = (qd > 0 && qd < p && lane < k) ? F[lane * P + qd] : 0.0;
    for (int j = 0; j < k; ++j) {
      const double di = dinv[j];
      const double lrj = (lane > j && lane < k) ? A[cofs(j, k) + lane - j] : 0.0;
      const double wj = bcast(wl, j) * di, uj = bcast(ul, j) * di;
      if (lane == j) {
        wl = wj;
        ul = uj;
      }
      wl -= lrj * wj;
      ul -= lrj * uj;
      for (int qd = 1; qd < p; ++qd) {
        const double gj = bcast(gl[qd], j) * di;
        if (lane == j) gl[qd] = gj;
        gl[qd] -= lrj * gj;
      }
    }
    if (p > 0) gl[0] = wl;
    // beta = (G^T G)^-1 G^T u, by a p x p Cholesky on every lane
    double beta[PA];
    double bsc = 0.0;
    bool bfail = false;
    if (p > 0) {
      double B[PA][PA], g[PA];
      for (int i = 0; i < p; ++i) {
        g[i] = wave_sum(gl[i] * ul);
        for (int j = 0; j <= i; ++j) B[i][j] = wave_sum(gl[i] * gl[j]);
      }
      bsc = B[0][0];
      for (int j = 0; j < p && !bfail; ++j) {
        double d = B[j][j];
        for (int l = 0; l < j; ++l) d -= B[j][l] * B[j][l];
        if (!(d > 0.0) || !isfinite(d)) {
          bfail = true;
          break;
        }
        B[j][j] = sqrt(d);
        for (int i = j + 1; i < p; ++i) {
          double s = B[i][j];
          for (int l = 0; l < j; ++l) s -= B[i][l] * B[j][l];
          B[i][j] = s / B[j][j];
        }
      }
      if (!bfail) {
        for (int i = 0; i < p; ++i) {
          double s = g[i];
          for (int l = 0; l < i; ++l) s -= B[i][l] * beta[l];
          beta[i] = s / B[i][i];
        }
        for (int i = p - 1; i >= 0; --i) {
          double s = beta[i];
          for (int l = i + 1; l < p; ++l) s -= B[l][i] * beta[l];
          beta[i] = s / B[i][i];
        }
      }
    }
    if (bfail) {
      if (lane == 0) a.fail_list[atomicAdd(a.fail_cnt, 1)] = ch;
      continue;
    }
    // backward solves with L^T: alpha = L^-T (u - G beta) (the dual
    // weights); for SK also C^-1 1 = L^-T w (the sum of the weights)
    double al_ = ul;
    for (int qd = 0; qd < p; ++qd) al_ -= gl[qd] * beta[qd];
    double rv_ = wl;
    if (lane >= k) al_ = rv_ = 0.0;
    for (int j = k - 1; j >= 0; --j) {
      const double di = dinv[j];
      // entry (j, lane) for lane < j
      const double ljr = lane < j ? A[cofs(lane, k) + j - lane] : 0.0;
      const double aj = bcast(al_, j) * di;
      if (lane == j) al_ = aj;
      al_ -= ljr * aj;
      if (sk) {
        const double rj = bcast(rv_, j) * di;
        if (lane == j) rv_ = rj;
        rv_ -= ljr * rj;
      }
    }
    if (lane == 0) atomicMax(&a.skappa[t], dbits(kap));
    const bool want_var = a.want_var && vo;
    // ---- cells of the chunk, four at a time ----
    constexpr int NB = 4;
    for (int i0 = 0; i0 < np; i0 += NB) {
      double d[NB], cl[NB], e1[NB], e2[NB], qn[NB], an[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        d[u] = INFINITY;
        cl[u] = 0.0;
        if (lane < k && i0 + u < np) {
          d[u] = dist2(cxl[i0 + u], cyl[i0 + u], xs[lane], ys[lane]);
          const double gm = vg_entry_t<PH>(vg, d[u], sk ? 1 : 0, a.mvg);
          cl[u] = sk ? gm : c0 - gm;
        }
        e1[u] = cl[u] * al_;
        e2[u] = cl[u] * rv_;
        qn[u] = 0.0;
        an[u] = 0.0;
      }
      if (want_var) {
        // y = L^-1 c per cell, one pass over the columns of L for the
        // four; q = |y|^2, a = w . y = 1^T C^-1 c
        double yl[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) yl[u] = cl[u];
        for (int j = 0; j < k; ++j) {
          const double lrj =
              (lane > j && lane < k) ? A[cofs(j, k) + lane - j] : 0.0;
          const double di = dinv[j];
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            const double yj = bcast(yl[u], j) * di;
            if (lane == j) yl[u] = yj;
            yl[u] -= lrj * yj;
          }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          qn[u] = yl[u] * yl[u];
          an[u] = wl * yl[u];
        }
      }
      // one pass of the butterfly for all sums
      for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          e1[u] += __shfl_xor(e1[u], off, 64);
          if (sk) e2[u] += __shfl_xor(e2[u], off, 64);
          if (want_var) {
            qn[u] += __shfl_xor(qn[u], off, 64);
            an[u] += __shfl_xor(an[u], off, 64);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        if (i0 + u >= np) break;
        const int c = cil[i0 + u];
        double e;
        if (sk) {
          if (a.skm) {
            e = mean + e1[u];
          } else if (neb_close1(e2[u])) {
            e = e1[u];
          } else {
            // the sum test fails: nearest station of the set
            // (interp/steps.py:418-427)
            double bd = d[u];
            int bj = lane < k ? lane : 0x7fffffff;
            wave_argmin(bd, bj);
            e = zs[bj < k ? bj : 0];
          }
        } else {
          e = e1[u] + beta[0];
          for (int qd = 1; qd < p; ++qd)
            e += a.cell_drift[(int64_t)(qd - 1) * a.C + c] * beta[qd];
        }
        if (lane == 0) {
          eo[c] = e;
          if (want_var)
            vo[c] = c0 - qn[u] + (an[u] - 1.0) * (an[u] - 2.0) / bsc;
        }
      }
    }
    }  // KRIG
  }
}

// ---- pseudo-inverse fallback ----------------------------------------------------
// The reference's bordered system of a failed chunk (interp/steps.py:
// 181-234), column-major P x P; rows and columns beyond k + p stay zero,
// which adds zero eigenvalues that the cutoff drops.
__global__ __launch_bounds__(256) void k_neb_pinv_assemble(
    NebArgs a, const int *__restrict__ list, int P, double *__restrict__ Kb) {
  const int f = blockIdx.x;
  const int ch = list[f];
  const int nch = *a.nch;
  const NebChunk q = neb_chunk(a, ch, nch);
  const int k = q.k, t = q.t;
  const int p = neb_p(a.method, a.m);
  const int sk = a.method == SPK_M_SK;
  const int32_t *nl = a.nb + ((int64_t)t * a.C + a.cells[q.p0]) * a.K;
  __shared__ int sid[kNebMax];
  for (int j = threadIdx.x; j < k; j += 256)
    sid[j] = a.idx[(int64_t)t * a.N + min(max(nl[j], 0), q.n - 1)];
  __syncthreads();
  const spk_vg &vg = a.vgs[t];
  double *Kf = Kb + (int64_t)f * P * P;
  for (int e = threadIdx.x; e < P * P; e += 256) {
    const int i = e % P, j = e / P;
    double v = 0.0;
    if (i < k && j < k) {
      v = vg_entry(vg, dist2(a.sx[sid[i]], a.sy[sid[i]], a.sx[sid[j]],
                             a.sy[sid[j]]),
                   sk, a.mvg);
    } else if (i < k && j < k + p) {
      v = j == k ? 1.0 : a.stn_drift[(int64_t)sid[i] * a.m + (j - k - 1)];
    } else if (j < k && i < k + p) {
      v = i == k ? 1.0 : a.stn_drift[(int64_t)sid[j] * a.m + (i - k - 1)];
    }
    Kf[e] = v;
  }
}

// One workgroup per failed chunk: pinv(K) = V diag(1/D) V^T over the
// eigenvalues above numpy's cutoff 1e-15 max|D|, then per cell lambda =
// pinv(K) rhs, the sum test and the reference's estimate and variance
// (interp/steps.py:403-435).  A failed eigendecomposition leaves the
// nearest station's value (interp/steps.py:363-384).
__global__ __launch_bounds__(256) void k_neb_pinv_cells(
    NebArgs a, const int *__restrict__ list, int P,
    const double *__restrict__ Vb, const double *__restrict__ Db,
    const int *__restrict__ infob) {
  extern __shared__ double lds[];
  double *Pi = lds;                 // [P x P]
  double *R = Pi + P * P;           // [64 cells x P] right-hand sides
  __shared__ double xs[kNebMax], ys[kNebMax], zs[kNebMax];
  __shared__ int sid[kNebMax];
  const int f = blockIdx.x, tid = threadIdx.x;
  const int ch = list[f];
  const int nch = *a.nch;
  const NebChunk q = neb_chunk(a, ch, nch);
  const int k = q.k, t = q.t;
  const int p = neb_p(a.method, a.m);
  const int np_ = k + p;
  const int sk = a.method == SPK_M_SK;
  const int32_t *nl = a.nb + ((int64_t)t * a.C + a.cells[q.p0]) * a.K;
  const double *V = Vb + (int64_t)f * P * P;
  const double *D = Db + (int64_t)f * P;
  for (int j = tid; j < k; j += 256) {
    const int pos = min(max(nl[j], 0), q.n - 1);
    const int s = a.idx[(int64_t)t * a.N + pos];
    sid[j] = s;
    xs[j] = a.sx[s];
    ys[j] = a.sy[s];
    zs[j] = a.zact[(int64_t)t * a.ldz + pos];
  }
  const double mx = fmax(fabs(D[0]), fabs(D[P - 1]));
  const double cut = 1e-15 * mx;
  const bool ok = infob[f] == 0;
  for (int e = tid; e < P * P; e += 256) {
    const int i = e % P, j = e / P;
    double s = 0.0;
    if (ok)
      for (int l = 0; l < P; ++l) {
        const double dl = D[l];
        if (fabs(dl) > cut)
          s += V[i + (int64_t)l * P] * (1.0 / dl) * V[j + (int64_t)l * P];
      }
    Pi[e] = s;
  }
  double mean = 0.0;
  for (int j = 0; j < k; ++j)
    mean += a.zact[(int64_t)t * a.ldz + min(max(nl[j], 0), q.n - 1)];
  mean /= k;
  if (tid == 0) {
    double mn = INFINITY;
    int dropped = 0;
    for (int l = 0; l < P; ++l) {
      if (fabs(D[l]) > cut)
        mn = fmin(mn, fabs(D[l]));
      else
        ++dropped;
    }
    const double kap = ok ? mx / mn : INFINITY;
    int fl = ok ? SPK_ST_PINV : SPK_ST_FAILED;
    if (!ok || dropped > P - np_ || !(kap <= 1e12)) fl |= SPK_ST_ILLCOND;
    atomicOr(&a.sflags[t], fl);
    if (ok) atomicMax(&a.skappa[t], dbits(kap));
  }
  __syncthreads();
  const spk_vg &vg = a.vgs[t];
  double *eo = a.est + (int64_t)t * a.C;
  double *vo = a.var ? a.var + (int64_t)t * a.C : nullptr;
  const bool want_var = a.want_var && vo;
  const int64_t ncell = q.p1 - q.p0;
  // thread per cell (at most 64 per chunk)
  if (tid < ncell) {
    const int c = a.cells[q.p0 + tid];
    const double x = a.cx[c], y = a.cy[c];
    double *r = R + tid * P;
    double bd = INFINITY;
    int bj = 0;
    for (int j = 0; j < k; ++j) {
      const double d = dist2(x, y, xs[j], ys[j]);
      if (d < bd) {
        bd = d;
        bj = j;
      }
      r[j] = vg_entry(vg, d, sk, a.mvg);
    }
    if (p > 0) r[k] = 1.0;
    for (int qd = 1; qd < p; ++qd)
      r[k + qd] = a.cell_drift[(int64_t)(qd - 1) * a.C + c];
    double lsum = 0.0, est = 0.0, vs = 0.0, lk = 0.0;
    for (int i = 0; i < np_; ++i) {
      double li = 0.0;
      for (int j = 0; j < np_; ++j) li += Pi[i + j * P] * r[j];
      if (i < k) {
        lsum += li;
        est += li * (a.skm && sk ? zs[i] - mean : zs[i]);
      }
      vs += li * r[i];
      if (i == k) lk = li;
    }
    if (a.skm && sk && ok) {
      eo[c] = mean + est;
    } else if (ok && neb_close1(lsum)) {
      eo[c] = est;
      if (want_var) vo[c] = vs + lk;
    } else {
      eo[c] = zs[bj];
      if (want_var) vo[c] = 0.0;
    }
  }
}

__global__ void k_neb_out(const StepSc *__restrict__ sc,
                          const int *__restrict__ sflags,
                          const unsigned long long *__restrict__ skappa,
                          int S, int32_t *__restrict__ info,
                          double *__restrict__ kappa) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  info[s] = (sc[s].n == 0 ? SPK_ST_NOSTNS : 0) | sflags[s];
  if (kappa) kappa[s] = __longlong_as_double((long long)skappa[s]);
}

__global__ void k_neb_reset(int *__restrict__ sflags,
                            unsigned long long *__restrict__ skappa, int S,
                            int *__restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) {
    sflags[s] = 0;
    skappa[s] = 0ull;
  }
  if (s < 4) cnt[s] = 0;
}

}  // namespace spk

using namespace spk;

static int neb_blas(rocblas_status st, const char *what) {
  if (st != rocblas_status_success)
    return set_error(SPK_EHIP, "%s failed (rocblas status %d)", what,
                     (int)st);
  return SPK_OK;
}

extern "C" {

int spk_interp_neb_steps_dev(spk_ctx *ctx, const spk_problem *prob,
                             const spk_opts *opts, const spk_neb *neb,
                             int64_t T, const double *vals, const spk_vg *vgs,
                             double *est, double *var, int32_t *info,
                             double *kappa) {
  if (!ctx || !prob || !opts || !neb) return set_error(SPK_EARG, "null argument");
  if (T < 0) return set_error(SPK_ESHAPE, "negative step count");
  if (T == 0) return SPK_OK;
  const int64_t N = prob->n_stns, C = prob->n_cells;
  if (N < 1 || C < 1) return set_error(SPK_ESHAPE, "need stations and cells");
  if (N > (int64_t)1 << 20)
    return set_error(SPK_ESHAPE, "more than 2^20 stations");
  if (C > ((int64_t)1 << 31) - 1)
    return set_error(SPK_ESHAPE, "more than 2^31 - 1 cells");
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
  const int sel = neb->sel;
  if (sel != SPK_NEB_NRST && sel != SPK_NEB_PIE)
    return set_error(SPK_EARG, "unknown neighbour selection %d", sel);
  const int K = neb->n_nebs;
  if (K < 1 || K > kNebMax)
    return set_error(SPK_EARG, "n_nebs must be in 1..%d", kNebMax);
  const int n_pies = neb->n_pies;
  if (sel == SPK_NEB_PIE) {
    if (n_pies < 1 || n_pies > K)
      return set_error(SPK_EARG, "need 0 < n_pies <= n_nebs "
                       "(interp/grps.py:35-37)");
    if (N < K)
      return set_error(SPK_ESHAPE, "pie needs at least n_nebs stations");
    if (pie_lds_bytes((int)N, n_pies, K) > 150 * 1024)
      return set_error(SPK_ESHAPE, "pie selection over %lld stations does "
                       "not fit the LDS", (long long)N);
  }
  const bool want_var = opts->want_var && method == SPK_M_OK && var;
  const int skm = method == SPK_M_SK && opts->sk_mean ? 1 : 0;
  SPK_HIP(hipSetDevice(ctx->device));
  hipStream_t sm = ctx->stream;

  double *sx, *sy, *cx, *cy;
  SPK_TRY(ws_get(ctx, "stn_xy", 2 * N * sizeof(double), (void **)&sx));
  sy = sx + N;
  SPK_TRY(ws_get(ctx, "cell_xy", 2 * C * sizeof(double), (void **)&cx));
  cy = cx + C;
  auto g1 = [](int64_t n) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)));
  };
  hipLaunchKernelGGL(k_recenter, g1(N), dim3(256), 0, sm, prob->stn_x, sx, N,
                     prob->origin_x);
  hipLaunchKernelGGL(k_recenter, g1(N), dim3(256), 0, sm, prob->stn_y, sy, N,
                     prob->origin_y);
  hipLaunchKernelGGL(k_recenter, g1(C), dim3(256), 0, sm, prob->cell_x, cx, C,
                     prob->origin_x);
  hipLaunchKernelGGL(k_recenter, g1(C), dim3(256), 0, sm, prob->cell_y, cy, C,
                     prob->origin_y);
  SPK_TRY(check_launch("k_recenter"));

  // batch size: selections [S C K] int32 plus about 40 bytes per cell
  const int64_t per_step = C * (int64_t)K * 4 + C * 40 + N * 20;
  const int64_t budget = (int64_t)4 << 30;
  int64_t S = std::max<int64_t>(1, std::min<int64_t>(budget / per_step, 64));
  S = std::min<int64_t>(S, (((int64_t)1 << 31) - 1) / C);
  S = std::max<int64_t>(1, std::min<int64_t>(S, T));
  const int64_t SC = S * C;
  const int64_t ldz = round_up(N, 64);
  int32_t *idx, *nbuf, *cell0, *cell1, *gs, *gst, *cf, *cid, *cst, *flist;
  unsigned long long *key0, *key1, *skap;
  double *zact, *vm;
  StepSc *sc;
  unsigned char *keep;
  int *cnt, *sfl;
  SPK_TRY(ws_get(ctx, "neb_idx", S * N * 4, (void **)&idx));
  SPK_TRY(ws_get(ctx, "neb_zact", S * ldz * 8, (void **)&zact));
  SPK_TRY(ws_get(ctx, "neb_vm", S * N * 8, (void **)&vm));
  SPK_TRY(ws_get(ctx, "neb_sc", S * sizeof(StepSc), (void **)&sc));
  SPK_TRY(ws_get(ctx, "neb_nb", SC * K * 4, (void **)&nbuf));
  SPK_TRY(ws_get(ctx, "neb_key", 2 * SC * 8, (void **)&key0));
  key1 = key0 + SC;
  SPK_TRY(ws_get(ctx, "neb_cell", 2 * SC * 4, (void **)&cell0));
  cell1 = cell0 + SC;
  SPK_TRY(ws_get(ctx, "neb_grp", 6 * SC * 4, (void **)&gs));
  gst = gs + SC;
  cf = gst + SC;
  cid = cf + SC;
  cst = cid + SC;
  flist = cst + SC;
  SPK_TRY(ws_get(ctx, "neb_keep", round_up(N, 256), (void **)&keep));
  SPK_TRY(ws_get(ctx, "neb_cnt", 64, (void **)&cnt));
  SPK_TRY(ws_get(ctx, "neb_sfl", S * 4 + 64, (void **)&sfl));
  SPK_TRY(ws_get(ctx, "neb_skap", S * 8 + 64, (void **)&skap));
  // cnt[0] chunks, cnt[1] failed chunks, cnt[2] bad bits
  int *nch = cnt, *fcnt = cnt + 1, *bad = cnt + 2;
  {
    const size_t need = 64;
    if (ctx->h_sc_bytes < need) {
      if (ctx->h_sc) SPK_HIP(hipHostFree(ctx->h_sc));
      ctx->h_sc = nullptr;
      ctx->h_sc_bytes = 0;
      SPK_HIP(hipHostMalloc(&ctx->h_sc, need));
      ctx->h_sc_bytes = need;
    }
  }
  int *h_cnt = static_cast<int *>(ctx->h_sc);
  // cells in Z order: the nrst selection tiles
  int32_t *perm, *mid;
  unsigned *mcode;
  double *bbd;
  SPK_TRY(ws_get(ctx, "neb_perm", 2 * C * 4, (void **)&perm));
  mid = perm + C;
  SPK_TRY(ws_get(ctx, "neb_mcode", 2 * C * 4 + 64, (void **)&mcode));
  bbd = (double *)(mcode + 2 * C + 8);
  size_t sort_bytes = 0, scan1 = 0, scan2 = 0, zsort = 0;
  SPK_HIP(hipcub::DeviceRadixSort::SortPairs(
      nullptr, sort_bytes, key0, key1, cell0, cell1, (int)SC, 0, 64, sm));
  SPK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, zsort, mcode,
                                             mcode + C, mid, perm, (int)C,
                                             0, 32, sm));
  SPK_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, scan1, gs, gst,
                                            hipcub::Max(), (int)SC, sm));
  SPK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan2, cf, cid, (int)SC,
                                           sm));
  void *tmp;
  const size_t tmp_bytes =
      std::max(std::max(sort_bytes, zsort), std::max(scan1, scan2));
  SPK_TRY(ws_get(ctx, "neb_cub", tmp_bytes, &tmp));
  const size_t pie_lds = pie_lds_bytes((int)N, n_pies, K);
  if (sel == SPK_NEB_PIE && pie_lds > 64 * 1024)
    SPK_HIP(hipFuncSetAttribute((const void *)k_neb_pie,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)pie_lds));
  auto select = [&](const int32_t *ix, const StepSc *s_, int Sb) -> int {
    ProfScope ps(ctx, "neb_select");
    if (sel == SPK_NEB_NRST)
      hipLaunchKernelGGL(k_neb_nrst, dim3((unsigned)ceil_div(C, kNebTile), Sb),
                         dim3(256), 0, sm, sx, sy, ix, N, s_, cx, cy, C,
                         (const int32_t *)perm, K, nbuf, key0, cell0);
    else
      hipLaunchKernelGGL(k_neb_pie, dim3((unsigned)C, Sb), dim3(64), pie_lds,
                         sm, sx, sy, ix, N, s_, cx, cy, C, K, n_pies, nbuf,
                         key0, cell0, bad);
    return check_launch("neighbour selection");
  };
  auto read_cnt = [&]() -> int {
    SPK_HIP(hipMemcpyAsync(h_cnt, cnt, 4 * sizeof(int), hipMemcpyDeviceToHost,
                           sm));
    SPK_HIP(hipStreamSynchronize(sm));
    return SPK_OK;
  };
  auto bad_error = [&](int hb) -> int {
    if (hb & 1)
      return set_error(SPK_EINDEX, "pie index out of range (an angle rounded "
                       "to 2*pi, pyx:868-870)");
    if (hb & 2)
      return set_error(SPK_ESHAPE, "a step has fewer active stations than "
                       "n_nebs (pie)");
    return SPK_OK;
  };

  if (sel == SPK_NEB_NRST) {
    hipLaunchKernelGGL(k_neb_bbox, dim3(1), dim3(1024), 0, sm, cx, cy, C, bbd);
    hipLaunchKernelGGL(k_neb_morton, g1(C), dim3(256), 0, sm, cx, cy, C,
                       (const double *)bbd, mcode, mid);
    SPK_TRY(check_launch("k_neb_morton"));
    size_t b = tmp_bytes;
    SPK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b, mcode, mcode + C, mid,
                                               perm, (int)C, 0, 32, sm));
  }
  // stations that any cell selects with every station present
  // (interp/steps.py:592-608)
  hipLaunchKernelGGL(k_neb_reset, dim3(1), dim3(64), 0, sm, sfl,
                     (unsigned long long *)nullptr, 0, cnt);
  hipLaunchKernelGGL(k_neb_all_init, g1(N), dim3(256), 0, sm, idx, N, sc);
  hipLaunchKernelGGL(k_neb_zero, g1(N), dim3(256), 0, sm, keep, N);
  SPK_TRY(check_launch("k_neb_all_init"));
  SPK_TRY(select(idx, sc, 1));
  const int k_all = (int)std::min<int64_t>(K, N);
  hipLaunchKernelGGL(k_neb_keep, g1(C * k_all), dim3(256), 0, sm, nbuf, C, K,
                     k_all, keep);
  SPK_TRY(check_launch("k_neb_keep"));
  SPK_TRY(read_cnt());
  SPK_TRY(bad_error(h_cnt[2]));

  const int p = neb_p(method, m);
  const int kfix = sel == SPK_NEB_PIE ? 1 : 0;
  // Pow or Hol components anywhere in the call: the kernel variant with
  // pow and sin (they cost registers, and so occupancy, when compiled in)
  bool ph = false;
  if (krig) {
    std::vector<spk_vg> hv((size_t)T);
    SPK_HIP(hipMemcpyAsync(hv.data(), vgs, T * sizeof(spk_vg),
                           hipMemcpyDeviceToHost, sm));
    SPK_HIP(hipStreamSynchronize(sm));
    for (const spk_vg &v : hv)
      for (int qd = 0; qd < v.n && qd < SPK_VG_MAX; ++qd)
        ph = ph || v.type[qd] == SPK_VG_POW || v.type[qd] == SPK_VG_HOL;
  }
  const int lds_krige = neb_lds_doubles(K, p) * 8;
  int hbad = 0;
  for (int64_t t0 = 0; t0 < T; t0 += S) {
    const int Sb = (int)std::min<int64_t>(S, T - t0);
    const int64_t tot = Sb * C;
    const spk_vg *gb = krig ? vgs + t0 : nullptr;
    hipLaunchKernelGGL(k_neb_reset, dim3((unsigned)ceil_div(Sb, 64)), dim3(64),
                       0, sm, sfl, skap, Sb, cnt);
    hipLaunchKernelGGL(k_neb_mask, g1(Sb * N), dim3(256), 0, sm, vals + t0 * N,
                       Sb * N, N, keep, vm);
    hipLaunchKernelGGL(k_compact, dim3(Sb), dim3(256), 0, sm, vm, N, idx, zact,
                       ldz, sc, gb, method, opts->min_var_thr, opts->min_vg_val,
                       (const int32_t *)nullptr);
    SPK_TRY(check_launch("k_compact"));
    SPK_TRY(select(idx, sc, Sb));
    {
      ProfScope ps(ctx, "neb_group");
      size_t b = tmp_bytes;
      SPK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b, key0, key1, cell0,
                                                 cell1, (int)tot, 0, 64, sm));
      hipLaunchKernelGGL(k_neb_mark, g1(tot), dim3(256), 0, sm, key1, cell1,
                         nbuf, sc, C, K, kfix, tot, gs);
      b = tmp_bytes;
      SPK_HIP(hipcub::DeviceScan::InclusiveScan(tmp, b, gs, gst, hipcub::Max(),
                                                (int)tot, sm));
      hipLaunchKernelGGL(k_neb_cflag, g1(tot), dim3(256), 0, sm, gst, tot, cf);
      b = tmp_bytes;
      SPK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, b, cf, cid, (int)tot, sm));
      hipLaunchKernelGGL(k_neb_cwrite, g1(tot), dim3(256), 0, sm, cf, cid, tot,
                         cst, nch);
      SPK_TRY(check_launch("neighbour groups"));
    }
    NebArgs a;
    a.sx = sx;
    a.sy = sy;
    a.N = N;
    a.idx = idx;
    a.zact = zact;
    a.ldz = ldz;
    a.sc = sc;
    a.vgs = gb;
    a.cx = cx;
    a.cy = cy;
    a.C = C;
    a.stn_drift = prob->stn_drift;
    a.cell_drift = prob->cell_drift;
    a.m = m;
    a.nb = nbuf;
    a.K = K;
    a.kfix = kfix;
    a.cells = cell1;
    a.cstart = cst;
    a.nch = nch;
    a.total = tot;
    a.method = method;
    a.want_var = want_var ? 1 : 0;
    a.skm = skm;
    a.thr = opts->min_var_thr;
    a.mvg = opts->min_vg_val;
    a.idw_exp = opts->idw_exp;
    a.est = est + t0 * C;
    a.var = want_var ? var + t0 * C : nullptr;
    a.sflags = sfl;
    a.skappa = skap;
    a.fail_list = flist;
    a.fail_cnt = fcnt;
    a.bad = bad;
    {
      ProfScope ps(ctx, "neb_krige");
      const unsigned grid =
          (unsigned)std::max<int64_t>(1, std::min<int64_t>(tot, 16384));
      const dim3 g(grid), b(64);
      if (!krig)
        hipLaunchKernelGGL((k_neb_krige<0, false, false>), g, b, lds_krige,
                           sm, a);
      else if (method == SPK_M_OK && ph)
        hipLaunchKernelGGL((k_neb_krige<1, true, true>), g, b, lds_krige, sm,
                           a);
      else if (method == SPK_M_OK)
        hipLaunchKernelGGL((k_neb_krige<1, false, true>), g, b, lds_krige, sm,
                           a);
      else if (method == SPK_M_EDK && ph)
        hipLaunchKernelGGL((k_neb_krige<kMaxP, true, true>), g, b, lds_krige,
                           sm, a);
      else if (method == SPK_M_EDK)
        hipLaunchKernelGGL((k_neb_krige<kMaxP, false, true>), g, b,
                           lds_krige, sm, a);
      else if (ph)
        hipLaunchKernelGGL((k_neb_krige<0, true, true>), g, b, lds_krige, sm,
                           a);
      else
        hipLaunchKernelGGL((k_neb_krige<0, false, true>), g, b, lds_krige, sm,
                           a);
      SPK_TRY(check_launch("k_neb_krige"));
    }
    SPK_TRY(read_cnt());
    hbad |= h_cnt[2];
    const int nf = h_cnt[1];
    prof_count(ctx, "neb_chunks", h_cnt[0]);
    prof_count(ctx, "neb_pinv_chunks", nf);
    if (nf > 0) {
      ProfScope ps(ctx, "neb_pinv");
      const int P = K + p;
      double *Kb, *Db, *Eb;
      int *ib;
      SPK_TRY(ws_get(ctx, "neb_pK", (size_t)nf * P * P * 8, (void **)&Kb));
      SPK_TRY(ws_get(ctx, "neb_pD", (size_t)nf * P * 16, (void **)&Db));
      Eb = Db + (size_t)nf * P;
      SPK_TRY(ws_get(ctx, "neb_pI", (size_t)nf * 4 + 64, (void **)&ib));
      hipLaunchKernelGGL(k_neb_pinv_assemble, dim3(nf), dim3(256), 0, sm, a,
                         (const int *)flist, P, Kb);
      SPK_TRY(check_launch("k_neb_pinv_assemble"));
      if (!ctx->blas) {
        rocblas_handle h = nullptr;
        SPK_TRY(neb_blas(rocblas_create_handle(&h), "rocblas_create_handle"));
        ctx->blas = h;
      }
      rocblas_handle h = (rocblas_handle)ctx->blas;
      SPK_TRY(neb_blas(rocblas_set_stream(h, sm), "rocblas_set_stream"));
      SPK_TRY(neb_blas(rocsolver_dsyevd_strided_batched(
                           h, rocblas_evect_original, rocblas_fill_lower, P,
                           Kb, P, (rocblas_stride)P * P, Db, P, Eb, P, ib, nf),
                       "rocsolver_dsyevd_strided_batched"));
      const size_t lds = ((size_t)P * P + 64 * (size_t)P) * 8;
      if (lds > 64 * 1024)
        SPK_HIP(hipFuncSetAttribute((const void *)k_neb_pinv_cells,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
      hipLaunchKernelGGL(k_neb_pinv_cells, dim3(nf), dim3(256), lds, sm, a,
                         (const int *)flist, P, (const double *)Kb,
                         (const double *)Db, (const int *)ib);
      SPK_TRY(check_launch("k_neb_pinv_cells"));
    }
    hipLaunchKernelGGL(k_neb_out, dim3((unsigned)ceil_div(Sb, 64)), dim3(64), 0,
                       sm, sc, sfl, skap, Sb, info + t0, kappa ? kappa + t0 : nullptr);
    SPK_TRY(check_launch("k_neb_out"));
  }
  if (hbad & 4)
    return set_error(SPK_EARG, "IDW weight sum below 1e-14 "
                     "(interp/steps.py:305)");
  SPK_TRY(bad_error(hbad));
  return SPK_OK;
}

}  // extern "C"
