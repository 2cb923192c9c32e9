// Variogram estimation and fitting for many steps at once
// (variograms/vgs.py:20-870, variograms/fitvgs.py:143-290).
//
//  spk_vgfit_pairs_dev : all station pairs (i > j) sorted by distance, once
//                        per station set (stable radix sort: equal distances
//                        keep the reference's generation order, i then j)
//  spk_vgfit_evg_dev   : per step, the pairs whose two stations have a
//                        value, the mdr distance clip (vgs.py:597-659) and
//                        bins of nk consecutive pairs with the robust or
//                        classic estimator (vgs.py:399-431, 495-520)
//  spk_vgfit_de_dev    : differential evolution (best1bin, dither
//                        U(0.5, 1) per generation, CR 0.7, Latin-hypercube
//                        start, stop when std(E) <= tol |mean(E)|) of every
//                        (step, candidate model) pair, one workgroup each;
//                        the objective is vgs.py:661-678 without weights.
#include <hipcub/hipcub.hpp>

#include "spk_internal.h"

namespace spk {

// ---- pairs ---------------------------------------------------------------
__global__ void k_pair_keys(const double *__restrict__ x,
                            const double *__restrict__ y, int64_t N,
                            unsigned long long *__restrict__ key,
                            int64_t *__restrict__ val) {
  const int64_t M = N * (N - 1) / 2;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < M;
       k += (int64_t)gridDim.x * blockDim.x) {
    // k = i (i - 1) / 2 + j, 0 <= j < i
    int64_t i = (int64_t)floor((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
    while (i * (i - 1) / 2 > k) --i;
    while ((i + 1) * i / 2 <= k) ++i;
    const int64_t j = k - i * (i - 1) / 2;
    // vgs.py:327-329: sqrt((x_i - x_j)^2 + (y_i - y_j)^2)
    const double dx = x[i] - x[j], dy = y[i] - y[j];
    double d;
    {
#pragma clang fp contract(off)
      d = sqrt(dx * dx + dy * dy);
    }
    key[k] = __double_as_longlong(d);  // d >= 0: bit order == value order
    val[k] = k;
  }
}

__global__ void k_pair_unpack(const unsigned long long *__restrict__ key,
                              const int64_t *__restrict__ val, int64_t M,
                              int32_t *__restrict__ pi, int32_t *__restrict__ pj,
                              double *__restrict__ pd) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < M;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = val[e];
    int64_t i = (int64_t)floor((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
    while (i * (i - 1) / 2 > k) --i;
    while ((i + 1) * i / 2 <= k) ++i;
    pi[e] = (int32_t)i;
    pj[e] = (int32_t)(k - i * (i - 1) / 2);
    pd[e] = __longlong_as_double((long long)key[e]);
  }
}

// ---- empirical variograms ----------------------------------------------------
// One workgroup per step.  list: per-step scratch of M int32 (indices into
// the sorted pair arrays of the kept pairs).
__global__ __launch_bounds__(256) void k_evg(
    const int32_t *__restrict__ pi, const int32_t *__restrict__ pj,
    const double *__restrict__ pd, int64_t M, const double *__restrict__ vals,
    int64_t N, double mdr, int nk, int classic, int max_bins,
    int32_t *__restrict__ list, double *__restrict__ bh,
    double *__restrict__ bg, int32_t *__restrict__ bn,
    int32_t *__restrict__ nbins) {
  const int s = blockIdx.x;
  const double *z = vals + (int64_t)s * N;
  int32_t *L = list + (int64_t)s * M;
  __shared__ int wcnt[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int64_t base = 0;
  for (int64_t b0 = 0; b0 < M; b0 += 256) {
    const int64_t e = b0 + tid;
    bool act = false;
    if (e < M) act = !isnan(z[pi[e]]) && !isnan(z[pj[e]]);
    const unsigned long long m = __ballot(act);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; ++w) off += wcnt[w];
    if (act) L[base + off + pre] = (int32_t)e;
    const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    __syncthreads();
    base += tot;
  }
  // base = active pairs.  Distance clip: argmin |d - mdr max| (first
  // occurrence) over the sorted active distances.
  __shared__ int64_t keep_s;
  if (tid == 0) {
    int64_t keep = base;
    if (base > 0) {
      const double act_max = pd[L[base - 1]];
      const double md = act_max * mdr;
      // first active position with d > md
      int64_t lo = 0, hi = base;
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (pd[L[mid]] > md) hi = mid; else lo = mid + 1;
      }
      int64_t idx;
      if (lo == 0) {
        idx = 0;
      } else if (lo == base) {
        idx = base - 1;
      } else {
        const double da = md - pd[L[lo - 1]], db = pd[L[lo]] - md;
        idx = (da <= db) ? lo - 1 : lo;
      }
      // first occurrence of that distance value
      const double dv = pd[L[idx]];
      {
        int64_t a = 0, b2 = idx;
        while (a < b2) {
          const int64_t mid = (a + b2) / 2;
          if (pd[L[mid]] < dv) a = mid + 1; else b2 = mid;
        }
        idx = a;
      }
      double clip_ul = pd[L[idx]];
      if (!(clip_ul - md > 0)) clip_ul = md;
      if (idx != base - 1) {
        // keep d < clip_ul
        int64_t a = 0, b2 = base;
        while (a < b2) {
          const int64_t mid = (a + b2) / 2;
          if (pd[L[mid]] < clip_ul) a = mid + 1; else b2 = mid;
        }
        keep = a;
      }
    }
    keep_s = keep;
  }
  __syncthreads();
  const int64_t keep = keep_s;
  int nb = (int)((keep + nk - 1) / nk);
  if (nb > max_bins) nb = max_bins;
  if (tid == 0) nbins[s] = nb;
  for (int b = tid; b < nb; b += 256) {
    const int64_t p0 = (int64_t)b * nk;
    const int cntb = (int)min((int64_t)nk, keep - p0);
    double sd = 0.0, sr = 0.0;
    for (int q = 0; q < cntb; ++q) {
      const int32_t e = L[p0 + q];
      sd += pd[e];
      sr += sqrt(fabs(z[pi[e]] - z[pj[e]]));
    }
    double g;
    if (!classic) {
      // vgs.py:414-422
      const double t = sr / cntb;
      const double t4 = (t * t) * (t * t);
      g = (t4 / (0.457 + (0.494 / cntb))) / 2.0;
    } else {
      // vgs.py:399-412: median of squared differences (insertion sort)
      double v[64];
      const int c2 = cntb < 64 ? cntb : 64;
      for (int q = 0; q < c2; ++q) {
        const int32_t e = L[p0 + q];
        const double dz = z[pi[e]] - z[pj[e]];
        double key = dz * dz;
        int w2 = q - 1;
        while (w2 >= 0 && v[w2] > key) {
          v[w2 + 1] = v[w2];
          --w2;
        }
        v[w2 + 1] = key;
      }
      g = (c2 & 1) ? v[c2 / 2] : 0.5 * (v[c2 / 2 - 1] + v[c2 / 2]);
    }
    bh[(int64_t)s * max_bins + b] = sd / cntb;
    bg[(int64_t)s * max_bins + b] = g;
    bn[(int64_t)s * max_bins + b] = cntb;
  }
}

// ---- differential evolution ----------------------------------------------------
struct Rng {
  unsigned long long s;
  __device__ explicit Rng(unsigned long long seed) : s(seed ? seed : 1ull) {}
  __device__ inline unsigned long long next() {
    // xorshift64*
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s * 2685821657736338717ull;
  }
  __device__ inline double uni() {  // [0, 1)
    return (double)(next() >> 11) * (1.0 / 9007199254740992.0);
  }
};

__device__ inline unsigned long long mix64(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// vgs.py:533-590 fit models (Sph and Lin switch at h > r)
__device__ inline double fit_model(int t, double h, double r, double s) {
  switch (t) {
    case SPK_VG_NUG: return s;
    case SPK_VG_SPH:
      return h > r ? s : s * ((1.5 * h) / r - (h * h * h) / (2.0 * r * r * r));
    case SPK_VG_EXP: return s * (1.0 - exp(-3.0 * (h / r)));
    case SPK_VG_LIN: return h > r ? s : s * (h / r);
    case SPK_VG_GAU: return s * (1.0 - exp(-3.0 * ((h * h) / (r * r))));
    case SPK_VG_POW: return s * pow(h, r);
    case SPK_VG_HOL: {
      if (!(h > 0)) return 0.0;
      const double x = (kPi * h) / r;
      return s * (1.0 - sin(x) / x);
    }
    default: return NAN;
  }
}

constexpr int kDeMaxD = 8;
constexpr int kDeMaxPop = 800;

// Population of a candidate with D parameters: popmul * D^2, at most
// kDeMaxPop (the host sizes the workgroup's LDS with the same function).
__host__ __device__ inline int de_pop(int popmul, int D) {
  const long long p = (long long)popmul * D * D;
  return p < kDeMaxPop ? (int)p : kDeMaxPop;
}

// Members per thread: member i belongs to thread i mod 256, as its k-th
// member with k = i / 256; its energy and its trial's energy stay in that
// thread's registers.
constexpr int kDeOwn = (kDeMaxPop + 255) / 256;

// LDS of one k_de workgroup for populations of up to pcap members with a
// row stride of ds parameters: population and trials (doubles); the
// Latin-hypercube permutation (pcap ints) shares the trials' space.
inline size_t de_lds_bytes(int pcap, int ds) {
  return (size_t)pcap * ds * 2 * sizeof(double);
}

// a[k] = v for a thread's k-th member (static indices keep a in registers)
__device__ inline void set_own(double (&a)[kDeOwn], int k, double v) {
#pragma unroll
  for (int j = 0; j < kDeOwn; ++j)
    if (j == k) a[j] = v;
}

__device__ double de_energy(const double *x, const double *lb,
                            const double *ub, const int *types, int ncomp,
                            const double *h, const double *g, int nb) {
  // the parameters are formed where they are used: an array of them,
  // indexed by the component, would live in scratch memory
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double hb = h[b];
    double v = 0.0;
    for (int c = 0; c < ncomp; ++c) {
      const int d = 2 * c;
      const double rg = lb[d] + x[d] * (ub[d] - lb[d]);
      const double sl = lb[d + 1] + x[d + 1] * (ub[d + 1] - lb[d + 1]);
      v += fit_model(types[c], hb, rg, sl);
    }
    const double r = g[b] - v;
    acc += r * r;
  }
  return acc;
}

__device__ inline double blk_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// (bv, bi) <- the smaller of (bv, bi) and (ov, oi) by value, then index;
// bi == INT_MAX marks an empty slot
__device__ inline void argmin_take(double &bv, int &bi, double ov, int oi) {
  if (oi != INT_MAX && (bi == INT_MAX || ov < bv || (ov == bv && oi < bi))) {
    bv = ov;
    bi = oi;
  }
}

// Index of the smallest energy of members i < P (each thread holds its own
// members' energies, en[k] of member threadIdx.x + 256 k) over the 256
// threads of the block: the first index among equal values, and NaN
// entries never chosen, which is what the scan
// `bi = 0; for i: if (en[i] < en[bi]) bi = i` returns (that scan keeps 0
// when the energy of member 0 is NaN).  Every thread gets the index and,
// in best, its energy.
__device__ inline int blk_argmin(const double (&en)[kDeOwn], int P,
                                 double *rv, int *ri, double *e0,
                                 double &best) {
  double bv = 0.0;
  int bi = INT_MAX;
#pragma unroll
  for (int k = 0; k < kDeOwn; ++k) {
    const int i = threadIdx.x + 256 * k;
    if (i < P) {
      const double v = en[k];
      if (!isnan(v) && (bi == INT_MAX || v < bv)) {
        bv = v;
        bi = i;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    argmin_take(bv, bi, ov, oi);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    rv[threadIdx.x >> 6] = bv;
    ri[threadIdx.x >> 6] = bi;
  }
  if (threadIdx.x == 0) *e0 = en[0];
  __syncthreads();
  bv = rv[0];
  bi = ri[0];
  for (int w = 1; w < 4; ++w) argmin_take(bv, bi, rv[w], ri[w]);
  if (isnan(*e0) || bi == INT_MAX) {
    best = *e0;
    return 0;
  }
  best = bv;
  return bi;
}

// One workgroup per (step, candidate).  cand[c]: {ncomp, type[4],
// fixed_first}; with fixed_first the first component's (range, sill) are
// pinned to the step's first bin (h0, g0), vgs.py:735-739.
__global__ __launch_bounds__(256) void k_de(
    const double *__restrict__ bh, const double *__restrict__ bg,
    const int32_t *__restrict__ nbins, int max_bins,
    const double *__restrict__ var, const int32_t *__restrict__ cand,
    int ncand, int maxiter, int popmul, double tol, double range_ub,
    unsigned long long seed, const double *__restrict__ init, double spread,
    int pcap, int ds, double *__restrict__ out_par,
    double *__restrict__ out_fun, int32_t *__restrict__ out_nit) {
  const int s = blockIdx.y, ci = blockIdx.x;
  const int tid = threadIdx.x;
  const int32_t *cd = cand + ci * 6;
  const int ncomp = cd[0];
  const int D = 2 * ncomp;
  const int P = de_pop(popmul, D);
  const int nb = nbins[s];
  const double *h = bh + (int64_t)s * max_bins;
  const double *g = bg + (int64_t)s * max_bins;
  // LDS sized by the launch for its largest population (pcap) and
  // parameter count (ds, the row stride), so that small candidates leave
  // room for more workgroups per CU
  extern __shared__ double de_lds[];
  double *pop = de_lds;
  double *trial = pop + (int64_t)pcap * ds;
  int *perm = (int *)trial;  // start only, before the first trial
  __shared__ double lb[kDeMaxD], ub[kDeMaxD];
  __shared__ int types[4];
  __shared__ double red[4];
  __shared__ double redv[4];
  __shared__ int redi[4];
  __shared__ double e0s;
  __shared__ double Fs;
  __shared__ int stop_s;
  double *opar = out_par + ((int64_t)s * ncand + ci) * kDeMaxD;
  if (nb <= 0 || P > pcap || D > ds) {
    if (tid == 0) {
      out_fun[(int64_t)s * ncand + ci] = NAN;
      out_nit[(int64_t)s * ncand + ci] = 0;
    }
    return;
  }
  if (tid == 0) {
    const double v = var[s];
    for (int c = 0; c < ncomp; ++c) {
      types[c] = cd[1 + c];
      if (types[c] == SPK_VG_POW) {
        lb[2 * c] = 1e-5;
        ub[2 * c] = 2.0;
        lb[2 * c + 1] = 1e-5;
        ub[2 * c + 1] = 2.0 * v;
      } else if (cd[5] && c == 0) {
        lb[0] = ub[0] = h[0];
        lb[1] = ub[1] = g[0];
      } else {
        lb[2 * c] = h[0];
        ub[2 * c] = range_ub;
        lb[2 * c + 1] = 1e-5;
        ub[2 * c + 1] = 2.0 * v;
      }
    }
  }
  Rng rng(mix64(seed ^ mix64(((unsigned long long)s << 20) ^
                             ((unsigned long long)ci << 8) ^ tid)));
  __syncthreads();
  // Latin hypercube start: per dimension a random permutation of P
  // segments plus a uniform offset inside each segment
  for (int d = 0; d < D; ++d) {
    if (tid == 0) {
      for (int i = 0; i < P; ++i) perm[i] = i;
      for (int i = P - 1; i > 0; --i) {
        const int j = (int)(rng.uni() * (i + 1));
        const int t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
      }
    }
    __syncthreads();
    if (!init) {
      for (int i = tid; i < P; i += 256)
        pop[i * ds + d] = (perm[i] + rng.uni()) / P;
    } else {
      // seeded start: member 0 is the given point; the others stratify a
      // log-box of half-width `spread` around it, clipped to the bounds
      const double w = ub[d] - lb[d];
      const double c = fmin(fmax(init[((int64_t)s * ncand + ci) * kDeMaxD + d],
                                 lb[d]), ub[d]);
      for (int i = tid; i < P; i += 256) {
        double u = 0.0;
        if (w > 0.0) {
          double v = c;
          if (i > 0)
            v = c * exp(spread * (2.0 * (perm[i] + rng.uni()) / P - 1.0));
          v = fmin(fmax(v, lb[d]), ub[d]);
          u = (v - lb[d]) / w;
        }
        pop[i * ds + d] = u;
      }
    }
    __syncthreads();
  }
  double en[kDeOwn], trial_e[kDeOwn];
#pragma unroll
  for (int k = 0; k < kDeOwn; ++k) en[k] = trial_e[k] = 0.0;
  for (int k = 0; tid + 256 * k < P; ++k) {
    const int i = tid + 256 * k;
    set_own(en, k, de_energy(&pop[i * ds], lb, ub, types, ncomp, h, g, nb));
  }
  __syncthreads();
  int nit = 0;
  double best;
  for (int it = 1; it <= maxiter; ++it) {
    nit = it;
    const int bi = blk_argmin(en, P, redv, redi, &e0s, best);
    if (tid == 0) Fs = 0.5 + 0.5 * rng.uni();
    __syncthreads();
    const double F = Fs;
    // trials and their energies first, selection after (deferred update)
    for (int k = 0; tid + 256 * k < P; ++k) {
      const int i = tid + 256 * k;
      int r0, r1;
      do { r0 = (int)(rng.uni() * P); } while (r0 == i);
      do { r1 = (int)(rng.uni() * P); } while (r1 == i || r1 == r0);
      const int fill = (int)(rng.uni() * D);
      double *tr = &trial[i * ds];
      for (int d = 0; d < D; ++d) {
        double v = pop[i * ds + d];
        if (d == fill || rng.uni() < 0.7) {
          v = pop[bi * ds + d] +
              F * (pop[r0 * ds + d] - pop[r1 * ds + d]);
          if (v < 0.0 || v > 1.0) v = rng.uni();
        }
        tr[d] = v;
      }
      set_own(trial_e, k, de_energy(tr, lb, ub, types, ncomp, h, g, nb));
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kDeOwn; ++k) {
      const int i = tid + 256 * k;
      if (i < P && trial_e[k] <= en[k]) {
        en[k] = trial_e[k];
        for (int d = 0; d < D; ++d) pop[i * ds + d] = trial[i * ds + d];
      }
    }
    __syncthreads();
    // convergence: std(E) <= tol * |mean(E)|
    double sm = 0.0;
#pragma unroll
    for (int k = 0; k < kDeOwn; ++k)
      if (tid + 256 * k < P) sm += en[k];
    const double mean = blk_sum(sm, red) / P;
    double sq = 0.0;
#pragma unroll
    for (int k = 0; k < kDeOwn; ++k) {
      if (tid + 256 * k < P) {
        const double dv = en[k] - mean;
        sq += dv * dv;
      }
    }
    const double sd = sqrt(blk_sum(sq, red) / P);
    if (tid == 0) stop_s = (sd <= tol * fabs(mean)) ? 1 : 0;
    __syncthreads();
    if (stop_s) break;
  }
  const int bi = blk_argmin(en, P, redv, redi, &e0s, best);
  if (tid == 0) {
    for (int d = 0; d < kDeMaxD; ++d)
      opar[d] = d < D ? lb[d] + pop[bi * ds + d] * (ub[d] - lb[d]) : 0.0;
    out_fun[(int64_t)s * ncand + ci] = best;
    out_nit[(int64_t)s * ncand + ci] = nit;
  }
}

// ---- spatial order of cells -------------------------------------------------
// Morton (Z-order) keys of the cells' coordinates quantised to 2^20 levels
// per axis over their bounding box; sorting by key makes runs of 64 / 256
// consecutive cells spatially compact tiles.
__global__ void k_bbox_part(const double *__restrict__ x,
                            const double *__restrict__ y, int64_t n,
                            double *__restrict__ part) {
  __shared__ double r[4][256];
  double a = INFINITY, b = -INFINITY, c = INFINITY, d = -INFINITY;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    a = fmin(a, x[i]);
    b = fmax(b, x[i]);
    c = fmin(c, y[i]);
    d = fmax(d, y[i]);
  }
  r[0][threadIdx.x] = a;
  r[1][threadIdx.x] = b;
  r[2][threadIdx.x] = c;
  r[3][threadIdx.x] = d;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r[0][threadIdx.x] = fmin(r[0][threadIdx.x], r[0][threadIdx.x + w]);
      r[1][threadIdx.x] = fmax(r[1][threadIdx.x], r[1][threadIdx.x + w]);
      r[2][threadIdx.x] = fmin(r[2][threadIdx.x], r[2][threadIdx.x + w]);
      r[3][threadIdx.x] = fmax(r[3][threadIdx.x], r[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = r[threadIdx.x][0];
}

__device__ inline unsigned long long spread_bits(unsigned long long v) {
  v &= 0xffffffffull;
  v = (v | (v << 16)) & 0x0000ffff0000ffffull;
  v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}

__global__ void k_morton(const double *__restrict__ x,
                         const double *__restrict__ y, int64_t n,
                         const double *__restrict__ part, int nparts,
                         unsigned long long *__restrict__ key,
                         int32_t *__restrict__ val) {
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int k = 0; k < nparts; ++k) {
    x0 = fmin(x0, part[4 * k]);
    x1 = fmax(x1, part[4 * k + 1]);
    y0 = fmin(y0, part[4 * k + 2]);
    y1 = fmax(y1, part[4 * k + 3]);
  }
  const double span = fmax(fmax(x1 - x0, y1 - y0), 1e-300);
  const double sc = 1048575.0 / span;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    double qx = (x[i] - x0) * sc, qy = (y[i] - y0) * sc;
    qx = fmin(fmax(qx, 0.0), 1048575.0);  // NaN coordinates sort first
    qy = fmin(fmax(qy, 0.0), 1048575.0);
    const unsigned long long ix = (unsigned long long)qx;
    const unsigned long long iy = (unsigned long long)qy;
    key[i] = spread_bits(ix) | (spread_bits(iy) << 1);
    val[i] = (int32_t)i;
  }
}

__global__ void k_gather_xy(const double *__restrict__ x,
                            const double *__restrict__ y,
                            const int32_t *__restrict__ perm, int64_t n,
                            double *__restrict__ px, double *__restrict__ py) {
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    px[i] = x[perm[i]];
    py[i] = y[perm[i]];
  }
}

int cell_order(spk_ctx *ctx, const double *cx, const double *cy, int64_t C,
               int32_t *perm, double *pcx, double *pcy) {
  if (C > INT32_MAX) return set_error(SPK_ESHAPE, "too many cells");
  constexpr int kParts = 256;
  hipStream_t st = ctx->stream;
  double *part;
  unsigned long long *k0, *k1;
  int32_t *v0;
  SPK_TRY(ws_get(ctx, "co_part", kParts * 4 * 8, (void **)&part));
  SPK_TRY(ws_get(ctx, "co_k0", C * 8, (void **)&k0));
  SPK_TRY(ws_get(ctx, "co_k1", C * 8, (void **)&k1));
  SPK_TRY(ws_get(ctx, "co_v0", C * 4, (void **)&v0));
  const unsigned g = (unsigned)std::min<int64_t>(ceil_div(C, 256), 4096);
  hipLaunchKernelGGL(k_bbox_part, dim3(kParts), dim3(256), 0, st, cx, cy, C,
                     part);
  hipLaunchKernelGGL(k_morton, dim3(g), dim3(256), 0, st, cx, cy, C, part,
                     kParts, k0, v0);
  SPK_TRY(check_launch("k_morton"));
  size_t tmp_bytes = 0;
  SPK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, v0,
                                             perm, (int)C, 0, 40, st));
  void *tmp;
  SPK_TRY(ws_get(ctx, "co_tmp", tmp_bytes, &tmp));
  SPK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, v0, perm,
                                             (int)C, 0, 40, st));
  hipLaunchKernelGGL(k_gather_xy, dim3(g), dim3(256), 0, st, cx, cy, perm, C,
                     pcx, pcy);
  return check_launch("k_gather_xy");
}

}  // namespace spk

using namespace spk;

extern "C" {

int spk_vgfit_pairs_dev(spk_ctx *ctx, const double *x, const double *y,
                        int64_t N, int32_t *pi, int32_t *pj, double *pd) {
  if (!ctx || !x || !y || !pi || !pj || !pd)
    return set_error(SPK_EARG, "null argument");
  if (N < 2) return set_error(SPK_ESHAPE, "need at least two stations");
  if (N > 46000) return set_error(SPK_ESHAPE, "too many stations for pairs");
  SPK_HIP(hipSetDevice(ctx->device));
  const int64_t M = N * (N - 1) / 2;
  unsigned long long *k0, *k1;
  int64_t *v0, *v1;
  SPK_TRY(ws_get(ctx, "vgk0", M * 8, (void **)&k0));
  SPK_TRY(ws_get(ctx, "vgk1", M * 8, (void **)&k1));
  SPK_TRY(ws_get(ctx, "vgv0", M * 8, (void **)&v0));
  SPK_TRY(ws_get(ctx, "vgv1", M * 8, (void **)&v1));
  const unsigned g = (unsigned)std::min<int64_t>(ceil_div(M, 256), 65536);
  hipLaunchKernelGGL(k_pair_keys, dim3(g), dim3(256), 0, ctx->stream, x, y, N,
                     k0, v0);
  SPK_TRY(check_launch("k_pair_keys"));
  size_t tmp_bytes = 0;
  SPK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, v0,
                                             v1, (int)M, 0, 64, ctx->stream));
  void *tmp;
  SPK_TRY(ws_get(ctx, "vgtmp", tmp_bytes, &tmp));
  SPK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, v0, v1,
                                             (int)M, 0, 64, ctx->stream));
  hipLaunchKernelGGL(k_pair_unpack, dim3(g), dim3(256), 0, ctx->stream, k1, v1,
                     M, pi, pj, pd);
  ctx->pairs_sorted = true;
  return check_launch("k_pair_unpack");
}

int spk_vgfit_evg_dev(spk_ctx *ctx, const int32_t *pi, const int32_t *pj,
                      const double *pd, int64_t M, int64_t N,
                      const double *vals, int64_t T, double mdr, int nk,
                      int classic, int max_bins, double *bin_h, double *bin_g,
                      int32_t *bin_n, int32_t *nbins) {
  if (!ctx || !pi || !pj || !pd || !vals || !bin_h || !bin_g || !bin_n ||
      !nbins)
    return set_error(SPK_EARG, "null argument");
  if (!(mdr > 0 && mdr <= 1)) return set_error(SPK_EARG, "mdr outside (0, 1]");
  if (nk <= 0) return set_error(SPK_EARG, "nk must be > 0");
  if (classic && nk > 64)
    return set_error(SPK_EARG, "the classic estimator supports nk <= 64");
  if (T <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  const int64_t S = std::min<int64_t>(T, 256);
  int32_t *list;
  SPK_TRY(ws_get(ctx, "vglist", S * M * 4, (void **)&list));
  for (int64_t t0 = 0; t0 < T; t0 += S) {
    const int Sb = (int)std::min<int64_t>(S, T - t0);
    hipLaunchKernelGGL(k_evg, dim3(Sb), dim3(256), 0, ctx->stream, pi, pj, pd,
                       M, vals + t0 * N, N, mdr, nk, classic, max_bins, list,
                       bin_h + t0 * max_bins, bin_g + t0 * max_bins,
                       bin_n + t0 * max_bins, nbins + t0);
    SPK_TRY(check_launch("k_evg"));
  }
  return SPK_OK;
}

static int de_launch(spk_ctx *ctx, const double *bin_h, const double *bin_g,
                     const int32_t *nbins, int max_bins, const double *var,
                     int64_t T, const int32_t *cands, int n_cands, int maxiter,
                     int popmul, double tol, double range_ub, uint64_t seed,
                     const double *init, double spread, double *params,
                     double *fun, int32_t *nit) {
  if (!ctx || !bin_h || !bin_g || !nbins || !var || !cands || !params ||
      !fun || !nit)
    return set_error(SPK_EARG, "null argument");
  if (n_cands <= 0 || n_cands > 64) return set_error(SPK_EARG, "bad n_cands");
  if (popmul <= 0 || maxiter <= 0) return set_error(SPK_EARG, "bad DE sizes");
  for (int c = 0; c < n_cands; ++c)
    if (cands[c * 6] < 1 || cands[c * 6] > 4)
      return set_error(SPK_EARG, "candidate %d: 1..4 components", c);
  if (init && !(spread >= 0.0 && spread <= 20.0))
    return set_error(SPK_EARG, "spread outside [0, 20]");
  if (T <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  int32_t *dc;
  SPK_TRY(ws_get(ctx, "vgcand", n_cands * 6 * 4, (void **)&dc));
  SPK_TRY(h2d(ctx, dc, cands, n_cands * 6 * 4));
  // LDS rows of the largest candidate: stride ds parameters, pcap members
  int ds = 0, pcap = 0;
  for (int c = 0; c < n_cands; ++c) {
    const int D = 2 * cands[c * 6];
    ds = std::max(ds, D);
    pcap = std::max(pcap, de_pop(popmul, D));
  }
  const size_t lds = de_lds_bytes(pcap, ds);
  if (lds > 64 * 1024)
    SPK_HIP(hipFuncSetAttribute((const void *)k_de,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  for (int64_t t0 = 0; t0 < T; t0 += 65535) {
    const int Tb = (int)std::min<int64_t>(65535, T - t0);
    hipLaunchKernelGGL(k_de, dim3(n_cands, Tb), dim3(256), lds, ctx->stream,
                       bin_h + t0 * max_bins, bin_g + t0 * max_bins,
                       nbins + t0, max_bins, var + t0, dc, n_cands, maxiter,
                       popmul, tol, range_ub, (unsigned long long)seed,
                       init ? init + t0 * n_cands * kDeMaxD : nullptr, spread,
                       pcap, ds, params + t0 * n_cands * kDeMaxD,
                       fun + t0 * n_cands,
                       nit + t0 * n_cands);
    SPK_TRY(check_launch("k_de"));
  }
  return SPK_OK;
}

int spk_vgfit_de_dev(spk_ctx *ctx, const double *bin_h, const double *bin_g,
                     const int32_t *nbins, int max_bins, const double *var,
                     int64_t T, const int32_t *cands, int n_cands, int maxiter,
                     int popmul, double tol, double range_ub, uint64_t seed,
                     double *params, double *fun, int32_t *nit) {
  return de_launch(ctx, bin_h, bin_g, nbins, max_bins, var, T, cands, n_cands,
                   maxiter, popmul, tol, range_ub, seed, nullptr, 0.0, params,
                   fun, nit);
}

int spk_vgfit_de_seeded_dev(spk_ctx *ctx, const double *bin_h,
                            const double *bin_g, const int32_t *nbins,
                            int max_bins, const double *var, int64_t T,
                            const int32_t *cands, int n_cands, int maxiter,
                            int popmul, double tol, double range_ub,
                            uint64_t seed, const double *init, double spread,
                            double *params, double *fun, int32_t *nit) {
  if (!init) return set_error(SPK_EARG, "null argument");
  return de_launch(ctx, bin_h, bin_g, nbins, max_bins, var, T, cands, n_cands,
                   maxiter, popmul, tol, range_ub, seed, init, spread, params,
                   fun, nit);
}

}  // extern "C"
