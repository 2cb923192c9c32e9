// IDW as a masked contraction on FP64 MFMA (interp/steps.py:293-313,
// pyx:784-808).
//
// The reference, per cell of a time group: d <- d / max(d), w_i = 1 / d_i^p
// over the group's active stations, then per step sum(w z) / sum(w).  The
// factor max(d)^p is common to numerator and denominator, so with
// W[c, i] = d_ci^-p over all N stations and the step's mask M[t, i] (1 where
// station i has a value) the fields are
//     num = (Z . M) W^T,   den = M W^T,   est[t, c] = num[t, c] / den[t, c],
// two GEMMs over the stations that share W, which is generated on the fly
// (once per block and chunk of stations, used for 32 steps x 2 sums).  The
// results differ from the reference's by rounding only (the sums run in
// another order, without the normalisation).
//
// Orientation: the MFMA output rows are steps and its columns cells, so the
// numerator and denominator of one (step, cell) land in the same lane and
// register of two accumulator tiles.  Block tile 32 steps x 128 cells, four
// waves of 32 steps x 32 cells (2 x 2 tiles of 16 x 16, for each of num and
// den), stations in chunks of 16 staged through LDS.
//
// A cell at zero distance from a station has an infinite weight: the
// reference then returns NaN for the steps where that station has a value.
// Such cells (rare) are recomputed per step in the reference's arithmetic.
#include "spk_krige.h"

namespace spk {

typedef double idw_d4 __attribute__((ext_vector_type(4)));

__device__ inline idw_d4 idw_mfma(double a, double b, idw_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One (cell, step) in the reference's arithmetic: w = 1 / (d / dmax)^p over
// the step's active stations (pyx:784-808), sum(w z) / sum(w).
__device__ double idw_cell_ref(double x, double y, const double *sx,
                               const double *sy, const int32_t *my,
                               const double *z, int n, double p, int *bad) {
#pragma clang fp contract(off)
  double dmax = 0.0;
  for (int j = 0; j < n; ++j) dmax = fmax(dmax, dist2(x, y, sx[my[j]], sy[my[j]]));
  double ws = 0.0, wz = 0.0;
  for (int j = 0; j < n; ++j) {
    double d = dist2(x, y, sx[my[j]], sy[my[j]]);
    if (dmax > 0) d = d / dmax;
    const double wj = 1.0 / (p == 2.0 ? d * d : pow(d, p));
    ws = ws + wj;
    wz = wz + wj * z[j];
  }
  if (!(ws >= 1e-14)) atomicOr(bad, 1);  // interp/steps.py:305
  return wz / ws;
}

// grid (ceil(C / 128), ceil(S / 32)); vals [S x N] the launch's steps (NaN =
// no value), sc [S] their records (k_compact), est [S x C].
__global__ __launch_bounds__(256, 2) void k_idw_mfma(
    const double *__restrict__ sx, const double *__restrict__ sy, int64_t N,
    const double *__restrict__ vals, int S, const StepSc *__restrict__ sc,
    const int32_t *__restrict__ idx, const double *__restrict__ zact,
    int64_t ldz, const double *__restrict__ cx, const double *__restrict__ cy,
    int64_t C, double p, double *__restrict__ est, int *__restrict__ bad) {
  constexpr int BS = 32, BC = 128, KC = 16;
  __shared__ double Ws[KC][BC + 2];       // W[station][cell]
  __shared__ double As[2][BS][KC + 1];    // z . m and m, [step][station]
  __shared__ double stx[KC], sty[KC];
  __shared__ int coinc[BC];               // cell at a station
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t c0 = blockIdx.x * (int64_t)BC;
  const int t0 = blockIdx.y * BS;
  // the thread's cell for the W tile
  const int wc = tid & (BC - 1);
  const int64_t gc = c0 + wc;
  const double x = gc < C ? cx[gc] : 0.0, y = gc < C ? cy[gc] : 0.0;
  if (tid < BC) coinc[tid] = 0;
  idw_d4 acc[2][2][2];  // [num/den][step tile][cell tile]
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[q][a][b] = idw_d4{0, 0, 0, 0};
  for (int k0 = 0; k0 < N; k0 += KC) {
    __syncthreads();  // previous chunk's readers are done
    if (tid < KC) {
      const int64_t i = k0 + tid;
      stx[tid] = i < N ? sx[i] : 0.0;
      sty[tid] = i < N ? sy[i] : 0.0;
    }
    // A: 32 steps x 16 stations, each value giving z . m and m
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = tid + 256 * u;
      const int s = e >> 4, k = e & 15;
      const int t = t0 + s;
      const int64_t i = k0 + k;
      double v = NAN;
      if (t < S && i < N) v = vals[(int64_t)t * N + i];
      const bool has = !isnan(v);
      As[0][s][k] = has ? v : 0.0;
      As[1][s][k] = has ? 1.0 : 0.0;
    }
    __syncthreads();
    // W: 16 stations x 128 cells, thread: cell wc, stations (tid >> 7) + 2 u
#pragma unroll
    for (int u = 0; u < KC / 2; ++u) {
      const int k = (tid >> 7) + 2 * u;
      double w = 0.0;
      if (k0 + k < N && gc < C) {
#pragma clang fp contract(off)
        const double dx = x - stx[k], dy = y - sty[k];
        const double s2 = (dx * dx) + (dy * dy);
        if (s2 > 0.0)
          w = p == 2.0 ? 1.0 / s2 : 1.0 / pow(sqrt(s2), p);
        else
          coinc[wc] = 1;  // infinite weight: per-cell path below
      }
      Ws[k][wc] = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      const int k = kk * 4 + lg;
      double af[2][2], bf[2];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int a = 0; a < 2; ++a) af[q][a] = As[q][a * 16 + li][k];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Ws[k][wave * 32 + b * 16 + li];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[q][a][b] = idw_mfma(af[q][a], bf[b], acc[q][a][b]);
    }
  }
  __syncthreads();
  // epilogue: lane (lg, li) holds rows lg + 4 r of each tile, column li
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int cl = wave * 32 + b * 16 + li;
    const int64_t c = c0 + cl;
    if (c >= C) continue;
    const bool at_stn = coinc[cl] != 0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int t = t0 + a * 16 + lg + 4 * r;
        if (t >= S) continue;
        const StepSc st = sc[t];
        if (st.status) continue;  // special steps: k_special
        double e;
        if (at_stn) {
          e = idw_cell_ref(cx[c], cy[c], sx, sy, idx + (int64_t)t * N,
                           zact + (int64_t)t * ldz, st.n, p, bad);
        } else {
          const double num = acc[0][a][b][r], den = acc[1][a][b][r];
          // the reference's check on the normalised weight sum (at least
          // one, when finite) fails only for a non-finite or zero sum
          if (!(den > 0.0) || !isfinite(den)) atomicOr(bad, 1);
          e = num / den;
        }
        est[(int64_t)t * C + c] = e;
      }
  }
}

}  // namespace spk
