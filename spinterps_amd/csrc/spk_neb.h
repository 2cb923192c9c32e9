// Device helpers of neighbour selection (interp/grps.py:141-247,
// pyx:811-890), shared by the standalone selection kernels (spk_neb.hip)
// and the neighbourhood interpolation path (spk_nebk.hip).
#pragma once

#include "spk_internal.h"

namespace spk {

// Strict order on (distance, station index): equal distances are ordered
// by index.
__device__ inline bool lex_less(double a, int ia, double b, int ib) {
  return (a < b) || (a == b && ia < ib);
}

// Wave-wide arg-min of (d, i) pairs; every lane returns the winner.
__device__ inline void wave_argmin(double &d, int &i) {
  for (int off = 32; off > 0; off >>= 1) {
    const double od = __shfl_xor(d, off, 64);
    const int oi = __shfl_xor(i, off, 64);
    if (lex_less(od, oi, d, i)) {
      d = od;
      i = oi;
    }
  }
}

// Angular sector of station (rx, ry) seen from (x, y), pyx:850-868.
__device__ inline long long pie_of(double x, double y, double rx, double ry,
                                   long long n_pies) {
  const double two_pi = 2.0 * kPi;
  const double xd = rx - x, yd = ry - y;
  double ang;
  if (!(xd != 0.0)) {
    ang = 0.0;
  } else {
    ang = atan(yd / xd);
    if (xd < 0 && yd > 0)
      ang = kPi + ang;
    else if (xd < 0 && yd < 0)
      ang = kPi + ang;
    else if (xd > 0 && yd < 0)
      ang = two_pi + ang;
  }
  return (long long)(ang * (double)n_pies / two_pi);
}

// LDS of pie_select for n stations, n_pies sectors and k selections.
__host__ __device__ inline size_t pie_lds_bytes(int n, int n_pies, int k) {
  return (size_t)(n + 2 * n_pies) * 8 + (size_t)(n + 2 * n_pies + k) * 4;
}

// pie for one destination point, run by one 64-thread workgroup: every
// station gets its distance rank inside its angular sector, stations are
// taken rank by rank, nearest first within a rank (interp/grps.py:168-247
// with min_dist_thresh = -1, pyx:846-889).  xy(j) returns station j's
// coordinates.  smem holds pie_lds_bytes(n, n_pies, k).  Writes the first
// min(k, ranked) selections to sel (LDS, selection order) and returns how
// many were written; a sector index outside [0, n_pies) sets bit 1 of bad.
template <typename XY>
__device__ int pie_select(int lane, double x, double y, int n, int k,
                          int n_pies, XY xy, char *smem, int **sel_out,
                          int *bad) {
  double *sd = (double *)smem;       // [n]
  double *cur_d = sd + n;            // [n_pies]
  double *cand_d = cur_d + n_pies;   // [n_pies]
  int *sp = (int *)(cand_d + n_pies);  // [n]
  int *cur_i = sp + n;               // [n_pies]
  int *sel = cur_i + n_pies;         // [k]
  int *cand = sel + k;               // [n_pies]
  *sel_out = sel;
  for (int j = lane; j < n; j += 64) {
    const double2 r = xy(j);
    sd[j] = dist2(x, y, r.x, r.y);
    const long long p = pie_of(x, y, r.x, r.y, n_pies);
    if (p < 0 || p >= n_pies) atomicOr(bad, 1);
    sp[j] = (int)min(max(p, 0ll), (long long)n_pies - 1);
  }
  for (int p = lane; p < n_pies; p += 64) {
    cur_d[p] = -INFINITY;
    cur_i[p] = -1;
  }
  __syncthreads();
  int nsel = 0;
  while (nsel < k) {
    int ncand = 0;
    for (int p = 0; p < n_pies; ++p) {
      const double pd = cur_d[p];
      const int pi = cur_i[p];
      double bd = INFINITY;
      int bi = 0x7fffffff;
      for (int j = lane; j < n; j += 64) {
        if (sp[j] != p) continue;
        const double d = sd[j];
        if (lex_less(pd, pi, d, j) && lex_less(d, j, bd, bi)) {
          bd = d;
          bi = j;
        }
      }
      wave_argmin(bd, bi);
      __syncthreads();
      if (bi != 0x7fffffff) {
        if (lane == 0) {
          cand[ncand] = bi;
          cand_d[ncand] = bd;
          cur_d[p] = bd;
          cur_i[p] = bi;
        }
        ++ncand;
      }
      __syncthreads();
    }
    if (ncand == 0) break;
    if (lane == 0) {
      // this rank's stations by (distance, index)
      for (int a = 1; a < ncand; ++a) {
        const int vi = cand[a];
        const double vd = cand_d[a];
        int b = a - 1;
        while (b >= 0 && lex_less(vd, vi, cand_d[b], cand[b])) {
          cand[b + 1] = cand[b];
          cand_d[b + 1] = cand_d[b];
          --b;
        }
        cand[b + 1] = vi;
        cand_d[b + 1] = vd;
      }
      for (int a = 0; a < ncand && nsel + a < k; ++a) sel[nsel + a] = cand[a];
    }
    nsel += ncand;
    __syncthreads();
  }
  return min(nsel, k);
}

}  // namespace spk
