// Neighbour selection for all destination cells (interp/grps.py:141-247)
// and the single-point sel_equidist_refs mirror (pyx:811-890).
//
// One wavefront per destination cell.  Selection is by repeated wave-wide
// arg-min with a strict lexicographic cursor on (key, station index), so the
// k smallest keys come out in order without sorting all stations; equal
// distances are ordered by station index (the reference's np.argsort is an
// unstable quicksort there, and such ties have measure zero for continuous
// coordinates).
#include "spk_neb.h"

namespace spk {

// nrst: the k nearest stations, indices sorted ascending
// (interp/grps.py:147-166).
__global__ __launch_bounds__(64) void k_nrst(
    const double *__restrict__ dx, const double *__restrict__ dy,
    int64_t n_dst, const double *__restrict__ rx,
    const double *__restrict__ ry, int n_refs, int k,
    int64_t *__restrict__ out) {
  const int64_t c = blockIdx.x;
  if (c >= n_dst) return;
  extern __shared__ int sel[];
  const int lane = threadIdx.x;
  const double x = dx[c], y = dy[c];
  double cd = -INFINITY;
  int ci = -1;
  for (int r = 0; r < k; ++r) {
    double bd = INFINITY;
    int bi = 0x7fffffff;
    for (int j = lane; j < n_refs; j += 64) {
      const double d = dist2(x, y, rx[j], ry[j]);
      if (lex_less(cd, ci, d, j) && lex_less(d, j, bd, bi)) {
        bd = d;
        bi = j;
      }
    }
    wave_argmin(bd, bi);
    cd = bd;
    ci = bi;
    if (lane == 0) sel[r] = bi;
  }
  __syncthreads();
  // rank of each selected index among the selected -> ascending output
  for (int r = lane; r < k; r += 64) {
    const int v = sel[r];
    int rank = 0;
    for (int q = 0; q < k; ++q) rank += sel[q] < v;
    out[c * k + rank] = v;
  }
}

// pie: stations ranked by distance inside their angular sector, taken rank
// by rank, nearest first within a rank (interp/grps.py:168-247 with
// min_dist_thresh = -1, pyx:846-889).
__global__ __launch_bounds__(64) void k_pie(
    const double *__restrict__ dx, const double *__restrict__ dy,
    int64_t n_dst, const double *__restrict__ rx,
    const double *__restrict__ ry, int n_refs, int k, int n_pies,
    int64_t *__restrict__ out, int *__restrict__ bad) {
  const int64_t c = blockIdx.x;
  if (c >= n_dst) return;
  extern __shared__ double smem[];
  const int lane = threadIdx.x;
  int *sel = nullptr;
  const int nsel = pie_select(
      lane, dx[c], dy[c], n_refs, k, n_pies,
      [&](int j) { return make_double2(rx[j], ry[j]); }, (char *)smem, &sel,
      bad);
  if (nsel < k) {
    atomicOr(bad, 2);
    return;
  }
  for (int r = lane; r < k; r += 64) {
    const int v = sel[r];
    int rank = 0;
    for (int q = 0; q < k; ++q) rank += sel[q] < v;
    out[c * k + rank] = v;
  }
}

// Single destination point, pyx:811-890.
__global__ __launch_bounds__(256) void k_sel_equidist(
    double x, double y, const double *__restrict__ rx,
    const double *__restrict__ ry, int n, long long n_pies, double thr,
    long long flag, double *__restrict__ dists, double *__restrict__ tem,
    long long *__restrict__ sel, unsigned long long *__restrict__ pie_idx,
    unsigned long long *__restrict__ pie_cts, int *__restrict__ bad) {
  const int tid = threadIdx.x;
  __shared__ double mind;
  __shared__ int mini;
  __shared__ int last_pie;
  for (int j = tid; j < n; j += 256) {
    sel[j] = flag;
    dists[j] = dist2(x, y, rx[j], ry[j]);
  }
  __syncthreads();
  if (tid == 0) {
    mind = INFINITY;
    mini = -1;
    for (int j = 0; j < n; ++j)
      if (dists[j] <= thr && dists[j] < mind) {
        mind = dists[j];
        mini = j;
      }
    if (mind < INFINITY) sel[mini] = 0;
  }
  __syncthreads();
  if (mind < INFINITY) return;
  for (long long p = tid; p < n_pies; p += 256) pie_cts[p] = 0;
  __syncthreads();
  for (int j = tid; j < n; j += 256) {
    const long long p = pie_of(x, y, rx[j], ry[j], n_pies);
    if (p < 0 || p >= n_pies) {
      atomicOr(bad, 1);
      continue;
    }
    pie_idx[j] = (unsigned long long)p;
    atomicAdd(&pie_cts[p], 1ull);
  }
  __syncthreads();
  if (*bad) return;
  // rank inside the pie by (distance, index); stations with an infinite
  // distance are never ranked (pyx:886-887)
  for (int j = tid; j < n; j += 256) {
    const unsigned long long p = pie_idx[j];
    const double d = dists[j];
    if (!(d < INFINITY)) continue;
    long long r = 0;
    for (int q = 0; q < n; ++q)
      if (pie_idx[q] == p && lex_less(dists[q], q, d, j)) ++r;
    sel[j] = r;
  }
  if (tid == 0) {
    int lp = -1;
    for (long long p = 0; p < n_pies; ++p)
      if (pie_cts[p]) lp = (int)p;
    last_pie = lp;
  }
  __syncthreads();
  if (last_pie >= 0)
    for (int j = tid; j < n; j += 256)
      tem[j] = (pie_idx[j] == (unsigned long long)last_pie) ? dists[j]
                                                           : INFINITY;
}

}  // namespace spk

using namespace spk;

extern "C" {

int spk_neb_idxs_nrst(spk_ctx *ctx, const double *dst_xs, const double *dst_ys,
                      int64_t n_dst, const double *ref_xs,
                      const double *ref_ys, int64_t n_refs, int64_t n_nebs,
                      int64_t *out) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (n_nebs <= 0) return set_error(SPK_EARG, "n_nebs must be > 0");
  if (n_refs <= 0 || n_dst <= 0)
    return set_error(SPK_ESHAPE, "need stations and destinations");
  if (n_refs > (1 << 30)) return set_error(SPK_ESHAPE, "too many stations");
  const int k = (int)std::min<int64_t>(n_nebs, n_refs);
  if (k > 8192) return set_error(SPK_ESHAPE, "n_nebs above 8192");
  SPK_HIP(hipSetDevice(ctx->device));
  char *d = nullptr;
  const size_t bd = 2 * n_dst * 8, br = 2 * n_refs * 8;
  SPK_TRY(ws_get(ctx, "io0", bd + br + n_dst * k * 8, (void **)&d));
  double *ddx = (double *)d, *ddy = ddx + n_dst;
  double *drx = ddx + 2 * n_dst, *dry = drx + n_refs;
  int64_t *dout = (int64_t *)(d + bd + br);
  SPK_TRY(h2d(ctx, ddx, dst_xs, n_dst * 8));
  SPK_TRY(h2d(ctx, ddy, dst_ys, n_dst * 8));
  SPK_TRY(h2d(ctx, drx, ref_xs, n_refs * 8));
  SPK_TRY(h2d(ctx, dry, ref_ys, n_refs * 8));
  hipLaunchKernelGGL(k_nrst, dim3((unsigned)n_dst), dim3(64), k * sizeof(int),
                     ctx->stream, ddx, ddy, n_dst, drx, dry, (int)n_refs, k,
                     dout);
  SPK_TRY(check_launch("k_nrst"));
  return d2h(ctx, out, dout, n_dst * k * 8);
}

int spk_neb_idxs_pie(spk_ctx *ctx, const double *dst_xs, const double *dst_ys,
                     int64_t n_dst, const double *ref_xs, const double *ref_ys,
                     int64_t n_refs, int64_t n_nebs, int64_t n_pies,
                     int64_t *out) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (n_pies <= 0 || n_nebs < n_pies)
    return set_error(SPK_EARG, "need 0 < n_pies <= n_nebs "
                     "(interp/grps.py:35-37)");
  if (n_refs <= 0 || n_dst <= 0)
    return set_error(SPK_ESHAPE, "need stations and destinations");
  if (n_nebs > n_refs)
    return set_error(SPK_ESHAPE, "n_nebs %lld > %lld stations",
                     (long long)n_nebs, (long long)n_refs);
  if (n_refs > (1 << 20) || n_pies > (1 << 20))
    return set_error(SPK_ESHAPE, "pie selection over %lld stations does not "
                     "fit the LDS", (long long)n_refs);
  const size_t lds = pie_lds_bytes((int)n_refs, (int)n_pies, (int)n_nebs);
  if (lds > 150 * 1024)
    return set_error(SPK_ESHAPE, "pie selection over %lld stations does not "
                     "fit the LDS", (long long)n_refs);
  SPK_HIP(hipSetDevice(ctx->device));
  char *d = nullptr;
  const size_t bd = 2 * n_dst * 8, br = 2 * n_refs * 8;
  SPK_TRY(ws_get(ctx, "io0", bd + br + n_dst * n_nebs * 8 + 64, (void **)&d));
  double *ddx = (double *)d, *ddy = ddx + n_dst;
  double *drx = ddx + 2 * n_dst, *dry = drx + n_refs;
  int64_t *dout = (int64_t *)(d + bd + br);
  int *bad = (int *)(d + bd + br + n_dst * n_nebs * 8);
  SPK_HIP(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
  SPK_TRY(h2d(ctx, ddx, dst_xs, n_dst * 8));
  SPK_TRY(h2d(ctx, ddy, dst_ys, n_dst * 8));
  SPK_TRY(h2d(ctx, drx, ref_xs, n_refs * 8));
  SPK_TRY(h2d(ctx, dry, ref_ys, n_refs * 8));
  if (lds > 64 * 1024)
    SPK_HIP(hipFuncSetAttribute((const void *)k_pie,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  hipLaunchKernelGGL(k_pie, dim3((unsigned)n_dst), dim3(64), lds, ctx->stream,
                     ddx, ddy, n_dst, drx, dry, (int)n_refs, (int)n_nebs,
                     (int)n_pies, dout, bad);
  SPK_TRY(check_launch("k_pie"));
  int hb = 0;
  SPK_TRY(d2h(ctx, &hb, bad, sizeof(int)));
  if (hb & 1)
    return set_error(SPK_EINDEX, "pie index out of range (an angle rounded "
                     "to 2*pi, pyx:868-870)");
  if (hb & 2) return set_error(SPK_ESHAPE, "fewer ranked stations than n_nebs");
  return d2h(ctx, out, dout, n_dst * n_nebs * 8);
}

int spk_sel_equidist_refs(spk_ctx *ctx, double dst_x, double dst_y,
                          const double *ref_xs, const double *ref_ys,
                          int64_t n_refs, uint64_t n_pies,
                          double min_dist_thresh, int64_t not_neb_flag,
                          double *dists, double *tem_ref_sel_dists,
                          int64_t *ref_sel_pie_idxs, uint64_t *ref_pie_idxs,
                          uint64_t *ref_pie_cts) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (n_refs <= 0) return SPK_OK;
  if (n_pies == 0) return set_error(SPK_EARG, "n_pies must be > 0");
  SPK_HIP(hipSetDevice(ctx->device));
  char *d = nullptr;
  const size_t nb = n_refs * 8;
  SPK_TRY(ws_get(ctx, "io0", 6 * nb + n_pies * 8 + 64, (void **)&d));
  double *drx = (double *)d, *dry = drx + n_refs, *dd = dry + n_refs,
         *dtem = dd + n_refs;
  long long *dsel = (long long *)(dtem + n_refs);
  unsigned long long *dpi = (unsigned long long *)(dsel + n_refs);
  unsigned long long *dct = dpi + n_refs;
  int *bad = (int *)(dct + n_pies);
  SPK_HIP(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
  SPK_TRY(h2d(ctx, drx, ref_xs, nb));
  SPK_TRY(h2d(ctx, dry, ref_ys, nb));
  // arrays the kernel leaves untouched keep the caller's contents
  SPK_TRY(h2d(ctx, dtem, tem_ref_sel_dists, nb));
  SPK_TRY(h2d(ctx, dpi, ref_pie_idxs, nb));
  SPK_TRY(h2d(ctx, dct, ref_pie_cts, n_pies * 8));
  hipLaunchKernelGGL(k_sel_equidist, dim3(1), dim3(256), 0, ctx->stream,
                     dst_x, dst_y, drx, dry, (int)n_refs, (long long)n_pies,
                     min_dist_thresh, (long long)not_neb_flag, dd, dtem, dsel,
                     dpi, dct, bad);
  SPK_TRY(check_launch("k_sel_equidist"));
  int hb = 0;
  SPK_TRY(d2h(ctx, &hb, bad, sizeof(int)));
  if (hb)
    return set_error(SPK_EINDEX, "pie index out of range (an angle rounded "
                     "to 2*pi, pyx:868-870)");
  SPK_TRY(d2h(ctx, dists, dd, nb));
  SPK_TRY(d2h(ctx, tem_ref_sel_dists, dtem, nb));
  SPK_TRY(d2h(ctx, ref_sel_pie_idxs, dsel, nb));
  SPK_TRY(d2h(ctx, ref_pie_idxs, dpi, nb));
  return d2h(ctx, ref_pie_cts, dct, n_pies * 8);
}

}  // extern "C"
