// Internal declarations shared by the libspk translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/spk.h"

namespace spk {

// ---- error state ------------------------------------------------------
int set_error(int code, const char *fmt, ...);

#define SPK_HIP(expr)                                                    \
  do {                                                                   \
    hipError_t e_ = (expr);                                              \
    if (e_ != hipSuccess)                                                \
      return ::spk::set_error(SPK_EHIP, "%s failed: %s (%s:%d)", #expr,  \
                              hipGetErrorString(e_), __FILE__, __LINE__);\
  } while (0)

#define SPK_TRY(expr)                                                    \
  do {                                                                   \
    int rc_ = (expr);                                                    \
    if (rc_ != SPK_OK) return rc_;                                       \
  } while (0)

// ---- device workspace ----------------------------------------------------
struct Buffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace spk

struct spk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // second stream (default priority) for factorisations that overlap the
  // cell contractions, and the events that order the two
  hipStream_t aux_stream = nullptr;
  // third stream (default priority): float32 host copies of finished
  // batches (spk_interp_steps_f32_dev)
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // named scratch buffers, grown on demand, freed by spk_ctx_trim/destroy
  std::vector<std::pair<std::string, spk::Buffer>> bufs;
  std::vector<hipEvent_t> events;
  std::vector<std::string> stage_names;
  int n_marks = 0;
  std::vector<double> last_ms;
  // set by spk_vgfit_pairs_dev: spk_simulate_fields_dev refuses to run on
  // a context that has sorted station pairs (that order stalled with no
  // cause found; simulations before any pair sort are unaffected)
  bool pairs_sorted = false;
  // pseudo-inverse fallback of failed steps: a rocBLAS handle (rocSOLVER
  // eigendecompositions and the reassembly GEMM) and pinned host copies
  // of the step records, read after each batch's Cholesky
  void *blas = nullptr;
  void *h_sc = nullptr;
  size_t h_sc_bytes = 0;
  // tile autotuner: mode (SPK_TUNE_*), the variant picked per (kernel
  // family, size class) with the timings it was picked on, and variants
  // forced by the caller
  int tune_mode = SPK_TUNE_AUTO;
  struct TuneEntry {
    std::string kernel;
    int cls = 0;
    int choice = 0;
    std::vector<float> ms;
  };
  std::vector<TuneEntry> tuned;
  std::vector<std::pair<std::string, int>> forced;
  // captured launch sequences (hipGraph) of the batched factorisation,
  // each keyed by every argument and grid of its launches
  bool graphs_on = true;
  struct GraphEntry {
    std::string tag;
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<GraphEntry> graphs;
};

namespace spk {

// Device buffer of at least `bytes`, cached under `name` in the context.
int ws_get(spk_ctx *ctx, const char *name, size_t bytes, void **out);
// Pinned-free helpers for the host-pointer API.
int h2d(spk_ctx *ctx, void *dst, const void *src, size_t bytes);
int d2h(spk_ctx *ctx, void *dst, const void *src, size_t bytes);
int sync(spk_ctx *ctx);
int check_launch(const char *what);
// Stage timing (events recorded on the context stream).
void mark_reset(spk_ctx *ctx);
int mark(spk_ctx *ctx, const char *name);
int mark_collect(spk_ctx *ctx);

// Tile autotuner.  run(v, timing) launches variant v of a kernel family
// (every variant writes the same results); with timing set it may launch
// a representative part of the grid.  The variant is the caller's forced
// one, the one already picked for (kernel, cls), or, the first time the
// class is seen with tuning on, each variant is launched three times on
// st (the context's other stream drained first), the faster of the last
// two runs (timed by events) counts, the fastest variant is kept, and it
// is then launched in full.  With tuning off the variant is dflt.
int tune_launch(spk_ctx *ctx, const char *kernel, int cls, hipStream_t st,
                int ncand, int dflt, const std::function<int(int, bool)> &run);

// Launches fn's kernels on st through a hipGraph: captured from fn the
// first time (tag, key) is seen on the context, replayed afterwards.  fn
// may only launch kernels on st (no copies, memsets, allocations or host
// waits), and key must hold every argument and grid size of those
// launches.  With graphs off (spk_ctx_set_graphs) fn runs directly.
int graph_run(spk_ctx *ctx, const char *tag, const std::vector<uint64_t> &key,
              hipStream_t st, const std::function<int()> &fn);
// Destroys the context's graphs (both streams are synchronised first).
int graphs_clear(spk_ctx *ctx);
// Appends values (pointers, integers, doubles) to a graph key.
struct GraphKey {
  std::vector<uint64_t> v;
  template <typename T>
  GraphKey &operator<<(const T &x) {
    static_assert(sizeof(T) <= 8, "key fields are at most 8 bytes");
    uint64_t u = 0;
    memcpy(&u, &x, sizeof(T));
    v.push_back(u);
    return *this;
  }
};

// Host-side parse of a variogram string (pyx:174-184).
int parse_vg(const char *s, spk_vg *out);

// ---- device helpers --------------------------------------------------------
constexpr double kPi = 3.14159265358979311600;  // M_PI as a double

// Euclidean distance with the reference's operation order and no fused
// multiply-add: ((dx*dx) + (dy*dy)) ** 0.5 (pyx:141).
__host__ __device__ inline double dist2(double x1, double y1, double x2,
                                       double y2) {
#pragma clang fp contract(off)
  const double dx = x1 - x2;
  const double dy = y1 - y2;
  const double s = (dx * dx) + (dy * dy);
  return sqrt(s);
}

// One variogram component gamma(h; r, s), pyx:38-83 (same operation order).
// PH = false leaves out the Pow and Hol models (NaN for them), whose pow and
// sin cost registers in kernels that never see them.
template <bool PH>
__host__ __device__ inline double vg_comp_t(int type, double h, double r,
                                           double s) {
#pragma clang fp contract(off)
  switch (type) {
    case SPK_VG_RNG:
      return h;
    case SPK_VG_NUG:
      return s;
    case SPK_VG_SPH: {
      if (h >= r) return s;
      const double a = (1.5 * h) / r;
      const double b = (h * h * h) / (2.0 * (r * r * r));
      return s * (a - b);
    }
    case SPK_VG_EXP:
      return s * (1.0 - exp((-3.0 * h) / r));
    case SPK_VG_LIN:
      if (h > r) return s;
      return s * (h / r);
    case SPK_VG_GAU:
      return s * (1.0 - exp(-3.0 * ((h * h) / (r * r))));
    case SPK_VG_POW:
      if constexpr (PH) return s * pow(h, r);
      return NAN;
    case SPK_VG_HOL: {
      if constexpr (PH) {
        if (h == 0.0) return 0.0;
        const double a = (kPi * h) / r;
        return s * (1.0 - (sin(a) / a));
      }
      return NAN;
    }
    default:
      return NAN;
  }
}

__host__ __device__ inline double vg_comp(int type, double h, double r,
                                         double s) {
  return vg_comp_t<true>(type, h, r, s);
}

// fill_vg_var_arr's per-entry value (pyx:162-216): components summed in
// order from 0.0, covariance form s - gamma when covar_flag, then entries
// <= min_vg_val set to 0.
template <bool PH>
__host__ __device__ inline double vg_entry_t(const spk_vg &vg, double h,
                                            int covar_flag,
                                            double min_vg_val) {
#pragma clang fp contract(off)
  double acc = 0.0;
  const double mult = covar_flag ? 1.0 : 0.0;
  const double sign = covar_flag ? -1.0 : 1.0;
  for (int k = 0; k < vg.n; ++k)
    acc += (mult * vg.sill[k]) +
           (sign * vg_comp_t<PH>(vg.type[k], h, vg.range[k], vg.sill[k]));
  if (acc <= min_vg_val) acc = 0.0;
  return acc;
}

__host__ __device__ inline double vg_entry(const spk_vg &vg, double h,
                                          int covar_flag, double min_vg_val) {
  return vg_entry_t<true>(vg, h, covar_flag, min_vg_val);
}

// A variogram with per-component constants folded in (reciprocals instead
// of divisions) for the batched path's bulk cell evaluation.  Values agree
// with vg_entry to a few ulp; the plain FP64 path keeps vg_entry.
struct VgFast {
  int n;
  int type[SPK_VG_MAX];
  double s[SPK_VG_MAX], r[SPK_VG_MAX], c1[SPK_VG_MAX], c2[SPK_VG_MAX];
};

__host__ __device__ inline void vg_fast_init(const spk_vg &vg, VgFast &f) {
  f.n = vg.n;
  for (int k = 0; k < vg.n; ++k) {
    const double r = vg.range[k];
    f.type[k] = vg.type[k];
    f.s[k] = vg.sill[k];
    f.r[k] = r;
    f.c1[k] = 0.0;
    f.c2[k] = 0.0;
    switch (vg.type[k]) {
      case SPK_VG_SPH:
        f.c1[k] = 1.5 / r;
        f.c2[k] = 0.5 / (r * r * r);
        break;
      case SPK_VG_EXP:
        f.c1[k] = -3.0 / r;
        break;
      case SPK_VG_GAU:
        f.c1[k] = -3.0 / (r * r);
        break;
      case SPK_VG_LIN:
        f.c1[k] = 1.0 / r;
        break;
      case SPK_VG_HOL:
        f.c1[k] = kPi / r;
        break;
      default:
        break;
    }
  }
}

__device__ inline double vg_fast_entry(const VgFast &f, double h,
                                       int covar_flag, double min_vg_val) {
#pragma clang fp contract(off)
  double acc = 0.0;
  for (int k = 0; k < f.n; ++k) {
    const double s = f.s[k];
    double g;
    switch (f.type[k]) {
      case SPK_VG_RNG: g = h; break;
      case SPK_VG_NUG: g = s; break;
      case SPK_VG_SPH:
        g = (h >= f.r[k]) ? s : s * (h * f.c1[k] - (h * h * h) * f.c2[k]);
        break;
      case SPK_VG_EXP: g = s * (1.0 - exp(h * f.c1[k])); break;
      case SPK_VG_GAU: g = s * (1.0 - exp((h * h) * f.c1[k])); break;
      case SPK_VG_LIN: g = (h > f.r[k]) ? s : s * (h * f.c1[k]); break;
      case SPK_VG_POW: g = s * pow(h, f.r[k]); break;
      case SPK_VG_HOL: {
        if (h == 0.0) {
          g = 0.0;
        } else {
          const double a = h * f.c1[k];
          g = s * (1.0 - sin(a) / a);
        }
        break;
      }
      default: g = NAN;
    }
    acc += covar_flag ? (s - g) : g;
  }
  if (acc <= min_vg_val) acc = 0.0;
  return acc;
}

// vg_fast_entry for records whose components are all Rng / Nug / Sph / Exp
// / Gau / Lin (vg_simple); the power and hole-effect branches are left
// out, which keeps register use low in the cell kernels.
__device__ inline double vg_simple_entry(const VgFast &f, double h,
                                         int covar_flag, double min_vg_val) {
#pragma clang fp contract(off)
  double acc = 0.0;
  for (int k = 0; k < f.n; ++k) {
    const double s = f.s[k];
    double g;
    switch (f.type[k]) {
      case SPK_VG_RNG: g = h; break;
      case SPK_VG_NUG: g = s; break;
      case SPK_VG_SPH:
        g = (h >= f.r[k]) ? s : s * (h * f.c1[k] - (h * h * h) * f.c2[k]);
        break;
      case SPK_VG_EXP: g = s * (1.0 - exp(h * f.c1[k])); break;
      case SPK_VG_GAU: g = s * (1.0 - exp((h * h) * f.c1[k])); break;
      case SPK_VG_LIN: g = (h > f.r[k]) ? s : s * (h * f.c1[k]); break;
      default: g = NAN;
    }
    acc += covar_flag ? (s - g) : g;
  }
  if (acc <= min_vg_val) acc = 0.0;
  return acc;
}

// One bounded model (spherical, exponential or Gaussian) with or without a
// nugget, the common form of fitted variograms, so that the cell kernels'
// station loops can be specialised on the model with no component loop
// inside.  The operations are those of vg_simple_entry for these
// components (its sum of two terms commutes exactly), and neither
// contracts a product and a sum into one rounding, so both give the same
// values.
struct VgOne {
  double nug, s, r, c1, c2;
};
__device__ inline int vg_one_of(const VgFast &f, VgOne &o) {
  int model = -1, nugs = 0;
  o.nug = 0.0;
  for (int k = 0; k < f.n; ++k) {
    const int t = f.type[k];
    if (t == SPK_VG_NUG) {
      ++nugs;
      o.nug = f.s[k];
    } else if (model < 0 &&
               (t == SPK_VG_SPH || t == SPK_VG_EXP || t == SPK_VG_GAU)) {
      model = k;
    } else {
      return -1;
    }
  }
  if (model < 0 || nugs > 1) return -1;
  o.s = f.s[model];
  o.r = f.r[model];
  o.c1 = f.c1[model];
  o.c2 = f.c2[model];
  return f.type[model];
}
template <int MT>
__device__ inline double vg_one(const VgOne &o, double h, double min_vg_val) {
#pragma clang fp contract(off)
  double g;
  if constexpr (MT == SPK_VG_SPH)
    g = (h >= o.r) ? o.s : o.s * (h * o.c1 - (h * h * h) * o.c2);
  else if constexpr (MT == SPK_VG_EXP)
    g = o.s * (1.0 - exp(h * o.c1));
  else
    g = o.s * (1.0 - exp((h * h) * o.c1));
  double acc = o.nug + g;
  if (acc <= min_vg_val) acc = 0.0;
  return acc;
}

__host__ __device__ inline bool vg_simple(const spk_vg &vg) {
  for (int k = 0; k < vg.n; ++k)
    if (vg.type[k] == SPK_VG_POW || vg.type[k] == SPK_VG_HOL) return false;
  return true;
}

// True when every component is bounded by its sill, so that
// c0 = sum of sills turns the semivariogram matrix into a covariance.
inline bool vg_bounded(const spk_vg &vg) {
  for (int k = 0; k < vg.n; ++k) {
    const int t = vg.type[k];
    if (t == SPK_VG_RNG || t == SPK_VG_POW || t == SPK_VG_HOL) return false;
  }
  return true;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace spk
