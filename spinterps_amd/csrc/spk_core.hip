// Context, error state, variogram parsing and the elementwise kernels that
// back the host-pointer mirrors of the cyth API (cyth/__init__.py:9-21).
#include <algorithm>
#include <cstring>
#include <string>

#include <mutex>

#include "spk_krige.h"

namespace spk {

// Scratch memory of the allocation-free entry point
// spk_idw_steps_noalloc_dev (layout in spk_krige.h, about 23 MB).  Each
// device the library runs on holds its own copy until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory:
// only the address resolved below (per device, with that device current)
// is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch[kMaxDevices] = {};
static int g_scratch_rc[kMaxDevices] = {};

// Once per device, with the device current.
static void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device] || g_scratch_rc[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_rc[device] = e != hipSuccess ? (int)e : -1;
    return;
  }
  g_scratch[device] = static_cast<unsigned char *>(p);
}

int noalloc_scratch(int device, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices || !g_scratch[device])
    return set_error(SPK_EHIP,
                     "scratch of the allocation-free entry point not "
                     "resolved on device %d (hip error %d)", device,
                     (device >= 0 && device < kMaxDevices)
                         ? g_scratch_rc[device] : 0);
  *out = g_scratch[device];
  return SPK_OK;
}

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SPK_EHIP, "launch of %s failed: %s", what,
                     hipGetErrorString(e));
  return SPK_OK;
}

int ws_get(spk_ctx *ctx, const char *name, size_t bytes, void **out) {
  if (bytes == 0) bytes = 256;
  for (auto &kv : ctx->bufs) {
    if (kv.first == name) {
      if (kv.second.bytes >= bytes) {
        *out = kv.second.ptr;
        return SPK_OK;
      }
      // The streams may still use the old buffer.
      SPK_HIP(hipStreamSynchronize(ctx->stream));
      if (ctx->aux_stream) SPK_HIP(hipStreamSynchronize(ctx->aux_stream));
      if (ctx->copy_stream) SPK_HIP(hipStreamSynchronize(ctx->copy_stream));
      SPK_HIP(hipFree(kv.second.ptr));
      kv.second.ptr = nullptr;
      kv.second.bytes = 0;
      void *p = nullptr;
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess)
        return set_error(SPK_ENOMEM, "hipMalloc(%zu) for %s: %s", bytes, name,
                         hipGetErrorString(e));
      kv.second.ptr = p;
      kv.second.bytes = bytes;
      *out = p;
      return SPK_OK;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    return set_error(SPK_ENOMEM, "hipMalloc(%zu) for %s: %s", bytes, name,
                     hipGetErrorString(e));
  ctx->bufs.push_back({std::string(name), Buffer{p, bytes}});
  *out = p;
  return SPK_OK;
}

int h2d(spk_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return SPK_OK;
  SPK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return SPK_OK;
}

int d2h(spk_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return SPK_OK;
  SPK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SPK_HIP(hipStreamSynchronize(ctx->stream));
  return SPK_OK;
}

int sync(spk_ctx *ctx) {
  SPK_HIP(hipStreamSynchronize(ctx->stream));
  return SPK_OK;
}

void mark_reset(spk_ctx *ctx) {
  ctx->n_marks = 0;
  ctx->stage_names.clear();
}

int mark(spk_ctx *ctx, const char *name) {
  if ((int)ctx->events.size() <= ctx->n_marks) {
    hipEvent_t ev;
    SPK_HIP(hipEventCreate(&ev));
    ctx->events.push_back(ev);
  }
  SPK_HIP(hipEventRecord(ctx->events[ctx->n_marks], ctx->stream));
  ctx->stage_names.push_back(name);
  ctx->n_marks++;
  return SPK_OK;
}

int mark_collect(spk_ctx *ctx) {
  ctx->last_ms.clear();
  if (ctx->n_marks < 2) return SPK_OK;
  SPK_HIP(hipEventSynchronize(ctx->events[ctx->n_marks - 1]));
  for (int i = 1; i < ctx->n_marks; ++i) {
    float ms = 0.f;
    SPK_HIP(hipEventElapsedTime(&ms, ctx->events[i - 1], ctx->events[i]));
    ctx->last_ms.push_back(ms);
  }
  return SPK_OK;
}

// ---- tile autotuner ----------------------------------------------------------
// The kernel families with tile variants, and how many each has.
struct TuneFamily {
  const char *name;
  int variants;
};
static const TuneFamily kTuneFamilies[] = {{"gemm_plain_f64", 3},
                                           {"gemm_tri_bf16", 2}};

static int tune_family_variants(const char *name) {
  for (const auto &f : kTuneFamilies)
    if (!strcmp(f.name, name)) return f.variants;
  return 0;
}

int tune_launch(spk_ctx *ctx, const char *kernel, int cls, hipStream_t st,
                int ncand, int dflt, const std::function<int(int, bool)> &run) {
  for (const auto &f : ctx->forced)
    if (f.first == kernel) return run(f.second, false);
  if (ctx->tune_mode == SPK_TUNE_OFF || ncand < 2) return run(dflt, false);
  for (const auto &e : ctx->tuned)
    if (e.kernel == kernel && e.cls == cls) return run(e.choice, false);
  // time the variants alone: the context's other stream (a factorisation
  // running beside the contraction) is drained first
  if (ctx->aux_stream && ctx->aux_stream != st)
    SPK_HIP(hipStreamSynchronize(ctx->aux_stream));
  hipEvent_t ev[3];
  for (auto &e : ev) SPK_HIP(hipEventCreate(&e));
  spk_ctx::TuneEntry ent;
  ent.kernel = kernel;
  ent.cls = cls;
  int rc = SPK_OK;
  float best = INFINITY;
  for (int v = 0; v < ncand && rc == SPK_OK; ++v) {
    // the first run warms caches (and code objects), the faster of the
    // next two is kept
    if (rc == SPK_OK) rc = run(v, true);
    if (rc == SPK_OK && hipEventRecord(ev[0], st) != hipSuccess) rc = SPK_EHIP;
    if (rc == SPK_OK) rc = run(v, true);
    if (rc == SPK_OK && hipEventRecord(ev[1], st) != hipSuccess) rc = SPK_EHIP;
    if (rc == SPK_OK) rc = run(v, true);
    if (rc == SPK_OK && hipEventRecord(ev[2], st) != hipSuccess) rc = SPK_EHIP;
    float ms = 0.f, ms2 = 0.f;
    if (rc == SPK_OK && (hipEventSynchronize(ev[2]) != hipSuccess ||
                         hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess ||
                         hipEventElapsedTime(&ms2, ev[1], ev[2]) != hipSuccess))
      rc = SPK_EHIP;
    if (rc != SPK_OK) break;
    ms = std::min(ms, ms2);
    ent.ms.push_back(ms);
    if (ms < best) {
      best = ms;
      ent.choice = v;
    }
  }
  for (auto &e : ev) hipEventDestroy(e);
  if (rc == SPK_EHIP)
    return set_error(SPK_EHIP, "tuning %s: %s", kernel,
                     hipGetErrorString(hipGetLastError()));
  SPK_TRY(rc);
  ctx->tuned.push_back(ent);
  return run(ent.choice, false);
}

// ---- captured launch sequences ------------------------------------------
// At most this many graphs per context; the cache is emptied when full.
constexpr size_t kMaxGraphs = 32;

int graphs_clear(spk_ctx *ctx) {
  if (ctx->graphs.empty()) return SPK_OK;
  // a replay may still run on either stream
  SPK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->aux_stream) SPK_HIP(hipStreamSynchronize(ctx->aux_stream));
  for (auto &g : ctx->graphs)
    if (g.exec) hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
  return SPK_OK;
}

int graph_run(spk_ctx *ctx, const char *tag, const std::vector<uint64_t> &key,
              hipStream_t st, const std::function<int()> &fn) {
  // the null stream cannot be captured
  if (!ctx->graphs_on || st == nullptr) return fn();
  for (const auto &g : ctx->graphs)
    if (g.tag == tag && g.key == key) {
      SPK_HIP(hipGraphLaunch(g.exec, st));
      return SPK_OK;
    }
  if (ctx->graphs.size() >= kMaxGraphs) SPK_TRY(graphs_clear(ctx));
  SPK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = fn();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != SPK_OK || e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    if (rc != SPK_OK) return rc;
    return set_error(SPK_EHIP, "capture of %s: %s", tag, hipGetErrorString(e));
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess)
    return set_error(SPK_EHIP, "instantiating %s: %s", tag,
                     hipGetErrorString(ei));
  ctx->graphs.push_back({tag, key, ex});
  SPK_HIP(hipGraphLaunch(ex, st));
  return SPK_OK;
}

// ---- variogram string parsing (pyx:174-184) ------------------------------

static std::string strip(const std::string &s) {
  // bytes.strip() removes ASCII whitespace
  size_t b = 0, e = s.size();
  while (b < e && isspace((unsigned char)s[b])) ++b;
  while (e > b && isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

static std::vector<std::string> split(const std::string &s, char sep) {
  std::vector<std::string> out;
  size_t start = 0;
  for (size_t i = 0; i <= s.size(); ++i) {
    if (i == s.size() || s[i] == sep) {
      out.push_back(s.substr(start, i - start));
      start = i + 1;
    }
  }
  return out;
}

// Python float(): strict decimal/exponent/inf/nan text with surrounding
// whitespace allowed.
static bool py_float(const std::string &txt, double *out) {
  std::string t = strip(txt);
  if (t.empty()) return false;
  std::string low = t;
  for (auto &c : low) c = (char)tolower((unsigned char)c);
  std::string body = low;
  double sign = 1.0;
  if (!body.empty() && (body[0] == '+' || body[0] == '-')) {
    if (body[0] == '-') sign = -1.0;
    body = body.substr(1);
  }
  if (body == "inf" || body == "infinity") {
    *out = sign * INFINITY;
    return true;
  }
  if (body == "nan") {
    *out = NAN;
    return true;
  }
  for (char c : body)
    if (!(isdigit((unsigned char)c) || c == '.' || c == 'e' || c == '+' ||
          c == '-' || c == '_'))
      return false;
  char *end = nullptr;
  double v = strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0') return false;
  *out = v;
  return true;
}

static int vg_type_of(const std::string &name) {
  static const char *names[] = {"Rng", "Nug", "Sph", "Exp",
                                "Lin", "Gau", "Pow", "Hol"};
  for (int i = 0; i < 8; ++i)
    if (name == names[i]) return i;
  return -1;
}

int parse_vg(const char *s, spk_vg *out) {
  if (!s) return set_error(SPK_EBADMODEL, "null variogram string");
  std::memset(out, 0, sizeof(*out));
  std::string str(s);
  if (str == "nan") return set_error(SPK_EBADMODEL, "NaN VG!");
  auto comps = split(str, '+');
  if ((int)comps.size() > SPK_VG_MAX)
    return set_error(SPK_EBADMODEL, "more than %d components in '%s'",
                     SPK_VG_MAX, s);
  double sill_sum = 0.0, rmax = 0.0;
  for (auto &c0 : comps) {
    std::string c = strip(c0);
    auto sp = split(c, ' ');
    if (sp.size() != 2)
      return set_error(SPK_EBADMODEL, "component '%s' of '%s' is not "
                       "'<sill> <Name>(<range>)'", c.c_str(), s);
    auto pr = split(sp[1], '(');
    if (pr.size() != 2)
      return set_error(SPK_EBADMODEL, "component '%s' of '%s' has no single "
                       "'('", c.c_str(), s);
    std::string name = pr[0];
    std::string rng = split(pr[1], ')')[0];
    int ty = vg_type_of(name);
    if (ty < 0)
      return set_error(SPK_EBADMODEL, "unknown variogram model '%s' in '%s'",
                       name.c_str(), s);
    double sill, range;
    if (!py_float(sp[0], &sill) || !py_float(rng, &range))
      return set_error(SPK_EBADMODEL, "bad number in component '%s' of '%s'",
                       c.c_str(), s);
    const int k = out->n++;
    out->type[k] = ty;
    out->sill[k] = sill;
    out->range[k] = std::max(1e-5, range);
    sill_sum += sill;
    rmax = std::max(rmax, range);
  }
  out->sill_sum = sill_sum;
  out->range_max_raw = rmax;
  return SPK_OK;
}

// ---- elementwise kernels --------------------------------------------------

__global__ void k_dists(const double *__restrict__ x1,
                        const double *__restrict__ y1, int64_t n1,
                        const double *__restrict__ x2,
                        const double *__restrict__ y2, int64_t n2,
                        double *__restrict__ out) {
  const int64_t total = n1 * n2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n2, j = e - i * n2;
    out[e] = dist2(x1[i], y1[i], x2[j], y2[j]);
  }
}

__global__ void k_vg_var(const double *__restrict__ d, int64_t rows,
                         int64_t cols, double *__restrict__ out, spk_vg vg,
                         int covar_flag, int diag_flag, double min_vg_val) {
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t h = e / cols, g = e - h * cols;
    // diag flag: the upper triangle (g >= h) is computed, the lower one is
    // a copy of its mirror entry (pyx:186-224).
    const double dist = (diag_flag && g < h) ? d[g * cols + h] : d[e];
    out[e] = vg_entry(vg, dist, covar_flag, min_vg_val);
  }
}

__global__ void k_gather2d(const double *__restrict__ arr, int64_t arr_cols,
                           const int64_t *__restrict__ ri, int64_t nr,
                           const int64_t *__restrict__ ci, int64_t nc,
                           double *__restrict__ out, int64_t out_cols) {
  const int64_t total = nr * nc;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / nc, j = e - i * nc;
    out[i * out_cols + j] = arr[ri[i] * arr_cols + ci[j]];
  }
}

__global__ void k_wts(const double *__restrict__ d, double *__restrict__ w,
                      int64_t n, double p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 1.0 / (p == 2.0 ? d[i] * d[i] : pow(d[i], p));
}

// Single-block ordered sum of a[i] (*b[i]); deterministic.
__global__ void k_sum(const double *__restrict__ a,
                      const double *__restrict__ b, int64_t n,
                      double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
    s += b ? a[i] * b[i] : a[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_theo(const double *__restrict__ h, int64_t n, int type,
                       double r, double s, double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] += vg_comp(type, h[i], r, s);
}

__global__ void k_nd_dists(const double *__restrict__ pts, int64_t n,
                           int64_t nd, double *__restrict__ out) {
  const int64_t total = n * (n - 1) / 2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    // e = i*(i-1)/2 + j with 0 <= j < i
    int64_t i = (int64_t)floor((1.0 + sqrt(1.0 + 8.0 * (double)e)) * 0.5);
    while (i * (i - 1) / 2 > e) --i;
    while ((i + 1) * i / 2 <= e) ++i;
    const int64_t j = e - i * (i - 1) / 2;
    double acc = 0.0;
    for (int64_t k = 0; k < nd; ++k) {
      const double t = pts[i * nd + k] - pts[j * nd + k];
      acc += t * t;
    }
    out[e] = sqrt(acc);
  }
}

static int grid_for(int64_t n) {
  int64_t g = ceil_div(n, 256);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

}  // namespace spk

using namespace spk;

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

const char *spk_last_error(void) { return g_err.c_str(); }
int spk_version(void) { return 1; }

int spk_ctx_create(int device, spk_ctx **out) {
  if (!out) return set_error(SPK_EARG, "null out");
  int n = 0;
  SPK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return set_error(SPK_EARG, "device %d not present (%d devices)", device, n);
  SPK_HIP(hipSetDevice(device));
  // the allocation-free entry point's scratch address on this device
  // (a failure only disables that entry point)
  resolve_scratch(device);
  spk_ctx *c = new spk_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_error(SPK_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  c->own_stream = true;
  *out = c;
  return SPK_OK;
}

int spk_ctx_trim(spk_ctx *ctx) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  SPK_HIP(hipSetDevice(ctx->device));
  SPK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->aux_stream) SPK_HIP(hipStreamSynchronize(ctx->aux_stream));
  if (ctx->copy_stream) SPK_HIP(hipStreamSynchronize(ctx->copy_stream));
  // the graphs hold the buffers' addresses
  SPK_TRY(graphs_clear(ctx));
  for (auto &kv : ctx->bufs)
    if (kv.second.ptr) SPK_HIP(hipFree(kv.second.ptr));
  ctx->bufs.clear();
  return SPK_OK;
}

int spk_ctx_destroy(spk_ctx *ctx) {
  if (!ctx) return SPK_OK;
  int rc = spk_ctx_trim(ctx);
  pinv_release(ctx);
  if (ctx->h_sc) hipHostFree(ctx->h_sc);
  for (auto ev : ctx->events) hipEventDestroy(ev);
  for (auto ev : ctx->ev)
    if (ev) hipEventDestroy(ev);
  if (ctx->aux_stream) hipStreamDestroy(ctx->aux_stream);
  if (ctx->copy_stream) hipStreamDestroy(ctx->copy_stream);
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return rc;
}

int spk_ctx_set_stream(spk_ctx *ctx, void *s) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  SPK_HIP(hipSetDevice(ctx->device));
  SPK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->own_stream) SPK_HIP(hipStreamDestroy(ctx->stream));
  ctx->stream = (hipStream_t)s;
  ctx->own_stream = false;
  return SPK_OK;
}

int spk_ctx_sync(spk_ctx *ctx) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  return sync(ctx);
}

int spk_ctx_set_graphs(spk_ctx *ctx, int on) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (on != 0 && on != 1) return set_error(SPK_EARG, "graphs %d (0 or 1)", on);
  ctx->graphs_on = on != 0;
  return SPK_OK;
}

int spk_ctx_set_tuning(spk_ctx *ctx, int mode) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (mode != SPK_TUNE_OFF && mode != SPK_TUNE_AUTO)
    return set_error(SPK_EARG, "tuning mode %d (SPK_TUNE_OFF or _AUTO)", mode);
  ctx->tune_mode = mode;
  return SPK_OK;
}

int spk_ctx_force_variant(spk_ctx *ctx, const char *family, int v) {
  if (!ctx || !family) return set_error(SPK_EARG, "null argument");
  const int nv = tune_family_variants(family);
  if (nv == 0) return set_error(SPK_EARG, "no kernel family '%s'", family);
  if (v >= nv)
    return set_error(SPK_EARG, "%s has variants 0..%d, not %d", family,
                     nv - 1, v);
  auto &f = ctx->forced;
  f.erase(std::remove_if(f.begin(), f.end(),
                         [&](const std::pair<std::string, int> &e) {
                           return e.first == family;
                         }),
          f.end());
  if (v >= 0) f.emplace_back(family, v);
  return SPK_OK;
}

int spk_ctx_tuning_report(spk_ctx *ctx, char *buf, int64_t cap,
                          int64_t *len) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  std::string s;
  char tmp[64];
  for (const auto &e : ctx->tuned) {
    s += e.kernel;
    snprintf(tmp, sizeof tmp, " %d %d", e.cls, e.choice);
    s += tmp;
    for (float ms : e.ms) {
      snprintf(tmp, sizeof tmp, " %.4f", ms);
      s += tmp;
    }
    s += '\n';
  }
  if (len) *len = (int64_t)s.size();
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
    memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return SPK_OK;
}

int spk_vg_parse(const char *vg_str, spk_vg *out) {
  if (!out) return set_error(SPK_EARG, "null out");
  return parse_vg(vg_str, out);
}

// misc.py:1074-1105, including its parse: the sill is the text before '('
// minus the three-letter name, the range the unclamped text inside ( ).
int spk_vg_full_nuggetness(const char *vg_str, double min_vg_val, int *out) {
  if (!vg_str || !out) return set_error(SPK_EARG, "null argument");
  std::string s(vg_str);
  if (s == "nan") {
    *out = 0;
    return SPK_OK;
  }
  auto comps = split(s, '+');
  double sill = 0.0, rng = 0.0;
  for (auto &c0 : comps) {
    std::string c = strip(c0);
    auto lp = split(c, '(');
    if (lp.size() < 2)
      return set_error(SPK_EBADMODEL, "bad component '%s'", c.c_str());
    std::string head = strip(lp[0]);
    if (head.size() < 3)
      return set_error(SPK_EBADMODEL, "bad component '%s'", c.c_str());
    double sv, rv;
    if (!py_float(head.substr(0, head.size() - 3), &sv) ||
        !py_float(split(lp[1], ')')[0], &rv))
      return set_error(SPK_EBADMODEL, "bad number in '%s'", c.c_str());
    sill += sv;
    rng = std::max(rng, rv);
  }
  *out = (sill <= min_vg_val || rng <= min_vg_val) ? 1 : 0;
  return SPK_OK;
}

// ---- host-pointer cyth mirrors -----------------------------------------

int spk_fill_dists_2d_mat(spk_ctx *ctx, const double *x1s, const double *y1s,
                          int64_t n1, const double *x2s, const double *y2s,
                          int64_t n2, double *dists) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (n1 < 0 || n2 < 0) return set_error(SPK_ESHAPE, "negative size");
  if (n1 == 0 || n2 == 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  double *d = nullptr;
  const size_t b1 = n1 * sizeof(double), b2 = n2 * sizeof(double);
  SPK_TRY(ws_get(ctx, "io0", 2 * b1 + 2 * b2 + n1 * n2 * sizeof(double),
                 (void **)&d));
  double *dx1 = d, *dy1 = d + n1, *dx2 = d + 2 * n1, *dy2 = d + 2 * n1 + n2;
  double *dout = d + 2 * n1 + 2 * n2;
  SPK_TRY(h2d(ctx, dx1, x1s, b1));
  SPK_TRY(h2d(ctx, dy1, y1s, b1));
  SPK_TRY(h2d(ctx, dx2, x2s, b2));
  SPK_TRY(h2d(ctx, dy2, y2s, b2));
  hipLaunchKernelGGL(k_dists, dim3(grid_for(n1 * n2)), dim3(256), 0,
                     ctx->stream, dx1, dy1, n1, dx2, dy2, n2, dout);
  SPK_TRY(check_launch("k_dists"));
  return d2h(ctx, dists, dout, n1 * n2 * sizeof(double));
}

int spk_fill_vg_var_arr(spk_ctx *ctx, const double *dists, int64_t rows,
                        int64_t cols, double *in_vars, int covar_flag,
                        int diag_mat_flag, const char *vg_models_str,
                        double min_vg_val) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (rows < 0 || cols < 0) return set_error(SPK_ESHAPE, "negative size");
  if (diag_mat_flag && rows != cols)
    return set_error(SPK_ESHAPE, "diag_mat_flag needs a square array, got "
                     "%lld x %lld", (long long)rows, (long long)cols);
  spk_vg vg;
  SPK_TRY(parse_vg(vg_models_str, &vg));
  if (rows == 0 || cols == 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  const size_t nb = rows * cols * sizeof(double);
  double *d = nullptr;
  SPK_TRY(ws_get(ctx, "io0", 2 * nb, (void **)&d));
  SPK_TRY(h2d(ctx, d, dists, nb));
  hipLaunchKernelGGL(k_vg_var, dim3(grid_for(rows * cols)), dim3(256), 0,
                     ctx->stream, d, rows, cols, d + rows * cols, vg,
                     covar_flag ? 1 : 0, diag_mat_flag ? 1 : 0, min_vg_val);
  SPK_TRY(check_launch("k_vg_var"));
  return d2h(ctx, in_vars, d + rows * cols, nb);
}

int spk_copy_2d_arr_at_idxs(spk_ctx *ctx, const double *arr, int64_t arr_rows,
                            int64_t arr_cols, const int64_t *row_idxs,
                            int64_t n_rows, const int64_t *col_idxs,
                            int64_t n_cols, double *subset_arr,
                            int64_t sub_rows, int64_t sub_cols) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (n_rows > sub_rows || n_cols > sub_cols)
    return set_error(SPK_ESHAPE, "subset array [%lld x %lld] smaller than the "
                     "index sets (%lld, %lld)", (long long)sub_rows,
                     (long long)sub_cols, (long long)n_rows,
                     (long long)n_cols);
  for (int64_t i = 0; i < n_rows; ++i)
    if (row_idxs[i] < 0 || row_idxs[i] >= arr_rows)
      return set_error(SPK_EINDEX, "row index %lld out of [0, %lld)",
                       (long long)row_idxs[i], (long long)arr_rows);
  for (int64_t j = 0; j < n_cols; ++j)
    if (col_idxs[j] < 0 || col_idxs[j] >= arr_cols)
      return set_error(SPK_EINDEX, "col index %lld out of [0, %lld)",
                       (long long)col_idxs[j], (long long)arr_cols);
  if (n_rows == 0 || n_cols == 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  const size_t ab = arr_rows * arr_cols * sizeof(double);
  const size_t sb = sub_rows * sub_cols * sizeof(double);
  char *d = nullptr;
  SPK_TRY(ws_get(ctx, "io0",
                 ab + sb + (n_rows + n_cols) * sizeof(int64_t), (void **)&d));
  double *darr = (double *)d;
  double *dsub = (double *)(d + ab);
  int64_t *dri = (int64_t *)(d + ab + sb);
  int64_t *dci = dri + n_rows;
  SPK_TRY(h2d(ctx, darr, arr, ab));
  SPK_TRY(h2d(ctx, dsub, subset_arr, sb));  // untouched entries keep values
  SPK_TRY(h2d(ctx, dri, row_idxs, n_rows * sizeof(int64_t)));
  SPK_TRY(h2d(ctx, dci, col_idxs, n_cols * sizeof(int64_t)));
  hipLaunchKernelGGL(k_gather2d, dim3(grid_for(n_rows * n_cols)), dim3(256), 0,
                     ctx->stream, darr, arr_cols, dri, n_rows, dci, n_cols,
                     dsub, sub_cols);
  SPK_TRY(check_launch("k_gather2d"));
  return d2h(ctx, subset_arr, dsub, sb);
}

int spk_fill_dists_one_pt(spk_ctx *ctx, double x, double y, const double *xs,
                          const double *ys, int64_t n, double *dists) {
  return spk_fill_dists_2d_mat(ctx, &x, &y, 1, xs, ys, n, dists);
}

int spk_fill_wts_and_sum(spk_ctx *ctx, const double *dists, double *wts,
                         int64_t n, double idw_exp, double *wts_sum) {
  if (!ctx || !wts_sum) return set_error(SPK_EARG, "null argument");
  *wts_sum = 0.0;
  if (n <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  double *d = nullptr;
  SPK_TRY(ws_get(ctx, "io0", (2 * n + 1) * sizeof(double), (void **)&d));
  SPK_TRY(h2d(ctx, d, dists, n * sizeof(double)));
  hipLaunchKernelGGL(k_wts, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d,
                     d + n, n, idw_exp);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, ctx->stream, d + n,
                     (const double *)nullptr, n, d + 2 * n);
  SPK_TRY(check_launch("k_wts/k_sum"));
  SPK_TRY(d2h(ctx, wts, d + n, n * sizeof(double)));
  return d2h(ctx, wts_sum, d + 2 * n, sizeof(double));
}

int spk_get_mults_sum(spk_ctx *ctx, const double *wts, const double *data,
                      int64_t n, double *out) {
  if (!ctx || !out) return set_error(SPK_EARG, "null argument");
  *out = 0.0;
  if (n <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  double *d = nullptr;
  SPK_TRY(ws_get(ctx, "io0", (2 * n + 1) * sizeof(double), (void **)&d));
  SPK_TRY(h2d(ctx, d, wts, n * sizeof(double)));
  SPK_TRY(h2d(ctx, d + n, data, n * sizeof(double)));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, ctx->stream, d, d + n, n,
                     d + 2 * n);
  SPK_TRY(check_launch("k_sum"));
  return d2h(ctx, out, d + 2 * n, sizeof(double));
}

int spk_fill_theo_vg_vals(spk_ctx *ctx, const char *vg_name,
                          const double *h_arr, int64_t n, double r, double s,
                          double *vg_arr) {
  if (!ctx || !vg_name) return set_error(SPK_EARG, "null argument");
  int ty = vg_type_of(strip(vg_name));
  if (ty < 0)
    return set_error(SPK_EBADMODEL, "unknown variogram model '%s'", vg_name);
  // pyx:112-115
  if (n <= 0) return set_error(SPK_EARG, "empty h_arr");
  if (!(s >= 0)) return set_error(SPK_EARG, "sill must be >= 0");
  if (!(r >= 0)) return set_error(SPK_EARG, "range must be >= 0");
  SPK_HIP(hipSetDevice(ctx->device));
  double *d = nullptr;
  SPK_TRY(ws_get(ctx, "io0", 2 * n * sizeof(double), (void **)&d));
  SPK_TRY(h2d(ctx, d, h_arr, n * sizeof(double)));
  SPK_TRY(h2d(ctx, d + n, vg_arr, n * sizeof(double)));
  hipLaunchKernelGGL(k_theo, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d,
                     n, ty, r, s, d + n);
  SPK_TRY(check_launch("k_theo"));
  return d2h(ctx, vg_arr, d + n, n * sizeof(double));
}

int spk_get_nd_dists(spk_ctx *ctx, const double *pts, int64_t n_pts,
                     int64_t n_dims, double *out) {
  if (!ctx) return set_error(SPK_EARG, "null ctx");
  if (n_pts < 0 || n_dims < 0) return set_error(SPK_ESHAPE, "negative size");
  const int64_t np = n_pts * (n_pts - 1) / 2;
  if (np <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  double *d = nullptr;
  SPK_TRY(ws_get(ctx, "io0", (n_pts * n_dims + np) * sizeof(double),
                 (void **)&d));
  SPK_TRY(h2d(ctx, d, pts, n_pts * n_dims * sizeof(double)));
  hipLaunchKernelGGL(k_nd_dists, dim3(grid_for(np)), dim3(256), 0, ctx->stream,
                     d, n_pts, n_dims, d + n_pts * n_dims);
  SPK_TRY(check_launch("k_nd_dists"));
  return d2h(ctx, out, d + n_pts * n_dims, np * sizeof(double));
}

int spk_clip_cast_dev(spk_ctx *ctx, const double *in64, float *out32,
                      int64_t n, int has_min, double min_cut, int has_max,
                      double max_cut);

int spk_last_timings(spk_ctx *ctx, double *ms, int max_n, int *n_out) {
  if (!ctx || !n_out) return set_error(SPK_EARG, "null argument");
  int n = (int)ctx->last_ms.size();
  *n_out = n;
  for (int i = 0; i < n && i < max_n; ++i) ms[i] = ctx->last_ms[i];
  return SPK_OK;
}

}  // extern "C"
