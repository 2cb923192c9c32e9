// Kernel timing by hipEvents around the heavy launches of a context
// (enabled by spk_prof_enable; read by spk_prof_get).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "spk_internal.h"

namespace spk {

struct Prof {
  bool on = false;
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<std::string, std::pair<double, int64_t>>> stats;
  // device counters of executed work: [0] 16-row chunks multiplied by the
  // FP32 triangular contraction, [1] 64-k chunks times plane products of
  // the int8 contraction, [2] 32-row chunks of the split-BF16 contraction
  unsigned long long *work = nullptr;
  // host-side counts of work items (e.g. neighbourhood chunks)
  std::vector<std::pair<std::string, double>> counts;
};

Prof &prof_of(spk_ctx *ctx);
hipEvent_t prof_event(Prof &pf);

// Adds v to the host-side count `name` while profiling is on.
inline void prof_count(spk_ctx *ctx, const char *name, double v) {
  Prof &pf = prof_of(ctx);
  if (!pf.on) return;
  for (auto &kv : pf.counts)
    if (kv.first == name) {
      kv.second += v;
      return;
    }
  pf.counts.push_back({name, v});
}

struct ProfScope {
  spk_ctx *ctx;
  Prof &pf;
  const char *name;
  hipStream_t st;
  hipEvent_t a{}, b{};
  ProfScope(spk_ctx *c, const char *nm, hipStream_t s = nullptr)
      : ctx(c), pf(prof_of(c)), name(nm), st(s ? s : c->stream) {
    if (pf.on) {
      a = prof_event(pf);
      (void)hipEventRecord(a, st);
    }
  }
  ~ProfScope() {
    if (pf.on) {
      b = prof_event(pf);
      (void)hipEventRecord(b, st);
      pf.pending.push_back({name, a, b});
    }
  }
};

}  // namespace spk
