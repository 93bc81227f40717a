// tune.cpp - the per-context autotuner of kernel variants.
//
// A family's variants compute the same thing in the same arithmetic order
// and differ in how the work is spread over the GPU (workgroup per plane or
// wave per line block, tile widths).  Which is faster depends on the shape:
// many planes fill the GPU either way, a handful of planes leaves it idle
// unless the lines of one plane are spread out.  The tuner measures instead
// of guessing: the first kReps * nvar calls of a (family, shape) run the
// variants in turn between two events; the call after the last trial waits
// for their end events (outside a graph capture; inside one the events are
// only queried) and fixes the variant with the smallest best time for that
// shape, for this context and for contexts created later in the process on
// the same device.
#include <algorithm>
#include <limits>
#include <mutex>

#include "kernels.hpp"

namespace g4 {

namespace {
// Decisions shared by all contexts of the process, keyed by device, family
// and shape: a new context (a batch group, a request lane) starts from what
// earlier contexts measured instead of trying the variants again.
std::mutex g_mu;
std::map<std::vector<int>, int> g_decided;
}  // namespace

int Context::tune_pick(int family, const std::vector<int>& shape, int nvar, hipEvent_t* a,
                       hipEvent_t* b) {
  *a = nullptr;
  *b = nullptr;
  if (family < 0 || family >= TUNE_COUNT || nvar <= 1) return 0;
  const int forced = tuner.forced[family];
  if (forced >= 0) return std::min(forced, nvar - 1);
  std::vector<int> key;
  key.reserve(shape.size() + 1);
  key.push_back(family);
  key.insert(key.end(), shape.begin(), shape.end());
  Tuner::Entry& e = tuner.table[key];
  if (e.nvar == 0) e.nvar = nvar;
  if (e.choice >= 0) return e.choice;
  std::vector<int> gkey{device};
  gkey.insert(gkey.end(), key.begin(), key.end());
  if (e.issued == 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_decided.find(gkey);
    if (it != g_decided.end()) {
      e.choice = it->second;
      return e.choice;
    }
  }
  if (capturing) return 0;  // no trials inside a graph: variant 0 until decided
  const int total = Tuner::kReps * e.nvar;
  if (e.issued < total) {
    // next trial: variants in turn, so every variant sees warm and cold caches
    Tuner::Trial t{e.issued % e.nvar, nullptr, nullptr};
    G4_HIP(hipEventCreate(&t.a));
    G4_HIP(hipEventCreate(&t.b));
    e.trials.push_back(t);
    ++e.issued;
    *a = t.a;
    *b = t.b;
    return t.variant;
  }
  // all trials issued: decide once they have all completed.  Outside a
  // capture the host waits for them here, once per shape (a few launches'
  // time), rather than running variant 0 until they happen to be done: the
  // host may be many steps ahead of the device, and variant 0 can be several
  // times slower than the best on shapes with few planes.
  for (const Tuner::Trial& t : e.trials) {
    if (!capturing) G4_HIP(hipEventSynchronize(t.b));
    const hipError_t q = hipEventQuery(t.b);
    if (q == hipErrorNotReady) return 0;  // provisional: variant 0
    G4_HIP(q);
  }
  std::vector<float> best(static_cast<size_t>(e.nvar), std::numeric_limits<float>::max());
  for (const Tuner::Trial& t : e.trials) {
    float ms = 0.f;
    G4_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    best[static_cast<size_t>(t.variant)] = std::min(best[static_cast<size_t>(t.variant)], ms);
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  e.trials.clear();
  e.choice = static_cast<int>(std::min_element(best.begin(), best.end()) - best.begin());
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_decided[gkey] = e.choice;
  }
  return e.choice;
}

void Context::tune_settle() {
  for (auto& kv : tuner.table) {
    Tuner::Entry& e = kv.second;
    if (e.choice >= 0 || e.issued < Tuner::kReps * e.nvar) continue;
    hipEvent_t a = nullptr, b = nullptr;
    std::vector<int> shape(kv.first.begin() + 1, kv.first.end());
    (void)tune_pick(kv.first[0], shape, e.nvar, &a, &b);  // decides if complete
  }
}

void Context::tune_release() {
  for (auto& kv : tuner.table) {
    for (const Tuner::Trial& t : kv.second.trials) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    kv.second.trials.clear();
  }
}

}  // namespace g4

using namespace g4;

extern "C" {

g4h_status g4h_context_set_variant(g4h_context* ctx, int family, int variant) {
  G4_API_BEGIN
  require(ctx, "ctx is null");
  require(family >= 0 && family < TUNE_COUNT, "unknown kernel family");
  require(variant >= -1, "variant must be >= -1 (-1: autotune)");
  require(family != TUNE_SPLIT || variant != 0, "workgroups per slice must be >= 1");
  require(family != TUNE_BATCH || variant <= 2, "batch variant must be -1 .. 2");
  require(family != TUNE_ADI || variant <= 3, "ADI variant must be -1 .. 3");
  ctx->tuner.forced[family] = variant;
  G4_API_END
}

g4h_status g4h_context_get_variant(g4h_context* ctx, int family, const int* shape, int nshape,
                                   int* variant) {
  G4_API_BEGIN
  require(ctx && variant, "null argument");
  require(family >= 0 && family < TUNE_COUNT, "unknown kernel family");
  require(nshape >= 0 && (nshape == 0 || shape), "invalid shape");
  std::vector<int> key{family};
  for (int i = 0; i < nshape; ++i) key.push_back(shape[i]);
  auto it = ctx->tuner.table.find(key);
  *variant = (it == ctx->tuner.table.end()) ? -1 : it->second.choice;
  G4_API_END
}

}  // extern "C"
