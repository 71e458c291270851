// tune.cpp — see tune.h.
#include "tune.h"

#include <math.h>
#include <stdlib.h>

#include <map>
#include <mutex>

namespace gft {

bool Key::operator<(const Key &o) const {
  if (nmech != o.nmech) return nmech < o.nmech;
  if (scalar != o.scalar) return scalar < o.scalar;
  if (models != o.models) return models < o.models;
  return cells_bucket < o.cells_bucket;
}

static std::mutex g_mu;
static std::map<Key, int> g_cache;

int cached(const Key &key) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(key);
  return it == g_cache.end() ? 0 : it->second;
}

int tune(const Key &key, hipStream_t stream, int trial_steps, const std::function<bool(int)> &prepare,
         const std::function<hipError_t(int, int)> &launch, double *best_us) {
  if (int c = cached(key)) return c;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 0;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return 0;
  }
  int best = 0, n = 0;
  double best_t = INFINITY;
  for (int th : kCandidates) {
    if (!prepare(th)) continue;
    if (launch(th, n++) != hipSuccess) continue;  // warm-up (code load, caches)
    bool ok = hipEventRecord(e0, stream) == hipSuccess;
    for (int k = 0; ok && k < trial_steps; k++) ok = launch(th, n++) == hipSuccess;
    ok = ok && hipEventRecord(e1, stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0;
    if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0 && ms < best_t) {
      best_t = ms;
      best = th;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
    if (best_us) *best_us = best_t * 1e3 / trial_steps;
  }
  return best;
}

}  // namespace gft
