// tune.cpp — see tune.h.
#include "tune.h"

#include <math.h>
#include <stdlib.h>

#include <map>
#include <mutex>
#include <tuple>

namespace gft {

bool Key::operator<(const Key &o) const {
  return std::tie(what, nmech, scalar, models, cells_bucket) <
         std::tie(o.what, o.nmech, o.scalar, o.models, o.cells_bucket);
}

std::vector<int> shape_candidates() {
  std::vector<int> out;
  for (int th : {1, 2, 4, 8})
    for (int w : {4, 2, 1}) out.push_back(Shape{th, w}.code());
  return out;
}

static std::mutex g_mu;
static std::map<Key, int> g_cache;

int cached(const Key &key) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(key);
  return it == g_cache.end() ? 0 : it->second;
}

int tune(const Key &key, hipStream_t stream, int trial_steps, const std::vector<int> &candidates,
         const std::function<bool(int)> &prepare, const std::function<hipError_t(int, int)> &launch,
         const std::function<double(int)> &work, double *best_us) {
  if (int c = cached(key)) return c;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 0;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return 0;
  }
  int best = 0, n = 0;
  double best_rate = 0, best_ms = 0;
  for (int c : candidates) {
    if (!prepare(c)) continue;
    if (launch(c, n++) != hipSuccess) continue;  // warm-up (code load, caches)
    bool ok = hipEventRecord(e0, stream) == hipSuccess;
    for (int k = 0; ok && k < trial_steps; k++) ok = launch(c, n++) == hipSuccess;
    ok = ok && hipEventRecord(e1, stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0;
    if (!ok || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0) continue;
    const double rate = (work ? work(c) : 1.0) / ms;
    if (rate > best_rate) {
      best_rate = rate;
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
    if (best_us) *best_us = best_ms * 1e3 / trial_steps;
  }
  return best;
}

}  // namespace gft
