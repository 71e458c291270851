// tune.cpp — see tune.h.
#include "tune.h"

#include <math.h>
#include <stdlib.h>

#include <algorithm>
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
    for (int w : {4, 2, 1}) out.push_back(Shape{th, w, 1}.code());
  for (int th : {8, 16, 32})
    for (int w : {4, 2, 1}) out.push_back(Shape{th, w, 2}.code());
  return out;
}

static std::mutex g_mu;
static std::map<Key, int> g_cache;

int cached(const Key &key) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(key);
  return it == g_cache.end() ? 0 : it->second;
}

int tune(const Key &key, hipStream_t stream, int trial_steps, const std::vector<int> &candidates, int preferred,
         const std::function<bool(int)> &prepare, const std::function<hipError_t(int, int)> &launch,
         const std::function<double(int)> &work, double *best_us) {
  if (int c = cached(key)) return c;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 0;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return 0;
  }
  const size_t nc = candidates.size();
  std::vector<std::vector<float>> ms(nc);
  int n = 0;
  for (int round = 0; round < kRounds; round++)
    for (size_t i = 0; i < nc; i++) {
      const int c = candidates[i];
      if (!prepare(c)) continue;
      if (launch(c, n++) != hipSuccess) continue;  // warm-up (code load, caches)
      bool ok = hipEventRecord(e0, stream) == hipSuccess;
      for (int k = 0; ok && k < trial_steps; k++) ok = launch(c, n++) == hipSuccess;
      ok = ok && hipEventRecord(e1, stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      float t = 0;
      if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess && t > 0) ms[i].push_back(t);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int best = 0;
  double best_rate = 0, best_ms = 0, preferred_rate = 0, preferred_ms = 0;
  for (size_t i = 0; i < nc; i++) {
    if (ms[i].size() != (size_t)kRounds) continue;  // failed in some round
    std::sort(ms[i].begin(), ms[i].end());
    const double med = ms[i][kRounds / 2];
    const double rate = (work ? work(candidates[i]) : 1.0) / med;
    if (candidates[i] == preferred) {
      preferred_rate = rate;
      preferred_ms = med;
    }
    if (rate > best_rate) {
      best_rate = rate;
      best_ms = med;
      best = candidates[i];
    }
  }
  if (preferred_rate > 0 && best_rate <= preferred_rate * (1.0 + kMargin)) {
    best = preferred;
    best_ms = preferred_ms;
  }
  if (best) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
    if (best_us) *best_us = best_ms * 1e3 / trial_steps;
  }
  return best;
}

}  // namespace gft
