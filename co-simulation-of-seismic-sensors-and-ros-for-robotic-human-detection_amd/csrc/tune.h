// tune.h — the tile-shape autotuner.
//
// The step kernel's speed depends on the tile height mostly through how many
// loads each SIMD keeps in flight, which depends on the grid sizes, the
// number of models in a launch and the attenuation variant (register use).
// Rather than model that, the tuner times the real launch: a few steps at
// each candidate height on the caller's own models, keeping the fastest.
// Results are cached in-process by (variant, models, grid-size bucket).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <functional>

namespace gft {

// Candidate tile heights: the fully unrolled kernels.
constexpr int kCandidates[] = {1, 2, 4, 8};

struct Key {
  int nmech, scalar, models;
  int64_t cells_bucket;  // log2 of the total cell count
  bool operator<(const Key &o) const;
};

// prepare(th) re-plans the launch for tile height th (returns false to skip
// the candidate); launch(th, n) enqueues step n on `stream`.  Returns the
// fastest th (0 if every candidate failed).  `trial_steps` steps per
// candidate, after one untimed warm-up step.  The state is left advanced and
// the launch planned for the last candidate: callers re-plan and reset.
int tune(const Key &key, hipStream_t stream, int trial_steps, const std::function<bool(int)> &prepare,
         const std::function<hipError_t(int, int)> &launch, double *best_us);

// Cached result for a key, or 0.
int cached(const Key &key);

}  // namespace gft
