// tune.h — the launch-shape autotuner.
//
// The step kernel's speed depends on the tile height mostly through how many
// loads each SIMD keeps in flight, which depends on the grid sizes, the
// number of models in a launch and the attenuation variant (register use);
// the tiles per workgroup change how the waves of a CU are scheduled; and a
// batch's chunk size decides whether its state stays in the last-level cache.
// Rather than model that, the tuner times the real launches: a few steps of
// each candidate on the caller's own models, keeping the fastest.  Results
// are cached in-process by (what is tuned, variant, models, grid-size bucket).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <functional>
#include <vector>

namespace gft {

// A launch shape: time steps per launch (span 1: the one-step kernel, whose
// tile heights 1, 2, 4, 8 are fully unrolled; span 2: the two-step kernel,
// any tile height), tile height and tiles (wavefronts) per workgroup (1, 2,
// 4), packed in one int.
struct Shape {
  int th = 1, waves = 4, span = 1;
  int code() const { return th | (waves << 8) | (span << 12); }
  static Shape of(int code) { return Shape{code & 0xff, (code >> 8) & 0xf, (code >> 12) & 0xf}; }
};

// Every candidate shape: the one-step kernel's, then the two-step kernel's
// (tile heights 8, 16, 32: its halo is three rows above and below, so short
// tiles repeat much of their work).
std::vector<int> shape_candidates();

// Candidate chunk budgets of a batch, MB of state per chunk.
constexpr int kBudgetsMB[] = {96, 128, 160, 192, 224};

struct Key {
  int what;  // 0: launch shape, 1: chunk budget, 2: launch shape of an all-together batch
  int nmech, scalar, models;
  int64_t cells_bucket;  // log2 of the total cell count
  bool operator<(const Key &o) const;
};

// prepare(c) re-plans the launch for candidate c (returns false to skip it);
// launch(c, k) enqueues the k-th trial launch on `stream` (reading buffer
// k & 1); work(c) is the algorithmic bytes of one launch under c (null:
// every candidate does the same work, so time decides).  The candidates are timed in kRounds interleaved rounds of
// `trial_steps` steps (after one untimed warm-up step each time), and each is
// scored by its median round.  Returns the best candidate by bytes per second,
// except that `preferred` (the untuned choice) is kept unless another beats
// it by more than kMargin: trial timings vary by a few per cent, and a choice
// made on noise can cost more over a long run than it saves.  0 if every
// candidate failed.  The state is left advanced and the launch planned for
// the last candidate: callers re-plan and reset.
constexpr int kRounds = 3;
constexpr double kMargin = 0.03;
int tune(const Key &key, hipStream_t stream, int trial_steps, const std::vector<int> &candidates, int preferred,
         const std::function<bool(int)> &prepare, const std::function<hipError_t(int, int)> &launch,
         const std::function<double(int)> &work, double *best_us);

// Cached result for a key, or 0.
int cached(const Key &key);

}  // namespace gft
