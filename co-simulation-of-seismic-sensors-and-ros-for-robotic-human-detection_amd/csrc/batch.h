// batch.h — many shots advanced together on one GPU (see run.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "sim.h"

namespace gfb {

// The launch plan of one set of models advanced together: one kernel variant
// (attenuation form), one tile height, models ordered longest run first so
// the running models are always a prefix of the tile list.
struct Group {
  int nmech = 0;
  bool scalar = true;
  int span = 1;                      // time steps per launch (1 or 2)
  int th = 0;                        // tile rows shared by every model of the group
  int waves = 4;                     // tiles (wavefronts) per workgroup
  std::vector<int> order;            // shot indices, longest run first
  std::vector<int64_t> prefix;       // tiles of the first k models
  gfk::GridDesc *d_descs = nullptr;
  int64_t descs_cap = 0;
  uint16_t *d_tile_model = nullptr;  // model of each tile
  int64_t tile_model_cap = 0;
  // two-step groups: the tile indices split by gfk::step2_interior, each kind
  // model by model ([interior tiles][the others]), so that the interior
  // tiles run in the kernel instance without the general path
  uint32_t *d_tile_list = nullptr;
  int64_t tile_list_cap = 0;
  std::vector<int64_t> in_prefix, bd_prefix;  // entries of each kind of the first k models
  int32_t *d_nbase = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_models = -1;             // active-model count the graph was captured for
};

// A group whose combined state fits the last-level cache, with its progress:
// every model of a chunk is at step `done` (or finished).
struct Chunk {
  Group G;
  int variant = 0;
  int64_t done = 0;
  int64_t max_steps = 0;
  int buf = 0;   // the buffer holding the chunk's current wavefields
  int mbuf = 0;  // the buffer holding its memory variables and stress psi
};

struct Batch {
  std::vector<std::unique_ptr<gfs::Sim>> sims;  // null for shots that failed set-up
  std::vector<int32_t> st, steps;
  std::vector<std::string> msg;
  std::vector<std::vector<int32_t>> cols;
  std::vector<std::vector<double>> snapped;
  int first_err = GF_OK;
  std::string first_msg;
  hipStream_t stream = nullptr;  // the stream the batch runs on
  int tile_rows = 0;             // tile height of the last group planned
  int waves = 0;                 // tiles per workgroup of the last group planned
  int span = 0;                  // time steps per launch of the last group planned
  double budget_mb = 0;          // chunk budget the batch was cut with (0: all together)
  double kernel_ms = 0;          // step-loop device time since the last reset (HIP events)
  int64_t launches = 0;          // step kernels launched since the last reset
  double bytes = 0;              // algorithmic HBM bytes of the steps run since the last reset
  double pass_bytes = 0;         // the same per launch: each launch reads and writes the state once
  std::vector<Chunk> chunks;
  bool planned = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;

  ~Batch();
  void setup(const gf_shot *shots, int32_t n, const std::vector<std::vector<double>> &ranges);
  // Group the set-up models into cache-sized chunks, pick each variant's tile
  // height (tuning it on long runs) and upload the launch tables.  Once.
  int plan();
  // Advance every chunk to ceil(longest run x num / den) steps (num == den:
  // to the end).  Chunks run one after another, each in graph blocks.
  int advance(int64_t num, int64_t den);
  // Every model back to step 0 (zero state, cleared probes), timing cleared.
  int reset();
  bool finished() const;
  int run() { return advance(1, 1); }
  void probes(int i, double *peak);
  int fail_shot(int i, int code, const std::string &m);
};

// Statistics of the calling thread's last batch (gf_batch_stats).
void record_stats(const Batch &b);
// Per-shot dt, t0, cells and snapped ranges of a set-up batch.
void describe_results(const Batch &B, gf_result *results);
// Probes, traces and statuses of a finished batch, as gf_run_batch returns them.
int finish_results(Batch &B, gf_result *results, int32_t *statuses);

}  // namespace gfb
