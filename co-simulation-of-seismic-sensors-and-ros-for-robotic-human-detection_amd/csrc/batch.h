// batch.h — many shots advanced together on one GPU (see run.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "sim.h"

namespace gfb {

struct Batch {
  std::vector<std::unique_ptr<gfs::Sim>> sims;  // null for shots that failed set-up
  std::vector<int32_t> st, steps;
  std::vector<std::string> msg;
  std::vector<std::vector<int32_t>> cols;
  std::vector<std::vector<double>> snapped;
  int first_err = GF_OK;
  std::string first_msg;
  hipStream_t stream = nullptr;  // the stream the batch ran on
  int tile_rows = 0;             // tile height of the last group run
  double kernel_ms = 0;          // step-loop device time (HIP events)
  int64_t launches = 0;          // step kernels launched
  double bytes = 0;              // algorithmic HBM bytes of all steps run

  void setup(const gf_shot *shots, int32_t n, const std::vector<std::vector<double>> &ranges);
  int run();
  void probes(int i, double *peak);
  int fail_shot(int i, int code, const std::string &m);
};

// Statistics of the calling thread's last batch (gf_batch_stats).
void record_stats(const Batch &b);

}  // namespace gfb
