// batch_driver.cpp — drives the library's batched and single-shot paths
// through the C ABI against the recording HIP stub (hipstub.cpp), built with
// AddressSanitizer.  Test infrastructure only.
//
//   batch_driver chunked   config-3 soils in cache-sized chunks, tuner on
//   batch_driver together  the same soils all advanced together, tuner on
//   batch_driver mixed     scalar SLS, GSLS and elastic models with failing
//                          shots in one batch, then a bank generation
//   batch_driver sim       the step API: create, drive, step, probe, reuse
//   batch_driver tuned     chunk budget and launch shape both tuned (the
//                          trial chunks re-planned between trial launches)
//   batch_driver handle    the batch as a handle: advanced in pieces, reset
//                          half-way and after its end, run again, read back
//
// Every HIP call the library makes is checked by the stub (see there); this
// program checks what comes back through the ABI.  It prints one summary line
// and exits 0, or aborts at the first violation.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"

extern "C" void hipstub_stats(long out[7]);

namespace {

int fails = 0;

void expect(bool ok, const char *what) {
  if (!ok) {
    fprintf(stderr, "batch_driver: %s (last error: %s)\n", what, gf_last_error());
    fails++;
  }
}

struct Shot {
  gf_soil soil;
  gf_model model;
  std::vector<double> force, ranges, traces, snapped;
  int steps;
};

// Config-3 soil `index` as the workload builds it, with `steps` steps.
void c3_shot(Shot &s, uint32_t index, int steps, int variant) {
  memset(&s.model, 0, sizeof s.model);
  expect(gf_soil_sample(0, index, &s.soil) == GF_OK, "gf_soil_sample");
  s.model.stack = s.soil.layers;
  s.model.n_layers = s.soil.n_layers;
  s.model.ref_freq_hz = 30;
  s.model.max_range_m = s.soil.max_range_m;
  s.model.depth_m = s.soil.depth_m;
  s.model.spacing_m = s.soil.spacing_m;
  s.model.pml_cells = s.soil.pml_cells;
  s.model.dominant_freq_hz = 40;
  if (variant == 'X') {
    s.model.gsls_mechanisms = 3;
    s.model.gsls_fmin_hz = 5;
    s.model.gsls_fmax_hz = 150;
  } else if (variant == 'R') {
    expect(gf_visco_fit(s.soil.layers[s.soil.n_layers - 1].qs, 30, &s.model.relax) == GF_OK, "gf_visco_fit");
  } else {  // elastic
    for (int l = 0; l < s.soil.n_layers; l++) s.soil.layers[l].qs = s.soil.layers[l].qp = 0;
  }
  double dt = 0;
  int32_t nr = 0, nz = 0;
  expect(gf_model_grid(&s.model, &dt, &nr, &nz) == GF_OK, "gf_model_grid");
  s.steps = steps;
  s.force.resize(steps);
  expect(gf_ricker(1000, 40, dt, steps, s.force.data()) == GF_OK, "gf_ricker");
  s.ranges = {1.0, 2.0, 5.0, 12.0, 40.0};
}

std::vector<gf_shot> shots_of(std::vector<Shot> &v) {
  std::vector<gf_shot> out(v.size());
  for (size_t i = 0; i < v.size(); i++) {
    Shot &s = v[i];
    s.traces.assign(std::max<size_t>(1, s.ranges.size() * (size_t)std::max(s.steps, 0)), -1.0);
    s.snapped.assign(s.ranges.size(), -1.0);
    out[i] = gf_shot{&s.model, s.force.data(), (int64_t)s.force.size(), s.ranges.data(), (int32_t)s.ranges.size(),
                     s.steps};
  }
  return out;
}

void run_batch(std::vector<Shot> &v, std::vector<int32_t> &status) {
  std::vector<gf_shot> sh = shots_of(v);
  std::vector<gf_result> res(v.size());
  for (size_t i = 0; i < v.size(); i++) {
    memset(&res[i], 0, sizeof res[i]);
    res[i].traces = v[i].traces.data();
    res[i].snapped = v[i].snapped.data();
  }
  status.assign(v.size(), -1);
  gf_run_batch(sh.data(), (int32_t)sh.size(), res.data(), status.data());
  for (size_t i = 0; i < v.size(); i++) expect(res[i].status == status[i], "result status mirrors statuses[]");
}

int scenario_batch(bool chunked) {
  setenv("GEOFDTD_CACHE_MB", chunked ? "60" : "0", 1);
  setenv("GEOFDTD_TUNE", "1", 1);
  std::vector<Shot> v(12);
  for (int i = 0; i < 12; i++) c3_shot(v[i], (uint32_t)i, 230 + 37 * i, 'X');
  std::vector<int32_t> st;
  run_batch(v, st);
  for (size_t i = 0; i < v.size(); i++) expect(st[i] == GF_OK, "config-3 shot runs");
  gf_batch_stats_t bs;
  expect(gf_batch_stats(&bs) == GF_OK && bs.models == 12, "batch stats");
  int64_t want = 0;
  for (auto &s : v) want += s.steps;
  (void)want;
  return 0;
}

int scenario_tuned() {
  unsetenv("GEOFDTD_CACHE_MB");
  unsetenv("GEOFDTD_TILE_ROWS");
  setenv("GEOFDTD_TUNE", "1", 1);
  std::vector<Shot> v(14);
  for (int i = 0; i < 14; i++) c3_shot(v[i], (uint32_t)(80 + i), 140 + 41 * i, i % 4 == 3 ? 'R' : 'X');
  std::vector<int32_t> st;
  run_batch(v, st);
  for (size_t i = 0; i < v.size(); i++) expect(st[i] == GF_OK, "tuned batch shot runs");
  gf_batch_stats_t bs;
  expect(gf_batch_stats(&bs) == GF_OK && bs.budget_mb >= 96 && bs.budget_mb <= 224 &&
             (bs.waves == 1 || bs.waves == 2 || bs.waves == 4),
         "tuned budget and shape reported");
  return 0;
}

int scenario_mixed() {
  setenv("GEOFDTD_CACHE_MB", "40", 1);
  setenv("GEOFDTD_TUNE", "1", 1);
  std::vector<Shot> v(7);
  const int variants[7] = {'X', 'R', 'E', 'X', 'R', 'X', 'E'};
  for (int i = 0; i < 7; i++) c3_shot(v[i], (uint32_t)(20 + i), 120 + 61 * i, variants[i]);
  v[3].ranges.push_back(1e4);  // outside the interior: this shot fails, the others run
  v[5].steps = 0;              // no steps: fails as Run does
  std::vector<int32_t> st;
  run_batch(v, st);
  for (int i = 0; i < 7; i++) {
    const bool bad = i == 3 || i == 5;
    expect(bad ? st[i] == GF_ERR_INVALID : st[i] == GF_OK, "per-shot status");
  }
  // bank generation over the same kind of batch (its own chunking and spectra)
  std::vector<Shot> b(4);
  for (int i = 0; i < 4; i++) c3_shot(b[i], (uint32_t)(40 + i), 200 + 50 * i, i % 2 ? 'R' : 'X');
  std::vector<gf_shot> sh = shots_of(b);
  for (auto &s : sh) s.n_ranges = 0, s.ranges = nullptr;
  gf_bank_spec spec;
  memset(&spec, 0, sizeof spec);
  spec.sample_rate_hz = 500;
  spec.samples = 256;
  spec.max_freq_hz = 100;
  spec.r_min_m = 1;
  spec.r_max_m = 20;
  spec.count = 8;
  const int bins = spec.samples / 2 + 1;
  std::vector<float> spectra((size_t)4 * spec.count * bins * 2);
  std::vector<double> snapped((size_t)4 * spec.count);
  std::vector<gf_result> res(4);
  memset(res.data(), 0, sizeof(gf_result) * 4);
  expect(gf_bank_generate(sh.data(), 4, &spec, spectra.data(), snapped.data(), res.data()) == GF_OK,
         "bank generation");
  return 0;
}

int scenario_sim() {
  std::vector<Shot> v(1);
  c3_shot(v[0], 7, 700, 'X');
  gf_sim *s = nullptr;
  expect(gf_sim_create(&v[0].model, &s) == GF_OK, "gf_sim_create");
  if (!s) return 1;
  expect(gf_sim_drive(s, v[0].force.data(), (int64_t)v[0].force.size()) == GF_OK, "gf_sim_drive");
  std::vector<int32_t> cols = {3, 10, 40};
  std::vector<double> out(cols.size() * 300);
  expect(gf_sim_step_n(s, 300, cols.data(), (int32_t)cols.size(), out.data()) == GF_OK, "gf_sim_step_n");
  expect(gf_sim_step_n(s, 150, cols.data(), 2, out.data()) == GF_OK, "gf_sim_step_n, fewer receivers");
  double vz = 0, speed = 0;
  expect(gf_sim_surface_velocity(s, 5, &vz) == GF_OK, "gf_sim_surface_velocity");
  expect(gf_sim_max_speed(s, &speed) == GF_OK, "gf_sim_max_speed");
  int32_t th = 0;
  expect(gf_sim_autotune(s, 4, &th) == GF_OK && th > 0, "gf_sim_autotune");
  expect(gf_sim_step_n(s, 250, cols.data(), (int32_t)cols.size(), out.data()) == GF_OK, "gf_sim_step_n after tune");
  gf_sim_destroy(s);
  // one shot through gf_run
  std::vector<gf_shot> sh = shots_of(v);
  gf_result r;
  memset(&r, 0, sizeof r);
  r.traces = v[0].traces.data();
  r.snapped = v[0].snapped.data();
  expect(gf_run(&sh[0], &r) == GF_OK, "gf_run");
  return 0;
}

int scenario_handle() {
  setenv("GEOFDTD_CACHE_MB", "50", 1);
  setenv("GEOFDTD_TUNE", "1", 1);
  std::vector<Shot> v(9);
  for (int i = 0; i < 9; i++) c3_shot(v[i], (uint32_t)(60 + i), 180 + 53 * i, i % 3 == 2 ? 'R' : 'X');
  std::vector<gf_shot> sh = shots_of(v);
  gf_batch *b = nullptr;
  expect(gf_batch_create(sh.data(), (int32_t)sh.size(), nullptr, &b) == GF_OK && b, "gf_batch_create");
  if (!b) return 1;
  std::vector<gf_result> res(v.size());
  for (size_t i = 0; i < v.size(); i++) {
    memset(&res[i], 0, sizeof res[i]);
    res[i].traces = v[i].traces.data();
    res[i].snapped = v[i].snapped.data();
  }
  std::vector<int32_t> st(v.size());
  expect(gf_batch_advance(b, 1, 3) == GF_OK, "advance to a third");
  expect(gf_batch_results(b, res.data(), st.data()) != GF_OK, "results refused before the end");
  expect(gf_batch_reset(b) == GF_OK, "reset half-way");
  for (int k = 1; k <= 7; k++) expect(gf_batch_advance(b, k, 7) == GF_OK, "advance in sevenths");
  expect(gf_batch_advance(b, 1, 1) == GF_OK, "advance at the end is a no-op");
  expect(gf_batch_results(b, res.data(), st.data()) == GF_OK, "results");
  expect(gf_batch_reset(b) == GF_OK, "reset after the end");
  expect(gf_batch_advance(b, 1, 1) == GF_OK, "second run");
  expect(gf_batch_results(b, res.data(), st.data()) == GF_OK, "results of the second run");
  gf_batch_destroy(b);
  // with a bank spec
  gf_bank_spec spec;
  memset(&spec, 0, sizeof spec);
  spec.sample_rate_hz = 500;
  spec.samples = 256;
  spec.max_freq_hz = 100;
  spec.r_min_m = 1;
  spec.r_max_m = 20;
  spec.count = 8;
  for (auto &s : sh) s.n_ranges = 0, s.ranges = nullptr;
  expect(gf_batch_create(sh.data(), (int32_t)sh.size(), &spec, &b) == GF_OK && b, "gf_batch_create with a bank");
  if (!b) return 1;
  const int bins = spec.samples / 2 + 1;
  std::vector<float> spectra((size_t)v.size() * spec.count * bins * 2);
  std::vector<double> snapped((size_t)v.size() * spec.count);
  for (int k = 1; k <= 4; k++) expect(gf_batch_advance(b, k, 4) == GF_OK, "bank batch in quarters");
  expect(gf_batch_spectra(b, spectra.data(), snapped.data(), res.data()) == GF_OK, "bank spectra");
  expect(gf_batch_reset(b) == GF_OK, "bank batch reset");
  expect(gf_batch_advance(b, 1, 1) == GF_OK, "bank batch again");
  expect(gf_batch_spectra(b, spectra.data(), snapped.data(), res.data()) == GF_OK, "bank spectra again");
  gf_batch_destroy(b);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const std::string what = argc > 1 ? argv[1] : "chunked";
  if (what == "chunked") scenario_batch(true);
  else if (what == "together") scenario_batch(false);
  else if (what == "mixed") scenario_mixed();
  else if (what == "sim") scenario_sim();
  else if (what == "handle") scenario_handle();
  else if (what == "tuned") scenario_tuned();
  else {
    fprintf(stderr, "unknown scenario %s\n", what.c_str());
    return 2;
  }
  long s[7];
  hipstub_stats(s);
  printf("batch_driver %s: %s; %ld allocations, %ld frees, %ld live, %ld copies, %ld kernel launches checked, "
         "%ld graph replays, %ld synchronizations\n",
         what.c_str(), fails ? "FAILED" : "clean", s[0], s[1], s[6], s[2], s[3], s[4], s[5]);
  return fails ? 1 : 0;
}
