/*
 * geofdtd.h — C ABI of libgeofdtd, the MI355X (gfx950) axisymmetric
 * viscoelastic FDTD solver.
 *
 * The entry points mirror the Go package internal/fdtd of the reference
 * (Ryan-Hancock/Co-Simulation-of-Seismic-Sensors-and-ROS-for-Robotic-Human-Detection)
 * one for one, so that a cgo shim can keep that package's API unchanged:
 * each declaration below names the Go symbol it replaces.
 *
 * Conventions
 *   - Every function returns an int status (GF_OK == 0).  On failure,
 *     gf_last_error() returns a thread-local message with the same text the
 *     Go code would have put in its error value.
 *   - All buffers are caller-owned.  The library copies what it needs at call
 *     time and keeps no caller pointer after returning (the cgo pointer rule).
 *   - Handles are opaque and single-threaded, like *fdtd.Sim; distinct handles
 *     may be used from distinct threads.  Calls block the calling thread.
 *   - Results leave the library as float64, like the reference's
 *     []units.Velocity; the GPU computes in float32 with float32 field state.
 */
#ifndef GEOFDTD_H
#define GEOFDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GF_ABI_VERSION 1

enum gf_status {
  GF_OK = 0,
  GF_ERR_INVALID = 1,  /* a model, shot or stack the reference rejects     */
  GF_ERR_DIVERGED = 2, /* "fdtd: diverged at step %d of %d" (run.go:118-124) */
  GF_ERR_DEVICE = 3,   /* HIP runtime failure, or no gfx950 device          */
  GF_ERR_ARGUMENT = 4  /* null pointer, negative size, bad handle           */
};

/* layer.Layer (internal/layer/layer.go:24-36).  The last entry of a stack is
 * the half-space and its thickness is ignored.  qp is an extension used only
 * by the GSLS attenuation model (gf_model.gsls_mechanisms > 0); zero there
 * means qp == qs.  qs == 0 means lossless, as in the reference. */
typedef struct gf_layer {
  double thickness_m;
  double vp_mps;
  double vs_mps;
  double density_kgm3;
  double qs;
  double qp;
} gf_layer;

/* visco.SLS (internal/visco/visco.go:42-47). */
typedef struct gf_sls {
  double tau_sigma;
  double tau_eps;
} gf_sls;

/* fdtd.Model (internal/fdtd/model.go:37-76).  Zero values select the same
 * defaults as the reference: ref_freq 30 Hz, pml_cells 20, pml_reflection
 * 1e-4, courant 0.5, dominant_freq "middle of the resolved band".
 *
 * Extension (not in the reference; zero reproduces the reference exactly):
 *   gsls_mechanisms  0: one scalar SLS `relax` over the whole model, Qp = Qs
 *                       (the reference physics).
 *                    N in [1, 8]: per-layer attenuation from gf_layer.qs/qp,
 *                       N relaxation mechanisms fitted to constant Q over
 *                       [gsls_fmin_hz, gsls_fmax_hz] (default 5-150 Hz),
 *                       separate P and S relaxation.  `relax` is ignored. */
typedef struct gf_model {
  const gf_layer *stack;
  int32_t n_layers;
  gf_sls relax;
  double ref_freq_hz;
  double max_range_m;
  double depth_m;
  double spacing_m;
  int32_t pml_cells;
  double pml_reflection;
  double dominant_freq_hz;
  double courant;
  int32_t gsls_mechanisms;
  double gsls_fmin_hz;
  double gsls_fmax_hz;
} gf_model;

/* fdtd.Shot (internal/fdtd/run.go:49-61), minus Progress: a caller that wants
 * progress drives gf_sim_step_n in blocks and reports between them. */
typedef struct gf_shot {
  const gf_model *model;
  const double *force; /* one sample per step; samples past n_force are 0 */
  int64_t n_force;
  const double *ranges; /* receiver ranges, m */
  int32_t n_ranges;
  int32_t steps;
} gf_shot;

/* fdtd.Result (internal/fdtd/run.go:72-86).  traces and snapped are
 * caller-owned: traces is [n_ranges][steps] row-major, snapped [n_ranges].
 * status and error carry this shot's outcome (the Go error text), so that a
 * batch can report each shot's error on its own. */
typedef struct gf_result {
  double dt;
  double t0;
  int32_t nr;
  int32_t nz;
  double peak;
  double *traces;
  double *snapped;
  int32_t status;
  char error[244];
} gf_result;

/* Last error message of the calling thread ("" after a success). */
const char *gf_last_error(void);
int gf_abi_version(void);

/* visco.Fit (internal/visco/visco.go:55-68). */
int gf_visco_fit(double q, double f0_hz, gf_sls *out);

/* fdtd.Model.SpacingFor (internal/fdtd/model.go:121-136). */
int gf_spacing_for(const gf_model *m, double max_freq_hz,
                   double cells_per_wavelength, double *out_spacing_m);

/* fdtd.Ricker and fdtd.RickerSpectrum (internal/fdtd/run.go:19-46). */
int gf_ricker(double peak_n, double f0_hz, double dt, int32_t n, double *out);
int gf_ricker_spectrum(double peak_n, double f0_hz, double f_hz, double *re,
                       double *im);

/* The grid fdtd.New would build for a model (fdtd.go:45-64: validation,
 * nr, nz, dt) without allocating anything or touching a device. */
int gf_model_grid(const gf_model *m, double *dt, int32_t *nr, int32_t *nz);

/* layer.Stack.Validate (internal/layer/layer.go:53-73): GF_ERR_INVALID with
 * the reference's message ("layer 0: Vs must be positive, got -1 m/s", ...)
 * for a physically impossible stack.  What layer.LoadGolden checks before a
 * model exists (cmd/geofdtd reads its medium that way). */
int gf_stack_validate(const gf_layer *stack, int32_t n_layers);

/* ---- *fdtd.Sim ------------------------------------------------------- */
typedef struct gf_sim gf_sim;

/* fdtd.New (internal/fdtd/fdtd.go:45-97).  Allocates the grid in HBM and
 * builds the coefficient tables in FP64 on the device. */
int gf_sim_create(const gf_model *m, gf_sim **out);
void gf_sim_destroy(gf_sim *s);

/* Return a simulation to the state New leaves it in (all fields, memory
 * variables and C-PML psi zero, step 0), keeping its allocation, tables,
 * source and receivers.  No reference counterpart: it lets a caller run many
 * shots of one model without re-allocating, as a benchmark does. */
int gf_sim_reset(gf_sim *s);

/* Sim.Dt and Sim.Cells (fdtd.go:118-121). */
int gf_sim_info(const gf_sim *s, double *dt, int32_t *nr, int32_t *nz);
/* Sim.Time (fdtd.go:125). */
int gf_sim_time(const gf_sim *s, double *out);
/* Sim.Drive (fdtd.go:129).  Copies the series (the Go version retains it). */
int gf_sim_drive(gf_sim *s, const double *force, int64_t n);
/* Sim.Column (fdtd.go:142-146). */
int gf_sim_column(const gf_sim *s, double r_m, int32_t *col, double *snapped_m);
/* Sim.Step (fdtd.go:149-156). */
int gf_sim_step(gf_sim *s);
/* n calls of Sim.Step, recording Sim.SurfaceVelocity at each of cols after
 * every step into out[c * n + k] (out may be NULL when ncols == 0). */
int gf_sim_step_n(gf_sim *s, int32_t n, const int32_t *cols, int32_t ncols,
                  double *out);
/* Sim.SurfaceVelocity, Sim.RadialVelocity, Sim.MaxSpeed (fdtd.go:330-360). */
int gf_sim_surface_velocity(gf_sim *s, int32_t col, double *out);
int gf_sim_radial_velocity(gf_sim *s, int32_t col, double *out);
int gf_sim_max_speed(gf_sim *s, double *out);

/* ---- fdtd.Run ------------------------------------------------------- */

/* fdtd.Run (internal/fdtd/run.go:88-127): one shot, start to finish. */
int gf_run(const gf_shot *shot, gf_result *res);

/* Many independent shots advanced together on one GPU (no reference
 * counterpart as an API; it replaces running fdtd.Run once per shot in a
 * worker pool, cmd/geobank/main.go:128-167).  Each result equals what gf_run
 * returns for that shot.  A shot that diverges reports GF_ERR_DIVERGED in
 * statuses[i] and does not stop the others; the call returns GF_OK if the
 * batch ran, and the first per-shot error otherwise. */
int gf_run_batch(const gf_shot *shots, int32_t n_shots, gf_result *results,
                 int32_t *statuses);

/* Timing of the calling thread's last gf_run_batch, gf_bank_generate or
 * gf_batch_advance (since the batch's last reset): the device time of its
 * step loop (HIP events around it on the solver's stream), the step kernels
 * launched, the algorithmic HBM bytes of every step of every model (each
 * state array read and written once per step), the launch shape (tile
 * height, tiles per workgroup, time steps per launch), the models run, the
 * chunk budget in MB (0: every model advanced together), and the bytes of
 * one read and one write of every running model's state per launch (equal
 * to algorithmic_bytes with one step per launch, about half of it with
 * two). */
typedef struct gf_batch_stats_t {
  double kernel_ms;
  int64_t step_launches;
  double algorithmic_bytes;
  int32_t tile_rows;
  int32_t models;
  int32_t waves;
  int32_t steps_per_launch;
  double budget_mb;
  double pass_bytes;
} gf_batch_stats_t;

int gf_batch_stats(gf_batch_stats_t *out);

/* ---- device-resident benchmarking -------------------------------------
 * The timed region of a benchmark must start with inputs already in HBM, so
 * these split gf_run into set-up, advance and read-back.  Receivers record
 * into a device buffer of capacity_steps samples each. */
int gf_sim_set_receivers(gf_sim *s, const int32_t *cols, int32_t ncols,
                         int32_t capacity_steps);
/* Advance n steps on the device without host synchronisation per step.
 * use_graph != 0 replays the steps from captured HIP graphs. */
int gf_sim_advance(gf_sim *s, int32_t n, int32_t use_graph);
/* Wait for the device, then report the elapsed time of every step kernel
 * launched since the last call, measured with HIP events on the solver's
 * stream, and how many step kernels that covers. */
int gf_sim_sync_timing(gf_sim *s, double *kernel_ms, int64_t *kernel_launches);
/* Copy the recorded receiver samples out: out[c * steps_recorded + k]. */
int gf_sim_read_traces(gf_sim *s, double *out, int32_t *steps_recorded);
/* Tile shape.  The step kernel's tile height (grid rows per wavefront) is
 * picked by timing each candidate on this model for trial_steps steps; the
 * simulation is then reset to step 0 (zero state, cleared receivers).
 * gf_sim_set_tile_rows fixes it by hand (1, 2, 4, 8 select unrolled kernels,
 * other values a row-marching kernel).  No reference counterpart. */
int gf_sim_autotune(gf_sim *s, int32_t trial_steps, int32_t *tile_rows);
int gf_sim_set_tile_rows(gf_sim *s, int32_t tile_rows);
/* Time steps per kernel launch (1: one step per launch; 2: the temporally
 * blocked kernel, which reads and writes the state once per two steps) with
 * its tile height, and the shape in use (also after gf_sim_autotune, which
 * tries both).  Results do not depend on the shape beyond FP32 rounding of
 * contracted products.  No reference counterpart. */
int gf_sim_set_launch(gf_sim *s, int32_t steps_per_launch, int32_t tile_rows);
int gf_sim_launch_shape(const gf_sim *s, int32_t *steps_per_launch, int32_t *tile_rows,
                        int32_t *waves);
/* Algorithmic HBM bytes one step moves (each state array read and written
 * once, plus the C-PML strips): the roofline numerator. */
int gf_sim_step_bytes(const gf_sim *s, double *bytes);

/* ---- Green's-function banks -------------------------------------------
 * bank.Header of internal/bank/format.go:115-131 as a flat struct.  Strings
 * are NUL-terminated; config_hash, created and notes are omitted from the
 * file when NULL or empty (their omitempty tags), solver never is. */
typedef struct gf_bank_header {
  const char *solver;
  const char *config_hash;
  const char *created;
  const char *notes;
  const gf_layer *medium; /* written as layer.Layer JSON; NULL/0 writes null */
  int32_t n_medium;
  double sample_rate_hz; /* bin k is at k * sample_rate_hz / samples */
  int32_t samples;       /* positive and even; bins = samples / 2 + 1 */
  double r_min_m;        /* bank.RangeGrid: count ranges, min to max */
  double r_max_m;
  int32_t count;
  const char *component;
  const char *units;
} gf_bank_header;

/* bank.Write (internal/bank/format.go:215-240): format v1, the header as Go's
 * encoding/json writes it, payload data[count][samples/2+1][2] as
 * little-endian float32 (re, im), range-major.  Validates as
 * bank.Header.Validate (format.go:142-165), with its messages. */
int gf_bank_write(const char *path, const gf_bank_header *h, const float *data);
/* The header JSON gf_bank_write would write (for inspection and tests). */
int gf_bank_header_json(const gf_bank_header *h, char *out, int32_t cap);

/* The bank grid a generated bank is reduced onto. */
typedef struct gf_bank_spec {
  double sample_rate_hz;
  int32_t samples;
  double max_freq_hz; /* bins above it, and bin 0, are left zero */
  double r_min_m;
  double r_max_m;
  int32_t count;
} gf_bank_spec;

/* Fill banks with the FDTD: one shot per soil model, advanced together as in
 * gf_run_batch, with receivers at the spec's range grid (shot ranges are
 * ignored).  Each shot's traces stay in HBM and are reduced on the device to
 * G(f) = V(f) / (i w W(f)) -- vertical surface displacement per unit vertical
 * point force, m/N, positive down -- at bins 1..max_freq of the spec's
 * frequency grid.  spectra is [n_shots][count][samples/2+1][2] float32;
 * snapped [n_shots][count] receives the ranges the receivers sat at (grid
 * columns); results (may be NULL) receive per-shot dt, peak, status, error.
 * No reference counterpart: the reference fills banks by wavenumber
 * integration per bin (cmd/geobank/main.go:128-167); this is the same file
 * filled by the FDTD. */
int gf_bank_generate(const gf_shot *shots, int32_t n_shots, const gf_bank_spec *spec, float *spectra,
                     double *snapped, gf_result *results);

/* The same batch as a handle, so a caller can run it in pieces: the bank
 * generator's job loop (cmd/geobank/main.go:128-167, one fdtd.Run per soil
 * in a worker pool) split into set-up, advance and read-back.  No single
 * reference function.
 *   gf_batch_create  sets every shot up (fdtd.New's checks and errors per
 *                    shot, as gf_run_batch) and plans the launches: models
 *                    grouped by kernel variant into chunks whose state fits
 *                    the last-level cache, tile heights tuned on long runs.
 *                    With spec != NULL the receivers sit on the bank's range
 *                    grid (shot ranges ignored) and gf_batch_spectra applies.
 *   gf_batch_advance advances every chunk to ceil(its longest run x num /
 *                    den) steps (num >= den: to the end), chunk after chunk;
 *                    gf_batch_stats then reports the work since the last
 *                    reset.
 *   gf_batch_reset   returns every model to step 0 (zero state, cleared
 *                    divergence probes) to run the batch again.
 *   gf_batch_results / gf_batch_spectra  once the batch has run to its end:
 *                    what gf_run_batch / gf_bank_generate return. */
typedef struct gf_batch gf_batch;
int gf_batch_create(const gf_shot *shots, int32_t n_shots,
                    const gf_bank_spec *spec, gf_batch **out);
int gf_batch_advance(gf_batch *b, int32_t num, int32_t den);
int gf_batch_reset(gf_batch *b);
int gf_batch_results(gf_batch *b, gf_result *results, int32_t *statuses);
int gf_batch_spectra(gf_batch *b, float *spectra, double *snapped,
                     gf_result *results);
void gf_batch_destroy(gf_batch *b);

/* The reduction gf_bank_generate applies, on caller traces: for each of
 * `rows` traces of n samples at t0 + k dt (float32, row-major) and the force
 * series that drove them (F[k] at k dt), G(f) = V(f) / (i w W(f)) at each of
 * nbins frequencies, out[row][bin][2] in FP64.  Runs on the device. */
int gf_spectrum(const float *traces, int32_t rows, int32_t n, double dt, double t0, const double *force,
                int64_t n_force, const double *freqs, int32_t nbins, double *out);

/* ---- bank lookup: bank.Bank's interpolation on the device --------------
 * (internal/bank/interp.go).  The payload is copied to HBM once and the
 * log-magnitude and range-unwrapped phase tables of Bank.prepare are built
 * there (one thread per bin walking the range axis); lookups run one thread
 * per output value.  A handle is single-threaded, like *bank.Bank's lazily
 * built tables. */
typedef struct gf_bank_dev gf_bank_dev;

/* bank.New and Set of a whole payload [count][samples/2+1][2] float32,
 * range-major (format.go:167-194), then Bank.prepare (interp.go:64-101).
 * The header is validated as bank.Header.Validate (format.go:142-165). */
int gf_bank_dev_open(const gf_bank_header *h, const float *payload, gf_bank_dev **out);
void gf_bank_dev_close(gf_bank_dev *b);

/* Bank.Response (interp.go:29-54) at n points (r_m[i], bin[i]): out [n][2]
 * (re, im), linear in log-magnitude and unwrapped phase between the two
 * bracketing grid ranges.  The first point whose Go call would fail sets the
 * error ("bank: bin %d out of range [0, %d)", "bank: range %g m is outside
 * the bank's %g to %g m") and nothing is written. */
int gf_bank_dev_response(gf_bank_dev *b, const double *r_m, const int32_t *bin, int64_t n, double *out);

/* Bank.Response at every bin for each of n ranges: out [n][samples/2+1][2],
 * the spectra a synthesis at those ranges reads. */
int gf_bank_dev_spectra(gf_bank_dev *b, const double *r_m, int32_t n, double *out);

/* The impulse responses runtime synthesis convolves with, from the bank:
 * sensing.FromBank's VerticalVelocityResponse (propagation.go:89-107: i w
 * times Bank.Response on the bank's own frequency grid) through
 * dsp.CausalImpulseResponseAt with the bank's transform size
 * (dsp.go:232-260: taper over the top fifth of the band, inverse real
 * transform, first `length` samples).  sensor (may be NULL: 1) multiplies
 * each bin, [samples/2+1][2], as the engine multiplies by the geophone's
 * response.  fs_hz must put every bin on the bank's grid.  out [n][length];
 * the first range whose Go call fails sets the error. */
int gf_bank_dev_impulse(gf_bank_dev *b, const double *r_m, int32_t n, int32_t length, double fs_hz,
                        const double *sensor, double *out);

/* Kernel time (HIP events) and algorithmic HBM bytes of the last lookup on
 * this handle: table reads (two range rows of log-magnitude and phase per
 * value, FP64) plus the FP64 complex values written. */
int gf_bank_dev_stats(const gf_bank_dev *b, double *kernel_ms, double *bytes);

/* bank.RangeNyquist and Header.CheckRangeSampling (interp.go:120-141):
 * whether the range grid is fine enough to unwrap phase up to max_freq_hz
 * for the header's medium. */
double gf_range_nyquist(double slowest_velocity_mps, double max_freq_hz);
int gf_bank_check_range_sampling(const gf_bank_header *h, double max_freq_hz);

/* ---- streaming convolution (the Go package internal/stream) ------------
 * stream.Convolver for many receivers at once: each receiver convolves its
 * own impulse response with its input, delivered a chunk at a time, and the
 * output equals the whole-record convolution truncated to the input length.
 * On the device the state is a ring of each receiver's past input in HBM and
 * each output sample is the direct sum over the response (LDS-tiled); the
 * reference's partitioned overlap-add is the same linear map, so results
 * agree to rounding (the reference's own bound: 1e-9 of the peak). */
typedef struct gf_convolver gf_convolver;

/* stream.NewConvolver (stream.go:72-107): impulse [n_recv][len] (one
 * response per receiver, all the same length), chunk samples per Process. */
int gf_convolver_create(const double *impulse, int32_t n_recv, int32_t len, int32_t chunk, gf_convolver **out);
void gf_convolver_destroy(gf_convolver *c);

/* ChunkSize and Partitions (stream.go:109-112), and the receiver count. */
int gf_convolver_info(const gf_convolver *c, int32_t *chunk, int32_t *partitions, int32_t *n_recv);

/* Convolver.Process (stream.go:120-162) for every receiver: in and out are
 * [n_recv][len_in] and [n_recv][len_out]; both lengths must equal the chunk
 * size ("stream: chunk length mismatch: in %d, out %d, want %d"); in and out
 * may alias. */
int gf_convolver_process(gf_convolver *c, const double *in, int32_t len_in, double *out, int32_t len_out);

/* Convolver.Reset (stream.go:164-171). */
int gf_convolver_reset(gf_convolver *c);

/* Convolver.ProcessAll (stream.go:175-186) for every receiver: in and out
 * [n_recv][n]; the input is padded to whole chunks, processed from the
 * current state in one pass on the device, and the state left as after those
 * chunks. */
int gf_convolver_process_all(gf_convolver *c, const double *in, int64_t n, double *out);

/* Kernel time (HIP events) of the last Process or ProcessAll. */
int gf_convolver_stats(const gf_convolver *c, double *kernel_ms);

/* ---- runtime synthesis from a bank (the Go package internal/sensing) ----
 * sensing.Engine over sensing.FromBank, for n_recv receivers at once (the
 * reference runs one Engine per receiver, at the origin).  Each contact gets
 * a voice: its vertical force convolved with the impulse response at its
 * range, rounded to RangeQuantum = 0.02 m (sensing.go:49, 178); the bank
 * carries no radial-force response (propagation.go:110-112), so the
 * reference's second convolver per voice adds exactly zero and is not run.
 * The schedule (source.Walk and its contacts) stays with the caller, which
 * warms the responses, admits each contact as its chunk begins and samples
 * its force; the engine keeps the responses and voices in HBM and mixes a
 * chunk for every receiver in one launch.  A handle is single-threaded, and
 * the bank handle must outlive it. */
typedef struct gf_engine gf_engine;

/* NewEngineWith(res, sch, chunk, FromBank(bank)) (sensing.go:117-170):
 * receivers at recv_xy [n_recv][2] (m), chunk samples per Next at fs_hz (the
 * resolved sampling rate), ir_len = ceil((far / c_R + 2) fs) as the Go code
 * derives it from the schedule's furthest range, capped here at three
 * quarters of the bank's transform (sensing.go:141-148); sensor (may be
 * NULL: 1) the sensor's response at the bank's bins, [samples/2+1][2].
 * max_contact_samples bounds one voice's force record.  Errors: "sensing:
 * chunk size must be positive, got %d", "sensing: bank is nil". */
int gf_engine_create(const gf_bank_dev *bank, int32_t n_recv, const double *recv_xy, int32_t chunk, double fs_hz,
                     int32_t ir_len, const double *sensor, int32_t max_contact_samples, gf_engine **out);
void gf_engine_destroy(gf_engine *e);

/* Engine.warm (sensing.go:176-184): the responses for contacts at
 * contact_xy [n][2], for every receiver, built in one batch; the first range
 * whose Go call would fail sets its error (the bank's range check, the
 * frequency-grid check). */
int gf_engine_warm(gf_engine *e, const double *contact_xy, int32_t n);

/* Engine.admit for one contact (sensing.go:295-330): a voice starting at the
 * next chunk, whose vertical force is force_z[m] at sample emitted + m
 * (m < n; zero after), running until ceil(end_s fs) + ir_len.  Errors:
 * "sensing: contact %q at the receiver's own position"; at most 64 voices
 * live at once (the reference has no bound; a walker keeps about four):
 * "geofdtd: more than 64 voices at once", the engine left as it was. */
int gf_engine_admit(gf_engine *e, const char *id, double x, double y, const double *force_z, int32_t n,
                    double end_s);

/* Engine.Next (sensing.go:246-289): out [n_recv][chunk] float32, the sum of
 * the active voices; then retire and advance.  len must equal the chunk
 * ("sensing: output length %d, want %d"). */
int gf_engine_next(gf_engine *e, float *out, int32_t len);

/* soil.HalfSpace (internal/soil/soil.go:25-40): a homogeneous half-space.
 * qp 0 means 9/4 qs; it is validated but not otherwise read. */
typedef struct gf_halfspace {
  double vp_mps;
  double vs_mps;
  double density_kgm3;
  double qs;
  double qp;
} gf_halfspace;

/* soil.HalfSpace.RayleighVelocity (soil.go:136-163): the Rayleigh root by
 * bisection, after Validate (soil.go:72-89, the same texts). */
int gf_rayleigh_velocity(const gf_halfspace *h, double *out_mps);

/* NewEngine (sensing.go:110-113): the engine over sensing.Analytic
 * (propagation.go:38-62), green.HalfSpaceGF{Soil: *h, RefFreq: ref_freq_hz}
 * (0: 30 Hz), whose responses carry the radial-force component too
 * (green.go:375-388), so each voice convolves both its vertical force and its
 * force along the contact-to-receiver line.  The responses are built on the
 * device at the transform size CausalImpulseResponseAt picks when the
 * propagation names none, nfft = NextPow2(4 ir_len) (dsp.go:240-242), with no
 * cap on ir_len; sensor (may be NULL) at that transform's bins,
 * [nfft/2+1][2].  Voices are admitted with gf_engine_admit_xyz. */
int gf_engine_create_analytic(const gf_halfspace *h, double ref_freq_hz, int32_t n_recv, const double *recv_xy,
                              int32_t chunk, double fs_hz, int32_t ir_len, const double *sensor,
                              int32_t max_contact_samples, gf_engine **out);

/* Engine.admit for one contact of an analytic engine: the contact's force
 * (fx, fy, fz)[m] at sample emitted + m, m < n. */
int gf_engine_admit_xyz(gf_engine *e, const char *id, double x, double y, const double *fx, const double *fy,
                        const double *fz, int32_t n, double end_s);

/* The transform size of the engine's responses (the bank's samples, or the
 * analytic engine's nfft). */
int gf_engine_transform(const gf_engine *e, int32_t *nfft);

/* Samples emitted (StartTime x fs), ActiveVoices, responses cached, and the
 * impulse-response length in use. */
int gf_engine_info(const gf_engine *e, int64_t *emitted, int32_t *active_voices, int32_t *responses,
                   int32_t *ir_len);

/* Kernel time (HIP events) of the last Next's mix and of every mix so far,
 * the multiply-adds those mixes did (per receiver, output and voice, the taps
 * where force and response are both defined), the total time of response
 * builds, and responses built. */
int gf_engine_stats(const gf_engine *e, double *mix_ms, double *mix_ms_total, int64_t *macs, double *build_ms,
                    int64_t *built);

/* ---- seeded soil models (batched bank generation) ----------------------
 * Soil model `index` of the stream keyed by `seed`, drawn with Philox4x32-10
 * (Salmon et al., SC'11): 1-3 layers over a half-space, each with Vs
 * U[120,400] m/s, Vp/Vs U[1.7,2.5], rho U[1600,2100] kg/m^3, Qs U[10,80],
 * Qp = 2 Qs, thickness U[0.5,5] m; cell size h = 0.9 min(Vs) / (100 Hz * 10)
 * (ten cells per shortest Rayleigh wavelength at 100 Hz); a 20-cell C-PML;
 * an interior of about 218 m by 108 m so that the grid is 600-2048 by
 * 300-1024 cells.  The same (seed, index) always gives the same model.
 * No reference counterpart (the reference's sweep axes, internal/sweep,
 * sample parameters of its analytic model, not grids). */
typedef struct gf_soil {
  gf_layer layers[4];
  int32_t n_layers;
  double spacing_m;
  double max_range_m;
  double depth_m;
  int32_t pml_cells;
} gf_soil;

int gf_soil_sample(uint64_t seed, uint32_t index, gf_soil *out);

/* Soil model `index` of the layered family the surrogate is trained on,
 * drawn around config 2 (two layers over a half-space): Vs1 U[120,200],
 * Vs2 U[200,320], Vs3 U[320,480] m/s; Qs1 U[10,40], Qs2 U[20,60], Qs3
 * U[40,100], Qp = 2 Qs; thicknesses U[0.5,3] and U[2,6] m; one Vp/Vs U[1.7,
 * 2.5]; densities rho0, rho0+150, rho0+300 with rho0 U[1600,1800]; h = 0.9
 * Vs1 / (100 Hz * 10); an interior of 80 m by 30 m (receivers out to 50 m
 * stay 30 m from the C-PML) and a 20-cell C-PML.
 * params receives the 10 numbers that define the model, in the order
 * {Vs1, Vs2, Vs3, Qs1, Qs2, Qs3, h1, h2, Vp/Vs, rho0}.  Philox counter word 2
 * is 1, so these draws never coincide with gf_soil_sample's.  No reference
 * counterpart. */
int gf_soil_sample_layered(uint64_t seed, uint32_t index, gf_soil *out, double params[10]);

/* ---- the Green's-function surrogate -----------------------------------
 * No reference counterpart: a network standing in for the bank lookup,
 * trained on spectra the batched FDTD generates, with the FDTD as its
 * fallback.  Inputs per query: 15 features (the 10 numbers of a layered soil,
 * the log range, the range, and the shear travel times r / Vs of the three
 * layers, each scaled to about [-1, 1] by the caller).  Output
 * per query: 512 complex bins, either in the normalized domain the network
 * is trained in (gf_sur_forward) or as a physical spectrum (gf_sur_predict:
 * times scale[q] and e^{-i 2 pi f delay[q]}, bin k at k * bin_hz).
 * Network: an encoder of six 512-wide GELU layers; a decoder of a dense
 * projection to 64 positions x 32 channels and two transposed convolutions
 * (kernel = stride) to 512 positions x 2; every product on BF16 MFMA with
 * FP32 accumulation;
 * FP32 master weights and Adam.  The loss is the relative spectral L2,
 * mean over queries of |o - t|^2 / |t|^2. */
typedef struct gf_sur gf_sur;

typedef struct gf_sur_config {
  int32_t max_batch; /* queries per launch chain; a multiple of 8, at most 16384 */
  uint64_t seed;     /* weight initialisation */
  double band_lo_hz; /* gf_sur_predict leaves bins outside [lo, hi] zero; */
  double band_hi_hz; /* hi <= 0 keeps every bin */
  int32_t polar;     /* outputs per bin: 0 (re, im); 1 (log-amplitude, phase) */
} gf_sur_config;

/* One GEMM layer: weights [n_out][k_in] at w_offset and bias_channels
 * biases at b_offset in the flat FP32 parameter vector; rows_per_query GEMM
 * rows per query (decoder positions); gelu 0 for the linear output layer. */
typedef struct gf_sur_layer_t {
  int32_t k_in, n_out, gelu, rows_per_query, bias_channels;
  int64_t w_offset, b_offset;
} gf_sur_layer_t;

typedef struct gf_sur_train_opts {
  int32_t steps;
  int32_t batch;        /* a multiple of 8, at most max_batch */
  double lr, lr_min;    /* cosine decay from lr to lr_min */
  double beta1, beta2, eps; /* Adam; zero selects 0.9, 0.999, 1e-8 */
  uint64_t seed;        /* minibatch order */
} gf_sur_train_opts;

/* Device time of the calling handle's last forward/predict (HIP events
 * around its kernels, inputs already in HBM) and of its last training run
 * (HIP events from its first step's kernels to its last, the training set
 * already in HBM). */
typedef struct gf_sur_stats_t {
  double forward_ms;
  int64_t forward_queries;
  double train_ms;
  int64_t train_steps;
  int64_t train_queries;
  double flops_per_query; /* forward GEMM flops, algorithmic */
} gf_sur_stats_t;

int gf_sur_create(const gf_sur_config *c, gf_sur **out);
void gf_sur_destroy(gf_sur *h);
int gf_sur_layers(const gf_sur *h, int32_t *n);
int gf_sur_layer(const gf_sur *h, int32_t l, gf_sur_layer_t *out);
int gf_sur_param_count(const gf_sur *h, int64_t *n);
int gf_sur_get_params(gf_sur *h, float *out);
/* Replace the parameters (and reset the Adam state). */
int gf_sur_set_params(gf_sur *h, const float *in);
/* x [n][15] -> out [n][1024], the normalized output. */
int gf_sur_forward(gf_sur *h, const float *x, int32_t n, float *out);
/* x [n][15] -> out [n][512][2], physical spectra. */
int gf_sur_predict(gf_sur *h, const float *x, const double *scale, const double *delay, double bin_hz, int32_t n,
                   float *out);
/* Mean loss over n queries (a multiple of 8, at most max_batch) against
 * targets t [n][1024], and its gradient (flat, like the parameters).  w NULL:
 * relative L2 |o - t|^2 / |t|^2 per query; otherwise the weighted squares
 * sum_i w_i (o_i - t_i)^2 with weights w [n][1024]. */
int gf_sur_loss_grad(gf_sur *h, const float *x, const float *t, const float *w, int32_t n, double *loss,
                     float *grad);
/* opts->steps Adam steps on minibatches of the n samples, with the loss of
 * gf_sur_loss_grad (w may be NULL); losses (may be NULL) receives each step's
 * mean loss. */
int gf_sur_train(gf_sur *h, const float *x, const float *t, const float *w, int64_t n, const gf_sur_train_opts *opts,
                 float *losses);
int gf_sur_stats(const gf_sur *h, gf_sur_stats_t *out);

/* The surrogate's GEMM kernel on its own, for testing: C[M][N] (FP32) =
 * sum_k A(m,k) B(k,n) with BF16 operands as raw bits; A(m,k) = A[m*lda+k]
 * (a_trans 0) or A[k*lda+m] (a_trans 1); B(k,n) = B[n*ldb+k] (b_nk 1) or
 * B[k*ldb+n] (b_nk 0); K split `splits` ways and summed.  K, lda, ldb must be
 * multiples of 8, and M (a_trans) and N (!b_nk) too. */
int gf_gemm_bf16(int32_t M, int32_t N, int32_t K, const uint16_t *A, int64_t lda, int32_t a_trans, const uint16_t *B,
                 int64_t ldb, int32_t b_nk, int32_t splits, float *C);

/* gf_gemm_bf16 for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * reference counterpart.  A, B and C are device pointers (A and B 16-byte
 * aligned), the work is enqueued on `stream` (a hipStream_t; NULL for the
 * null stream) and the call returns without waiting.  The split-K partial
 * sums use a file-scope device array of 4 Mi floats (16 MiB) that each GPU
 * the library runs on holds until the process ends; a call needing more
 * (ceil(K / k_split) * M * N floats) is refused with GF_ERR_ARGUMENT.  Calls
 * that overlap on one GPU share that array and race.  Allocates nothing. */
int gf_gemm_bf16_noalloc(int32_t M, int32_t N, int32_t K, const uint16_t *A, int64_t lda, int32_t a_trans,
                         const uint16_t *B, int64_t ldb, int32_t b_nk, int32_t splits, float *C, void *stream);

/* Philox4x32-10 block (counter, key) -> 4 words, for checking the sampler. */
void gf_philox4x32(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);

/* ---- frequency-wavenumber solver (the Go package internal/fk) ----------
 * The layered boundary-value solve and the Hankel quadrature run on the GPU
 * in float64, one wavenumber sample per thread and many frequencies per
 * launch.  The host checks arguments and chooses grids exactly as the Go
 * code does, so errors carry the Go text.  A stack holds at most
 * GF_FK_MAX_LAYERS entries. */
#define GF_FK_MAX_LAYERS 16

/* fk.Medium (internal/fk/fk.go:35-53).  ref_freq_hz 0 -> 30 Hz; default_q 0
 * -> 30, used by layers whose qs is 0.  has_relax != 0 is Relax != nil: one
 * SLS applied to every layer in place of Kjartansson's constant-Q law.
 * gf_layer.qp is not read: the solver takes Qp = 9/4 Qs (fk.go:119-120). */
typedef struct gf_medium {
  const gf_layer *stack;
  int32_t n_layers;
  double ref_freq_hz;
  double default_q;
  int32_t has_relax;
  gf_sls relax;
} gf_medium;

/* fk.Integration (fk.go:309-331): samples 0 -> 20000 for one range, or the
 * grid's own choice for many; kmax_factor 0 -> 30. */
typedef struct gf_integration {
  int32_t samples;
  double kmax_factor;
} gf_integration;

/* fk.Grid (fk.go:439-443). */
typedef struct gf_fk_grid {
  double kmax;
  int32_t samples;
} gf_fk_grid;

/* Medium.SurfaceResponse (fk.go:276-307) at n points (k[i], f_hz[i]):
 * out [n][2] (re, im), m/N per unit Hankel measure.  The first point in order
 * that the Go call would reject sets the error. */
int gf_fk_surface_response(const gf_medium *m, const double *k, const double *f_hz, int64_t n, double *out);

/* Medium.StaticCoefficient (fk.go:347-353): k_far * SurfaceResponse(k_far, f). */
int gf_fk_static_coefficient(const gf_medium *m, double f_hz, double k_far, double out[2]);

/* Medium.GridFor (fk.go:456-489). */
int gf_fk_grid_for(const gf_medium *m, const double *ranges, int32_t n_ranges, double f_hz,
                   gf_integration opt, gf_fk_grid *out);

/* Medium.checkLayerPhase (fk.go:253-268), MaxLayerPhase 16. */
int gf_fk_check_layer_phase(const gf_medium *m, double f_hz);

/* Medium.VerticalDisplacement (fk.go:378-431): out (re, im), m/N. */
int gf_fk_vertical_displacement(const gf_medium *m, double r_m, double f_hz, gf_integration opt, double out[2]);

/* Medium.VerticalDisplacementMulti (fk.go:500-550) for n_freqs frequencies at
 * once: out [n_freqs][n_ranges][2].  Each frequency gets the grid and error
 * checks one Go call at that frequency would; the first frequency in order
 * whose Go call would fail sets the error, and nothing is written then. */
int gf_fk_vertical_displacement_multi(const gf_medium *m, const double *ranges, int32_t n_ranges,
                                      const double *f_hz, int32_t n_freqs, gf_integration opt, double *out);

/* synthesise in cmd/geofdtd (main.go:168-200): the f-k response to a Ricker
 * force (peak_n, f0_hz) at each range, as vertical velocity on the instants
 * t0 + i*dt, i < n, band-limited to band_hz, with `samples` wavenumber
 * samples (0: the grid's choice).  The medium is stack + relax with
 * has_relax set, as geofdtd -compare builds it.  out [n_ranges][n]. */
int gf_fk_synthesise(const gf_medium *m, const double *ranges, int32_t n_ranges, double peak_n, double f0_hz,
                     double band_hz, double dt, double t0, int32_t n, int32_t samples, double *out);

/* The wavenumber-integration fill of a bank, as cmd/geobank's build does it
 * (main.go:100-175): h validated as bank.Header.Validate, its range grid
 * checked with CheckRangeSampling at max_freq_hz, then bins 1..top (top =
 * int(max_freq / bin spacing), at most the last bin) from
 * VerticalDisplacementMulti at every grid range with the default
 * integration, all bins in one call; bin 0 and the bins above top stay zero.
 * payload [count][samples/2+1][2] float32, as bank.Set stores it.  A failing
 * bin reports "bin %d (%.3f Hz): " and the solver's error, for the lowest
 * failing bin (geobank's worker pool reports whichever finishes first). */
int gf_fk_bank_fill(const gf_medium *m, const gf_bank_header *h, double max_freq_hz, float *payload);

/* Counters of the last f-k integration on the calling thread (HIP events
 * on its stream): the per-frequency prologue (layer constants and the static
 * coefficient), the per-sample layered solve, the Hankel sums, and the fold
 * of the partial sums; and the work done. */
typedef struct gf_fk_stats_t {
  double kernel_ms;     /* prologue + solve + hankel + fold                  */
  double prologue_ms;
  double solve_ms;
  double hankel_ms;
  double fold_ms;
  int64_t samples;      /* wavenumber samples solved, over all frequencies  */
  int64_t weights;      /* samples x ranges Hankel terms J0(k r) summed      */
  int32_t frequencies;
  int32_t launches;
} gf_fk_stats_t;
int gf_fk_stats(gf_fk_stats_t *out);

#ifdef __cplusplus
}
#endif
#endif /* GEOFDTD_H */
