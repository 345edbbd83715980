/* moses2pdf.h -- C ABI of the MI355X-native adNOC_2Pb time-step engine.
 *
 * Drop-in for the reference's run path (MoSES_2PDF, landslide and inflow modes).  The
 * reference has no FFI: its boundary is the C++ class RunKernels
 * (MoSES_2PDF_landslide/adNOC_2Pb_runKernels.cuh:27-377), whose constructor parses
 * ./par_List and the data files (adNOC_2Pb_runKernels.cu:4-9, 11-304) and whose run()
 * allocates, runs the metric setup and the time loop and writes ./result2Pb/
 * (adNOC_2Pb_runKernels.cu:306-356).  Each entry point below names the reference code it
 * replaces.
 *
 * Conventions: every function returns 0 on success and a negative M2P_E* code on error
 * (the library never exits); m2p_last_error() gives the message.  A context owns all of
 * its device memory and orders all of its work on one HIP stream, taken round robin from
 * a pool of 8 per device (creating a stream costs about a millisecond, and the device
 * serves streams from 4 hardware queues): contexts sharing a stream also wait for each
 * other's queued work.  A context is not thread-safe; distinct contexts may be driven
 * from different host threads.  Buffers passed in are owned by
 * the caller.  State planes are laid out as in the reference: 7 planes of nxd*nyd
 * doubles, plane k holding u[k] at index row*nxd+col, with nxd = nx+6, nyd = ny+6
 * (three ghost cells per side), k = 0 hs, 1 hs*us, 2 hs*vs, 3 hf, 4 hf*uf, 5 hf*vf,
 * 6 phi.  All quantities are in the reference's scaled units (lengths x10, g = 1).
 */
#ifndef MOSES2PDF_H
#define MOSES2PDF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct m2p_ctx m2p_ctx;

enum { M2P_LANDSLIDE = 0, M2P_INFLOW = 1 };
/* Scalar type of the time step (m2p_params.precision).  M2P_FP64 is the reference's
 * arithmetic.  M2P_FP32 keeps the state and the per-step arithmetic in float (hardware
 * reciprocal and square root); the metric is computed in FP64 and rounded once, and the
 * CFL max, the dt control, the simulated time and the mass totals stay FP64. */
enum { M2P_FP64 = 0, M2P_FP32 = 1 };

enum {
  M2P_OK = 0,
  M2P_EINVAL = -1,   /* bad argument or malformed input */
  M2P_EIO = -2,      /* file missing or unwritable */
  M2P_EHIP = -3,     /* HIP runtime error */
  M2P_ENOMEM = -4,   /* device allocation failed */
  M2P_ESTATE = -5    /* call not valid in the context's state */
};

/* Run parameters.  The first seven are par_List lines 1-7 (adNOC_2Pb_runKernels.cu:21-40);
 * the next four (cfl .. dt_max) are constants the reference hard-codes
 * (adNOC_2Pb_kernels.cuh:16, adNOC_2Pb_kernels.cu:1710 and 2014-2017; inflow 2060) and
 * default to those values; the last three choose how the step is computed.  The reference
 * reads the first seven with sscanf and checks nothing; m2p_open and m2p_open_grid return
 * M2P_EINVAL with the reason in m2p_last_error unless every value is finite and
 * total_sim >= 0, each_out > 0, 0 <= delta0 < 90, Cd >= 0, N_R > 0, varTheta >= 0,
 * 0 <= phiS0 <= 1 (inflow: 0 < phiS0 <= 1, its injection divides by it), 0 < cfl <= 1,
 * 0 <= rho_ratio < 1, 0 < dt_min <= dt_max, and the three ints take the values listed. */
typedef struct m2p_params {
  double total_sim;   /* seconds */
  double each_out;    /* seconds */
  double delta0;      /* basal friction angle, degrees */
  double Cd;
  double N_R;
  double varTheta;
  double phiS0;
  double cfl;         /* reference 0.1 */
  double rho_ratio;   /* reference 1.420/2.60 */
  double dt_min;      /* reference 0.001 (landslide), 0.002 (inflow) */
  double dt_max;      /* reference 2.0 */
  int precision;      /* M2P_FP64 (default) or M2P_FP32 */
  int dry_skip;       /* 1 (default): tiles whose whole stencil is dry (every component of
                       * every cell zero) are skipped; their results are exactly those of
                       * computing them.  0: every tile is computed, as the reference does. */
  int metric_form;    /* 0 (default): when a set-up check finds the inverse Jacobian formed
                       * from svec and cvalue bit-identical to the stored metric planes
                       * (every uniform grid), the step forms it instead of loading it.
                       * 1: the step always loads the stored planes (what it does anyway
                       * on a grid that fails the check).  FP64 results are the same bit
                       * for bit either way; FP32 ones differ by float rounding.  No
                       * reference counterpart (it loads the planes). */
} m2p_params;

/* Fills *p with the reference's values for `mode` (par_List values of the shipped
 * examples, constants of the kernels). */
void m2p_default_params(int mode, m2p_params* p);

/* Replaces RunKernels::RunKernels(char* fname) (adNOC_2Pb_runKernels.cu:4-9): reads the
 * par_List at `par_list` (9 lines landslide, 11 lines inflow; '#' starts a comment) and
 * the files it names (relative paths resolve against the current directory, as in the
 * reference), parses them with the reference's precision rules and uploads them.
 * `overlay` is NULL or "key=value;key=value" over cfl, rho_ratio, dt_min, dt_max,
 * precision (fp64 | fp32), dry_skip (0 | 1) and metric_form (0 | 1); a number must be the
 * whole value. */
int m2p_open(const char* par_list, int mode, const char* overlay, m2p_ctx** out);

/* In-memory constructor (no files).  topo[NY*NX] and depth[NY*NX] hold the grid values
 * as a file would: line j, value i at [j*NX+i], in metres, already rounded as the
 * reference's parser rounds them (float for topography).  depth is ignored in inflow
 * mode.  phi (nullable, landslide only) is a per-cell solid fraction replacing phiS0 at
 * release cells.  Inflow mode takes the hydrograph tables as parsed: t[nsamp] seconds,
 * h/u/v[nsamp*nloc] (sample-major), loc[nloc*3] = (NO, X, Y) per location. */
int m2p_open_grid(int mode, int NX, int NY, double cellsize, const double* topo,
                  const double* depth, const double* phi, const m2p_params* par, int nsamp,
                  int nloc, const double* t, const double* h, const double* u,
                  const double* v, const double* loc, m2p_ctx** out);

/* Replaces RunKernels::run() (adNOC_2Pb_runKernels.cu:306-356, loop 646-1330): runs the
 * time loop to the last output time (at most 99999 steps) writing DEM.dat, 001.dat,
 * Info.dat, Time.dat and NNN.dat into `result_dir` in the reference's formats.
 * *wall_s (nullable) receives the wall time of the call. */
int m2p_run(m2p_ctx* ctx, const char* result_dir, double* wall_s);

/* m2p_run for at most max_steps steps from the context's current state, so a long run can
 * be split into segments (with m2p_checkpoint / m2p_restore in between).  *finished
 * (nullable) is 1 once the last output time has been written. */
int m2p_run_steps(m2p_ctx* ctx, const char* result_dir, int max_steps, int* finished,
                  double* wall_s);

/* Advances `nsteps` full time steps (predictor + corrector) with no output clipping and no
 * host synchronisation inside the loop.  No reference counterpart as a call: it is the
 * body of the kernelStep loop (adNOC_2Pb_runKernels.cu:646-1330) without the file output. */
int m2p_step(m2p_ctx* ctx, int nsteps);

/* Advances until the simulated time (scaled, as m2p_get_state reports it) reaches t_end,
 * the last step shortened to end on it, or until max_steps steps are done; *steps_done
 * (nullable) receives the count.  No reference counterpart: it replaces the output-time
 * clipping of the loop (adNOC_2Pb_runKernels.cu:983-988) for a single end time, so that
 * runs of different precision or dt can be compared at the same time.  The halting rules
 * of m2p_set_stop stay in force during the call (it also stops where they halt the
 * scenario) and are left unchanged by it. */
int m2p_run_to(m2p_ctx* ctx, double t_end, int max_steps, int* steps_done);

/* Waits for all work queued on the context's stream. */
int m2p_sync(m2p_ctx* ctx);

/* Copies the state (7*nxd*nyd doubles, layout above) and the simulated time (scaled) and
 * step count out.  Any pointer may be NULL. */
int m2p_get_state(m2p_ctx* ctx, double* u, double* t, int* nt);

/* Overwrites the state (7*nxd*nyd doubles, all cells including ghosts).  In landslide
 * mode the ghost cells of the momenta (planes 1, 2, 4, 5) are set to 0, the value every
 * step's ghost fill gives them before anything reads them. */
int m2p_set_state(m2p_ctx* ctx, const double* u);

int m2p_dims(const m2p_ctx* ctx, int* nx, int* ny, int* nxd, int* nyd);

/* Total solid and fluid volume over interior cells, sum of u[0]*Detmin and u[3]*Detmin,
 * summed in a fixed order with compensated (double-double) arithmetic. */
int m2p_mass(m2p_ctx* ctx, double* solid, double* fluid);

/* Last computed scalars: dt of the last corrector, the float-rounded max wave speed, and
 * the output counter `iter` (adNOC_2Pb_runKernels.cu:390). */
int m2p_scalars(m2p_ctx* ctx, double* dt, double* dmax, int* iter);

/* Copies a named nxd*nyd (or k*nxd*nyd) device field out, for verification: "cvalue",
 * "svec" (2 planes), "Detmin", "invJ11".."invJ23", "dxdxi11".."dxdxi22", "J13dxi",
 * "J23dxi", "J33dxi", "J13det", "J23det", "J33det", "uone45" (2 planes); and
 * "metric_uniform", one value: 1 when the step forms the inverse Jacobian from svec and
 * cvalue instead of loading it (the grid's coordinate derivatives are exactly the identity;
 * in FP64 a set-up check found the formed values bit-identical to the stored ones wherever
 * the step reads them, in FP32 they are formed in float), else 0 (also when
 * m2p_params.metric_form is 1).  Returns planes copied, or a negative code. */
int m2p_get_field(m2p_ctx* ctx, const char* name, double* out);

/* Mean device time per launch of each kernel family since the last reset, measured with
 * HIP events on the context's stream when profiling is enabled (work of another context
 * queued on the same pool stream in between would be counted too).  enable: 0 off, 1 every
 * family ("inflow", "ghost", "cfl", "dt", "stage0", "stage1", "vmax"), 2 + k only the
 * k-th of them in that order (fewer events in the stream: each pair of event records
 * between two kernels costs a few microseconds of a step). */
int m2p_profile(m2p_ctx* ctx, int enable);
int m2p_profile_get(m2p_ctx* ctx, const char* kernel, double* ms_per_launch, int* launches);

/* The context's last error message; with ctx == NULL, the calling thread's last error of
 * a call that takes no context (m2p_mass_noalloc). */
const char* m2p_last_error(const m2p_ctx* ctx);

/* Closes the context (removing it from its batch first) and frees its device memory. */
void m2p_close(m2p_ctx* ctx);

/* ---- Allocation-free variant (no reference counterpart) ----------------------------- */

/* m2p_mass for a state held in the caller's own device memory, allocating nothing: no
 * counterpart in the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * u: device pointer to a state (7 planes of (nx+6)*(ny+6) values, layout above), double
 * for precision M2P_FP64 and float for M2P_FP32; detmin: device pointer to the Detmin
 * plane ((nx+6)*(ny+6) doubles, as m2p_get_field "Detmin" gives it); out: device pointer
 * to 2 doubles that receive the solid and fluid totals, equal bit for bit to what
 * m2p_mass returns for that state; stream: a hipStream_t, or NULL for the null stream.
 * The call enqueues its two kernels on `stream` and returns without waiting (so it can be
 * captured into a graph).  Its block partials go to one fixed 16 KiB device array per
 * GPU (held until the process ends), whatever the grid size: calls that overlap on one
 * GPU race.  Rejects null pointers, a bad precision and grids outside 1 <= nx, ny,
 * nx*ny <= INT_MAX with M2P_EINVAL (message: m2p_last_error(NULL)). */
int m2p_mass_noalloc(const void* u, int precision, const double* detmin, int nx, int ny,
                     double* out, void* stream);

/* ---- Checkpoint / resume (no reference counterpart: its text outputs are lossy) ---- */

/* Writes the scenario's resumable state to `path` (binary, written to path.tmp and then
 * renamed): a header (magic "M2PCKPT1", version 2, mode, NX, NY, precision, output-time
 * count, P, and an FNV-1a hash of the padded topography, the run parameters and the
 * hydrograph tables), the device loop state (time, dt, output and inflow counters, step
 * count, halting rules, non-finite flag), the Time.dat history of m2p_run, the 7 state
 * planes of u^n and of the half-step buffer (its ghost rows feed the next step's
 * landslide corners), the FP32 limiter flags, and an FNV-1a checksum.  Stepping a context
 * restored from it gives bit-identical results to never having stopped. */
int m2p_checkpoint(m2p_ctx* ctx, const char* path);

/* Loads a checkpoint into a context opened on the same scenario (mode, grid, precision,
 * output times, topography, parameters and hydrograph tables must match; M2P_EINVAL
 * otherwise, with the reason in m2p_last_error).  m2p_run then continues the output
 * sequence where it stopped. */
int m2p_restore(m2p_ctx* ctx, const char* path);

/* ---- Halting rules and status (no reference counterpart) -------------------------- */

/* Sets when a scenario halts, counted from now: t_end (scaled time; <= 0: none; the step
 * that reaches it is shortened to end on it), vstop (flow speed, scaled units = m/s;
 * <= 0: none; the scenario takes one more step after the first corrector whose largest
 * flow speed over cells deeper than 1 mm falls below vstop, once some corrector has
 * reached vstop), max_steps (< 0: none).  The kernels of a halted scenario return at
 * once, so it can stay in a batch or a captured graph.  Clears an earlier halt. */
int m2p_set_stop(m2p_ctx* ctx, double t_end, double vstop, int max_steps);

/* Switches dry-tile skipping (m2p_params.dry_skip) on (1) or off (0) from the next step.
 * Results are the same either way; skipping only leaves out tiles whose inputs are all
 * zero.  No reference counterpart. */
int m2p_set_dry_skip(m2p_ctx* ctx, int on);

/* Simulated time (scaled), steps done and whether the scenario has halted. */
int m2p_status(m2p_ctx* ctx, double* t, int* nt, int* halted);

/* Whether the state has gone non-finite.  The reference's dt reduction drops a NaN wave
 * speed through fmaxf (adNOC_2Pb_kernels.cu:1988, 1996), and its speed floor and face
 * maxima drop it before that (Flux5/6Kernel, adNOC_2Pb_kernels.cu:1762-1807), so a run
 * that blows up keeps stepping NaNs with no sign of it.  Here the CFL pass of every step
 * also marks a face whose one-sided wave speeds are not all finite, in the same
 * reduction as dt (which still ignores it, so results are unchanged): *flag is 1 once
 * that has happened, and *first_nt the step count of the first state it was seen in (the
 * state after that many steps held a NaN or an Inf; -1 while *flag is 0).  The CFL pass
 * reads u^n, so a state is seen by the step after the one that made it.  m2p_set_state
 * clears the flag; a checkpoint carries it. */
int m2p_nonfinite(m2p_ctx* ctx, int* flag, int* first_nt);

/* Device bytes a scenario of this shape holds once open (*steady) and at most while it
 * is being set up (*peak), for admission by HBM footprint. */
int m2p_footprint(int mode, int NX, int NY, int precision, int nsamp, int nloc,
                  long long* steady, long long* peak);

/* ---- Batched ensembles (no reference counterpart) --------------------------------- */

/* A batch steps many scenarios of one precision, of any grid sizes, with one launch per
 * phase of the time step for all of them (inflow, ghost fill, CFL box and ring, dt,
 * predictor, ghost fill, corrector), captured as a HIP graph of `graph_steps` steps.
 * Each member keeps its own buffers, dt and halting rules; its results are bit-identical
 * to stepping it alone with m2p_step.  While a scenario is in a batch, m2p_step,
 * m2p_run, m2p_run_to and m2p_set_state on it return M2P_ESTATE; the read functions work
 * (they order after the batch's queued steps). */
typedef struct m2p_batch m2p_batch;

int m2p_batch_create(int precision, m2p_batch** out);
/* Returns the member's index (>= 0) or an error code.  The caller keeps ownership. */
int m2p_batch_add(m2p_batch* b, m2p_ctx* ctx);
int m2p_batch_remove(m2p_batch* b, m2p_ctx* ctx);
int m2p_batch_size(const m2p_batch* b);
/* Status of every member, in order of addition, with one device-to-host copy for the whole
 * batch (no counterpart in the reference; for serving loops that poll many members): per
 * member the scaled time, the steps done and whether it has halted (as m2p_status), and
 * the non-finite flag and the step count of the first non-finite state, -1 if none (as
 * m2p_nonfinite).  n must equal m2p_batch_size; any array may be null. */
int m2p_batch_status(m2p_batch* b, int n, double* t, int* nt, int* halted, int* nonfinite,
                     int* nonfinite_nt);
/* Enqueues nsteps steps of every member (asynchronous; m2p_batch_sync waits). */
int m2p_batch_step(m2p_batch* b, int nsteps);
int m2p_batch_sync(m2p_batch* b);
/* Autotuner: times, for trial_steps steps on the batch itself, every compiled stage-kernel
 * variant (tile height x waves per SIMD) with XCD banding and largest-first member order
 * on; then, for the fastest, member order as added and banding off, each kept if faster;
 * then the CFL launch's block count, doubling from 4096 and halving while faster
 * (2048 .. 32768).  The trials start from the state one step on, restored after each;
 * every member's state, time and step count are put back at the end, and a failed search
 * leaves the schedule as it was.  A text table of the timings goes to report (nullable). */
int m2p_batch_tune(m2p_batch* b, int trial_steps, char* report, int report_len);
/* Manual schedule: keys "ty", "waves", "remap", "order", "graph_steps" (0: no graph),
 * "cfl_blocks" (>= 64; most blocks of the CFL launch, default 4096), "plan_cap" (>= 1;
 * with dry-tile skipping, the stage launch has a block for each of the first plan_cap
 * entries of each of its 8 tile lists, and a second launch of 512 blocks takes the rest,
 * at about half the speed per tile and on fewer blocks; default 4096, so up to 32768
 * listed tiles per stage run at full speed; raise it for batches that are mostly wet).
 * Results do not depend on any of them.
 * The (ty, waves) pair must be a compiled variant; m2p_batch_step reports M2P_EINVAL
 * otherwise. */
int m2p_batch_set(m2p_batch* b, const char* key, int value);
int m2p_batch_get(const m2p_batch* b, const char* key);
const char* m2p_batch_last_error(const m2p_batch* b);
/* Detaches the members (they keep their state and can be stepped alone) and frees the
 * batch. */
void m2p_batch_destroy(m2p_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* MOSES2PDF_H */
