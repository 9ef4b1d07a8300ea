/*
 * euler2d_hip.h -- C ABI of the MI355X-native unsplit MUSCL-Hancock Godunov
 * solver (libeuler2d_hip.so).
 *
 * This is the drop-in boundary for the reference's solver API,
 * pkestene/euler2d_kokkos:
 *   - HydroParams / HydroSettings + ConfigMap .ini parsing
 *       (src/HydroParams.h:107-265, src/HydroParams.cpp:19-224,
 *        config/ConfigMap.h:25-47, config/inih/INIReader.h:47-67)
 *   - class HydroRun<device_t> (src/HydroRun.h:44-134): constructor and
 *     problem initialisation (:142-216), compute_dt (:227-251),
 *     godunov_unsplit (:258-364), make_boundaries (:388-399)
 *   - the time loop of main() (src/main.cpp:87-145)
 *
 * Conventions
 *   - Every function returns 0 on success and a negative E2D_ERR_* code on
 *     failure; nothing calls exit() or throws across the ABI.  The message of
 *     the last failure on the calling thread is returned by e2d_last_error().
 *   - Reals are double; every real-valued .ini parameter is parsed as float
 *     (strtof) and then widened, as the reference does.
 *   - A handle owns its device buffers and one HIP stream.  Calls on one handle
 *     must be serialised by the caller; different handles may be used from
 *     different threads.
 *   - Variable order in every state array: 0 rho, 1 E, 2 rho*u, 3 rho*v
 *     (src/HydroParams.h:27-34, ID=0, IP=IE=1, IU=2, IV=3).
 */
#ifndef EULER2D_HIP_H_
#define EULER2D_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define E2D_OK 0
#define E2D_ERR_ARG (-1)         /* invalid argument or handle */
#define E2D_ERR_IO (-2)          /* file could not be opened */
#define E2D_ERR_HIP (-3)         /* HIP runtime failure / no device */
#define E2D_ERR_ALLOC (-4)       /* device or host allocation failed */
#define E2D_ERR_UNSUPPORTED (-5) /* parameter combination not supported */

/* state selectors (HydroRun members U, U2) */
#define E2D_U 0
#define E2D_U2 1

/* host array layouts for e2d_copy_state / e2d_upload_state */
#define E2D_LAYOUT_RIGHT 0     /* [i][j][var], ghosts included (reference OpenMP View layout) */
#define E2D_LAYOUT_LEFT 1      /* [var][j][i], ghosts included (reference CUDA/HIP View layout) */
#define E2D_LAYOUT_INTERIOR 2  /* [var][j][i], interior nx*ny only */

/* numeric precision of a run */
#define E2D_PRECISION_FP64 0
#define E2D_PRECISION_FP32_SPLIT 1 /* hi+lo FP32 state, FP32 arithmetic */
#define E2D_PRECISION_FP32 2       /* all-float comparator (reference USE_DOUBLE=OFF) */

/* step schedules of the FP64 path */
#define E2D_SCHEDULE_FUSED 0    /* one fused kernel per step (default) */
#define E2D_SCHEDULE_PASSES 1   /* reference pass structure: BC, copy, prim, fluxes, update */
#define E2D_SCHEDULE_DIRECTIONAL 2 /* implementationVersion 1: stored slopes, x then y
                                      trace/flux + update passes (HydroRun.h:334-351) */
#define E2D_SCHEDULE_ATOMIC 3   /* implementationVersion 2: fluxes added to both neighbours
                                   with FP64 atomics (HydroRun.h:352-357); last bits vary */

/*
 * POD mirror of HydroParams + HydroSettings + blast + shocked-bubble fields
 * (src/HydroParams.h:107-265).  Derived fields (imax, jmax, isize, jsize, dx,
 * dy, gamma6, smallp, smallpp) are filled by e2d_params_init().
 * Extension fields, whose defaults keep the reference's behaviour:
 *   honor_riemann       0: HLLC always (the reference ignores `riemann`,
 *                       src/HydroRunFunctors.h:567,631); 1: use riemannSolverType
 *                       ([hydro] honor_riemann)
 *   implode_pressure_in high-state pressure of the implode problem
 *                       ([implode] pressure_in, default 1; the reference
 *                       hard-codes 1, src/HydroRunFunctors.h:1390)
 *   kin_frac            split precision only: a cell whose FP32 pressure is
 *                       below kin_frac * (gamma0 - 1) * E (internal energy
 *                       below that fraction of the total) has its primitive
 *                       variables converted again in FP64 from hi + lo
 *                       ([hydro] kin_frac, default 0.1; 0 never converts again)
 */
typedef struct e2d_params {
  /* [run] */
  int32_t nStepmax;
  int32_t nOutput;
  int32_t enableOutput;
  int32_t ioVTK;
  int32_t ioHDF5;
  int32_t ghostWidth; /* always 2 */
  double tEnd;
  /* [mesh] */
  int32_t nx, ny;
  int32_t imin, imax, jmin, jmax;
  int32_t isize, jsize;
  int32_t boundary_type_xmin, boundary_type_xmax; /* 1 reflect, 2 zero-gradient, other periodic */
  int32_t boundary_type_ymin, boundary_type_ymax;
  double xmin, xmax, ymin, ymax, dx, dy;
  /* [hydro] */
  double gamma0, gamma6, cfl, slope_type, smallr, smallc, smallp, smallpp;
  int32_t niter_riemann;
  int32_t riemannSolverType; /* 0 approx, 1 hll, 2 hllc */
  int32_t problemType;       /* 0 implode, 1 blast, 2 four_quadrant, 3 discontinuity, 4 shocked_bubble */
  int32_t implementationVersion;
  int32_t honor_riemann;
  int32_t blast_nbins;
  /* [blast] */
  double blast_radius, blast_center_x, blast_center_y;
  double blast_density_in, blast_density_out;
  double blast_pressure_in, blast_pressure_out;
  double blast_total_energy_inside;
  /* [shocked_bubble] (ShockedBubbleParams, src/HydroParams.h:134-147) */
  double sb_bubble_radius, sb_bubble_center_x, sb_bubble_center_y;
  double sb_bubble_density, sb_bubble_pressure;
  double sb_preshock_density, sb_preshock_pressure;
  double sb_postshock_density, sb_postshock_pressure, sb_postshock_velocity;
  double sb_shock_loc;
  /* [implode] */
  double implode_pressure_in;
  /* split-precision FP64 primitive-conversion threshold */
  double kin_frac;
} e2d_params;

typedef struct e2d_run e2d_run;       /* opaque solver handle (HydroRun) */
typedef struct e2d_config e2d_config; /* opaque parsed .ini (ConfigMap) */

/* ---------------------------------------------------------------- library */
const char *e2d_last_error(void);
int e2d_version(void);
int e2d_params_size(void); /* sizeof(e2d_params), for FFI layout checks */
/* name of device 0 and its CU count; E2D_ERR_HIP when no GPU is visible */
int e2d_device_info(char *name, int name_len, int *num_cu);

/* ------------------------------------------------------- ConfigMap (.ini) */
/* ConfigMap(std::string) -- config/ConfigMap.cpp:20-23, INIReader.cpp:17-20.
 * parse_error receives ini_parse's result: 0, the first bad line, or -1 when
 * the file cannot be opened (the reference then silently uses defaults). */
int e2d_config_open(const char *path, e2d_config **out, int *parse_error);
int e2d_config_close(e2d_config *cfg);
/* INIReader::getString -- INIReader.cpp:36-42; returns the full length */
int e2d_config_get_string(const e2d_config *cfg, const char *section, const char *name,
                          const char *default_value, char *out, int out_len);
/* INIReader::getInteger -- INIReader.cpp:54-64 (strtol, base 0) */
int e2d_config_get_integer(const e2d_config *cfg, const char *section, const char *name,
                           long default_value, long *out);
/* ConfigMap::getFloat -- config/ConfigMap.cpp:31-40 (strtof) */
int e2d_config_get_float(const e2d_config *cfg, const char *section, const char *name,
                         float default_value, float *out);
/* ConfigMap::getBool -- config/ConfigMap.cpp:57-76 */
int e2d_config_get_bool(const e2d_config *cfg, const char *section, const char *name,
                        int default_value, int *out);

/* ------------------------------------------------------------ HydroParams */
/* HydroParams::setup(ConfigMap&) -- src/HydroParams.cpp:42-156 */
int e2d_params_setup(const e2d_config *cfg, e2d_params *out);
/* ConfigMap(path) + HydroParams::setup in one call (parse_error may be NULL) */
int e2d_params_from_ini(const char *path, e2d_params *out, int *parse_error);
/* HydroParams::init -- src/HydroParams.cpp:160-190 (derived fields) */
int e2d_params_init(e2d_params *p);
/* HydroParams::print -- src/HydroParams.cpp:195-224 */
int e2d_params_print(const e2d_params *p);
/* checks what the kernels assume: nx,ny >= 2, ghostWidth 2, dx,dy > 0,
 * gamma0 > 1, slope_type in {0,1,2}, riemannSolverType in {0,1,2} */
int e2d_params_validate(const e2d_params *p);

/* --------------------------------------------------------------- HydroRun */
/* HydroRun(HydroParams&, ConfigMap&) -- src/HydroRun.h:142-216, minus the
 * problem initialisation (see e2d_init_problem).  Copies the params. */
int e2d_create(const e2d_params *p, int precision, e2d_run **out);
int e2d_destroy(e2d_run *run);
/* Init<Problem>Functor::apply(params, U); deep_copy(U2, U) -- HydroRun.h:185-214 */
int e2d_init_problem(e2d_run *run);
/* HydroRun::compute_dt(useU) -- HydroRun.h:227-251 (synchronous) */
int e2d_compute_dt(e2d_run *run, int useU, double *dt);
/* HydroRun::make_boundaries(U or U2) -- HydroRun.h:388-399 */
int e2d_make_boundaries(e2d_run *run, int which);
/* HydroRun::godunov_unsplit(nStep, dt) -- HydroRun.h:258-364: in = U for even
 * nStep else U2, out = the other; asynchronous on the handle's stream */
int e2d_godunov_unsplit(e2d_run *run, int nStep, double dt);
/* The main() time loop (src/main.cpp:87-145) without output: on entry *nStep
 * and *t hold the loop state (0, 0 for a fresh run; the startup dt and
 * boundary fills of main.cpp:87-91 run when *nStep == 0).  Runs until
 * t >= tEnd, nStep >= nStepmax, or max_steps iterations (max_steps <= 0: no
 * cap).  dt, the tEnd clamp and t += dt stay on the device (no host round
 * trip per step); returns the new t and nStep. */
int e2d_run_loop(e2d_run *run, int max_steps, double *t, int *nStep);

/* Bytes of the static device array that holds e2d_run_loop_noalloc's scratch
 * (one copy per GPU, kept until the process ends). */
#define E2D_NOALLOC_SCRATCH_BYTES 16384
/* e2d_run_loop on caller-owned state arrays, allocating nothing.  No
 * counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or other
 * allocating routine; the loop's scratch (control rings, loop-start partial
 * maxima, arrival counter) lives in a static device array of
 * E2D_NOALLOC_SCRATCH_BYTES per GPU, whose size does not depend on the grid.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (two host threads, or two streams) race, so serialise them.
 *   U, U2     the two state arrays (device memory, distinct), each 4 planes of
 *             isize x jsize cells (ghosts included) in this library's device
 *             layout for `precision`: element (v, j, i) at index
 *             v * plane + j * pitch + i, an element being a double (FP64), a
 *             (hi, lo) pair of floats (split, value hi + lo) or a float
 *             (FP32); aligned to the element size
 *   pitch     elements between rows (>= isize); plane: elements between
 *             planes (>= pitch * jsize)
 *   seg_rows  the step kernel's tile (0: 16; results do not depend on it)
 *   stream    a hipStream_t (NULL: the default stream)
 * Otherwise as e2d_run_loop: *t and *nStep hold the loop state on entry and
 * on return; at *nStep == 0 the ghost cells of both arrays are filled first;
 * runs until tEnd, nStepmax or max_steps iterations (<= 0: no cap).  The
 * step's input is U for even nStep, else U2.  Synchronous. */
int e2d_run_loop_noalloc(const e2d_params *params, int precision, void *U, void *U2,
                         long long pitch, long long plane, int seg_rows, int max_steps,
                         double *t, int *nStep, void *stream);
/* device <-> host copies of U / U2 in one of the E2D_LAYOUT_* layouts */
int e2d_copy_state(e2d_run *run, int which, double *host, int layout);
int e2d_upload_state(e2d_run *run, int which, const double *host, int layout);
/* Kokkos::deep_copy(dst, src) of a whole state array (HydroRun.h:214,302),
 * asynchronous on the handle's stream */
int e2d_copy_array(e2d_run *run, int src, int dst);
/* FP64 interior sums of the four conserved variables of U / U2 */
int e2d_conserved_sums(e2d_run *run, int which, double sums[4]);
/* accumulated device time in seconds of the five timers of HydroRun.h:74-75:
 * boundaries, godunov, primitives, fluxes, update (the fused schedule books
 * its whole step under godunov) */
int e2d_timers(e2d_run *run, double out[5]);
/* Interior cells of U / U2 whose internal energy is below kin_frac of the
 * total energy (no reference counterpart).  For the split precision these are
 * exactly the cells whose primitive variables the step converts again in FP64
 * (FP32 pressure below kin_frac * (gamma0 - 1) * E); for FP64 the same test in
 * FP64.  Synchronous. */
int e2d_kinetic_cells(e2d_run *run, int which, long long *count);
/* E2D_SCHEDULE_*.  e2d_create picks the schedule of the params'
 * implementationVersion: 0 -> FUSED (same bits as the reference's version 0),
 * 1 -> DIRECTIONAL, 2 -> ATOMIC.  Every schedule but FUSED is FP64 only. */
int e2d_set_schedule(e2d_run *run, int schedule);
int e2d_get_schedule(e2d_run *run, int *schedule);
/* rows marched by one wavefront of the fused step kernel (1..4096); the
 * wavefront covers 60 output columns.  Results do not depend on it. */
int e2d_set_tile(e2d_run *run, int seg_rows);
int e2d_get_tile(e2d_run *run, int *seg_rows);
/* Times `steps` loop iterations for each candidate seg_rows on the current
 * state (restored afterwards, t and nStep untouched), in two interleaved
 * passes, and keeps the candidate with the fastest timing.
 * candidates == NULL or ncand <= 0: a fixed set plus, for k = 1..4, the
 * smallest tile whose wavefronts fit k per SIMD of the device. */
int e2d_autotune(e2d_run *run, const int *candidates, int ncand, int steps, int *best_seg_rows,
                 double *best_ms_per_step);
int e2d_synchronize(e2d_run *run);
/* Device-side timing of `nsteps` loop iterations (e2d_run_loop semantics) with
 * HIP events on the handle's stream; returns the elapsed milliseconds and the
 * number of steps actually executed.  Used by benchmarks.  The loop's step
 * launches are replayed from captured graphs of 240 or 30 launches where the budget
 * allows (single launches otherwise); a graph missing for the current tile is captured and
 * instantiated (a few milliseconds of host work) before the start event is
 * recorded, so it never falls inside the timed region.  The loop start is one
 * launch (the first step's CFL rate, and at nStep == 0 the ghost cells of U
 * and U2), and the end needs no copy command: the kernels also store the loop
 * state in pinned host memory. */
int e2d_time_loop(e2d_run *run, int nsteps, double *elapsed_ms, int *executed, double *t,
                  int *nStep);

/* ------------------------------------------------------------- output */
/* HydroRun::saveVTK (src/HydroRun.h:502-609): ASCII VTK ImageData file
 * <output_dir>/<output_prefix>_<7-digit step>.vti with the interior cells of
 * U / U2, i fastest, arrays rho, E, mx, my (Float32 for E2D_PRECISION_FP32,
 * else Float64) */
int e2d_save_vtk(e2d_run *run, int which, const char *output_dir, const char *output_prefix,
                 int step);
/* the same file from a host array in E2D_LAYOUT_INTERIOR order ([var][j][i],
 * nx*ny per variable) */
int e2d_write_vti(const e2d_params *p, const double *interior, int use_float32,
                  const char *output_dir, const char *output_prefix, int step);
/* ComputeRadialProfileFunctor::apply (src/ComputeRadialProfileFunctor.h:83-128)
 * without the file output: bin centres (i+0.5)*dr, dr = half-diagonal/nbins,
 * and the mean density of every cell (ghosts included) whose distance from the
 * box centre falls in the bin (NaN for an empty bin); counts may be NULL.
 * Cells past the last bin are left out (the reference writes outside its
 * histogram for them). */
int e2d_radial_profile(e2d_run *run, int which, int nbins, double *distances, double *density,
                       int64_t *counts);
/* cnpy::npy_save of a 1-D double array (src/cnpy/cnpy.h:129-187,343-376):
 * numpy format 1.0, header padded to a multiple of 16 bytes */
int e2d_write_npy(const char *path, const double *data, int64_t n);

/* ------------------------------------------------------------------ batch */
/* Many independent problems advanced together (no reference counterpart: the
 * reference runs one grid per process, src/main.cpp:86).  Every instance keeps
 * the HydroRun semantics above -- its own grid size, problem, dt, t, nStep --
 * and all share the precision, the Riemann solver and the step kernel.  A run
 * is one persistent launch (each instance steps as soon as its own previous
 * step is done) or one launch per step for all instances, replayed from a
 * captured hipGraph: see e2d_batch_set_launch. */
typedef struct e2d_batch e2d_batch;
int e2d_batch_create(const e2d_params *params, int count, int precision, e2d_batch **out);
int e2d_batch_destroy(e2d_batch *batch);
/* Init<Problem> of every instance, boundary fill, and the startup dt of each
 * instance (main.cpp:87-91); resets t and nStep to 0 */
int e2d_batch_init(e2d_batch *batch);
int e2d_batch_set_tile(e2d_batch *batch, int seg_rows);
/* How e2d_batch_run launches the steps (no reference counterpart: the
 * reference runs one grid per process, src/main.cpp:86).
 *   E2D_BATCH_LAUNCH_PERSISTENT  all the run's steps in one launch: every
 *                                instance is assigned to one XCD, whose
 *                                wavefronts take its work items from that
 *                                XCD's queue and start its step n + 1 as soon
 *                                as its own step n has completed; use_graph
 *                                is ignored
 *   E2D_BATCH_LAUNCH_STEPS       one launch per step for all instances,
 *                                replayed from hipGraphs when use_graph != 0
 * A new batch takes the persistent launch when its instances fill the XCD
 * queues evenly (the fullest queue holds at most 10% more cells than the
 * mean), else one launch per step (a few large grids would otherwise leave
 * XCDs idle); e2d_batch_get_launch reports the choice.
 * spin_limit bounds each dependency wait of the persistent launch (polls of
 * the completion count, each about a microsecond; <= 0: the default, 2^20,
 * about a second); a wait that runs past it ends
 * the launch and e2d_batch_run returns E2D_ERR_HIP, leaving the instances
 * part-way (run e2d_batch_init).  Results are the same bits either way. */
#define E2D_BATCH_LAUNCH_STEPS 0
#define E2D_BATCH_LAUNCH_PERSISTENT 1
int e2d_batch_set_launch(e2d_batch *batch, int mode, int spin_limit);
int e2d_batch_get_launch(e2d_batch *batch, int *mode);
/* XCD queues of the persistent launch (one per XCD the device showed at
 * e2d_batch_create) and their imbalance: the fullest queue's cells over the
 * mean (1.0 = even). */
int e2d_batch_queue_info(e2d_batch *batch, int *nxcd, double *imbalance);
/* Runs up to max_steps batched steps (instances whose loop has ended stay put).
 * use_graph != 0 replays captured hipGraphs of the step launch.  Returns the
 * device time in milliseconds and the number of instances still active.
 * E2D_ERR_ARG until e2d_batch_init has run (as for e2d_batch_autotune). */
int e2d_batch_run(e2d_batch *batch, int max_steps, int use_graph, double *elapsed_ms,
                  int *active_left);
int e2d_batch_status(e2d_batch *batch, int instance, double *t, int *nStep);
/* current state (U if the instance's nStep is even, else U2) of one instance */
int e2d_batch_copy_state(e2d_batch *batch, int instance, double *host, int layout);
/* total interior cells, wavefronts per step, instance count */
int e2d_batch_info(e2d_batch *batch, long long *cells, int *waves, int *count);
/* Times each candidate seg_rows over `steps` batched steps, in two
 * interleaved passes, and keeps the fastest; returns it.  Until
 * e2d_batch_set_launch fixes the launch mode, and when the instances fill the
 * XCD queues evenly, each candidate is timed in both launch modes and the
 * fastest (mode, seg_rows) pair is kept (e2d_batch_get_launch reports the
 * mode); otherwise the batch's launch mode is used.  A candidate whose run saw
 * instances reach the end of their loops is timed again after e2d_batch_init.
 * Restoring the state afterwards is the caller's business: call
 * e2d_batch_init. */
int e2d_batch_autotune(e2d_batch *batch, const int *candidates, int ncand, int steps,
                       int *best_seg_rows, double *best_ms_per_step);

#ifdef __cplusplus
}
#endif

#endif /* EULER2D_HIP_H_ */
