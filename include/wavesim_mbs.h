/* wavesim_mbs.h -- C ABI of the MI355X (gfx950) modified Born series solver.
 *
 * The library replaces the native loop RunWaveSimAlgorithmMex that
 * WaveSimBase.run_algorithm calls (reference WaveSimBase.m:126-155, call at
 * WaveSimBase.m:145-146); the MATLAB-branch loop it must reproduce is
 * WaveSimBase.m:157-197 with the stopping rule of Simulation.m:193-217.
 *
 * Arrays are complex single precision in MATLAB column-major order
 * [y, x, z, pol] (y contiguous), exactly as mxArrays hold them.  Positions and
 * ROIs are 1-based and inclusive, as in the reference.  Every function returns
 * WSIM_OK (0) or a negative error code; wsim_last_error() describes the last
 * failure on the calling thread.  No C++ exception crosses this boundary.
 */
#ifndef WAVESIM_MBS_H
#define WAVESIM_MBS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WSIM_ABI_VERSION 1

enum {
  WSIM_OK = 0,
  WSIM_ERR_ARG = -1,          /* invalid argument (shape, null pointer, range) */
  WSIM_ERR_HIP = -2,          /* HIP runtime error (no device, launch failure) */
  WSIM_ERR_UNSUPPORTED = -3,  /* a size/shape the kernels do not handle */
  WSIM_ERR_NOMEM = -4,        /* device allocation failed */
  WSIM_ERR_ABORTED = -5,      /* the callback asked to stop */
  WSIM_ERR_DIVERGED = -6      /* the stop rule saw a diverging iteration (results are readable;
                                 the rhs reports converged = -1) */
};

typedef struct {
  float re, im;
} wsim_cfloat;

typedef struct {
  double re, im;
} wsim_cdouble;

/* One wiggle descriptor (WaveSimBase.m:300-316): Fourier coordinates already
 * shifted by a quarter pixel and divided by sqrt(epsilon), and the real-space
 * phase ramps.  pxe/gx have Nx entries, pye/gy Ny, pze/gz Nz. */
typedef struct {
  const float* pxe;
  const float* pye;
  const float* pze;
  const wsim_cfloat* gx;
  const wsim_cfloat* gy;
  const wsim_cfloat* gz;
} wsim_wiggle;

/* One source patch, i.e. one entry of Source.positions/Source.values
 * (Source.m:17-21) after Simulation.exec shifted and cropped it to the grid
 * (Simulation.m:154-156): 1-based [y x z pol] position of its first element,
 * its extent, and its values in column-major order. */
typedef struct {
  double pos[4];
  int64_t size[4];
  const wsim_cfloat* values;
} wsim_source_patch;

/* Progress callback (the role of SimulationCallback.m, invoked by the MEX).
 * Called as Simulation.next calls its callback (Simulation.m:213): after
 * every iteration that is a multiple of callback_interval and after the last
 * one, with the diff_energy history so far (the iterations run in batches
 * that end on those multiples).  Return nonzero to stop the solve
 * (WSIM_ERR_ABORTED). */
typedef int (*wsim_callback)(const char* type, int64_t it, const double* diff_energy,
                             int64_t n, double energy_threshold, void* user);

/* ---------------------------------------------------------------------------
 * Drop-in for RunWaveSimAlgorithmMex (WaveSimBase.m:145-146):
 *   [E, diff_energy] = RunWaveSimAlgorithmMex(N, gamma, Nwiggles, wiggles, source,
 *        epsilon, k02e, roi, maxIter, energy_threshold, callbackString, callback_interval)
 *  N         [Ny Nx Nz Npol] (Npol 1 = WaveSim, 3 = WaveSimVector)
 *  gamma     Ny*Nx*Nz complex (WaveSimBase.m:73)
 *  roi       2x4 column-major doubles: roi[2*c] = start, roi[2*c+1] = end of
 *            column c in [y x z pol] (Simulation.m:117-124)
 *  E_out     N[0]*N[1]*N[2]*N[3] complex: the primed field E' = gamma*E on the
 *            full grid; the caller divides by gamma and crops (WaveSimBase.m:150-153)
 *  diff_energy_out  one entry per iteration (capacity diff_energy_cap; entries
 *            past it are not written); *n_diff_energy receives the number of
 *            iterations run.  The cap max_iter applies at the first multiple of
 *            Nwiggles at or after it (Simulation.m:202), so up to
 *            ceil(max_iter / nwiggles) * nwiggles iterations run (at most 2 000 000,
 *            rounded up likewise).
 * The loop follows WaveSimBase.m:157-197 exactly: source i/eps/Nw added during
 * the first Nw iterations, alpha = 0.7, energy refreshed at iterations
 * 1, 2, 2+interval, ... with interval = min(8, callback_interval)
 * (Simulation.m:61, 135), stop tested at multiples of Nw. */
int wsim_run_algorithm(const double N[4], const wsim_cfloat* gamma, int nwiggles,
                       const wsim_wiggle* wiggles, int npatches, const wsim_source_patch* patches,
                       float epsilon, wsim_cfloat k02e, const double roi[8], double max_iter,
                       double energy_threshold, const char* callback_type, double callback_interval,
                       wsim_callback cb, void* user, wsim_cfloat* E_out, double* diff_energy_out,
                       int64_t diff_energy_cap, int64_t* n_diff_energy);

/* ---------------------------------------------------------------------------
 * Allocation-free variant of wsim_run_algorithm.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code that
 * must not call hipMalloc, such as tight serving loops or harnesses that
 * forbid allocation).  It calls no hipMalloc, hipFree, stream or graph
 * creation: its work fields (dE', T, S), tables, state and histories live in
 * a static 64 MiB device array that every GPU loading the library holds
 * until the process ends.
 *  gamma_dev  device array, Ny*Nx*Nz complex (read only)
 *  E_out_dev  device array, N[0]*N[1]*N[2]*N[3] complex: receives E' (full grid)
 *  stream     hipStream_t the kernels run on (NULL: the null stream); the
 *             call waits for the stream's earlier work first and returns
 *             when the solve is done
 * Other arguments and results as wsim_run_algorithm (host tables, patches,
 * diff_energy_out).  Limits, refused with WSIM_ERR_UNSUPPORTED: 3 * Npol *
 * Ny*Nx*Nz * 8 bytes of work fields within the 64 MiB (plus small tables;
 * e.g. 1152^2 and 128^3 scalar fit, 320^3 does not) and at most 100 000
 * iterations.  Calls on one device share the array: within a process they
 * run one at a time (a lock); the array keeps the last call's contents, and
 * every call clears what must start at zero. */
int wsim_run_algorithm_noalloc(const double N[4], const wsim_cfloat* gamma_dev, int nwiggles,
                               const wsim_wiggle* wiggles, int npatches, const wsim_source_patch* patches,
                               float epsilon, wsim_cfloat k02e, const double roi[8], double max_iter,
                               double energy_threshold, const char* callback_type, double callback_interval,
                               wsim_callback cb, void* user, wsim_cfloat* E_out_dev, double* diff_energy_out,
                               int64_t diff_energy_cap, int64_t* n_diff_energy, void* stream);

/* ---------------------------------------------------------------------------
 * Handle API: the same loop with device-resident buffers, for repeated solves
 * and for nrhs right-hand sides that share gamma (one launch per pass covers
 * every rhs; each rhs has its own energy history and stop flag).  There is no
 * reference counterpart beyond the loop above: the reference solves several
 * sources by calling exec repeatedly (examples/pulse2d.m:31-40). */
typedef struct wsim_solver wsim_solver;

typedef struct {
  int64_t N[4];         /* [Ny Nx Nz Npol] */
  int nrhs;             /* right-hand sides sharing gamma */
  int nwiggles;
  float epsilon;
  wsim_cfloat k02e;
  int64_t roi[2][4];    /* 1-based inclusive [start; end] of [y x z pol] */
  int energy_interval;  /* energy_calculation_interval (Simulation.m:61) */
} wsim_solver_desc;

/* Supported grids (every entry point): at most 2^29 voxels (Ny*Nx*Nz; one
 * complex64 component below 4 GiB), axis lengths whose prime factors are
 * 2, 3, 5, 7 and 11 only (the Stockham radices; the reference's
 * SimulationGrid.efficient_size pads non-periodic axes to such lengths,
 * SimulationGrid.m:64-80, but leaves periodic axes as given, so a periodic
 * axis of, e.g., 13 or 101 samples is refused with WSIM_ERR_UNSUPPORTED),
 * and vector fields (Npol 3) with at least two axes longer than 1. */
int wsim_solver_create(wsim_solver** out, const wsim_solver_desc* desc, const wsim_cfloat* gamma,
                       const wsim_wiggle* wiggles);
/* Replace the source of one rhs (host patches; the library scales by i/eps/Nw). */
int wsim_solver_set_source(wsim_solver* s, int rhs, int npatches, const wsim_source_patch* patches);
/* Run the loop from E' = dE' = 0 for every rhs with a nonzero source.
 * batch = iterations per launch batch (0 = automatic; with a callback and
 * callback_interval >= 1 the batch is the largest divisor of the interval up
 * to 256); the batch is captured once into a HIP graph and replayed.  max_iter >= 1; one solve runs at most
 * 2 000 000 iterations (rounded up to a multiple of Nwiggles), the length of
 * the device-side energy history. */
int wsim_solver_solve(wsim_solver* s, int64_t max_iter, double energy_threshold, int64_t batch,
                      const char* callback_type, double callback_interval, wsim_callback cb,
                      void* user);
/* Per-rhs outcome: iterations run, converged flag (1 converged, 0 stopped at
 * the cap, -1 stopped by the failure exit), diff_energy history (up to cap
 * entries) and the every-iteration ROI energy (may be null).
 * Failure exit (no counterpart in the reference, which would run such a
 * solve to its cap): the control step stops a rhs whose ROI increment energy
 * is not finite (NaN or Inf in gamma or the source, overflow), or has grown
 * over each of the last two wiggle cycles to above 100 times the first
 * iteration's energy (or the fixed reference energy) -- a diverging series,
 * e.g. with epsilon forced below the convergence bound (WaveSimBase.m:63).
 * The solve then returns WSIM_ERR_DIVERGED. */
int wsim_solver_result(wsim_solver* s, int rhs, int64_t* n_iter, int* converged, double* diff_energy,
                       double* energy_every, int64_t cap);
/* Copy E' (primed, full grid, N[0]*N[1]*N[2]*N[3] entries) of one rhs to the host. */
int wsim_solver_get_field(wsim_solver* s, int rhs, wsim_cfloat* E_out);
/* Device pointers: which = 0 E', 1 dE', 2 gamma, 3 work field T, 4 scaled source. */
int wsim_solver_device_ptr(wsim_solver* s, int which, int rhs, void** dptr);
/* One application of the propagation operator of wiggle w (WaveSim.m:15-36 /
 * WaveSimVector.m:19-51) to a host field; in and out hold N[0..3] entries. */
int wsim_solver_propagate(wsim_solver* s, int wiggle, const wsim_cfloat* in, wsim_cfloat* out);
/* Average duration (ms) of each kernel of one iteration, timed with HIP events
 * on the solver's stream over `reps` iterations; bytes[i] = its algorithmic HBM
 * bytes per launch; names[i] gets a static string.  With D >= 2 axes the last
 * pass also runs the next iteration's first forward pass ("pass_inv_mix_fwd")
 * and the separate first pass runs only once per solve (not listed). */
int wsim_solver_profile(wsim_solver* s, int reps, int max_kernels, double* ms, double* bytes,
                        const char** names, int* nkernels);
/* FFT plan of every non-trivial axis, as text "y:5,4,16/16/512;x:5,16,4/16/0/xcd;..."
 * (radices of the Stockham stages in forward order / workgroup columns,
 * 0 = widest that fits / optional workgroup size, a multiple of 64 up to
 * 512 (strided axes: also a multiple of the workgroup columns), 0 = default /
 * optional "xcd": each of the 8 XCDs takes one contiguous range of the
 * axis's tiles instead of every 8th tile).  Radices are 2, 3, 4, 5, 7, 8,
 * 11, 16 (16 not for 3-component fields), and, for scalar fields in launch
 * shapes that have a plan-specialised kernel (wsim_solver_fixed_launches),
 * the composite register radices 9, 12, 20, 24, 32, 36.  A plan changes
 * speed only.  A plan whose launches cannot be built returns an error and
 * leaves the previous plan in force.
 * wsim_solver_tune times whole iterations of this problem for candidate radix
 * orders, tile widths, workgroup sizes and tile ranges, axis by axis, keeps
 * the fastest and reports it (the reference picks FFT-friendly sizes only,
 * SimulationGrid.m:64-80; the per-size plan choice has no counterpart). */
int wsim_solver_tune(wsim_solver* s, int reps, char* plan_out, int plan_cap, double* ms_per_iteration);
int wsim_solver_get_plan(wsim_solver* s, char* plan_out, int plan_cap);
int wsim_solver_set_plan(wsim_solver* s, const char* plan);
/* Number of the pass launches of one iteration that run kernels compiled with
 * the FFT plan as compile-time constants (built for the shipped tuned plans;
 * other shapes run the run-time-plan kernels, with the same results; the
 * environment variable WSIM_FIXED=0 turns them off).  No reference
 * counterpart (a property of this build's kernels). */
int wsim_solver_fixed_launches(wsim_solver* s);
/* A new medium on the same grid and wiggle count: gamma, epsilon, k02e and
 * the wiggle tables are replaced (as a new WaveSim would compute them,
 * WaveSimBase.m:40-82); every rhs loses its source, set them again.  Lets one
 * solver (and its device buffers) serve a sequence of media. */
int wsim_solver_set_medium(wsim_solver* s, const wsim_cfloat* gamma, float epsilon, wsim_cfloat k02e,
                           const wsim_wiggle* wiggles);
/* Warm start (no counterpart in the reference): the next solves start rhs
 * from the field E0 (unprimed, N[0..3] entries, host or device memory)
 * instead of 0.  With W0 = gamma E0: E' starts at W0 and during the first Nw
 * iterations, alongside the source portions i/eps/Nw S of the reference loop
 * (WaveSimBase.m:171-172), the correction W0/Nw enters the increment
 * recurrence of WaveSimBase.m:96-124 through each wiggle in turn:
 *     dE'_k = M_w (dE'_{k-1} + W0/Nw) - W0/Nw + B_w S/Nw,
 * so E0 = 0 is exactly the reference loop and a good E0 leaves only small
 * increments.  Rhs without a warm start have W0 = 0.
 * wsim_solver_clear_warm_start returns to the reference start. */
int wsim_solver_set_warm_start(wsim_solver* s, int rhs, const wsim_cfloat* E0, int on_device);
int wsim_solver_clear_warm_start(wsim_solver* s);
/* Stop-test denominator of one rhs (default <= 0: diff_energy(1) as in
 * Simulation.m:198).  A fixed value makes the relative residual independent
 * of the start, for comparing warm and zero starts at equal accuracy. */
int wsim_solver_set_reference_energy(wsim_solver* s, int rhs, double energy);
int wsim_solver_destroy(wsim_solver* s);

/* ---- complex128 solvers ---------------------------------------------------
 * The reference runs in complex128 when single_precision is false
 * (Simulation.m:31, data_array Simulation.m:221-252): gamma, the wiggle
 * tables, the sources and the fields are double, and the loop of
 * WaveSimBase.m:157-197 runs in double on the GPU (its MATLAB branch; the
 * MEX of wsim_run_algorithm takes single-precision arrays only,
 * WaveSimBase.m:139-146).  These entry points create such a solver; solve,
 * result, device_ptr, profile, tune, plans, reference energy, groups and
 * destroy are the wsim_solver_* functions above (device pointers then address
 * complex128 elements).  The complex64 set_source / get_field / propagate /
 * set_medium / set_warm_start refuse a complex128 solver (WSIM_ERR_ARG or
 * WSIM_ERR_UNSUPPORTED); new media and warm starts are complex64 only.
 * Grids: at most 2^28 voxels (one complex128 component below 4 GiB). */
typedef struct {
  const double* pxe;
  const double* pye;
  const double* pze;
  const wsim_cdouble* gx;
  const wsim_cdouble* gy;
  const wsim_cdouble* gz;
} wsim_wiggle_f64;

typedef struct {
  double pos[4];
  int64_t size[4];
  const wsim_cdouble* values;
} wsim_source_patch_f64;

typedef struct {
  int64_t N[4];         /* [Ny Nx Nz Npol] */
  int nrhs;
  int nwiggles;
  double epsilon;
  wsim_cdouble k02e;
  int64_t roi[2][4];    /* 1-based inclusive [start; end] of [y x z pol] */
  int energy_interval;
} wsim_solver_desc_f64;

int wsim_solver_create_f64(wsim_solver** out, const wsim_solver_desc_f64* desc, const wsim_cdouble* gamma,
                           const wsim_wiggle_f64* wiggles);
/* patches scaled by i/eps/Nw in double (WaveSimBase.m:171-172) */
int wsim_solver_set_source_f64(wsim_solver* s, int rhs, int npatches, const wsim_source_patch_f64* patches);
int wsim_solver_get_field_f64(wsim_solver* s, int rhs, wsim_cdouble* E_out);
int wsim_solver_propagate_f64(wsim_solver* s, int wiggle, const wsim_cdouble* in, wsim_cdouble* out);
/* bytes of one complex element of the solver's fields: 8 (complex64) or 16 (complex128) */
int wsim_solver_element_size(wsim_solver* s);

/* ---- groups of independent problems -------------------------------------
 * Replaces a loop of exec calls over different simulations (the reference's
 * way of running several problems: examples/pulse2d.m:31-40,
 * examples/phase_conjugation.m:48-56) by one schedule: the solvers' (any
 * grid sizes, sources, media) passes run as shared launches, each problem
 * owning a range of workgroups, with one control launch for all problems per
 * iteration.  Every problem keeps its own state, stop test and iteration
 * count, and reads back through wsim_solver_result / wsim_solver_get_field
 * exactly as after wsim_solver_solve (identical results).  The group
 * references the solvers: destroy it before them.  Solvers must be on one
 * device; a solver may belong to several groups but not twice to one. */
typedef struct wsim_group wsim_group;
int wsim_group_create(wsim_group** out, wsim_solver* const* solvers, int n);
/* max_iter[i], energy_threshold[i]: as wsim_solver_solve, per solver */
int wsim_group_solve(wsim_group* g, const int64_t* max_iter, const double* energy_threshold, int64_t batch);
int wsim_group_destroy(wsim_group* g);

/* ---------------------------------------------------------------------------
 * Problem set-up without MATLAB or Python: the constructor of WaveSim /
 * WaveSimVector (WaveSimBase.m:40-82 with Medium.m:15-72, SimulationGrid.m:
 * 20-80, apply_edge_filters WaveSimBase.m:217-258, compute_wiggles /
 * wiggle_descriptor WaveSimBase.m:272-316, wiggle_perm.m:1-29), from a
 * refractive-index map to what wsim_solver_create takes, with the reference's
 * double-precision arithmetic (results cast to single as data_array does,
 * Simulation.m:221-252).  Only the ARL boundary (the reference's default;
 * the PBL types are deprecated, Medium.m:86).
 *
 * wsim_problem_query gives the grid, ROI and wiggle count; the caller then
 * allocates gamma (N[0]*N[1]*N[2] values, column-major [y x z]) and, per
 * wiggle, pxe/gx with N[1], pye/gy with N[0] and pze/gz with N[2] entries, and
 * wsim_problem_build fills them and the rest of info.  The result of
 * build is the solver_desc fields N, nwiggles, epsilon, k02e and roi. */
typedef struct {
  int64_t shape[3];           /* refractive-index map [ny nx nz] (1 for absent axes) */
  double pixel_size;          /* opt.pixel_size */
  double lambda;              /* opt.lambda (0: the reference's default 1) */
  int64_t boundary_widths[3]; /* opt.boundary_widths in pixels; 0 = periodic axis */
  double epsilon;             /* > 0: forced (opt.epsilon); 0: calculate_epsilon */
  double epsilonmin;          /* opt.epsilonmin (0: 3) */
  int acc[3];                 /* ACC per axis (1/0); all three -1: on for the non-periodic axes */
  int vector;                 /* 0: WaveSim, 1: WaveSimVector (3 field components) */
  int has_roi;                /* 1: roi below, 1-based inclusive [y x z pol] within the map */
  int64_t roi[2][4];
} wsim_problem_desc;

typedef struct {
  int64_t N[4];               /* grid [Ny Nx Nz Npol] */
  int64_t roi[2][4];          /* 1-based inclusive within the grid */
  int64_t Bl[3], Br[3];       /* padding before / after the map per axis (Medium.m:67-68) */
  int nwiggles;
  float epsilon;              /* single precision, as the loop uses it */
  wsim_cfloat k02e;           /* k0c^2/epsilon + i */
  double k0c;
  double iterations_per_cycle;
} wsim_problem_info;

/* Wiggle tables to be filled (the same arrays as wsim_wiggle, writable). */
typedef struct {
  float* pxe;
  float* pye;
  float* pze;
  wsim_cfloat* gx;
  wsim_cfloat* gy;
  wsim_cfloat* gz;
} wsim_wiggle_out;

int wsim_problem_query(const wsim_problem_desc* desc, wsim_problem_info* info);
/* n: the refractive-index map (shape[0]*shape[1]*shape[2] complex values in
 * double precision, as MATLAB holds it, column-major); imaginary parts must
 * be >= 0 (no gain, Medium.m:44). */
int wsim_problem_build(const wsim_problem_desc* desc, const wsim_cdouble* n, wsim_problem_info* info,
                       wsim_cfloat* gamma, wsim_wiggle_out* wiggles);

const char* wsim_last_error(void);
int wsim_abi_version(void);
/* Host-side self checks of the index arithmetic the kernels rely on. */
int wsim_selftest(void);

#ifdef __cplusplus
}
#endif
#endif /* WAVESIM_MBS_H */
