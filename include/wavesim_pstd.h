/* wavesim_pstd.h -- C ABI of the MI355X pseudospectral time-domain (PSTD)
 * solver, the reference's comparison method (PSTD.m), on the axis-pass FFT
 * kernels of the MBS solver (include/wavesim_mbs.h: error codes, wsim_cfloat,
 * wsim_source_patch, wsim_last_error, supported grid sizes).
 *
 * One step (PSTD.m:108-111):  E_next = c2 E + c1 E_prev + c3 (lap E + A S),
 * lap E = ifftn(-|p|^2 fftn(E)), A = the CW source amplitude of the step
 * (PSTD.default_source, PSTD.m:130-137: exp(-2 pi i N) ramped up over 100
 * optical cycles, N = step / iterations_per_cycle).  The step energy
 * |crop(E_next - E exp(i(angle A - angle A_prev)))|^2 / |A|^2 is refreshed as
 * Simulation.next does (steps 1, 2, 2+interval, ...), and the run stops when
 * real(A) > 0.9 and the energy relative to step 1's is below the threshold,
 * or at max_iterations (Simulation.m:193-217, PSTD.m:113-121).  Scalar fields
 * only, as in the reference.  No counterpart entry point exists in the
 * reference (PSTD.run_algorithm is MATLAB code); the arguments are what its
 * constructor computes (PSTD.m:40-69). */
#ifndef WAVESIM_PSTD_H
#define WAVESIM_PSTD_H

#include <stdint.h>

#include "wavesim_mbs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wsim_pstd wsim_pstd;

typedef struct {
  int64_t N[3];                 /* grid [Ny Nx Nz] */
  int64_t roi[2][4];            /* 1-based inclusive [start; end] of [y x z pol] (pol 1..1) */
  double iterations_per_cycle;  /* lambda / dt (PSTD.m:53) */
  int energy_interval;          /* min(energy_calculation_interval, callback_interval) (Simulation.m:135) */
} wsim_pstd_desc;

/* c1, c2, c3: Ny*Nx*Nz coefficients (PSTD.m:57-62), column-major [y x z];
 * py, px, pz: the grid's Fourier coordinates (SimulationGrid py_range,
 * px_range, pz_range; the Laplacian is -(px^2 + py^2 + pz^2), PSTD.m:65-68). */
int wsim_pstd_create(wsim_pstd** out, const wsim_pstd_desc* desc, const float* c1, const float* c2,
                     const float* c3, const float* py, const float* px, const float* pz);
/* The source after Simulation.exec shifted and cropped it (Simulation.m:154-156). */
int wsim_pstd_set_source(wsim_pstd* p, int npatches, const wsim_source_patch* patches);
/* Steps from E = E_prev = 0 until the stop (max_iterations <= 0 or inf: no
 * cap but the 2 000 000-step history).  *n_iter = steps run. */
int wsim_pstd_run(wsim_pstd* p, double max_iterations, double energy_threshold, int64_t* n_iter,
                  int* converged);
/* E after the last step on the full grid (before the reference's final phase
 * factor exp(-i angle A_next), PSTD.m:125-126) and the step energies. */
int wsim_pstd_result(wsim_pstd* p, wsim_cfloat* E_out, double* diff_energy, int64_t cap);
int wsim_pstd_destroy(wsim_pstd* p);
/* Launches of one step that run plan-specialised pass kernels (the grid's
 * shipped plan compiled with its FFT plan as constants; no counterpart in the
 * reference, whose PSTD.m calls fftn/ifftn): >= 0, or a negative error code. */
int wsim_pstd_fixed_launches(wsim_pstd* p);

#ifdef __cplusplus
}
#endif
#endif /* WAVESIM_PSTD_H */
