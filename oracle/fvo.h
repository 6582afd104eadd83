/* The oracle: a plain-C, FP64 restatement of the reference's uniform-grid residual and time step.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker. The product (fenice_amd) never calls it.
 *
 * It follows the reference's two-pass path (face states -> Riemann solver -> flux divergence, then the
 * viscous term), which the reference's tests hold bit-identical to its fused row/plane pipelines
 * (cpp/tests/test_core.cpp:704-811, 817-1111, 1112-1196, 1669-1777). Every operation is written in the
 * reference's order and the file is compiled with -ffp-contract=off, so results are meant to be the
 * reference's bits. State layout: AoS, 5 doubles per cell, cell (k*ny + j)*nx + i.
 */
#ifndef FVO_H
#define FVO_H

#include <stddef.h>

#include "../include/fvg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dt_length of the mesh (Simulation.cpp:223-227). */
double fvo_dt_length(const fvg_config* cfg);

/* Riemann flux of one face (RiemannSolver::face_flux, Schemes.cpp:202-208): F[5], *smax. */
void fvo_face_flux(const fvg_config* cfg, const double* UL, const double* UR, int axis, double* F, double* smax);
void fvo_face_flux_batch(const fvg_config* cfg, const double* UL, const double* UR, const int* axis, size_t n, double* F,
                         double* smax);

/* Simulation::residual (Simulation.cpp:811-1354) for the uniform mesh: R (AoS, 5 per cell) and the
 * largest wave speed. half_dt > 0: the MUSCL-Hancock predictor (TimeScheme.hpp:92). *reverted (may be
 * NULL) receives the number of predicted states left as they were. Returns 0, or a negative value for a
 * configuration the oracle does not restate. */
int fvo_residual(const fvg_config* cfg, const double* U, double* R, double half_dt, double* smax, long long* reverted);
/* flux_quadrature = fourth: the residual's branch counts (faces reconstructed, admissibility fallbacks, good, pt,
 * full, pt without full, good without pt), for tests that check which branches a state exercises */
int fvo_fourth_stats(const fvg_config* cfg, const double* U, long long out[7]);

/* TimeScheme::step of cfg->time_scheme (TimeScheme.cpp:110-319) on U in place. lagged_s: the
 * MUSCL-Hancock wave speed of the previous step (in/out; 0 = none yet). Returns 0, or -1 when the
 * reference's check_progress would throw (the state is left as the step made it). */
int fvo_step(const fvg_config* cfg, double* U, double t, double t_max, double dt_max, double* lagged_s, double* dt,
             double* smax);

/* Simulation::volume_averages (SimulationRun.cpp:45-64): sequential sums / n. */
void fvo_volume_averages(const fvg_config* cfg, const double* U, double out[3]);

/* Number of threads the oracle's OpenMP loops use (0: the OpenMP default). */
void fvo_set_threads(int n);
/* 1: WENO5-Z face values take the linear weights (upwind-5) everywhere; 0: the reference's WENO5-Z (default). */
void fvo_set_weno_linear(int on);
int fvo_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
