#pragma once

/* Extensions of this implementation to libbfm's API (no reference
 * counterpart): solver settings and statistics of bfm_sim_run, which solves
 * on the GPU with preconditioned conjugate gradients instead of the
 * reference's dense band LU.
 *
 * Defaults: point Jacobi, FP64, rtol 1e-10, max_iter 0 (= 20 x DOFs).
 * bfm_sim_run solves several instances with constant body forces in one
 * batched launch when the solver is point Jacobi or block Jacobi (instances
 * of at most 32768 nodes) or multigrid (any size), in FP64.
 * A run that stops at the rounding floor of the residual (above rtol but no
 * longer decreasing) is accepted: it is the attainable accuracy. */

#include <bfm/sim.h>

typedef struct {
	int precond;     /* 0 point Jacobi, 1 2x2 block Jacobi, 2 FP32 multigrid V-cycle
	                    inside FP64 flexible CG, 3 learned network (flexible CG),
	                    4 MINRES with |diagonal| Jacobi, 5 multigrid V-cycle plus
	                    the learned network's correction (flexible CG);
	                    axisymmetric sims and systems always use 4 (their K is
	                    indefinite) */
	int precision;   /* 0 FP64, 1 FP32 matrix inside FP64 refinement */
	double rtol;     /* ||b - K u|| / ||b|| target, FP64 */
	size_t max_iter; /* 0 = 20 x DOFs */
} bfm_solver_t;

typedef struct {
	int status;            /* 0 converged, 1 stopped at the rounding floor */
	size_t iterations;
	double rel_residual;   /* final ||b - K u|| / ||b||, FP64 */
	double solve_ms;
	double total_ms;       /* upload, renumbering, assembly, conditions, solve */
} bfm_solve_stats_t;

int bfm_sim_set_solver(bfm_sim_t* sim, bfm_solver_t const* solver);
int bfm_sim_get_solver(bfm_sim_t* sim, bfm_solver_t* solver);

/* Parameters of the learned preconditioner (precond 3 or 5), copied: the FP32
 * vector of bfmx_model_get_params (n_params must equal bfmx_model_n_params).
 * NULL clears them. */
int bfm_sim_set_model(bfm_sim_t* sim, float const* params, size_t n_params);

/* Statistics of instance `index` of the last bfm_sim_run. */
int bfm_sim_get_stats(bfm_sim_t* sim, size_t index, bfm_solve_stats_t* stats);
