/* Internal helpers of the libbfm host layer. */
#pragma once

#include <bfm/ez.h>
#include <bfm/gpu.h>
#include <bfm/sim.h>
#include <bfm/system.h>
#include <bfmx.h>

#define BFM_ERROR(state, ...) bfm_set_error((state), __FILE__, __func__, __LINE__, __VA_ARGS__)

/* Records a failure in state->err (the message lives in thread-local
 * storage owned by the library). */
void bfm_set_error(bfm_state_t* state, char const* file, char const* func, size_t line, char const* fmt, ...);

/* Solver settings attached to a sim (defaults when none were set). */
bfm_solver_t bfm_sim_solver(bfm_sim_t const* sim);
void bfm_sim_forget(bfm_sim_t const* sim);
int bfm_sim_store_stats(bfm_sim_t const* sim, size_t n, bfm_solve_stats_t const* stats);

/* The object behind bfm_matrix_t.bsr.handle: the device system and the DOF
 * order the host sees (renumbered by bfm_system_renumber or original). */
typedef struct {
	bfmx_system_t* sys;
	size_t* perm;      /* perm[original] = exposed, NULL when not renumbered */
	size_t* inv_perm;  /* inv_perm[exposed] = original */
	bfm_solver_t solver;
	bfmx_stats_t stats;
} bfm_bsr_t;

double bfm_bsr_get(bfm_matrix_t* A, size_t i, size_t j);
int bfm_bsr_solve(bfm_matrix_t* A, bfm_vec_t* y);
size_t bfm_bsr_bandwidth(bfm_matrix_t* A);
void bfm_bsr_destroy(bfm_matrix_t* A);

/* Default solver settings of the host layer. */
bfm_solver_t bfm_solver_default(void);

/* Runs the device solver with the host layer's policy: a stop at the
 * rounding floor of the residual is accepted (return 1), breakdown and
 * iteration limits fail (-1). */
int bfm_device_solve(bfm_state_t* state, bfmx_system_t* sys, double const* b, double* u, bfm_solver_t const* solver,
                     bfmx_stats_t* stats);

/* Builds the device system of an instance: mesh upload and renumbering,
 * assembly with the summed body forces, conditions in list order.
 * b_out (2 n_nodes, may be NULL) receives the condition-applied load. */
int bfm_planar_device_system(bfmx_system_t** out, bfm_instance_t* instance, size_t n_forces,
                             bfm_force_t** forces, int stress, double* b_out);
