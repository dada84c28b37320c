/* Simulations (libbfm/src/sim.c).
 *
 * bfm_sim_run solves every instance on the GPU: upload and renumbering,
 * assembly, conditions in list order, preconditioned CG to the solver's
 * tolerance, displacements written to instance->effects in the original
 * node order (sim.c:103-135 with the dense band LU replaced). Unlike the
 * reference, which ignores its solver's status (sim.c:122), a breakdown or
 * an unconverged solve makes the run fail with a message in state->err.
 *
 * Solver settings and per-instance statistics are kept beside the sim in a
 * small registry keyed by its address, so the bfm_sim_t layout (and every
 * caller that allocates it) is unchanged. */

#include <pthread.h>
#include <string.h>
#include <time.h>

#include "internal.h"

typedef struct {
	bfm_sim_t const* sim;
	bfm_solver_t solver;
	bfm_solve_stats_t* stats;
	size_t n_stats;
} sim_entry_t;

static pthread_mutex_t reg_lock = PTHREAD_MUTEX_INITIALIZER;
static sim_entry_t* reg;
static size_t reg_n, reg_cap;

static sim_entry_t* reg_find(bfm_sim_t const* sim, int create) {
	for (size_t i = 0; i < reg_n; i++) {
		if (reg[i].sim == sim) {
			return &reg[i];
		}
	}
	if (!create) {
		return NULL;
	}
	if (reg_n == reg_cap) {
		size_t const cap = reg_cap ? 2 * reg_cap : 16;
		sim_entry_t* grown = realloc(reg, cap * sizeof *grown);
		if (!grown) {
			return NULL;
		}
		reg = grown;
		reg_cap = cap;
	}
	sim_entry_t* e = &reg[reg_n++];
	memset(e, 0, sizeof *e);
	e->sim = sim;
	e->solver = bfm_solver_default();
	return e;
}

bfm_solver_t bfm_sim_solver(bfm_sim_t const* sim) {
	pthread_mutex_lock(&reg_lock);
	sim_entry_t const* e = reg_find(sim, 0);
	bfm_solver_t const s = e ? e->solver : bfm_solver_default();
	pthread_mutex_unlock(&reg_lock);
	return s;
}

void bfm_sim_forget(bfm_sim_t const* sim) {
	pthread_mutex_lock(&reg_lock);
	for (size_t i = 0; i < reg_n; i++) {
		if (reg[i].sim == sim) {
			free(reg[i].stats);
			reg[i] = reg[--reg_n];
			break;
		}
	}
	pthread_mutex_unlock(&reg_lock);
}

int bfm_sim_store_stats(bfm_sim_t const* sim, size_t n, bfm_solve_stats_t const* stats) {
	pthread_mutex_lock(&reg_lock);
	sim_entry_t* e = reg_find(sim, 1);
	int rv = -1;
	if (e) {
		bfm_solve_stats_t* copy = malloc((n ? n : 1) * sizeof *copy);
		if (copy) {
			memcpy(copy, stats, n * sizeof *copy);
			free(e->stats);
			e->stats = copy;
			e->n_stats = n;
			rv = 0;
		}
	}
	pthread_mutex_unlock(&reg_lock);
	return rv;
}

int bfm_sim_set_solver(bfm_sim_t* sim, bfm_solver_t const* solver) {
	if (!(solver->rtol > 0) || solver->precond < 0 || solver->precond > 2 || solver->precision < 0 ||
	    solver->precision > 1) {
		BFM_ERROR(sim->state, "invalid solver settings");
		return -1;
	}
	pthread_mutex_lock(&reg_lock);
	sim_entry_t* e = reg_find(sim, 1);
	if (e) {
		e->solver = *solver;
	}
	pthread_mutex_unlock(&reg_lock);
	return e ? 0 : -1;
}

int bfm_sim_get_solver(bfm_sim_t* sim, bfm_solver_t* solver) {
	*solver = bfm_sim_solver(sim);
	return 0;
}

int bfm_sim_get_stats(bfm_sim_t* sim, size_t index, bfm_solve_stats_t* stats) {
	pthread_mutex_lock(&reg_lock);
	sim_entry_t const* e = reg_find(sim, 0);
	int rv = -1;
	if (e && index < e->n_stats) {
		*stats = e->stats[index];
		rv = 0;
	}
	pthread_mutex_unlock(&reg_lock);
	return rv;
}

int bfm_sim_create(bfm_sim_t* sim, bfm_state_t* state, bfm_sim_kind_t kind) {
	memset(sim, 0, sizeof *sim);
	sim->state = state;
	sim->kind = kind;
	return 0;
}

int bfm_sim_destroy(bfm_sim_t* sim) {
	bfm_state_t* const state = sim->state;
	if (sim->instances) {
		state->free(sim->instances);
	}
	if (sim->forces) {
		state->free(sim->forces);
	}
	sim->instances = NULL;
	sim->forces = NULL;
	bfm_sim_forget(sim);
	return 0;
}

int bfm_sim_set_n_instances(bfm_sim_t* sim, size_t n_instances) {
	bfm_state_t* const state = sim->state;
	if (sim->instances) {
		state->free(sim->instances);
	}
	sim->n_instances = n_instances;
	size_t const size = n_instances * sizeof *sim->instances;
	sim->instances = state->alloc(size ? size : 1);
	if (!sim->instances) {
		return -1;
	}
	memset(sim->instances, 0, size);
	return 0;
}

int bfm_sim_add_instance(bfm_sim_t* sim, bfm_instance_t* instance) {
	bfm_instance_t** grown = sim->state->realloc(sim->instances, (sim->n_instances + 1) * sizeof *grown);
	if (!grown) {
		return -1;
	}
	sim->instances = grown;
	sim->instances[sim->n_instances++] = instance;
	return 0;
}

int bfm_sim_set_n_forces(bfm_sim_t* sim, size_t n_forces) {
	bfm_state_t* const state = sim->state;
	if (sim->forces) {
		state->free(sim->forces);
	}
	sim->n_forces = n_forces;
	size_t const size = n_forces * sizeof *sim->forces;
	sim->forces = state->alloc(size ? size : 1);
	if (!sim->forces) {
		return -1;
	}
	memset(sim->forces, 0, size);
	return 0;
}

int bfm_sim_add_force(bfm_sim_t* sim, bfm_force_t* force) {
	bfm_force_t** grown = sim->state->realloc(sim->forces, (sim->n_forces + 1) * sizeof *grown);
	if (!grown) {
		return -1;
	}
	sim->forces = grown;
	sim->forces[sim->n_forces++] = force;
	return 0;
}

static double now_ms(void) {
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return 1e3 * (double) ts.tv_sec + 1e-6 * (double) ts.tv_nsec;
}

static int run_planar(bfm_sim_t* sim, int stress) {
	bfm_state_t* const state = sim->state;
	bfm_solver_t solver = bfm_sim_solver(sim);
	if (stress == BFMX_AXISYMMETRIC) {
		solver.precond = BFMX_PRECOND_MINRES;   /* the reference's axisymmetric K is indefinite: no CG */
	}
	bfm_solve_stats_t* stats = state->alloc((sim->n_instances ? sim->n_instances : 1) * sizeof *stats);
	if (!stats) {
		BFM_ERROR(state, "out of memory");
		return -1;
	}
	memset(stats, 0, sim->n_instances * sizeof *stats);
	int rv = 0;
	for (size_t i = 0; i < sim->n_instances && rv == 0; i++) {
		bfm_instance_t* const instance = sim->instances[i];
		double const t0 = now_ms();
		bfmx_system_t* sys = NULL;
		if (bfm_planar_device_system(&sys, instance, sim->n_forces, sim->forces, stress, NULL) < 0) {
			rv = -1;
			break;
		}
		bfmx_stats_t st;
		memset(&st, 0, sizeof st);
		int const r = bfm_device_solve(state, sys, NULL, instance->effects, &solver, &st);
		bfmx_system_destroy(sys);
		if (r < 0) {
			rv = -1;
			break;
		}
		stats[i].status = r;
		stats[i].iterations = (size_t) st.iterations;
		stats[i].rel_residual = st.rel_residual;
		stats[i].solve_ms = st.solve_ms;
		stats[i].total_ms = now_ms() - t0;
	}
	bfm_sim_store_stats(sim, sim->n_instances, stats);
	state->free(stats);
	return rv;
}

int bfm_sim_run(bfm_sim_t* sim) {
	switch (sim->kind) {
	case BFM_SIM_KIND_NONE:
		return 0;
	case BFM_SIM_KIND_PLANAR_STRAIN:
		return run_planar(sim, 0);
	case BFM_SIM_KIND_PLANAR_STRESS:
		return run_planar(sim, 1);
	case BFM_SIM_KIND_AXISYMMETRIC_STRAIN:
		return run_planar(sim, BFMX_AXISYMMETRIC);
	}
	return -1;
}
