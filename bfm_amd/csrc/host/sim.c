/* Simulations (libbfm/src/sim.c).
 *
 * bfm_sim_run solves every instance on the GPU: upload and renumbering,
 * assembly, conditions in list order, preconditioned CG to the solver's
 * tolerance, displacements written to instance->effects in the original
 * node order (sim.c:103-135 with the dense band LU replaced). Unlike the
 * reference, which ignores its solver's status (sim.c:122), a breakdown or
 * an unconverged solve makes the run fail with a message in state->err.
 *
 * Several instances with constant body forces and a point-Jacobi,
 * block-Jacobi or multigrid solver are assembled and solved together in one
 * batch (bfmx_batch_create_planar / bfmx_batch_solve) instead of one device
 * system after another.
 *
 * Solver settings, the learned preconditioner's parameters and per-instance
 * statistics are kept beside the sim in a small registry keyed by its
 * address, so the bfm_sim_t layout (and every caller that allocates it) is
 * unchanged. */

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "internal.h"

typedef struct {
	bfm_sim_t const* sim;
	bfm_solver_t solver;
	bfm_solve_stats_t* stats;
	size_t n_stats;
	float* model;           /* learned preconditioner's parameters (copy), or NULL */
	size_t n_model;
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
			free(reg[i].model);
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
	if (!(solver->rtol > 0) || solver->precond < 0 || solver->precond > BFMX_PRECOND_AMG_LEARNED ||
	    solver->precision < 0 || solver->precision > 1) {
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

int bfm_sim_set_model(bfm_sim_t* sim, float const* params, size_t n_params) {
	size_t want = 0;
	if (bfmx_model_n_params(&want) != BFMX_OK || (params && n_params != want)) {
		BFM_ERROR(sim->state, "the learned preconditioner has %zu parameters, %zu given", want, n_params);
		return -1;
	}
	float* copy = NULL;
	if (params) {
		copy = malloc(n_params * sizeof *copy);
		if (!copy) {
			BFM_ERROR(sim->state, "out of memory");
			return -1;
		}
		memcpy(copy, params, n_params * sizeof *copy);
	}
	pthread_mutex_lock(&reg_lock);
	sim_entry_t* e = reg_find(sim, 1);
	if (e) {
		free(e->model);
		e->model = copy;
		e->n_model = params ? n_params : 0;
	}
	pthread_mutex_unlock(&reg_lock);
	if (!e) {
		free(copy);
		return -1;
	}
	return 0;
}

/* The sim's learned preconditioner as a device model, or NULL with an error
 * when none was given. */
static bfmx_model_t* sim_model(bfm_sim_t const* sim) {
	bfmx_model_t* m = NULL;
	pthread_mutex_lock(&reg_lock);
	sim_entry_t const* e = reg_find(sim, 0);
	if (e && e->model && bfmx_model_create(&m, 0) == BFMX_OK && bfmx_model_set_params(m, e->model) != BFMX_OK) {
		bfmx_model_destroy(m);
		m = NULL;
	}
	int const had = e && e->model;
	pthread_mutex_unlock(&reg_lock);
	if (!m) {
		if (had) {
			BFM_ERROR(sim->state, "%s", bfmx_last_error());
		} else {
			BFM_ERROR(sim->state, "the learned preconditioner needs parameters (bfm_sim_set_model)");
		}
	}
	return m;
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

/* Largest instance (nodes) that the batched path solves with point or block
 * Jacobi: those solves run one workgroup per instance, which suits many small
 * problems, not a few large ones. */
#define BATCH_MAX_NODES 32768

/* Whether bfm_sim_run takes the batched path: several instances, the batched
 * solvers (point / block Jacobi, or multigrid for any size), FP64, a planar
 * formulation and body forces that are constant over the mesh. */
static int batchable(bfm_sim_t const* sim, int stress, bfm_solver_t const* solver) {
	if (sim->n_instances < 2 || stress == BFMX_AXISYMMETRIC || solver->precision != BFMX_PRECISION_FP64) {
		return 0;
	}
	if (solver->precond != BFMX_PRECOND_JACOBI && solver->precond != BFMX_PRECOND_BLOCK_JACOBI &&
	    solver->precond != BFMX_PRECOND_AMG) {
		return 0;
	}
	for (size_t k = 0; k < sim->n_forces; k++) {
		if (sim->forces[k]->kind == BFM_FORCE_KIND_FUNKY || sim->forces[k]->dim != 2) {
			return 0;
		}
	}
	for (size_t i = 0; i < sim->n_instances; i++) {
		bfm_mesh_t const* const mesh = sim->instances[i]->obj->mesh;
		if (mesh->dim != 2 || (mesh->kind != BFM_ELEM_KIND_SIMPLEX && mesh->kind != BFM_ELEM_KIND_QUAD)) {
			return 0;
		}
		if (solver->precond != BFMX_PRECOND_AMG && mesh->n_nodes > BATCH_MAX_NODES) {
			return 0;
		}
	}
	return 1;
}

/* Every instance of the sim assembled and solved in one batch
 * (bfmx_batch_create_planar + bfmx_batch_solve): the instance loop of
 * libbfm/src/sim.c:103-135 as one launch. */
static int run_batched(bfm_sim_t* sim, int stress, bfm_solver_t const* solver, bfm_solve_stats_t* stats) {
	bfm_state_t* const state = sim->state;
	size_t const n = sim->n_instances;
	double gx = 0, gy = 0;
	for (size_t k = 0; k < sim->n_forces; k++) {
		if (sim->forces[k]->kind == BFM_FORCE_KIND_LINEAR) {
			gx += sim->forces[k]->linear.force.data[0];
			gy += sim->forces[k]->linear.force.data[1];
		}
	}
	bfmx_instance_t* rec = calloc(n, sizeof *rec);
	void** owned = calloc(5 * n, sizeof *owned);   /* edge nodes, edge flags, kinds, values, flags per instance */
	bfmx_stats_t* st = calloc(n, sizeof *st);
	bfmx_batch_t* b = NULL;
	int rv = -1;
	double const t0 = now_ms();
	if (!rec || !owned || !st) {
		BFM_ERROR(state, "out of memory");
		goto out;
	}
	for (size_t i = 0; i < n; i++) {
		bfm_instance_t const* const inst = sim->instances[i];
		bfm_mesh_t const* const mesh = inst->obj->mesh;
		bfm_material_t const* const mat = inst->obj->material;
		size_t const nc = inst->n_conditions, nn = mesh->n_nodes, ne = mesh->n_edges;
		size_t* en = malloc((2 * ne + 1) * sizeof *en);
		uint8_t* eb = malloc(ne + 1);
		int32_t* ck = malloc((nc + 1) * sizeof *ck);
		double* cv = malloc((nc + 1) * sizeof *cv);
		uint8_t* cf = malloc(nc * nn + 1);
		owned[5 * i] = en;
		owned[5 * i + 1] = eb;
		owned[5 * i + 2] = ck;
		owned[5 * i + 3] = cv;
		owned[5 * i + 4] = cf;
		if (!en || !eb || !ck || !cv || !cf) {
			BFM_ERROR(state, "out of memory");
			goto out;
		}
		for (size_t e = 0; e < ne; e++) {
			en[2 * e] = mesh->edges[e].nodes[0];
			en[2 * e + 1] = mesh->edges[e].nodes[1];
			eb[e] = mesh->edges[e].elems[1] == -1;
		}
		for (size_t c = 0; c < nc; c++) {
			bfm_condition_t const* const cond = inst->conditions[c];
			ck[c] = (int32_t) cond->kind;
			cv[c] = cond->value;
			for (size_t j = 0; j < nn; j++) {
				cf[c * nn + j] = cond->nodes[j] ? 1 : 0;
			}
		}
		rec[i].kind = (int32_t) mesh->kind;
		rec[i].stress = stress;
		rec[i].n_nodes = nn;
		rec[i].coords = mesh->coords;
		rec[i].n_elems = mesh->n_elems;
		rec[i].elems = mesh->elems;
		rec[i].n_edges = ne;
		rec[i].edge_nodes = en;
		rec[i].edge_boundary = eb;
		rec[i].E = mat->E;
		rec[i].nu = mat->nu;
		rec[i].rho = mat->rho;
		rec[i].elem_E = NULL;
		rec[i].gx = gx;
		rec[i].gy = gy;
		rec[i].n_conditions = nc;
		rec[i].cond_kind = ck;
		rec[i].cond_value = cv;
		rec[i].cond_flags = cf;
	}
	if (bfmx_batch_create_planar(&b, n, rec, 1) != BFMX_OK) {
		BFM_ERROR(state, "%s", bfmx_last_error());
		goto out;
	}
	double ms = 0;
	int const rc = bfmx_batch_solve(b, solver->precond, solver->rtol, solver->max_iter, NULL, st, &ms);
	if (rc != BFMX_OK && rc != BFMX_E_MAXITER && rc != BFMX_E_STAGNATION && rc != BFMX_E_BREAKDOWN) {
		BFM_ERROR(state, "%s", bfmx_last_error());
		goto out;
	}
	for (size_t i = 0; i < n; i++) {
		bfm_instance_t* const inst = sim->instances[i];
		stats[i].iterations = (size_t) st[i].iterations;
		stats[i].rel_residual = st[i].rel_residual;
		stats[i].solve_ms = ms;
		if (st[i].status == BFMX_OK) {
			if (bfmx_batch_get_u(b, i, inst->effects) != BFMX_OK) {
				BFM_ERROR(state, "%s", bfmx_last_error());
				goto out;
			}
			continue;
		}
		/* an instance the batch left above rtol (a tolerance below its
		 * rounding floor, where the batched solvers have no stall test) is
		 * solved again on its own, with the single-system policy: a stop at
		 * the floor is accepted */
		bfmx_system_t* own = NULL;
		bfmx_stats_t one;
		memset(&one, 0, sizeof one);
		if (bfmx_batch_system(b, i, &own) != BFMX_OK) {
			BFM_ERROR(state, "%s", bfmx_last_error());
			goto out;
		}
		int const r = bfm_device_solve(state, own, NULL, inst->effects, solver, &one);
		if (r < 0) {
			goto out;
		}
		stats[i].status = r;
		stats[i].iterations += (size_t) one.iterations;
		stats[i].rel_residual = one.rel_residual;
		stats[i].solve_ms += one.solve_ms;
	}
	double const total = now_ms() - t0;
	for (size_t i = 0; i < n; i++) {
		stats[i].total_ms = total;
	}
	rv = 0;
out:
	if (b) {
		bfmx_batch_destroy(b);
	}
	if (owned) {
		for (size_t k = 0; k < 5 * n; k++) {
			free(owned[k]);
		}
	}
	free(owned);
	free(rec);
	free(st);
	return rv;
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
	if (batchable(sim, stress, &solver)) {
		int const r = run_batched(sim, stress, &solver, stats);
		if (r == 0) {
			bfm_sim_store_stats(sim, sim->n_instances, stats);
		}
		state->free(stats);
		return r;
	}
	bfmx_model_t* model = NULL;
	if ((solver.precond == BFMX_PRECOND_LEARNED || solver.precond == BFMX_PRECOND_AMG_LEARNED) &&
	    !(model = sim_model(sim))) {
		state->free(stats);
		return -1;
	}
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
		if (model && bfmx_system_set_model(sys, model) != BFMX_OK) {
			BFM_ERROR(state, "%s", bfmx_last_error());
			bfmx_system_destroy(sys);
			rv = -1;
			break;
		}
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
	if (model) {
		bfmx_model_destroy(model);
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
