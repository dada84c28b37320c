/* Linear systems (libbfm/src/system.c).
 *
 * bfm_system_create and bfm_system_renumber keep the reference's dense host
 * behaviour for systems built by hand (dense A, RCM, band conversion,
 * system.c:5-81). bfm_system_create_planar_strain/stress assemble on the
 * GPU instead: the instance's mesh is uploaded and renumbered, the element
 * matrices are assembled and every condition is applied there, in list
 * order, with the reference's semantics (system.c:427-529). A is then a
 * BFM_MATRIX_KIND_BSR matrix owning the device system; b is the
 * condition-applied load copied back in the original DOF order. */

#include <math.h>
#include <string.h>

#include "internal.h"

bfm_solver_t bfm_solver_default(void) {
	bfm_solver_t s = {BFMX_PRECOND_JACOBI, BFMX_PRECISION_FP64, 1e-10, 0};
	return s;
}

int bfm_system_create(bfm_system_t* system, bfm_state_t* state, size_t n) {
	memset(system, 0, sizeof *system);
	system->state = state;
	system->n = n;
	bfm_perm_create(&system->perm, state, n);
	if (bfm_matrix_full_create(&system->A, state, BFM_MATRIX_MAJOR_ROW, n) < 0) {
		return -1;
	}
	if (bfm_vec_create(&system->b, state, n) < 0) {
		bfm_matrix_destroy(&system->A);
		return -1;
	}
	return 0;
}

int bfm_system_destroy(bfm_system_t* system) {
	bfm_perm_destroy(&system->perm);
	bfm_matrix_destroy(&system->A);
	if (system->b.data) {
		bfm_vec_destroy(&system->b);
	}
	return 0;
}

int bfm_system_renumber(bfm_system_t* system) {
	bfm_state_t* const state = system->state;
	if (bfm_perm_rcm(&system->perm, &system->A) < 0) {
		return -1;
	}
	if (system->A.kind == BFM_MATRIX_KIND_BSR) {
		/* the device already holds the matrix in this order: expose it */
		bfm_bsr_t* const h = system->A.bsr.handle;
		size_t const m = system->n;
		if (!h->perm) {
			h->perm = state->alloc(m * sizeof *h->perm + 1);
			h->inv_perm = state->alloc(m * sizeof *h->inv_perm + 1);
			if (!h->perm || !h->inv_perm) {
				BFM_ERROR(state, "cannot allocate the exposed permutation");
				return -1;
			}
		}
		memcpy(h->perm, system->perm.perm, m * sizeof *h->perm);
		memcpy(h->inv_perm, system->perm.inv_perm, m * sizeof *h->inv_perm);
		return bfm_perm_perm_vec(&system->perm, &system->b, false);
	}
	if (bfm_perm_perm_matrix(&system->perm, &system->A, false) < 0 ||
	    bfm_perm_perm_vec(&system->perm, &system->b, false) < 0) {
		return -1;
	}
	size_t const k = bfm_matrix_bandwidth(&system->A);
	bfm_matrix_t band;
	if (bfm_matrix_band_create(&band, state, system->A.major, system->A.m, k) < 0) {
		return -1;
	}
	if (bfm_matrix_copy(&band, &system->A) < 0) {
		bfm_matrix_destroy(&band);
		return -1;
	}
	bfm_matrix_destroy(&system->A);
	system->A = band;
	return 0;
}

/* ------------------------------------------------------------ device path */

/* Body force per unit mass at every node: the sum of the forces, linear ones
 * as constants, callbacks evaluated once per node position (the reference
 * evaluates them at the node positions too, once per element and Gauss
 * point, system.c:190-203). node_force is NULL when every force is constant;
 * then (gx, gy) carries their sum. */
static int body_forces(bfm_state_t* state, bfm_mesh_t* mesh, size_t n_forces, bfm_force_t** forces,
                       double** node_force, double* gx, double* gy) {
	*gx = *gy = 0;
	*node_force = NULL;
	int funky = 0;
	for (size_t k = 0; k < n_forces; k++) {
		if (forces[k]->dim != 2) {
			BFM_ERROR(state, "force %zu has dimension %zu, the planar path needs 2", k, forces[k]->dim);
			return -1;
		}
		if (forces[k]->kind == BFM_FORCE_KIND_LINEAR) {
			*gx += forces[k]->linear.force.data[0];
			*gy += forces[k]->linear.force.data[1];
		} else if (forces[k]->kind == BFM_FORCE_KIND_FUNKY) {
			funky = 1;
		}
	}
	if (!funky) {
		return 0;
	}
	size_t const n = mesh->n_nodes;
	double* f = state->alloc(2 * n * sizeof *f + 1);
	if (!f) {
		BFM_ERROR(state, "cannot allocate node forces");
		return -1;
	}
	double pos_d[2], val_d[2];
	bfm_vec_t pos = {state, 2, pos_d}, val = {state, 2, val_d};
	for (size_t i = 0; i < n; i++) {
		f[2 * i] = *gx;
		f[2 * i + 1] = *gy;
		pos_d[0] = mesh->coords[2 * i];
		pos_d[1] = mesh->coords[2 * i + 1];
		for (size_t k = 0; k < n_forces; k++) {
			if (forces[k]->kind != BFM_FORCE_KIND_FUNKY) {
				continue;
			}
			val_d[0] = val_d[1] = 0;
			forces[k]->funky.func(forces[k], &pos, &val, forces[k]->funky.data);
			f[2 * i] += val_d[0];
			f[2 * i + 1] += val_d[1];
		}
	}
	*node_force = f;
	return 0;
}

#define BFMX_TRY(state, call)                                        \
	do {                                                             \
		if ((call) != BFMX_OK) {                                     \
			BFM_ERROR((state), "%s", bfmx_last_error());             \
			goto fail;                                               \
		}                                                            \
	} while (0)

int bfm_planar_device_system(bfmx_system_t** out, bfm_instance_t* instance, size_t n_forces, bfm_force_t** forces,
                             int stress, double* b_out) {
	bfm_state_t* const state = instance->state;
	bfm_obj_t* const obj = instance->obj;
	bfm_mesh_t* const mesh = obj->mesh;
	bfm_material_t* const material = obj->material;
	*out = NULL;
	if (mesh->dim != 2) {
		BFM_ERROR(state, "planar systems need a 2-D mesh (dim %zu)", mesh->dim);
		return -1;
	}
	if (mesh->kind != BFM_ELEM_KIND_SIMPLEX && mesh->kind != BFM_ELEM_KIND_QUAD) {
		BFM_ERROR(state, "planar systems need P1 triangles or Q1 quads (kind %d)", (int) mesh->kind);
		return -1;
	}
	bfmx_system_t* sys = NULL;
	double* node_force = NULL;
	size_t* edge_nodes = NULL;
	uint8_t* edge_bnd = NULL;
	uint8_t* flags = NULL;
	double gx, gy;
	if (body_forces(state, mesh, n_forces, forces, &node_force, &gx, &gy) < 0) {
		return -1;
	}
	BFMX_TRY(state, bfmx_system_create(&sys, (int) mesh->kind, mesh->n_nodes, mesh->coords, mesh->n_elems,
	                                   mesh->elems, 1));
	if (mesh->n_edges) {
		edge_nodes = state->alloc(2 * mesh->n_edges * sizeof *edge_nodes);
		edge_bnd = state->alloc(mesh->n_edges);
		if (!edge_nodes || !edge_bnd) {
			BFM_ERROR(state, "cannot allocate the edge list");
			goto fail;
		}
		for (size_t e = 0; e < mesh->n_edges; e++) {
			edge_nodes[2 * e] = mesh->edges[e].nodes[0];
			edge_nodes[2 * e + 1] = mesh->edges[e].nodes[1];
			edge_bnd[e] = mesh->edges[e].elems[1] == -1;
		}
		BFMX_TRY(state, bfmx_system_set_edges(sys, mesh->n_edges, edge_nodes, edge_bnd));
	}
	BFMX_TRY(state, bfmx_system_assemble(sys, stress, material->E, material->nu, material->rho, NULL, node_force,
	                                     gx, gy));
	flags = state->alloc(mesh->n_nodes + 1);
	if (!flags) {
		BFM_ERROR(state, "cannot allocate condition flags");
		goto fail;
	}
	for (size_t c = 0; c < instance->n_conditions; c++) {
		bfm_condition_t const* const cond = instance->conditions[c];
		for (size_t i = 0; i < mesh->n_nodes; i++) {
			flags[i] = cond->nodes[i] ? 1 : 0;
		}
		BFMX_TRY(state, bfmx_system_apply_condition(sys, (int) cond->kind, cond->value, flags));
	}
	if (b_out) {
		BFMX_TRY(state, bfmx_system_get_rhs(sys, b_out));
	}
	state->free(node_force);
	state->free(edge_nodes);
	state->free(edge_bnd);
	state->free(flags);
	*out = sys;
	return 0;
fail:
	bfmx_system_destroy(sys);
	state->free(node_force);
	state->free(edge_nodes);
	state->free(edge_bnd);
	state->free(flags);
	return -1;
}

int bfm_device_solve(bfm_state_t* state, bfmx_system_t* sys, double const* b, double* u, bfm_solver_t const* solver,
                     bfmx_stats_t* stats) {
	int const rc = bfmx_system_solve(sys, b, u, solver->precond, solver->precision, solver->rtol, solver->max_iter,
	                                 stats);
	if (rc == BFMX_OK) {
		return 0;
	}
	if (rc == BFMX_E_STAGNATION && bfmx_system_get_u(sys, u) == BFMX_OK) {
		return 1;
	}
	BFM_ERROR(state, "%s", bfmx_last_error());
	return -1;
}

static int create_planar(bfm_system_t* system, bfm_instance_t* instance, size_t n_forces, bfm_force_t** forces,
                         int stress) {
	bfm_state_t* const state = instance->state;
	bfm_mesh_t* const mesh = instance->obj->mesh;
	size_t const n = mesh->n_nodes * mesh->dim;
	memset(system, 0, sizeof *system);
	system->state = state;
	system->n = n;
	bfm_perm_create(&system->perm, state, n);
	if (bfm_vec_create(&system->b, state, n) < 0) {
		return -1;
	}
	bfm_bsr_t* h = state->alloc(sizeof *h);
	if (!h) {
		BFM_ERROR(state, "out of memory");
		bfm_vec_destroy(&system->b);
		return -1;
	}
	memset(h, 0, sizeof *h);
	h->solver = bfm_solver_default();
	if (stress == BFMX_AXISYMMETRIC) {
		h->solver.precond = BFMX_PRECOND_MINRES;   /* the reference's axisymmetric K is indefinite */
	}
	if (bfm_planar_device_system(&h->sys, instance, n_forces, forces, stress, system->b.data) < 0) {
		state->free(h);
		bfm_vec_destroy(&system->b);
		return -1;
	}
	system->A.state = state;
	system->A.kind = BFM_MATRIX_KIND_BSR;
	system->A.major = BFM_MATRIX_MAJOR_ROW;
	system->A.m = n;
	size_t nnzb = 0;
	bfmx_system_info(h->sys, NULL, &nnzb, NULL, NULL);
	system->A.bsr.nnzb = nnzb;
	system->A.bsr.handle = h;
	return 0;
}

int bfm_system_create_planar_strain(bfm_system_t* system, bfm_instance_t* instance, size_t n_forces,
                                    bfm_force_t** forces) {
	return create_planar(system, instance, n_forces, forces, 0);
}

int bfm_system_create_planar_stress(bfm_system_t* system, bfm_instance_t* instance, size_t n_forces,
                                    bfm_force_t** forces) {
	return create_planar(system, instance, n_forces, forces, 1);
}

/* libbfm/src/system.c:539-622: the device system with the axisymmetric
 * formulation (the reference's element formulas and condition loop, kept as
 * written), solved by MINRES. */
int bfm_system_create_axisymmetric_strain(bfm_system_t* system, bfm_instance_t* instance, size_t n_forces,
                                          bfm_force_t** forces) {
	return create_planar(system, instance, n_forces, forces, BFMX_AXISYMMETRIC);
}

/* ------------------------------------------------------------ BSR matrices */

double bfm_bsr_get(bfm_matrix_t* A, size_t i, size_t j) {
	bfm_bsr_t const* const h = A->bsr.handle;
	size_t const oi = h->perm ? h->inv_perm[i] : i;
	size_t const oj = h->perm ? h->inv_perm[j] : j;
	double v = BFM_NAN;
	if (bfmx_system_get_entry(h->sys, oi, oj, &v) != BFMX_OK) {
		BFM_ERROR(A->state, "%s", bfmx_last_error());
		return BFM_NAN;
	}
	return v;
}

int bfm_bsr_solve(bfm_matrix_t* A, bfm_vec_t* y) {
	bfm_bsr_t* const h = A->bsr.handle;
	bfm_state_t* const state = A->state;
	size_t const m = A->m;
	double* b = state->alloc(m * sizeof *b + 1);
	double* u = state->alloc(m * sizeof *u + 1);
	int rv = -1;
	if (!b || !u) {
		BFM_ERROR(state, "out of memory");
		goto out;
	}
	for (size_t i = 0; i < m; i++) {
		b[i] = h->perm ? y->data[h->perm[i]] : y->data[i];
	}
	if (bfm_device_solve(state, h->sys, b, u, &h->solver, &h->stats) < 0) {
		goto out;
	}
	for (size_t i = 0; i < m; i++) {
		if (h->perm) {
			y->data[h->perm[i]] = u[i];
		} else {
			y->data[i] = u[i];
		}
	}
	rv = 0;
out:
	state->free(b);
	state->free(u);
	return rv;
}

size_t bfm_bsr_bandwidth(bfm_matrix_t* A) {
	bfm_bsr_t const* const h = A->bsr.handle;
	size_t n = 0, nnzb = 0;
	bfmx_system_info(h->sys, &n, &nnzb, NULL, NULL);
	int64_t* rowptr = A->state->alloc((n + 1) * sizeof *rowptr);
	int32_t* col = A->state->alloc((nnzb + 1) * sizeof *col);
	size_t k = 0;
	if (rowptr && col && bfmx_system_get_bsr(h->sys, rowptr, col, NULL) == BFMX_OK) {
		for (size_t i = 0; i < n; i++) {
			for (int64_t t = rowptr[i]; t < rowptr[i + 1]; t++) {
				size_t const j = (size_t) col[t];
				size_t const d = i > j ? 2 * (i - j) + 1 : 2 * (j - i) + 1;
				k = d > k ? d : k;
			}
		}
	}
	A->state->free(rowptr);
	A->state->free(col);
	return k;
}

void bfm_bsr_destroy(bfm_matrix_t* A) {
	bfm_bsr_t* const h = A->bsr.handle;
	if (!h) {
		return;
	}
	bfmx_system_destroy(h->sys);
	A->state->free(h->perm);
	A->state->free(h->inv_perm);
	A->state->free(h);
	A->bsr.handle = NULL;
}
