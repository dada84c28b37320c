/* Body forces per unit mass (libbfm/src/force.c). bfm_force_eval of a
 * linear force returns 0 after copying its vector; the reference returns -1
 * there although the copy succeeded (force.c:66-73), a status its only
 * caller ignores (system.c:198). */

#include <string.h>

#include "internal.h"

int bfm_force_create(bfm_force_t* force, bfm_state_t* state, size_t dim) {
	memset(force, 0, sizeof *force);
	force->state = state;
	force->dim = dim;
	return 0;
}

int bfm_force_destroy(bfm_force_t* force) {
	if (force->kind == BFM_FORCE_KIND_LINEAR && force->linear.force.data && force->linear.force.state) {
		bfm_vec_destroy(&force->linear.force);
	}
	force->kind = BFM_FORCE_KIND_NONE;
	return 0;
}

int bfm_force_set_none(bfm_force_t* force) {
	force->kind = BFM_FORCE_KIND_NONE;
	return 0;
}

int bfm_force_set_linear(bfm_force_t* force, bfm_vec_t* vec) {
	force->kind = BFM_FORCE_KIND_LINEAR;
	memset(&force->linear, 0, sizeof force->linear);
	if (vec->n != force->dim) {
		BFM_ERROR(force->state, "force vector has %zu components, force dimension is %zu", vec->n, force->dim);
		return -1;
	}
	if (bfm_vec_create(&force->linear.force, vec->state, vec->n) < 0) {
		return -1;
	}
	return bfm_vec_copy(&force->linear.force, vec);
}

int bfm_force_set_funky(bfm_force_t* force, bfm_force_funky_func_t func, void* data) {
	force->kind = BFM_FORCE_KIND_FUNKY;
	force->funky.func = func;
	force->funky.data = data;
	return 0;
}

int bfm_force_eval(bfm_force_t* force, bfm_vec_t* pos, bfm_vec_t* force_ref) {
	if (force_ref->n != force->dim) {
		return -1;
	}
	switch (force->kind) {
	case BFM_FORCE_KIND_NONE:
		memset(force_ref->data, 0, force_ref->n * sizeof *force_ref->data);
		return 0;
	case BFM_FORCE_KIND_LINEAR:
		return bfm_vec_copy(force_ref, &force->linear.force);
	case BFM_FORCE_KIND_FUNKY:
		return force->funky.func(force, pos, force_ref, force->funky.data);
	}
	return -1;
}
