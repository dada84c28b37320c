/* Host vectors (libbfm/src/vec.c). */

#include <string.h>

#include "internal.h"

int bfm_vec_create(bfm_vec_t* vec, bfm_state_t* state, size_t n) {
	vec->state = state;
	vec->n = n;
	vec->data = state->alloc(n * sizeof *vec->data);
	if (!vec->data) {
		BFM_ERROR(state, "cannot allocate a vector of %zu doubles", n);
		return -1;
	}
	memset(vec->data, 0, n * sizeof *vec->data);
	return 0;
}

int bfm_vec_copy(bfm_vec_t* vec, bfm_vec_t* src) {
	memcpy(vec->data, src->data, src->n * sizeof *vec->data);
	return 0;
}

int bfm_vec_destroy(bfm_vec_t* vec) {
	vec->state->free(vec->data);
	vec->data = NULL;
	return 0;
}
