/* State, allocator injection and error reporting (libbfm/src/state.c). */

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "internal.h"

static _Thread_local char err_msg[512];

int bfm_state_create(bfm_state_t* state) {
	memset(state, 0, sizeof *state);
	state->alloc = malloc;
	state->realloc = realloc;
	state->free = free;
	return 0;
}

int bfm_state_destroy(bfm_state_t* state) {
	(void) state;
	return 0;
}

int bfm_set_alloc(bfm_state_t* state, bfm_alloc_t alloc) {
	state->alloc = alloc;
	return 0;
}

int bfm_set_realloc(bfm_state_t* state, bfm_realloc_t realloc_fn) {
	state->realloc = realloc_fn;
	return 0;
}

int bfm_set_free(bfm_state_t* state, bfm_free_t free_fn) {
	state->free = free_fn;
	return 0;
}

void bfm_set_error(bfm_state_t* state, char const* file, char const* func, size_t line, char const* fmt, ...) {
	if (!state) {
		return;
	}
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(err_msg, sizeof err_msg, fmt, ap);
	va_end(ap);
	state->err.has = true;
	state->err.msg = err_msg;
	state->err.file = file;
	state->err.func = func;
	state->err.line = line;
}

int bfm_err_print(bfm_state_t* state) {
	bfm_err_t const* err = &state->err;
	if (err->has) {
		printf("[BFM %s:%zu (%s)] %s\n", err->file, err->line, err->func, err->msg);
	}
	return 0;
}
