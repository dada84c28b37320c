#pragma once

/* Matrices (libbfm/src/bfm/matrix.h:10-127).
 *
 * FULL and BAND are host-resident, as in the reference: FULL is m*m dense,
 * BAND stores row i's entries j in [i-k, i+k] at data[i*2k + j]
 * (libbfm/src/matrix.c:208).
 *
 * BSR is appended by this implementation: a planar-elasticity stiffness
 * matrix resident in GPU memory as 2x2 blocks. Its union member has the same
 * size as the band member, so sizeof(bfm_matrix_t) is unchanged and callers
 * that allocate the struct themselves (cffi, ctypes) keep working.
 * bfm_matrix_get reads one entry back from the device; bfm_matrix_solve runs
 * the device conjugate-gradient solver. */

#include <math.h>
#include <stdbool.h>
#include <sys/types.h>

#include <bfm/bfm.h>
#include <bfm/math.h>

typedef enum {
	BFM_MATRIX_KIND_FULL,
	BFM_MATRIX_KIND_BAND,
	BFM_MATRIX_KIND_BSR,
} bfm_matrix_kind_t;

typedef enum {
	BFM_MATRIX_MAJOR_ROW,
	BFM_MATRIX_MAJOR_COLUMN,
} bfm_matrix_major_t;

typedef struct {
	double* data;
} bfm_matrix_full_t;

typedef struct {
	size_t k;
	double* data;
} bfm_matrix_band_t;

typedef struct {
	size_t nnzb;
	void* handle;
} bfm_matrix_bsr_t;

typedef struct {
	bfm_state_t* state;

	bfm_matrix_kind_t kind;
	bfm_matrix_major_t major;

	size_t m;

	union {
		bfm_matrix_full_t full;
		bfm_matrix_band_t band;
		bfm_matrix_bsr_t bsr;
	};
} bfm_matrix_t;

/* Dense m x m matrix, zero-initialised (matrix.h:54). 0 on success, -1 on failure. */
int bfm_matrix_full_create(bfm_matrix_t* matrix, bfm_state_t* state, bfm_matrix_major_t major, size_t m);

/* Band matrix with half-bandwidth k, zero-initialised (matrix.h:66). */
int bfm_matrix_band_create(bfm_matrix_t* matrix, bfm_state_t* state, bfm_matrix_major_t major, size_t m, size_t k);

int bfm_matrix_copy(bfm_matrix_t* matrix, bfm_matrix_t* src);

int bfm_matrix_destroy(bfm_matrix_t* matrix);

/* Entry (i, j); NaN when out of range (matrix.c:25-28). */
double bfm_matrix_get(bfm_matrix_t* matrix, size_t i, size_t j);

int bfm_matrix_set(bfm_matrix_t* matrix, size_t i, size_t j, double val);

int bfm_matrix_add(bfm_matrix_t* matrix, size_t i, size_t j, double val);

size_t bfm_matrix_bandwidth(bfm_matrix_t* matrix);

/* In-place LU without pivoting (matrix.c:73-121, 253-302). */
int bfm_matrix_lu(bfm_matrix_t* matrix);

/* Solve LUx = y in place on y (matrix.c:123-174, 351-404). */
int bfm_matrix_lu_solve(bfm_matrix_t* matrix, bfm_vec_t* y);

/* Solve Ax = y in place on y (matrix.c:531-541). */
int bfm_matrix_solve(bfm_matrix_t* matrix, bfm_vec_t* y);
