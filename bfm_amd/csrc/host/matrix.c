/* Matrices (libbfm/src/matrix.c).
 *
 * FULL and BAND stay host-resident with the reference's layouts and its
 * unpivoted LU (matrix.c:73-174, 253-302, 351-404); indices are size_t
 * throughout, so dense matrices above m = 46,340 do not overflow an int index
 * as the reference's do (matrix.c:30). BSR is the planar system resident on
 * the GPU: get reads one entry back, solve runs the device CG solver, lu is
 * a no-op and lu_solve solves. */

#include <math.h>
#include <string.h>

#include "internal.h"

/* ------------------------------------------------------------ full */

static inline size_t full_idx(bfm_matrix_t const* A, size_t i, size_t j) {
	return A->major == BFM_MATRIX_MAJOR_ROW ? i * A->m + j : i + j * A->m;
}

/* ------------------------------------------------------------ band
 * Entry (i, j), |i - j| <= k, of row-major band storage lives at j + 2k i,
 * i.e. row i occupies [i(2k+1) - k, i(2k+1) + k] (matrix.c:208). The
 * column-major variant mirrors it. */

static inline size_t band_idx(bfm_matrix_t const* A, size_t i, size_t j) {
	size_t const k2 = 2 * A->band.k;
	return A->major == BFM_MATRIX_MAJOR_ROW ? j + i * k2 : i + j * k2;
}

static inline int in_band(bfm_matrix_t const* A, size_t i, size_t j) {
	size_t const d = i > j ? i - j : j - i;
	return d <= A->band.k;
}

/* ------------------------------------------------------------ creation */

static void matrix_init(bfm_matrix_t* A, bfm_state_t* state, bfm_matrix_kind_t kind, bfm_matrix_major_t major,
                        size_t m) {
	A->state = state;
	A->kind = kind;
	A->major = major;
	A->m = m;
}

int bfm_matrix_full_create(bfm_matrix_t* matrix, bfm_state_t* state, bfm_matrix_major_t major, size_t m) {
	matrix_init(matrix, state, BFM_MATRIX_KIND_FULL, major, m);
	if (m && m > ((size_t) -1) / sizeof(double) / m) {
		BFM_ERROR(state, "a dense %zu x %zu matrix does not fit in memory", m, m);
		matrix->full.data = NULL;
		return -1;
	}
	size_t const size = m * m * sizeof *matrix->full.data;
	matrix->full.data = state->alloc(size ? size : 1);
	if (!matrix->full.data) {
		BFM_ERROR(state, "cannot allocate a dense %zu x %zu matrix", m, m);
		return -1;
	}
	memset(matrix->full.data, 0, size);
	return 0;
}

int bfm_matrix_band_create(bfm_matrix_t* matrix, bfm_state_t* state, bfm_matrix_major_t major, size_t m, size_t k) {
	matrix_init(matrix, state, BFM_MATRIX_KIND_BAND, major, m);
	matrix->band.k = k;
	size_t const size = m * (2 * k + 1) * sizeof *matrix->band.data;
	matrix->band.data = state->alloc(size ? size : 1);
	if (!matrix->band.data) {
		BFM_ERROR(state, "cannot allocate a band matrix (m %zu, k %zu)", m, k);
		return -1;
	}
	memset(matrix->band.data, 0, size);
	return 0;
}

int bfm_matrix_destroy(bfm_matrix_t* matrix) {
	switch (matrix->kind) {
	case BFM_MATRIX_KIND_FULL:
		matrix->state->free(matrix->full.data);
		matrix->full.data = NULL;
		return 0;
	case BFM_MATRIX_KIND_BAND:
		matrix->state->free(matrix->band.data);
		matrix->band.data = NULL;
		return 0;
	case BFM_MATRIX_KIND_BSR:
		bfm_bsr_destroy(matrix);
		return 0;
	}
	return -1;
}

/* ------------------------------------------------------------ access */

double bfm_matrix_get(bfm_matrix_t* matrix, size_t i, size_t j) {
	if (i >= matrix->m || j >= matrix->m) {
		return BFM_NAN;
	}
	switch (matrix->kind) {
	case BFM_MATRIX_KIND_FULL:
		return matrix->full.data[full_idx(matrix, i, j)];
	case BFM_MATRIX_KIND_BAND:
		return in_band(matrix, i, j) ? matrix->band.data[band_idx(matrix, i, j)] : 0;
	case BFM_MATRIX_KIND_BSR:
		return bfm_bsr_get(matrix, i, j);
	}
	return -1;
}

static int set_or_add(bfm_matrix_t* matrix, size_t i, size_t j, double val, int add) {
	if (i >= matrix->m || j >= matrix->m) {
		return -1;
	}
	switch (matrix->kind) {
	case BFM_MATRIX_KIND_FULL: {
		double* p = &matrix->full.data[full_idx(matrix, i, j)];
		*p = add ? *p + val : val;
		return 0;
	}
	case BFM_MATRIX_KIND_BAND: {
		if (!in_band(matrix, i, j)) {
			/* outside the band only (near-)zeros are accepted (matrix.c:221-222) */
			return fabs(val) < BFM_PIVOT_EPS ? 0 : -1;
		}
		double* p = &matrix->band.data[band_idx(matrix, i, j)];
		*p = add ? *p + val : val;
		return 0;
	}
	case BFM_MATRIX_KIND_BSR:
		BFM_ERROR(matrix->state, "entries of a device-resident stiffness matrix are read-only");
		return -1;
	}
	return -1;
}

int bfm_matrix_set(bfm_matrix_t* matrix, size_t i, size_t j, double val) {
	return set_or_add(matrix, i, j, val, 0);
}

int bfm_matrix_add(bfm_matrix_t* matrix, size_t i, size_t j, double val) {
	return set_or_add(matrix, i, j, val, 1);
}

size_t bfm_matrix_bandwidth(bfm_matrix_t* matrix) {
	switch (matrix->kind) {
	case BFM_MATRIX_KIND_FULL: {
		size_t k = 0;
		for (size_t i = 0; i < matrix->m; i++) {
			for (size_t j = 0; j < matrix->m; j++) {
				if (matrix->full.data[full_idx(matrix, i, j)] != 0) {
					size_t const d = i > j ? i - j : j - i;
					k = d > k ? d : k;
				}
			}
		}
		return k;
	}
	case BFM_MATRIX_KIND_BAND:
		return matrix->band.k;
	case BFM_MATRIX_KIND_BSR:
		return bfm_bsr_bandwidth(matrix);
	}
	return (size_t) -1;
}

int bfm_matrix_copy(bfm_matrix_t* matrix, bfm_matrix_t* src) {
	if (matrix->m != src->m) {
		return -1;
	}
	if (matrix->kind == BFM_MATRIX_KIND_BSR || src->kind == BFM_MATRIX_KIND_BSR) {
		BFM_ERROR(matrix->state, "device-resident matrices are not copied");
		return -1;
	}
	if (matrix->kind == BFM_MATRIX_KIND_FULL && src->kind == BFM_MATRIX_KIND_FULL && matrix->major == src->major) {
		memcpy(matrix->full.data, src->full.data, src->m * src->m * sizeof(double));
		return 0;
	}
	if (matrix->kind == BFM_MATRIX_KIND_BAND && src->kind == BFM_MATRIX_KIND_BAND) {
		if (matrix->band.k != src->band.k) {
			return -1;
		}
		memcpy(matrix->band.data, src->band.data, src->m * (2 * src->band.k + 1) * sizeof(double));
		return 0;
	}
	for (size_t i = 0; i < matrix->m; i++) {
		for (size_t j = 0; j < matrix->m; j++) {
			double const v = bfm_matrix_get(src, i, j);
			if (BFM_IS_NAN(v) || bfm_matrix_set(matrix, i, j, v) < 0) {
				return -1;
			}
		}
	}
	return 0;
}

/* ------------------------------------------------------------ direct solves */

static int full_lu(bfm_matrix_t* A) {
	size_t const m = A->m;
	for (size_t p = 0; p + 1 < m; p++) {
		double const pivot = A->full.data[full_idx(A, p, p)];
		if (BFM_IS_NAN(pivot) || fabs(pivot) < BFM_PIVOT_EPS) {
			return -1;
		}
		for (size_t i = p + 1; i < m; i++) {
			double const l = A->full.data[full_idx(A, i, p)] / pivot;
			A->full.data[full_idx(A, i, p)] = l;
			for (size_t j = p + 1; j < m; j++) {
				A->full.data[full_idx(A, i, j)] += -l * A->full.data[full_idx(A, p, j)];
			}
		}
	}
	return 0;
}

static int full_lu_solve(bfm_matrix_t* A, double* y) {
	size_t const m = A->m;
	for (size_t i = 0; i < m; i++) {
		for (size_t j = 0; j < i; j++) {
			y[i] -= A->full.data[full_idx(A, i, j)] * y[j];
		}
	}
	for (size_t i = m; i-- > 0;) {
		for (size_t j = i + 1; j < m; j++) {
			y[i] -= A->full.data[full_idx(A, i, j)] * y[j];
		}
		double const pivot = A->full.data[full_idx(A, i, i)];
		if (BFM_IS_NAN(pivot)) {
			return -1;
		}
		y[i] /= pivot;
	}
	return 0;
}

static int band_lu(bfm_matrix_t* A) {
	size_t const m = A->m, k = A->band.k;
	double* const a = A->band.data;
	for (size_t p = 0; p + 1 < m; p++) {
		double const pivot = a[band_idx(A, p, p)];
		if (BFM_IS_NAN(pivot) || fabs(pivot) < BFM_PIVOT_EPS) {
			return -1;
		}
		size_t const end = p + k + 1 < m ? p + k + 1 : m;
		for (size_t i = p + 1; i < end; i++) {
			double const l = a[band_idx(A, i, p)] / pivot;
			a[band_idx(A, i, p)] = l;
			for (size_t j = p + 1; j < end; j++) {
				a[band_idx(A, i, j)] += -l * a[band_idx(A, p, j)];
			}
		}
	}
	return 0;
}

static int band_lu_solve(bfm_matrix_t* A, double* y) {
	size_t const m = A->m, k = A->band.k;
	double const* const a = A->band.data;
	for (size_t p = 0; p < m; p++) {
		for (size_t i = p > k ? p - k : 0; i < p; i++) {
			y[p] -= a[band_idx(A, p, i)] * y[i];
		}
	}
	for (size_t p = m; p-- > 0;) {
		size_t const end = p + k + 1 < m ? p + k + 1 : m;
		for (size_t i = p + 1; i < end; i++) {
			y[p] -= y[i] * a[band_idx(A, p, i)];
		}
		double const pivot = a[band_idx(A, p, p)];
		if (BFM_IS_NAN(pivot) || pivot == 0) {
			return -1;
		}
		y[p] /= pivot;
	}
	return 0;
}

int bfm_matrix_lu(bfm_matrix_t* matrix) {
	switch (matrix->kind) {
	case BFM_MATRIX_KIND_FULL:
		return full_lu(matrix);
	case BFM_MATRIX_KIND_BAND:
		return band_lu(matrix);
	case BFM_MATRIX_KIND_BSR:
		return 0;
	}
	return -1;
}

int bfm_matrix_lu_solve(bfm_matrix_t* matrix, bfm_vec_t* y) {
	if (matrix->m != y->n) {
		return -1;
	}
	switch (matrix->kind) {
	case BFM_MATRIX_KIND_FULL:
		return full_lu_solve(matrix, y->data);
	case BFM_MATRIX_KIND_BAND:
		return band_lu_solve(matrix, y->data);
	case BFM_MATRIX_KIND_BSR:
		return bfm_bsr_solve(matrix, y);
	}
	return -1;
}

int bfm_matrix_solve(bfm_matrix_t* matrix, bfm_vec_t* y) {
	if (bfm_matrix_lu(matrix) < 0) {
		return -1;
	}
	return bfm_matrix_lu_solve(matrix, y);
}
