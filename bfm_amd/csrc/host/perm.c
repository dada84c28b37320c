/* Permutations and reverse Cuthill-McKee (libbfm/src/perm.c).
 *
 * perm[old] = new, inv_perm[new] = old. For dense matrices the RCM follows
 * the reference's rules (perm.c:117-331): degree = non-zeros of the row, the
 * start of each component is the last unvisited node of minimum degree,
 * neighbours are taken in index order and enqueued in a stable sort by
 * degree, the visit order is reversed. For a device-resident BSR matrix the
 * permutation is the node renumbering done at upload, expanded to DOFs. */

#include <string.h>

#include "internal.h"

int bfm_perm_create(bfm_perm_t* perm, bfm_state_t* state, size_t m) {
	memset(perm, 0, sizeof *perm);
	perm->state = state;
	perm->m = m;
	return 0;
}

int bfm_perm_destroy(bfm_perm_t* perm) {
	if (perm->has_perm) {
		perm->state->free(perm->perm);
		perm->state->free(perm->inv_perm);
	}
	perm->has_perm = false;
	return 0;
}

int bfm_perm_perm_matrix(bfm_perm_t* perm, bfm_matrix_t* matrix, bool inv) {
	if (!perm->has_perm || matrix->m != perm->m || matrix->kind != BFM_MATRIX_KIND_FULL) {
		return -1;
	}
	bfm_state_t* const state = perm->state;
	size_t const* const p = inv ? perm->inv_perm : perm->perm;
	size_t const m = matrix->m;
	bfm_matrix_t old;
	if (bfm_matrix_full_create(&old, state, matrix->major, m) < 0) {
		return -1;
	}
	memcpy(old.full.data, matrix->full.data, m * m * sizeof(double));
	for (size_t i = 0; i < m; i++) {
		for (size_t j = 0; j < m; j++) {
			bfm_matrix_set(matrix, p[i], p[j], bfm_matrix_get(&old, i, j));
		}
	}
	bfm_matrix_destroy(&old);
	return 0;
}

int bfm_perm_perm_vec(bfm_perm_t* perm, bfm_vec_t* vec, bool inv) {
	if (!perm->has_perm || vec->n != perm->m) {
		return -1;
	}
	bfm_state_t* const state = perm->state;
	size_t const* const p = inv ? perm->inv_perm : perm->perm;
	double* old = state->alloc(vec->n * sizeof *old + 1);
	if (!old) {
		BFM_ERROR(state, "cannot allocate a vector copy");
		return -1;
	}
	memcpy(old, vec->data, vec->n * sizeof *old);
	for (size_t i = 0; i < vec->n; i++) {
		vec->data[p[i]] = old[i];
	}
	state->free(old);
	return 0;
}

static int perm_alloc(bfm_perm_t* perm) {
	bfm_state_t* const state = perm->state;
	if (perm->has_perm) {
		state->free(perm->perm);
		state->free(perm->inv_perm);
		perm->has_perm = false;
	}
	perm->perm = state->alloc(perm->m * sizeof *perm->perm + 1);
	perm->inv_perm = state->alloc(perm->m * sizeof *perm->inv_perm + 1);
	if (!perm->perm || !perm->inv_perm) {
		state->free(perm->perm);
		state->free(perm->inv_perm);
		BFM_ERROR(state, "cannot allocate a permutation of %zu", perm->m);
		return -1;
	}
	return 0;
}

typedef struct {
	size_t i;
	size_t deg;
} rcm_entry_t;

int bfm_perm_rcm(bfm_perm_t* perm, bfm_matrix_t* A) {
	size_t const n = A->m;
	if (perm->m != n) {
		return -1;
	}
	if (A->kind == BFM_MATRIX_KIND_BSR) {
		bfm_bsr_t const* const h = A->bsr.handle;
		size_t const nn = n / 2;
		int64_t* np = perm->state->alloc(nn * sizeof *np + 1);
		if (!np || perm_alloc(perm) < 0) {
			perm->state->free(np);
			return -1;
		}
		bfmx_system_node_perm(h->sys, np);
		for (size_t i = 0; i < nn; i++) {
			for (size_t c = 0; c < 2; c++) {
				perm->perm[2 * i + c] = 2 * (size_t) np[i] + c;
				perm->inv_perm[2 * (size_t) np[i] + c] = 2 * i + c;
			}
		}
		perm->state->free(np);
		perm->has_perm = true;
		return 0;
	}
	bfm_state_t* const state = perm->state;
	size_t* degs = state->alloc(n * sizeof *degs + 1);
	size_t* queue = state->alloc(n * sizeof *queue + 1);
	rcm_entry_t* nb = state->alloc(n * sizeof *nb + 1);
	bool* visited = state->alloc(n * sizeof *visited + 1);
	int rv = -1;
	if (!degs || !queue || !nb || !visited || perm_alloc(perm) < 0) {
		goto out;
	}
	memset(visited, 0, n * sizeof *visited);
	for (size_t i = 0; i < n; i++) {
		degs[i] = 0;
		for (size_t j = 0; j < n; j++) {
			degs[i] += bfm_matrix_get(A, i, j) != 0;
		}
	}
	size_t head = 0, tail = 0, p = 0;
	while (p < n) {
		size_t start = 0, best = (size_t) -1;
		for (size_t i = 0; i < n; i++) {
			if (!visited[i] && degs[i] <= best) {
				best = degs[i];
				start = i;
			}
		}
		queue[tail++] = start;
		while (head != tail) {
			size_t const cur = queue[head++];
			visited[cur] = true;
			size_t cnt = 0;
			for (size_t i = 0; i < n; i++) {
				if (!visited[i] && bfm_matrix_get(A, cur, i) != 0) {
					visited[i] = true;
					size_t t = cnt++;   /* insertion keeps equal degrees in index order */
					while (t > 0 && nb[t - 1].deg > degs[i]) {
						nb[t] = nb[t - 1];
						t--;
					}
					nb[t] = (rcm_entry_t){i, degs[i]};
				}
			}
			for (size_t t = 0; t < cnt; t++) {
				queue[tail++] = nb[t].i;
			}
			perm->inv_perm[n - 1 - p++] = cur;
		}
	}
	for (size_t i = 0; i < n; i++) {
		perm->perm[perm->inv_perm[i]] = i;
	}
	perm->has_perm = true;
	rv = 0;
out:
	state->free(degs);
	state->free(queue);
	state->free(nb);
	state->free(visited);
	return rv;
}
