/* Meshes and their readers (libbfm/src/mesh.c).
 *
 * .lepl1110: nodes, boundary edges, triangles or quads, named domains
 * (lists of edge indices). An edge's elems[0] holds the index read in front
 * of it and elems[1] = -1, as in the reference (mesh.c:138-141).
 * Wavefront .obj: "v x y z" and "f a b c" (1-based) lines; the edge list is
 * built as the reference's compute_edges does (mesh.c:52-102): all element
 * edges, sorted by descending smaller node then ascending larger node (ties
 * in element order), a pair of opposite consecutive edges merged into one
 * with elems[1] set, and the last edge of the sorted list dropped unless it
 * merged. */

#include <stdio.h>
#include <string.h>

#include "internal.h"

int bfm_mesh_create(bfm_mesh_t* mesh, bfm_state_t* state, size_t dim, bfm_elem_kind_t kind) {
	memset(mesh, 0, sizeof *mesh);
	mesh->state = state;
	mesh->dim = dim;
	mesh->kind = kind;
	return 0;
}

int bfm_mesh_destroy(bfm_mesh_t* mesh) {
	bfm_state_t* const state = mesh->state;
	state->free(mesh->coords);
	state->free(mesh->elems);
	state->free(mesh->edges);
	for (size_t i = 0; i < mesh->n_domains && mesh->domains; i++) {
		state->free(mesh->domains[i].elements);
	}
	state->free(mesh->domains);
	mesh->coords = NULL;
	mesh->elems = NULL;
	mesh->edges = NULL;
	mesh->domains = NULL;
	return 0;
}

static void* zalloc(bfm_state_t* state, size_t size) {
	void* p = state->alloc(size ? size : 1);
	if (p) {
		memset(p, 0, size);
	}
	return p;
}

int bfm_mesh_read_lepl1110(bfm_mesh_t* mesh, bfm_state_t* state, char const* name) {
	bfm_mesh_create(mesh, state, 2, BFM_ELEM_KIND_SIMPLEX);
	FILE* fp = fopen(name, "r");
	if (!fp) {
		BFM_ERROR(state, "cannot open %s", name);
		return -1;
	}
	int rv = -1;
	size_t idx;
	char kind_str[16] = "";
	if (fscanf(fp, " Number of nodes %zu", &mesh->n_nodes) != 1) {
		BFM_ERROR(state, "%s: no node section", name);
		goto out;
	}
	mesh->coords = zalloc(state, 2 * mesh->n_nodes * sizeof *mesh->coords);
	if (!mesh->coords) {
		goto nomem;
	}
	for (size_t i = 0; i < mesh->n_nodes; i++) {
		if (fscanf(fp, " %zu : %lf %lf", &idx, &mesh->coords[2 * i], &mesh->coords[2 * i + 1]) != 3) {
			BFM_ERROR(state, "%s: bad node line %zu", name, i);
			goto out;
		}
	}
	if (fscanf(fp, " Number of edges %zu", &mesh->n_edges) != 1) {
		BFM_ERROR(state, "%s: no edge section", name);
		goto out;
	}
	mesh->edges = zalloc(state, mesh->n_edges * sizeof *mesh->edges);
	if (!mesh->edges) {
		goto nomem;
	}
	for (size_t i = 0; i < mesh->n_edges; i++) {
		bfm_edge_t* const e = &mesh->edges[i];
		if (fscanf(fp, " %zd : %zu %zu", &e->elems[0], &e->nodes[0], &e->nodes[1]) != 3) {
			BFM_ERROR(state, "%s: bad edge line %zu", name, i);
			goto out;
		}
		e->elems[1] = -1;
	}
	if (fscanf(fp, " Number of %15s %zu", kind_str, &mesh->n_elems) != 2) {
		BFM_ERROR(state, "%s: no element section", name);
		goto out;
	}
	if (strcmp(kind_str, "triangles") == 0) {
		mesh->kind = BFM_ELEM_KIND_SIMPLEX;
	} else if (strcmp(kind_str, "quads") == 0) {
		mesh->kind = BFM_ELEM_KIND_QUAD;
	} else {
		BFM_ERROR(state, "%s: unknown element kind '%s'", name, kind_str);
		goto out;
	}
	size_t const k = mesh->kind;
	mesh->elems = zalloc(state, mesh->n_elems * k * sizeof *mesh->elems);
	if (!mesh->elems) {
		goto nomem;
	}
	for (size_t i = 0; i < mesh->n_elems; i++) {
		size_t* const e = &mesh->elems[i * k];
		int const got = k == 3 ? fscanf(fp, " %zu : %zu %zu %zu", &idx, &e[0], &e[1], &e[2])
		                       : fscanf(fp, " %zu : %zu %zu %zu %zu", &idx, &e[0], &e[1], &e[2], &e[3]);
		if (got != (int) k + 1) {
			BFM_ERROR(state, "%s: bad element line %zu", name, i);
			goto out;
		}
	}
	if (fscanf(fp, " Number of domains %zu", &mesh->n_domains) != 1) {
		mesh->n_domains = 0;   /* the section is optional */
		rv = 0;
		goto out;
	}
	mesh->domains = zalloc(state, mesh->n_domains * sizeof *mesh->domains);
	if (!mesh->domains) {
		goto nomem;
	}
	for (size_t i = 0; i < mesh->n_domains; i++) {
		size_t id;
		if (fscanf(fp, " Domain : %zu", &id) != 1 || id >= mesh->n_domains) {
			BFM_ERROR(state, "%s: bad domain header %zu", name, i);
			goto out;
		}
		bfm_domain_t* const d = &mesh->domains[id];
		/* the name runs to the end of its line, trailing spaces included */
		if (fscanf(fp, " Name : %49[^\n]", d->name) != 1 ||
		    fscanf(fp, " Number of elements : %zu", &d->n_elements) != 1) {
			BFM_ERROR(state, "%s: bad domain %zu", name, id);
			goto out;
		}
		d->elements = zalloc(state, d->n_elements * sizeof *d->elements);
		if (!d->elements) {
			goto nomem;
		}
		for (size_t j = 0; j < d->n_elements; j++) {
			if (fscanf(fp, " %zu", &d->elements[j]) != 1) {
				BFM_ERROR(state, "%s: domain %zu is short", name, id);
				goto out;
			}
		}
	}
	rv = 0;
	goto out;
nomem:
	BFM_ERROR(state, "%s: out of memory", name);
out:
	fclose(fp);
	return rv;
}

static int edge_cmp(bfm_edge_t const* a, bfm_edge_t const* b) {
	size_t const ma = BFM_MIN(a->nodes[0], a->nodes[1]), mb = BFM_MIN(b->nodes[0], b->nodes[1]);
	if (ma != mb) {
		return ma > mb ? -1 : 1;
	}
	size_t const Ma = BFM_MAX(a->nodes[0], a->nodes[1]), Mb = BFM_MAX(b->nodes[0], b->nodes[1]);
	return (Ma > Mb) - (Ma < Mb);
}

/* Stable merge sort (ties keep element order). */
static void edge_sort(bfm_edge_t* e, bfm_edge_t* tmp, size_t n) {
	if (n < 2) {
		return;
	}
	size_t const h = n / 2;
	edge_sort(e, tmp, h);
	edge_sort(e + h, tmp, n - h);
	size_t i = 0, j = h, k = 0;
	while (i < h && j < n) {
		tmp[k++] = edge_cmp(&e[j], &e[i]) < 0 ? e[j++] : e[i++];
	}
	while (i < h) {
		tmp[k++] = e[i++];
	}
	while (j < n) {
		tmp[k++] = e[j++];
	}
	memcpy(e, tmp, n * sizeof *e);
}

static int compute_edges(bfm_mesh_t* mesh) {
	bfm_state_t* const state = mesh->state;
	size_t const k = mesh->kind, n = mesh->n_elems * k;
	bfm_edge_t* e = state->alloc(n * sizeof *e + 1);
	bfm_edge_t* tmp = state->alloc(n * sizeof *tmp + 1);
	if (!e || !tmp) {
		state->free(e);
		state->free(tmp);
		return -1;
	}
	for (size_t i = 0; i < mesh->n_elems; i++) {
		for (size_t j = 0; j < k; j++) {
			e[i * k + j].nodes[0] = mesh->elems[i * k + j];
			e[i * k + j].nodes[1] = mesh->elems[i * k + (j + 1) % k];
			e[i * k + j].elems[0] = (ssize_t) i;
			e[i * k + j].elems[1] = -1;
		}
	}
	edge_sort(e, tmp, n);
	state->free(tmp);
	size_t out = 0;
	for (size_t i = 1; i < n; i++) {
		e[out] = e[i - 1];
		if (e[i - 1].nodes[0] == e[i].nodes[1] && e[i - 1].nodes[1] == e[i].nodes[0]) {
			e[out].elems[1] = e[i].elems[0];
			i++;
		}
		out++;
	}
	mesh->n_edges = out;
	mesh->edges = state->realloc(e, (out ? out : 1) * sizeof *e);
	return mesh->edges ? 0 : -1;
}

int bfm_mesh_read_wavefront(bfm_mesh_t* mesh, bfm_state_t* state, char const* name, bool full) {
	bfm_mesh_create(mesh, state, full ? 3 : 2, BFM_ELEM_KIND_SIMPLEX);
	FILE* fp = fopen(name, "r");
	if (!fp) {
		BFM_ERROR(state, "cannot open %s", name);
		return -1;
	}
	size_t cap_n = 0, cap_e = 0;
	char line[1024];
	int rv = -1;
	while (fgets(line, sizeof line, fp)) {
		char tag[16];
		int off = 0;
		if (sscanf(line, "%15s%n", tag, &off) != 1) {
			continue;
		}
		if (!strcmp(tag, "v")) {
			if (mesh->n_nodes == cap_n) {
				cap_n = cap_n ? 2 * cap_n : 256;
				double* c = state->realloc(mesh->coords, cap_n * mesh->dim * sizeof *c);
				if (!c) {
					goto nomem;
				}
				mesh->coords = c;
			}
			double xyz[3] = {0, 0, 0};
			sscanf(line + off, "%lf %lf %lf", &xyz[0], &xyz[1], &xyz[2]);
			for (size_t d = 0; d < mesh->dim; d++) {
				mesh->coords[mesh->n_nodes * mesh->dim + d] = xyz[d];
			}
			mesh->n_nodes++;
		} else if (!strcmp(tag, "f")) {
			if (mesh->n_elems == cap_e) {
				cap_e = cap_e ? 2 * cap_e : 256;
				size_t* el = state->realloc(mesh->elems, cap_e * 3 * sizeof *el);
				if (!el) {
					goto nomem;
				}
				mesh->elems = el;
			}
			size_t a = 0, b = 0, c = 0;
			if (sscanf(line + off, "%zu %zu %zu", &a, &b, &c) != 3 || !a || !b || !c) {
				BFM_ERROR(state, "%s: unsupported face line: %s", name, line);
				goto out;
			}
			size_t* const e = &mesh->elems[mesh->n_elems++ * 3];
			e[0] = a - 1, e[1] = b - 1, e[2] = c - 1;
		}
	}
	if (compute_edges(mesh) < 0) {
		goto nomem;
	}
	rv = 0;
	goto out;
nomem:
	BFM_ERROR(state, "%s: out of memory", name);
out:
	fclose(fp);
	if (rv < 0) {
		state->free(mesh->coords);
		state->free(mesh->elems);
		mesh->coords = NULL;
		mesh->elems = NULL;
	}
	return rv;
}
