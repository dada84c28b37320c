#pragma once

/* Meshes (libbfm/src/bfm/mesh.h:5-54). coords holds dim values per node,
 * interleaved; elems holds kind node indices per element; edges are
 * boundary edges for .lepl1110 meshes and all element edges for .obj meshes. */

#include <stdlib.h>

#include <bfm/matrix.h>

typedef enum {
	BFM_ELEM_KIND_SIMPLEX = 3,
	BFM_ELEM_KIND_QUAD = 4,
	BFM_ELEM_KIND_QUADRATIC_TRIANGLE = 6,
} bfm_elem_kind_t;

typedef enum {
	BFM_PLANAR_STRAINS,
} bfm_problem_type_t;

typedef struct {
	size_t nodes[2];
	ssize_t elems[2];
} bfm_edge_t;

typedef struct {
	char name[50];
	size_t n_elements;
	size_t* elements;
} bfm_domain_t;

typedef struct {
	bfm_state_t* state;

	size_t dim;
	bfm_elem_kind_t kind;

	size_t n_elems;
	size_t n_nodes;
	size_t n_edges;

	double* coords;
	size_t* elems;
	bfm_edge_t* edges;

	size_t n_domains;
	bfm_domain_t* domains;
} bfm_mesh_t;

int bfm_mesh_create(bfm_mesh_t* mesh, bfm_state_t* state, size_t dim, bfm_elem_kind_t kind);
int bfm_mesh_destroy(bfm_mesh_t* mesh);

int bfm_mesh_read_lepl1110(bfm_mesh_t* mesh, bfm_state_t* state, char const* name);
int bfm_mesh_read_wavefront(bfm_mesh_t* mesh, bfm_state_t* state, char const* name, bool full);
