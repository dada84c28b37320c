#pragma once

/* Quadrature rules (libbfm/src/bfm/rule.h:5-24). points[i] is the i-th
 * integration point (dim coordinates), weights[i] its weight. */

#include <bfm/mesh.h>
#include <bfm/shape.h>

typedef struct {
	bfm_state_t* state;

	size_t dim;
	bfm_elem_kind_t kind;
	size_t n_points;

	double* weights;
	double** points;

	bfm_shape_t shape;
} bfm_rule_t;

int bfm_rule_create(bfm_rule_t* rule, bfm_state_t* state, size_t dim, bfm_elem_kind_t kind, size_t n_points);
int bfm_rule_destroy(bfm_rule_t* rule);

int bfm_rule_create_gauss_legendre(bfm_rule_t* rule, bfm_state_t* state, size_t dim, bfm_elem_kind_t kind);
