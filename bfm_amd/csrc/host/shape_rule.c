/* Shape functions and Gauss-Legendre rules (libbfm/src/shape.c, rule.c).
 *
 * P1: phi = (1 - xsi - eta, xsi, eta). Q1: nodes counter-clockwise from
 * (+1, +1), phi_j = (1 +- xsi)(1 +- eta) / 4. P2 fills its values but reports
 * failure, as the reference does (shape.c:28-37): the planar path accepts P1
 * and Q1 only. Rules: 3 points of weight 1/6 on the triangle; 2 x 2 points at
 * +-0.577350269189626 (the reference's literal) of weight 1 on the quad. */

#include <string.h>

#include "internal.h"

static int shape_phi(bfm_shape_t* shape, double* point, double* phi) {
	if (shape->dim != 2) {
		return -1;
	}
	double const x = point[0], e = point[1];
	switch (shape->kind) {
	case BFM_ELEM_KIND_SIMPLEX:
		phi[0] = 1 - x - e;
		phi[1] = x;
		phi[2] = e;
		return 0;
	case BFM_ELEM_KIND_QUAD:
		phi[0] = (1 + x) * (1 + e) / 4;
		phi[1] = (1 - x) * (1 + e) / 4;
		phi[2] = (1 - x) * (1 - e) / 4;
		phi[3] = (1 + x) * (1 - e) / 4;
		return 0;
	case BFM_ELEM_KIND_QUADRATIC_TRIANGLE:
		phi[0] = 1 - 3 * (x + e) + 2 * (x + e) * (x + e);
		phi[1] = x * (2 * x - 1);
		phi[2] = e * (2 * e - 1);
		phi[3] = 4 * x * (1 - x - e);
		phi[4] = 4 * x * e;
		phi[5] = 4 * e * (1 - x - e);
		return -1;
	}
	return -1;
}

static int shape_dphi(bfm_shape_t* shape, size_t wrt, double* point, double* d) {
	if (shape->dim != 2) {
		return -1;
	}
	double const x = point[0], e = point[1];
	switch (shape->kind) {
	case BFM_ELEM_KIND_SIMPLEX:
		d[0] = -1;
		d[1] = wrt == 0 ? 1 : 0;
		d[2] = wrt == 0 ? 0 : 1;
		return 0;
	case BFM_ELEM_KIND_QUAD:
		if (wrt == 0) {
			d[0] = (1 + e) / 4, d[1] = (-1 - e) / 4, d[2] = (-1 + e) / 4, d[3] = (1 - e) / 4;
		} else {
			d[0] = (1 + x) / 4, d[1] = (1 - x) / 4, d[2] = (-1 + x) / 4, d[3] = (-1 - x) / 4;
		}
		return 0;
	case BFM_ELEM_KIND_QUADRATIC_TRIANGLE:
		if (wrt == 0) {
			d[0] = -3 + 4 * (x + e), d[1] = 4 * x - 1, d[2] = 0;
			d[3] = 4 - 8 * x - 4 * e, d[4] = 4 * e, d[5] = -4 * e;
		} else {
			d[0] = -3 + 4 * (x + e), d[1] = 0, d[2] = 4 * e - 1;
			d[3] = -4 * x, d[4] = 4 * x, d[5] = 4 - 4 * x - 8 * e;
		}
		return -1;
	}
	return -1;
}

int bfm_shape_create(bfm_shape_t* shape, bfm_state_t* state, size_t dim, bfm_elem_kind_t kind) {
	shape->state = state;
	shape->dim = dim;
	shape->kind = kind;
	shape->phi = shape_phi;
	shape->dphi = shape_dphi;
	return 0;
}

int bfm_shape_destroy(bfm_shape_t* shape) {
	(void) shape;
	return 0;
}

static void rule_free(bfm_rule_t* rule) {
	bfm_state_t* const state = rule->state;
	if (rule->points) {
		for (size_t i = 0; i < rule->n_points; i++) {
			if (rule->points[i]) {
				state->free(rule->points[i]);
			}
		}
		state->free(rule->points);
	}
	if (rule->weights) {
		state->free(rule->weights);
	}
	rule->points = NULL;
	rule->weights = NULL;
}

int bfm_rule_create(bfm_rule_t* rule, bfm_state_t* state, size_t dim, bfm_elem_kind_t kind, size_t n_points) {
	memset(rule, 0, sizeof *rule);
	rule->state = state;
	rule->dim = dim;
	rule->kind = kind;
	rule->n_points = n_points;
	rule->weights = state->alloc(n_points * sizeof *rule->weights + 1);
	rule->points = state->alloc(n_points * sizeof *rule->points + 1);
	if (!rule->weights || !rule->points) {
		goto fail;
	}
	memset(rule->weights, 0, n_points * sizeof *rule->weights);
	memset(rule->points, 0, n_points * sizeof *rule->points);
	for (size_t i = 0; i < n_points; i++) {
		rule->points[i] = state->alloc(dim * sizeof **rule->points + 1);
		if (!rule->points[i]) {
			goto fail;
		}
		memset(rule->points[i], 0, dim * sizeof **rule->points);
	}
	if (bfm_shape_create(&rule->shape, state, dim, kind) < 0) {
		goto fail;
	}
	return 0;
fail:
	BFM_ERROR(state, "cannot allocate a %zu-point rule", n_points);
	rule_free(rule);
	return -1;
}

int bfm_rule_destroy(bfm_rule_t* rule) {
	rule_free(rule);
	bfm_shape_destroy(&rule->shape);
	return 0;
}

int bfm_rule_create_gauss_legendre(bfm_rule_t* rule, bfm_state_t* state, size_t dim, bfm_elem_kind_t kind) {
	if (dim != 2 || (kind != BFM_ELEM_KIND_SIMPLEX && kind != BFM_ELEM_KIND_QUAD)) {
		BFM_ERROR(state, "Gauss-Legendre rules exist for 2-D P1 and Q1 elements only");
		return -1;
	}
	if (bfm_rule_create(rule, state, dim, kind, (size_t) kind) < 0) {
		return -1;
	}
	if (kind == BFM_ELEM_KIND_SIMPLEX) {
		double const sixth = 1. / 6, two_thirds = 1 - 1. / 3;
		double const p[3][2] = {{sixth, sixth}, {two_thirds, sixth}, {sixth, two_thirds}};
		for (int i = 0; i < 3; i++) {
			rule->weights[i] = sixth;
			rule->points[i][0] = p[i][0];
			rule->points[i][1] = p[i][1];
		}
	} else {
		double const a = 0.577350269189626;
		double const p[4][2] = {{-a, a}, {-a, -a}, {a, -a}, {a, a}};
		for (int i = 0; i < 4; i++) {
			rule->weights[i] = 1;
			rule->points[i][0] = p[i][0];
			rule->points[i][1] = p[i][1];
		}
	}
	return 0;
}
