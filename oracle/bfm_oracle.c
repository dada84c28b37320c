/* CPU oracle for libbfm's planar-elasticity path. TEST INFRASTRUCTURE ONLY:
 * this file is the checker the GPU implementation is compared against. Only
 * tests, the smoke check and the benchmark's CPU-baseline leg load it; the
 * library in bfm_amd/ never links or calls it.
 *
 * It restates the reference algorithm (paths relative to the reference tree):
 *   orc_assemble_dense   create_planar + fill_elasticity_elem
 *                        (libbfm/src/system.c:109-227, 427-466), same loop
 *                        nesting and operation order, so the dense K and b are
 *                        bit-identical to libbfm's.
 *   orc_apply_conditions the ordered condition loop (system.c:358-526).
 *   orc_rcm              bfm_perm_rcm (libbfm/src/perm.c:117-331); the degree
 *                        sort is stable, as glibc 2.35's qsort (a merge sort
 *                        for arrays of this size) is.
 *   orc_bandwidth        matrix_full_bandwidth (libbfm/src/matrix.c:57-71).
 *   orc_band_*           band layout, unpivoted LU and substitutions
 *                        (matrix.c:196-302, 351-404).
 *   orc_solve_dense      run_elasticity (libbfm/src/sim.c:103-135).
 *   stress == 2          the axisymmetric formulation:
 *                        fill_axisymmetric_elem (system.c:228-356) with its
 *                        constants (the reference multiplies by 1 - 2 nu
 *                        where the plane-strain formula divides, line 242)
 *                        and its f_22 (second term without r, line 339), and
 *                        the condition loop of
 *                        bfm_system_create_axisymmetric_strain
 *                        (system.c:539-622: Neumann X/Y weighted by r1 + r2
 *                        as the reference forms them, Neumann normal/tangent
 *                        ignored).
 * and adds sparse counterparts for meshes whose dense matrix does not fit:
 *   orc_csr_*            the same element arithmetic accumulated into CSR in
 *                        the same element order (entries equal the dense
 *                        ones), the same ordered conditions, a CSR SpMV and an
 *                        OpenMP Jacobi-PCG. The PCG has no reference
 *                        counterpart; it is pinned by the dense direct solve
 *                        on small meshes.
 *
 * Pinning (tests/test_oracle.py): K, b and u against fixtures written by the
 * reference compiled from its own sources (oracle/_ref), the reference's
 * data/U.txt and data/V.txt, numpy's dense solve for the band LU, closed
 * forms (rigid-body null space, patch test, element area, Neumann totals,
 * D constants) and dense-vs-CSR equality. */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- quadrature
 * libbfm/src/rule.c:90-111 and shape.c:3-63. The quad abscissa is the
 * reference's truncated literal, not 1/sqrt(3). */

#define ORC_SIXTH (1. / 6)
#define ORC_THIRD (1. / 3)
#define ORC_SQRT_THIRD 0.577350269189626

static void orc_rule(int kind, int* n_points, double w[4], double pts[4][2]) {
	if (kind == 3) {
		*n_points = 3;
		double const p[3][2] = {{ORC_SIXTH, ORC_SIXTH}, {1 - ORC_THIRD, ORC_SIXTH}, {ORC_SIXTH, 1 - ORC_THIRD}};
		for (int i = 0; i < 3; i++) {
			w[i] = ORC_SIXTH;
			pts[i][0] = p[i][0];
			pts[i][1] = p[i][1];
		}
	} else {
		*n_points = 4;
		double const p[4][2] = {{-ORC_SQRT_THIRD, ORC_SQRT_THIRD}, {-ORC_SQRT_THIRD, -ORC_SQRT_THIRD}, {ORC_SQRT_THIRD, -ORC_SQRT_THIRD}, {ORC_SQRT_THIRD, ORC_SQRT_THIRD}};
		for (int i = 0; i < 4; i++) {
			w[i] = 1;
			pts[i][0] = p[i][0];
			pts[i][1] = p[i][1];
		}
	}
}

static void orc_shape(int kind, double xsi, double eta, double phi[4], double dxsi[4], double deta[4]) {
	if (kind == 3) {
		phi[0] = 1 - xsi - eta;
		phi[1] = xsi;
		phi[2] = eta;
		dxsi[0] = -1, dxsi[1] = 1, dxsi[2] = 0;
		deta[0] = -1, deta[1] = 0, deta[2] = 1;
	} else {
		phi[0] = (1 + xsi) * (1 + eta) / 4;
		phi[1] = (1 - xsi) * (1 + eta) / 4;
		phi[2] = (1 - xsi) * (1 - eta) / 4;
		phi[3] = (1 + xsi) * (1 - eta) / 4;
		dxsi[0] = (1 + eta) / 4;
		dxsi[1] = (-1 - eta) / 4;
		dxsi[2] = (-1 + eta) / 4;
		dxsi[3] = (1 - eta) / 4;
		deta[0] = (1 + xsi) / 4;
		deta[1] = (1 - xsi) / 4;
		deta[2] = (-1 + xsi) / 4;
		deta[3] = (-1 - xsi) / 4;
	}
}

/* D-matrix constants, libbfm/src/system.c:454-458. */
void orc_dconst(int stress, double E, double nu, double* a, double* b, double* c) {
	if (stress == 2) {   /* axisymmetric, system.c:241-243 as written */
		*a = E * (1 - nu) / (1 + nu) * (1 - 2 * nu);
		*b = E * nu / (1 + nu) / (1 - 2 * nu);
		*c = E / (2 * (1 + nu));
		return;
	}
	*a = !stress ? E * (1 - nu) / (1 + nu) / (1 - 2 * nu) : E / (1 - nu * nu);
	*b = !stress ? E * nu / (1 + nu) / (1 - 2 * nu) : E * nu / (1 - nu * nu);
	*c = E / (2 * (1 + nu));
}

/* Element kernel shared by the dense and CSR assemblies. For each Gauss
 * point it emits the body load and the 2x2 stiffness blocks through the
 * callbacks in exactly the order fill_elasticity_elem adds them
 * (libbfm/src/system.c:136-224). */
typedef void (*orc_add_fn)(void* ctx, size_t i, size_t j, double v);

/* Axisymmetric element (fill_axisymmetric_elem, system.c:228-356): r is the
 * x coordinate interpolated at the Gauss point. */
static void orc_element_axi(int kind, const size_t* map, const double* coords, double E, double nu, double rho,
                            size_t n_forces, const double* forces, double* b, orc_add_fn add, void* ctx) {
	double a, bb, c;
	orc_dconst(2, E, nu, &a, &bb, &c);

	double x[4], y[4];
	for (int j = 0; j < kind; j++) {
		x[j] = coords[map[j] * 2 + 0];
		y[j] = coords[map[j] * 2 + 1];
	}

	int n_points;
	double w[4], pts[4][2];
	orc_rule(kind, &n_points, w, pts);

	for (int g = 0; g < n_points; g++) {
		double phi[4], dxsi[4], deta[4];
		orc_shape(kind, pts[g][0], pts[g][1], phi, dxsi, deta);

		double dx_dxsi = 0, dx_deta = 0, dy_dxsi = 0, dy_deta = 0, r = 0;
		for (int j = 0; j < kind; j++) {
			dx_dxsi += x[j] * dxsi[j];
			dx_deta += x[j] * deta[j];
			dy_dxsi += y[j] * dxsi[j];
			dy_deta += y[j] * deta[j];
			r += x[j] * phi[j];
		}

		double const det_J = fabs(dx_dxsi * dy_deta - dx_deta * dy_dxsi);
		double dphi_dx[4], dphi_dy[4];

		for (int j = 0; j < kind; j++) {
			dphi_dx[j] = (dxsi[j] * dy_deta - deta[j] * dy_dxsi) / det_J;
			dphi_dy[j] = (deta[j] * dx_dxsi - dxsi[j] * dx_deta) / det_J;
		}

		for (int j = 0; j < kind; j++) {
			size_t const ii = 2 * map[j];
			for (size_t k = 0; k < n_forces; k++) {
				b[ii + 0] += det_J * w[g] * forces[2 * k + 0] * rho * phi[j] * r;
				b[ii + 1] += det_J * w[g] * forces[2 * k + 1] * rho * phi[j] * r;
			}
		}

		for (int j = 0; j < kind; j++) {
			size_t const ii = 2 * map[j];
			for (int k = 0; k < kind; k++) {
				size_t const jj = 2 * map[k];
				double const f11 = a * dphi_dx[j] * dphi_dx[k] * r + c * dphi_dy[j] * dphi_dy[k] * r +
				                   phi[j] * (bb * dphi_dx[k] + a * phi[k] / r) + dphi_dx[j] * bb * phi[k];
				double const f12 = bb * dphi_dx[j] * dphi_dy[k] * r + c * dphi_dy[j] * dphi_dx[k] * r + phi[j] * bb * dphi_dy[k];
				double const f21 = bb * dphi_dy[j] * dphi_dx[k] * r + c * dphi_dx[j] * dphi_dy[k] * r + dphi_dy[j] * bb * phi[k];
				double const f22 = a * dphi_dy[j] * dphi_dy[k] * r + c * dphi_dx[j] * dphi_dx[k];
				add(ctx, ii + 0, jj + 0, det_J * w[g] * f11);
				add(ctx, ii + 0, jj + 1, det_J * w[g] * f12);
				add(ctx, ii + 1, jj + 0, det_J * w[g] * f21);
				add(ctx, ii + 1, jj + 1, det_J * w[g] * f22);
			}
		}
	}
}

static void orc_element(int kind, const size_t* map, const double* coords, int stress, double E, double nu, double rho,
                        size_t n_forces, const double* forces, double* b, orc_add_fn add, void* ctx) {
	if (stress == 2) {
		orc_element_axi(kind, map, coords, E, nu, rho, n_forces, forces, b, add, ctx);
		return;
	}
	double a, bb, c;
	orc_dconst(stress, E, nu, &a, &bb, &c);

	double x[4], y[4];
	for (int j = 0; j < kind; j++) {
		x[j] = coords[map[j] * 2 + 0];
		y[j] = coords[map[j] * 2 + 1];
	}

	int n_points;
	double w[4], pts[4][2];
	orc_rule(kind, &n_points, w, pts);

	for (int g = 0; g < n_points; g++) {
		double phi[4], dxsi[4], deta[4];
		orc_shape(kind, pts[g][0], pts[g][1], phi, dxsi, deta);

		double dx_dxsi = 0, dx_deta = 0, dy_dxsi = 0, dy_deta = 0;
		for (int j = 0; j < kind; j++) {
			dx_dxsi += x[j] * dxsi[j];
			dx_deta += x[j] * deta[j];
			dy_dxsi += y[j] * dxsi[j];
			dy_deta += y[j] * deta[j];
		}

		double const det_J = fabs(dx_dxsi * dy_deta - dx_deta * dy_dxsi);
		double dphi_dx[4], dphi_dy[4];

		for (int j = 0; j < kind; j++) {
			dphi_dx[j] = (dxsi[j] * dy_deta - deta[j] * dy_dxsi) / det_J;
			dphi_dy[j] = (deta[j] * dx_dxsi - dxsi[j] * dx_deta) / det_J;
		}

		for (int j = 0; j < kind; j++) {
			size_t const ii = 2 * map[j];
			for (size_t k = 0; k < n_forces; k++) {
				b[ii + 0] += det_J * w[g] * forces[2 * k + 0] * rho * phi[j];
				b[ii + 1] += det_J * w[g] * forces[2 * k + 1] * rho * phi[j];
			}
		}

		for (int j = 0; j < kind; j++) {
			size_t const ii = 2 * map[j];
			for (int k = 0; k < kind; k++) {
				size_t const jj = 2 * map[k];
				double const f11 = a * dphi_dx[j] * dphi_dx[k] + c * dphi_dy[j] * dphi_dy[k];
				double const f12 = bb * dphi_dx[j] * dphi_dy[k] + c * dphi_dy[j] * dphi_dx[k];
				double const f21 = bb * dphi_dy[j] * dphi_dx[k] + c * dphi_dx[j] * dphi_dy[k];
				double const f22 = a * dphi_dy[j] * dphi_dy[k] + c * dphi_dx[j] * dphi_dx[k];
				add(ctx, ii + 0, jj + 0, det_J * w[g] * f11);
				add(ctx, ii + 0, jj + 1, det_J * w[g] * f12);
				add(ctx, ii + 1, jj + 0, det_J * w[g] * f21);
				add(ctx, ii + 1, jj + 1, det_J * w[g] * f22);
			}
		}
	}
}

/* ------------------------------------------------------------ dense path */

typedef struct {
	double* K;
	size_t m;
} orc_dense_t;

static void orc_dense_add(void* ctx, size_t i, size_t j, double v) {
	orc_dense_t* d = ctx;
	d->K[i * d->m + j] += v;
}

/* Dense assembly (system.c:427-466). K is m*m row-major with m = 2 n_nodes,
 * b has m entries; both are zeroed here. elem_E (optional) overrides E per
 * element, an extension used by the heterogeneous-modulus workload. */
int orc_assemble_dense(int kind, size_t n_nodes, const double* coords, size_t n_elems, const size_t* elems, int stress,
                       double E, double nu, double rho, const double* elem_E, size_t n_forces, const double* forces,
                       double* K, double* b) {
	if (kind != 3 && kind != 4) {
		return -1;
	}
	size_t const m = 2 * n_nodes;
	memset(K, 0, m * m * sizeof *K);
	memset(b, 0, m * sizeof *b);
	orc_dense_t d = {K, m};
	for (size_t e = 0; e < n_elems; e++) {
		orc_element(kind, elems + e * kind, coords, stress, elem_E ? elem_E[e] : E, nu, rho, n_forces, forces, b,
		            orc_dense_add, &d);
	}
	return 0;
}

/* apply_constraint (system.c:358-374) on a dense row-major matrix. */
static void orc_dense_constrain(double* K, double* b, size_t m, size_t d, double v) {
	for (size_t i = 0; i < m; i++) {
		b[i] -= v * K[i * m + d];
		K[i * m + d] = 0;
	}
	for (size_t i = 0; i < m; i++) {
		K[d * m + i] = 0;
	}
	K[d * m + d] = 1;
	b[d] = v;
}

/* Tangent of a boundary node for normal/tangent Dirichlet conditions
 * (system.c:394-420): half the sum of unit vectors from each adjacent
 * boundary-edge neighbour to the node. */
static void orc_node_tangent(size_t i, const double* coords, size_t n_edges, const size_t* edge_nodes,
                             const uint8_t* edge_boundary, double* tx, double* ty) {
	*tx = 0;
	*ty = 0;
	for (size_t j = 0; j < n_edges; j++) {
		if (!edge_boundary[j]) {
			continue;
		}
		size_t n2;
		if (edge_nodes[2 * j + 0] == i) {
			n2 = edge_nodes[2 * j + 1];
		} else if (edge_nodes[2 * j + 1] == i) {
			n2 = edge_nodes[2 * j + 0];
		} else {
			continue;
		}
		double const length = sqrt(pow(coords[i * 2 + 0] - coords[n2 * 2 + 0], 2) + pow(coords[i * 2 + 1] - coords[n2 * 2 + 1], 2));
		*tx += (coords[i * 2 + 0] - coords[n2 * 2 + 0]) / length / 2;
		*ty += (coords[i * 2 + 1] - coords[n2 * 2 + 1]) / length / 2;
	}
}

/* Neumann loads of one condition on b (system.c:477-522). Axisymmetric
 * (system.c:584-617): X/Y loads carry the factor r1 + r2 as the reference
 * forms each r (from one node's x and y), normal/tangent loads add nothing. */
static void orc_neumann(int kind, double value, const uint8_t* flags, const double* coords, size_t n_edges,
                        const size_t* edge_nodes, int axi, double* b) {
	if (axi && (kind == 4 || kind == 5)) {
		return;
	}
	for (size_t j = 0; j < n_edges; j++) {
		size_t const n1 = edge_nodes[2 * j + 0];
		size_t const n2 = edge_nodes[2 * j + 1];
		if (!flags[n1] || !flags[n2]) {
			continue;
		}
		if (kind == 2 || kind == 3) {
			double fac = 1;
			if (axi) {
				double const r1 = coords[n1 * 2 + 0] * (1 - 1 / sqrt(3)) / 2 + coords[n1 * 2 + 1] * (1 + 1 / sqrt(3)) / 2;
				double const r2 = coords[n2 * 2 + 0] * (1 - 1 / sqrt(3)) / 2 + coords[n2 * 2 + 1] * (1 + 1 / sqrt(3)) / 2;
				fac = r1 + r2;
			}
			double const c2 = pow(coords[n1 * 2 + 0] - coords[n2 * 2 + 0], 2) + pow(coords[n1 * 2 + 1] - coords[n2 * 2 + 1], 2);
			double const jac = sqrt(c2) / 2;
			size_t const shift = kind == 2 ? 0 : 1;
			if (axi) {
				b[n1 * 2 + shift] += fac * jac * value;
				b[n2 * 2 + shift] += fac * jac * value;
			} else {
				b[n1 * 2 + shift] += jac * value;
				b[n2 * 2 + shift] += jac * value;
			}
		} else {
			double const dx = coords[n1 * 2 + 0] - coords[n2 * 2 + 0];
			double const dy = coords[n1 * 2 + 1] - coords[n2 * 2 + 1];
			int const tangent = kind == 5;
			b[n1 * 2 + 0] += 0.5 * value * (tangent ? dx : -dy);
			b[n1 * 2 + 1] += 0.5 * value * (tangent ? dy : dx);
			b[n2 * 2 + 0] += 0.5 * value * (tangent ? dx : -dy);
			b[n2 * 2 + 1] += 0.5 * value * (tangent ? dy : dx);
		}
	}
}

/* The ordered condition loop of create_planar (system.c:470-526) or of
 * bfm_system_create_axisymmetric_strain (axi, system.c:575-619) on a dense
 * system. flags holds n_conditions * n_nodes bytes. */
static int orc_conditions_dense(size_t n_nodes, const double* coords, size_t n_edges, const size_t* edge_nodes,
                                const uint8_t* edge_boundary, size_t n_conditions, const int* kinds,
                                const double* values, const uint8_t* flags, int axi, double* K, double* b) {
	size_t const m = 2 * n_nodes;
	for (size_t c = 0; c < n_conditions; c++) {
		int const kind = kinds[c];
		double const v = values[c];
		const uint8_t* const f = flags + c * n_nodes;
		if (kind == 0 || kind == 1) {
			for (size_t j = 0; j < n_nodes; j++) {
				if (f[j]) {
					orc_dense_constrain(K, b, m, 2 * j + (size_t) kind, v);
				}
			}
		} else if (kind >= 2 && kind <= 5) {
			orc_neumann(kind, v, f, coords, n_edges, edge_nodes, axi, b);
		} else if (kind == 6 || kind == 7) {
			for (size_t i = 0; i < n_nodes; i++) {
				if (!f[i]) {
					continue;
				}
				double tx, ty;
				orc_node_tangent(i, coords, n_edges, edge_nodes, edge_boundary, &tx, &ty);
				orc_dense_constrain(K, b, m, 2 * i + 0, v * (kind == 7 ? tx : -ty));
				orc_dense_constrain(K, b, m, 2 * i + 1, v * (kind == 7 ? ty : tx));
			}
		} else {
			return -1;
		}
	}
	return 0;
}

int orc_apply_conditions_dense(size_t n_nodes, const double* coords, size_t n_edges, const size_t* edge_nodes,
                               const uint8_t* edge_boundary, size_t n_conditions, const int* kinds,
                               const double* values, const uint8_t* flags, double* K, double* b) {
	return orc_conditions_dense(n_nodes, coords, n_edges, edge_nodes, edge_boundary, n_conditions, kinds, values,
	                            flags, 0, K, b);
}

int orc_apply_conditions_dense_axi(size_t n_nodes, const double* coords, size_t n_edges, const size_t* edge_nodes,
                                   const uint8_t* edge_boundary, size_t n_conditions, const int* kinds,
                                   const double* values, const uint8_t* flags, double* K, double* b) {
	return orc_conditions_dense(n_nodes, coords, n_edges, edge_nodes, edge_boundary, n_conditions, kinds, values,
	                            flags, 1, K, b);
}

/* ------------------------------------------------------------------- RCM */

typedef struct {
	size_t i;
	size_t deg;
} orc_rcm_node_t;

/* Stable insertion sort by degree; neighbour lists hold at most a few dozen
 * entries. */
static void orc_sort_deg(orc_rcm_node_t* v, size_t n) {
	for (size_t i = 1; i < n; i++) {
		orc_rcm_node_t const t = v[i];
		size_t j = i;
		while (j > 0 && v[j - 1].deg > t.deg) {
			v[j] = v[j - 1];
			j--;
		}
		v[j] = t;
	}
}

/* Reverse Cuthill-McKee over the non-zero pattern of a dense matrix
 * (perm.c:117-331): degree = non-zeros in the row, start node = the last
 * unvisited node of minimum degree, neighbours visited in index order and
 * enqueued sorted by degree, result reversed. perm[old] = new. */
int orc_rcm(size_t n, const double* A, size_t* perm, size_t* inv_perm) {
	size_t* degs = calloc(n, sizeof *degs);
	size_t* queue = malloc(n * sizeof *queue);
	orc_rcm_node_t* to_sort = malloc(n * sizeof *to_sort);
	uint8_t* visited = calloc(n, 1);
	if (!degs || !queue || !to_sort || !visited) {
		free(degs), free(queue), free(to_sort), free(visited);
		return -1;
	}
	for (size_t i = 0; i < n; i++) {
		for (size_t j = 0; j < n; j++) {
			degs[i] += A[i * n + j] != 0;
		}
	}
	size_t qs = 0, qe = 0, p = 0, unvisited = n;
	while (unvisited && p < n) {
		size_t min_deg = (size_t) -1, start = 0;
		for (size_t i = 0; i < n; i++) {
			if (!visited[i] && degs[i] <= min_deg) {
				min_deg = degs[i];
				start = i;
			}
		}
		queue[qe++ % n] = start;
		while (qs != qe) {
			size_t const cur = queue[qs++ % n];
			if (!visited[cur]) {
				visited[cur] = 1;
				unvisited--;
			}
			size_t cnt = 0;
			for (size_t i = 0; i < n; i++) {
				if (visited[i] || A[cur * n + i] == 0) {
					continue;
				}
				to_sort[cnt].i = i;
				to_sort[cnt].deg = degs[i];
				cnt++;
				visited[i] = 1;
			}
			orc_sort_deg(to_sort, cnt);
			for (size_t i = 0; i < cnt; i++) {
				queue[qe++ % n] = to_sort[i].i;
			}
			inv_perm[n - p++ - 1] = cur;
		}
	}
	for (size_t i = 0; i < n; i++) {
		perm[inv_perm[i]] = i;
	}
	free(degs), free(queue), free(to_sort), free(visited);
	return 0;
}

/* bfm_perm_perm_matrix / perm_vec (perm.c:26-103): out[perm[i]][perm[j]] = A[i][j]. */
void orc_perm_matrix(size_t n, const size_t* perm, const double* A, double* out) {
	for (size_t i = 0; i < n; i++) {
		for (size_t j = 0; j < n; j++) {
			out[perm[i] * n + perm[j]] = A[i * n + j];
		}
	}
}

void orc_perm_vec(size_t n, const size_t* perm, const double* v, double* out) {
	for (size_t i = 0; i < n; i++) {
		out[perm[i]] = v[i];
	}
}

/* matrix_full_bandwidth (matrix.c:57-71). */
size_t orc_bandwidth(size_t n, const double* A) {
	size_t k = 0;
	for (size_t i = 0; i < n; i++) {
		for (size_t j = 0; j < n; j++) {
			if (A[i * n + j] != 0) {
				size_t const d = i > j ? i - j : j - i;
				k = d > k ? d : k;
			}
		}
	}
	return k;
}

/* ------------------------------------------------------------------ band
 * Row-major band storage of libbfm (matrix.c:208): entry (i, j) with
 * |i - j| <= k lives at data[j + i * 2k]; the array has m(2k+1) entries. */

static inline double* orc_band_at(double* band, size_t k, size_t i, size_t j) {
	return &band[j + i * 2 * k];
}

/* Dense to band copy (matrix.c:426-438 through matrix_band_set). */
int orc_band_from_dense(size_t m, size_t k, const double* A, double* band) {
	memset(band, 0, m * (2 * k + 1) * sizeof *band);
	for (size_t i = 0; i < m; i++) {
		for (size_t j = 0; j < m; j++) {
			double const v = A[i * m + j];
			size_t const d = i > j ? i - j : j - i;
			if (d > k) {
				if (fabs(v) >= 1e-20) {
					return -1;
				}
				continue;
			}
			*orc_band_at(band, k, i, j) = v;
		}
	}
	return 0;
}

/* The first `steps` pivot steps of the unpivoted band LU in place
 * (matrix.c:253-302); all of them when steps >= m - 1. */
int orc_band_lu_steps(size_t m, size_t k, double* band, size_t steps) {
	for (size_t p = 0; p + 1 < m && p < steps; p++) {
		double const pivot = *orc_band_at(band, k, p, p);
		if (pivot != pivot || fabs(pivot) < 1e-20) {
			return -1;
		}
		size_t const len = p + k + 1 < m ? p + k + 1 : m;
		for (size_t i = p + 1; i < len; i++) {
			double l = *orc_band_at(band, k, i, p);
			l /= pivot;
			*orc_band_at(band, k, i, p) = l;
			for (size_t j = p + 1; j < len; j++) {
				*orc_band_at(band, k, i, j) += -l * *orc_band_at(band, k, p, j);
			}
		}
	}
	return 0;
}

/* Unpivoted band LU in place (matrix.c:253-302). */
int orc_band_lu(size_t m, size_t k, double* band) {
	return orc_band_lu_steps(m, k, band, m);
}

/* Forward (unit L) and backward (U) substitution in place (matrix.c:351-404). */
int orc_band_lu_solve(size_t m, size_t k, double* band, double* y) {
	for (size_t p = 0; p < m; p++) {
		size_t const lo = p > k ? p - k : 0;
		for (size_t i = lo; i < p; i++) {
			y[p] -= *orc_band_at(band, k, p, i) * y[i];
		}
	}
	for (size_t pp = m; pp-- > 0;) {
		size_t const hi = pp + k + 1 < m ? pp + k + 1 : m;
		for (size_t i = pp + 1; i < hi; i++) {
			y[pp] -= y[i] * *orc_band_at(band, k, pp, i);
		}
		double const pivot = *orc_band_at(band, k, pp, pp);
		if (pivot != pivot || pivot == 0) {
			return -1;
		}
		y[pp] /= pivot;
	}
	return 0;
}

/* run_elasticity for one instance (sim.c:103-135): dense assembly, ordered
 * conditions, RCM renumbering, band LU, inverse permutation. u has 2 n_nodes
 * entries; K_out / b_out (optional, m*m and m) receive the condition-applied
 * system before renumbering; perm_out / k_out (optional) the RCM permutation
 * and half-bandwidth. */
int orc_solve_dense(int kind, size_t n_nodes, const double* coords, size_t n_elems, const size_t* elems, int stress,
                    double E, double nu, double rho, const double* elem_E, size_t n_forces, const double* forces,
                    size_t n_edges, const size_t* edge_nodes, const uint8_t* edge_boundary, size_t n_conditions,
                    const int* kinds, const double* values, const uint8_t* flags, double* u, double* K_out,
                    double* b_out, size_t* perm_out, size_t* k_out) {
	size_t const m = 2 * n_nodes;
	int rv = -1;
	double* K = malloc(m * m * sizeof *K);
	double* Kp = malloc(m * m * sizeof *Kp);
	double* b = malloc(m * sizeof *b);
	double* bp = malloc(m * sizeof *bp);
	size_t* perm = malloc(m * sizeof *perm);
	size_t* iperm = malloc(m * sizeof *iperm);
	double* band = NULL;
	if (!K || !Kp || !b || !bp || !perm || !iperm) {
		goto out;
	}
	if (orc_assemble_dense(kind, n_nodes, coords, n_elems, elems, stress, E, nu, rho, elem_E, n_forces, forces, K, b) < 0) {
		goto out;
	}
	if (orc_conditions_dense(n_nodes, coords, n_edges, edge_nodes, edge_boundary, n_conditions, kinds, values, flags,
	                         stress == 2, K, b) < 0) {
		goto out;
	}
	if (K_out) {
		memcpy(K_out, K, m * m * sizeof *K);
	}
	if (b_out) {
		memcpy(b_out, b, m * sizeof *b);
	}
	if (orc_rcm(m, K, perm, iperm) < 0) {
		goto out;
	}
	orc_perm_matrix(m, perm, K, Kp);
	orc_perm_vec(m, perm, b, bp);
	size_t const k = orc_bandwidth(m, Kp);
	band = malloc(m * (2 * k + 1) * sizeof *band);
	if (!band || orc_band_from_dense(m, k, Kp, band) < 0) {
		goto out;
	}
	if (perm_out) {
		memcpy(perm_out, perm, m * sizeof *perm);
	}
	if (k_out) {
		*k_out = k;
	}
	/* run_elasticity ignores the solve's return code (sim.c:122); the oracle reports it. */
	int const solved = orc_band_lu(m, k, band) == 0 && orc_band_lu_solve(m, k, band, bp) == 0;
	orc_perm_vec(m, iperm, bp, u);
	rv = solved ? 0 : -2;
out:
	free(K), free(Kp), free(b), free(bp), free(perm), free(iperm), free(band);
	return rv;
}

/* ------------------------------------------------------------------- CSR
 * Scalar CSR over DOFs with the node-graph pattern (every DOF pair of nodes
 * sharing an element), columns sorted. Entries are accumulated in the
 * reference's element and Gauss-point order, so each equals the dense entry. */

typedef struct {
	size_t m;
	int64_t* rowptr;
	int32_t* col;
	double* val;
} orc_csr_t;

static int64_t orc_csr_find(const orc_csr_t* A, size_t i, size_t j) {
	int64_t lo = A->rowptr[i], hi = A->rowptr[i + 1] - 1;
	while (lo <= hi) {
		int64_t const mid = (lo + hi) / 2;
		if ((size_t) A->col[mid] == j) {
			return mid;
		}
		if ((size_t) A->col[mid] < j) {
			lo = mid + 1;
		} else {
			hi = mid - 1;
		}
	}
	return -1;
}

static void orc_csr_add(void* ctx, size_t i, size_t j, double v) {
	orc_csr_t* A = ctx;
	A->val[orc_csr_find(A, i, j)] += v;
}

static int cmp_i32(const void* a, const void* b) {
	int32_t const x = *(const int32_t*) a, y = *(const int32_t*) b;
	return (x > y) - (x < y);
}

/* Number of non-zeros of the DOF pattern; call with rowptr/col NULL to size
 * the arrays, then again to fill them (rowptr m+1, col nnz). */
int64_t orc_csr_pattern(int kind, size_t n_nodes, size_t n_elems, const size_t* elems, int64_t* rowptr, int32_t* col) {
	/* node adjacency via per-node candidate lists */
	int64_t* cnt = calloc(n_nodes + 1, sizeof *cnt);
	if (!cnt) {
		return -1;
	}
	for (size_t e = 0; e < n_elems; e++) {
		for (int a = 0; a < kind; a++) {
			cnt[elems[e * kind + a] + 1] += kind;
		}
	}
	for (size_t i = 0; i < n_nodes; i++) {
		cnt[i + 1] += cnt[i];
	}
	int32_t* cand = malloc((size_t) cnt[n_nodes] * sizeof *cand);
	int64_t* fill = malloc(n_nodes * sizeof *fill);
	int64_t* nadj = malloc((n_nodes + 1) * sizeof *nadj);
	if (!cand || !fill || !nadj) {
		free(cnt), free(cand), free(fill), free(nadj);
		return -1;
	}
	memcpy(fill, cnt, n_nodes * sizeof *fill);
	for (size_t e = 0; e < n_elems; e++) {
		for (int a = 0; a < kind; a++) {
			size_t const na = elems[e * kind + a];
			for (int b = 0; b < kind; b++) {
				cand[fill[na]++] = (int32_t) elems[e * kind + b];
			}
		}
	}
	nadj[0] = 0;
	for (size_t i = 0; i < n_nodes; i++) {
		int32_t* const c = cand + cnt[i];
		size_t const len = (size_t) (cnt[i + 1] - cnt[i]);
		qsort(c, len, sizeof *c, cmp_i32);
		size_t u = 0;
		for (size_t t = 0; t < len; t++) {
			if (u == 0 || c[u - 1] != c[t]) {
				c[u++] = c[t];
			}
		}
		nadj[i + 1] = nadj[i] + (int64_t) u;
		fill[i] = (int64_t) u;
	}
	int64_t const nnz = 4 * nadj[n_nodes];
	if (rowptr && col) {
		rowptr[0] = 0;
		for (size_t i = 0; i < n_nodes; i++) {
			int64_t const len = fill[i];
			for (int r = 0; r < 2; r++) {
				int64_t const base = rowptr[2 * i + r];
				for (int64_t t = 0; t < len; t++) {
					int32_t const nj = cand[cnt[i] + t];
					col[base + 2 * t + 0] = 2 * nj + 0;
					col[base + 2 * t + 1] = 2 * nj + 1;
				}
				rowptr[2 * i + r + 1] = base + 2 * len;
			}
		}
	}
	free(cnt), free(cand), free(fill), free(nadj);
	return nnz;
}

/* CSR assembly, same arithmetic and order as orc_assemble_dense. */
int orc_csr_assemble(int kind, size_t n_nodes, const double* coords, size_t n_elems, const size_t* elems, int stress,
                     double E, double nu, double rho, const double* elem_E, size_t n_forces, const double* forces,
                     const int64_t* rowptr, const int32_t* col, double* val, double* b) {
	if (kind != 3 && kind != 4) {
		return -1;
	}
	size_t const m = 2 * n_nodes;
	orc_csr_t A = {m, (int64_t*) rowptr, (int32_t*) col, val};
	memset(val, 0, (size_t) rowptr[m] * sizeof *val);
	memset(b, 0, m * sizeof *b);
	for (size_t e = 0; e < n_elems; e++) {
		orc_element(kind, elems + e * kind, coords, stress, elem_E ? elem_E[e] : E, nu, rho, n_forces, forces, b,
		            orc_csr_add, &A);
	}
	return 0;
}

/* apply_constraint on CSR: column d is visited through the rows of d's
 * neighbours (the pattern is structurally symmetric). */
static void orc_csr_constrain(orc_csr_t* A, double* b, size_t d, double v) {
	for (int64_t t = A->rowptr[d]; t < A->rowptr[d + 1]; t++) {
		size_t const i = (size_t) A->col[t];
		int64_t const s = orc_csr_find(A, i, d);
		b[i] -= v * A->val[s];
		A->val[s] = 0;
	}
	for (int64_t t = A->rowptr[d]; t < A->rowptr[d + 1]; t++) {
		A->val[t] = 0;
	}
	A->val[orc_csr_find(A, d, d)] = 1;
	b[d] = v;
}

/* The ordered conditions on CSR; axi selects the axisymmetric loop. */
int orc_csr_apply_conditions(size_t n_nodes, const double* coords, size_t n_edges, const size_t* edge_nodes,
                             const uint8_t* edge_boundary, size_t n_conditions, const int* kinds, const double* values,
                             const uint8_t* flags, int axi, const int64_t* rowptr, const int32_t* col, double* val,
                             double* b) {
	orc_csr_t A = {2 * n_nodes, (int64_t*) rowptr, (int32_t*) col, val};
	for (size_t c = 0; c < n_conditions; c++) {
		int const kind = kinds[c];
		double const v = values[c];
		const uint8_t* const f = flags + c * n_nodes;
		if (kind == 0 || kind == 1) {
			for (size_t j = 0; j < n_nodes; j++) {
				if (f[j]) {
					orc_csr_constrain(&A, b, 2 * j + (size_t) kind, v);
				}
			}
		} else if (kind >= 2 && kind <= 5) {
			orc_neumann(kind, v, f, coords, n_edges, edge_nodes, axi, b);
		} else if (kind == 6 || kind == 7) {
			for (size_t i = 0; i < n_nodes; i++) {
				if (!f[i]) {
					continue;
				}
				double tx, ty;
				orc_node_tangent(i, coords, n_edges, edge_nodes, edge_boundary, &tx, &ty);
				orc_csr_constrain(&A, b, 2 * i + 0, v * (kind == 7 ? tx : -ty));
				orc_csr_constrain(&A, b, 2 * i + 1, v * (kind == 7 ? ty : tx));
			}
		} else {
			return -1;
		}
	}
	return 0;
}

/* y = A x (FP64, OpenMP over rows). */
void orc_csr_spmv(size_t m, const int64_t* rowptr, const int32_t* col, const double* val, const double* x, double* y) {
#pragma omp parallel for schedule(static)
	for (size_t i = 0; i < m; i++) {
		double s = 0;
		for (int64_t t = rowptr[i]; t < rowptr[i + 1]; t++) {
			s += val[t] * x[col[t]];
		}
		y[i] = s;
	}
}

static double orc_dot(size_t m, const double* a, const double* b) {
	double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
	for (size_t i = 0; i < m; i++) {
		s += a[i] * b[i];
	}
	return s;
}

int orc_num_threads(void) {
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
	omp_set_num_threads(n);
#else
	(void) n;
#endif
}

/* Jacobi-preconditioned CG from x = 0 until ||r||/||b|| <= rtol (recursive
 * residual) or max_iter iterations; returns the iteration count, -1 on
 * allocation failure, -2 on breakdown. This has no reference counterpart:
 * it is the CPU baseline at sizes libbfm's dense path cannot hold, and the
 * tight-tolerance solution for checks on mid-size meshes. */
int64_t orc_csr_pcg(size_t m, const int64_t* rowptr, const int32_t* col, const double* val, const double* b, double* x,
                    double rtol, int64_t max_iter, double* rel_res) {
	double* r = malloc(m * sizeof *r);
	double* z = malloc(m * sizeof *z);
	double* p = malloc(m * sizeof *p);
	double* q = malloc(m * sizeof *q);
	double* dinv = malloc(m * sizeof *dinv);
	if (!r || !z || !p || !q || !dinv) {
		free(r), free(z), free(p), free(q), free(dinv);
		return -1;
	}
	for (size_t i = 0; i < m; i++) {
		double d = 0;
		for (int64_t t = rowptr[i]; t < rowptr[i + 1]; t++) {
			if ((size_t) col[t] == i) {
				d = val[t];
			}
		}
		dinv[i] = 1 / d;
		x[i] = 0;
		r[i] = b[i];
		z[i] = dinv[i] * r[i];
		p[i] = z[i];
	}
	double const bnorm = sqrt(orc_dot(m, b, b));
	double rz = orc_dot(m, r, z);
	int64_t it = 0;
	double rr = bnorm * bnorm;
	int64_t rv = 0;
	if (bnorm == 0) {
		*rel_res = 0;
		goto done;
	}
	while (it < max_iter && sqrt(rr) > rtol * bnorm) {
		orc_csr_spmv(m, rowptr, col, val, p, q);
		double const pq = orc_dot(m, p, q);
		if (!(pq > 0) || !isfinite(pq)) {
			rv = -2;
			break;
		}
		double const alpha = rz / pq;
		double rzn = 0, rrn = 0;
#pragma omp parallel for reduction(+ : rzn, rrn) schedule(static)
		for (size_t i = 0; i < m; i++) {
			x[i] += alpha * p[i];
			r[i] -= alpha * q[i];
			z[i] = dinv[i] * r[i];
			rzn += r[i] * z[i];
			rrn += r[i] * r[i];
		}
		double const beta = rzn / rz;
		rz = rzn;
		rr = rrn;
#pragma omp parallel for schedule(static)
		for (size_t i = 0; i < m; i++) {
			p[i] = z[i] + beta * p[i];
		}
		it++;
	}
	*rel_res = sqrt(rr) / bnorm;
done:
	free(r), free(z), free(p), free(q), free(dinv);
	return rv < 0 ? rv : it;
}


/* ------------------------------------------------------- sparse band path
 * libbfm's direct path (RCM, band LU, substitution) with the matrix kept
 * sparse until it is laid out as a band: the same algorithm at sizes whose
 * dense m x m matrix cannot be held. Used as the direct CPU baseline at
 * scale; the LU is timed over a window of its pivot steps (the band LU does
 * the same arithmetic at every interior step, whatever the values). */

/* orc_rcm over a CSR matrix (perm.c:117-331 rules): degree = entries whose
 * value is non-zero (stored zeros, as left by Dirichlet elimination, do not
 * count), start = the last unvisited node of minimum degree, neighbours in
 * index order, enqueued sorted by degree, order reversed. perm[old] = new. */
int orc_rcm_csr(size_t n, const int64_t* rowptr, const int32_t* col, const double* val, size_t* perm,
                size_t* inv_perm) {
	size_t* degs = calloc(n, sizeof *degs);
	size_t* queue = malloc(n * sizeof *queue);
	orc_rcm_node_t* to_sort = NULL;
	uint8_t* visited = calloc(n, 1);
	size_t max_row = 0;
	for (size_t i = 0; i < n; i++) {
		size_t const len = (size_t) (rowptr[i + 1] - rowptr[i]);
		max_row = len > max_row ? len : max_row;
	}
	to_sort = malloc((max_row + 1) * sizeof *to_sort);
	if (!degs || !queue || !to_sort || !visited) {
		free(degs), free(queue), free(to_sort), free(visited);
		return -1;
	}
	for (size_t i = 0; i < n; i++) {
		for (int64_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
			degs[i] += val[e] != 0;
		}
	}
	size_t qs = 0, qe = 0, p = 0, unvisited = n;
	while (unvisited && p < n) {
		size_t min_deg = (size_t) -1, start = 0;
		for (size_t i = 0; i < n; i++) {
			if (!visited[i] && degs[i] <= min_deg) {
				min_deg = degs[i];
				start = i;
			}
		}
		queue[qe++ % n] = start;
		while (qs != qe) {
			size_t const cur = queue[qs++ % n];
			if (!visited[cur]) {
				visited[cur] = 1;
				unvisited--;
			}
			size_t cnt = 0;
			for (int64_t e = rowptr[cur]; e < rowptr[cur + 1]; e++) {
				size_t const i = (size_t) col[e];
				if (visited[i] || val[e] == 0) {
					continue;
				}
				to_sort[cnt].i = i;
				to_sort[cnt].deg = degs[i];
				cnt++;
				visited[i] = 1;
			}
			orc_sort_deg(to_sort, cnt);
			for (size_t i = 0; i < cnt; i++) {
				queue[qe++ % n] = to_sort[i].i;
			}
			inv_perm[n - p++ - 1] = cur;
		}
	}
	for (size_t i = 0; i < n; i++) {
		perm[inv_perm[i]] = i;
	}
	free(degs), free(queue), free(to_sort), free(visited);
	return 0;
}

/* matrix_full_bandwidth (matrix.c:57-71) of P A P^T over the non-zero values. */
size_t orc_bandwidth_csr(size_t n, const int64_t* rowptr, const int32_t* col, const double* val, const size_t* perm) {
	size_t k = 0;
	for (size_t i = 0; i < n; i++) {
		for (int64_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
			if (val[e] != 0) {
				size_t const a = perm[i], b = perm[col[e]];
				size_t const d = a > b ? a - b : b - a;
				k = d > k ? d : k;
			}
		}
	}
	return k;
}

/* Rows r0 .. r0 + nrows - 1 of P A P^T in libbfm's band storage of
 * half-bandwidth k, row and column indices counted from r0 (entries whose
 * column falls outside the window are left out); band has nrows (2k+1)
 * entries. -1 when a non-zero lies outside the band. */
int orc_band_window_csr(size_t n, const int64_t* rowptr, const int32_t* col, const double* val, const size_t* perm,
                        const size_t* inv_perm, size_t k, size_t r0, size_t nrows, double* band) {
	memset(band, 0, nrows * (2 * k + 1) * sizeof *band);
	for (size_t li = 0; li < nrows && r0 + li < n; li++) {
		size_t const o = inv_perm[r0 + li];
		for (int64_t e = rowptr[o]; e < rowptr[o + 1]; e++) {
			size_t const j = perm[col[e]];
			size_t const i = r0 + li;
			size_t const d = i > j ? i - j : j - i;
			if (d > k) {
				if (val[e] != 0) {
					return -1;
				}
				continue;
			}
			if (j < r0 || j >= r0 + nrows) {
				continue;
			}
			*orc_band_at(band, k, li, j - r0) = val[e];
		}
	}
	return 0;
}
