// Device-side element assembly and boundary conditions, one element, node or
// row per call: shared by the streaming kernels of assemble.hip (one lane per
// item, one launch per colour or condition) and the batched persistent
// kernel of batch.hip (one workgroup per system), so both produce bitwise the
// same matrix and load from the same mesh.
//
// Element arithmetic follows fill_elasticity_elem (libbfm/src/system.c:136-224
// in the reference tree): shape-function gradients and |det J| at each Gauss
// point, B^T D B per node pair summed over the points in registers, the body
// load rho f(x_j) phi_j; the element matrix is made exactly symmetric.
// Conditions follow create_planar's loop (system.c:358-526).
//
// Every function body sets `fp contract(on)`: a multiply-add is fused only
// where one source expression holds both, decided by the front end, so the
// rounding does not depend on the kernel the function is inlined into (the
// default `fast` contraction lets the back end fuse across statements,
// differently in differently shaped kernels).

#pragma once

#include <cmath>

#include "internal.h"

namespace bfmx {
namespace dev {

template <int KIND>
struct Rule;

// 3-point triangle rule (libbfm/src/rule.c:94-102).
template <>
struct Rule<3> {
	static constexpr int NG = 3;
	__device__ static constexpr double xi(int g) { return g == 1 ? 1.0 - 1.0 / 3 : 1.0 / 6; }
	__device__ static constexpr double eta(int g) { return g == 2 ? 1.0 - 1.0 / 3 : 1.0 / 6; }
	__device__ static constexpr double w(int) { return 1.0 / 6; }
	__device__ static void shape(double xs, double et, double* phi, double* dxs, double* det) {
#pragma clang fp contract(on)
		phi[0] = 1 - xs - et;
		phi[1] = xs;
		phi[2] = et;
		dxs[0] = -1, dxs[1] = 1, dxs[2] = 0;
		det[0] = -1, det[1] = 0, det[2] = 1;
	}
};

// 2x2 Gauss-Legendre rule with libbfm's abscissa literal (rule.c:92, 104-111).
template <>
struct Rule<4> {
	static constexpr int NG = 4;
	static constexpr double kA = 0.577350269189626;
	__device__ static constexpr double xi(int g) { return g < 2 ? -kA : kA; }
	__device__ static constexpr double eta(int g) { return (g == 0 || g == 3) ? kA : -kA; }
	__device__ static constexpr double w(int) { return 1.0; }
	__device__ static void shape(double xs, double et, double* phi, double* dxs, double* det) {
#pragma clang fp contract(on)
		phi[0] = (1 + xs) * (1 + et) / 4;
		phi[1] = (1 - xs) * (1 + et) / 4;
		phi[2] = (1 - xs) * (1 - et) / 4;
		phi[3] = (1 + xs) * (1 - et) / 4;
		dxs[0] = (1 + et) / 4;
		dxs[1] = (-1 - et) / 4;
		dxs[2] = (-1 + et) / 4;
		dxs[3] = (1 - et) / 4;
		det[0] = (1 + xs) / 4;
		det[1] = (1 - xs) / 4;
		det[2] = (-1 + xs) / 4;
		det[3] = (-1 - xs) / 4;
	}
};

// Adds a 2x2 block at block slot blk of row lane r (sliced-ELL values).
__device__ __forceinline__ void add_block(double* __restrict__ val, int blk, int r, double a00, double a01,
                                          double a10, double a11) {
#pragma clang fp contract(on)
	if (blk < 0) {
		return;   // pair missing from the pattern (cannot happen for a pattern built from these elements)
	}
	double2* top = reinterpret_cast<double2*>(val + 4 * static_cast<int64_t>(blk) - 2 * r);
	double2* bot = top + kSlice;
	double2 t = *top;
	double2 u = *bot;
	t.x += a00;
	t.y += a01;
	u.x += a10;
	u.y += a11;
	*top = t;
	*bot = u;
}

// Block slot of (row, target) in the sliced-ELL pattern, -1 when absent.
__device__ __forceinline__ int32_t find_slot(const int32_t* __restrict__ slice_ptr,
                                             const int32_t* __restrict__ row_len, const int32_t* __restrict__ col,
                                             int32_t row, int32_t target) {
	int32_t const base = slice_ptr[row >> 6] + (row & 63);
	int32_t const len = row_len[row];
	for (int k = 0; k < len; k++) {
		if (col[base + k * kSlice] == target) {
			return base + k * kSlice;
		}
	}
	return -1;
}

// Material of an assembly: D-matrix inputs, body load.
struct Material {
	const double* elemE;   // per element (colour order) or null
	double E, nu, rho;
	int stress;            // 0 plane strain, 1 plane stress, 2 axisymmetric
	const double* nforce;  // per node body acceleration (2 per node) or null
	double gx, gy;         // constant body acceleration when nforce is null
};

// Assembles element e (colour-order position) into val and b. slot(j, k)
// gives the block slot of (node j, node k) of the element.
template <int KIND, typename Slot>
__device__ __forceinline__ void assemble_element(int64_t e, const int32_t* __restrict__ el,
                                                 const double* __restrict__ xy, const Material& M, Slot slot,
                                                 double* __restrict__ val, double* __restrict__ b) {
#pragma clang fp contract(on)
	using R = Rule<KIND>;
	int n[KIND];
	double x[KIND], y[KIND];
#pragma unroll
	for (int j = 0; j < KIND; j++) {
		n[j] = el[e * KIND + j];
		double2 const p = reinterpret_cast<const double2*>(xy)[n[j]];
		x[j] = p.x;
		y[j] = p.y;
	}
	double const E = M.elemE ? M.elemE[e] : M.E;
	double const nu = M.nu;
	// D-matrix constants (system.c:454-458)
	double const a = !M.stress ? E * (1 - nu) / (1 + nu) / (1 - 2 * nu) : E / (1 - nu * nu);
	double const bb = !M.stress ? E * nu / (1 + nu) / (1 - 2 * nu) : E * nu / (1 - nu * nu);
	double const c = E / (2 * (1 + nu));

	double gxs[R::NG][KIND], gys[R::NG][KIND], phis[R::NG][KIND], djw[R::NG];
#pragma unroll
	for (int g = 0; g < R::NG; g++) {
		double phi[KIND], dxs[KIND], det[KIND];
		R::shape(R::xi(g), R::eta(g), phi, dxs, det);
		double dx_dxsi = 0, dx_deta = 0, dy_dxsi = 0, dy_deta = 0;
#pragma unroll
		for (int j = 0; j < KIND; j++) {
			dx_dxsi += x[j] * dxs[j];
			dx_deta += x[j] * det[j];
			dy_dxsi += y[j] * dxs[j];
			dy_deta += y[j] * det[j];
		}
		double const det_J = fabs(dx_dxsi * dy_deta - dx_deta * dy_dxsi);
#pragma unroll
		for (int j = 0; j < KIND; j++) {
			gxs[g][j] = (dxs[j] * dy_deta - det[j] * dy_dxsi) / det_J;
			gys[g][j] = (det[j] * dx_dxsi - dxs[j] * dx_deta) / det_J;
			phis[g][j] = phi[j];
		}
		djw[g] = det_J * R::w(g);
	}

	// body load: rho f(x_j) phi_j, f evaluated at the node (system.c:190-203)
#pragma unroll
	for (int j = 0; j < KIND; j++) {
		double fx = M.gx, fy = M.gy;
		if (M.nforce) {
			fx = M.nforce[2 * n[j]];
			fy = M.nforce[2 * n[j] + 1];
		}
		double bx = 0, by = 0;
#pragma unroll
		for (int g = 0; g < R::NG; g++) {
			bx += djw[g] * fx * M.rho * phis[g][j];
			by += djw[g] * fy * M.rho * phis[g][j];
		}
		double2* bp = reinterpret_cast<double2*>(b) + n[j];
		double2 v = *bp;
		v.x += bx;
		v.y += by;
		*bp = v;
	}

	// stiffness blocks (system.c:207-223); upper triangle, mirrored
#pragma unroll
	for (int j = 0; j < KIND; j++) {
#pragma unroll
		for (int k = j; k < KIND; k++) {
			double s11 = 0, s12 = 0, s21 = 0, s22 = 0;
#pragma unroll
			for (int g = 0; g < R::NG; g++) {
				double const xj = gxs[g][j], yj = gys[g][j], xk = gxs[g][k], yk = gys[g][k];
				s11 += djw[g] * (a * xj * xk + c * yj * yk);
				s12 += djw[g] * (bb * xj * yk + c * yj * xk);
				s21 += djw[g] * (bb * yj * xk + c * xj * yk);
				s22 += djw[g] * (a * yj * yk + c * xj * xk);
			}
			if (j == k) {
				s21 = s12;
			}
			add_block(val, slot(j, k, n), n[j] & 63, s11, s12, s21, s22);
			if (k != j) {
				add_block(val, slot(k, j, n), n[k] & 63, s11, s21, s12, s22);
			}
		}
	}
}

// Axisymmetric element (fill_axisymmetric_elem, system.c:228-356), r = x:
// the reference's constants (a multiplied by 1 - 2 nu where plane strain
// divides, system.c:241-243) and its f_11 .. f_22 (f_22's second term without
// r, system.c:339), Gauss-point sums in registers, upper triangle mirrored
// (f_21(k, j) equals f_12(j, k)); body load rho f phi_j r.
template <int KIND, typename Slot>
__device__ __forceinline__ void assemble_element_axi(int64_t e, const int32_t* __restrict__ el,
                                                     const double* __restrict__ xy, const Material& M, Slot slot,
                                                     double* __restrict__ val, double* __restrict__ b) {
#pragma clang fp contract(on)
	using R = Rule<KIND>;
	int n[KIND];
	double x[KIND], y[KIND];
#pragma unroll
	for (int j = 0; j < KIND; j++) {
		n[j] = el[e * KIND + j];
		double2 const p = reinterpret_cast<const double2*>(xy)[n[j]];
		x[j] = p.x;
		y[j] = p.y;
	}
	double const E = M.elemE ? M.elemE[e] : M.E;
	double const nu = M.nu;
	double const a = E * (1 - nu) / (1 + nu) * (1 - 2 * nu);
	double const bb = E * nu / (1 + nu) / (1 - 2 * nu);
	double const c = E / (2 * (1 + nu));

	double gxs[R::NG][KIND], gys[R::NG][KIND], phis[R::NG][KIND], djw[R::NG], rg[R::NG];
#pragma unroll
	for (int g = 0; g < R::NG; g++) {
		double phi[KIND], dxs[KIND], det[KIND];
		R::shape(R::xi(g), R::eta(g), phi, dxs, det);
		double dx_dxsi = 0, dx_deta = 0, dy_dxsi = 0, dy_deta = 0, r = 0;
#pragma unroll
		for (int j = 0; j < KIND; j++) {
			dx_dxsi += x[j] * dxs[j];
			dx_deta += x[j] * det[j];
			dy_dxsi += y[j] * dxs[j];
			dy_deta += y[j] * det[j];
			r += x[j] * phi[j];
		}
		double const det_J = fabs(dx_dxsi * dy_deta - dx_deta * dy_dxsi);
#pragma unroll
		for (int j = 0; j < KIND; j++) {
			gxs[g][j] = (dxs[j] * dy_deta - det[j] * dy_dxsi) / det_J;
			gys[g][j] = (det[j] * dx_dxsi - dxs[j] * dx_deta) / det_J;
			phis[g][j] = phi[j];
		}
		djw[g] = det_J * R::w(g);
		rg[g] = r;
	}

#pragma unroll
	for (int j = 0; j < KIND; j++) {
		double fx = M.gx, fy = M.gy;
		if (M.nforce) {
			fx = M.nforce[2 * n[j]];
			fy = M.nforce[2 * n[j] + 1];
		}
		double bx = 0, by = 0;
#pragma unroll
		for (int g = 0; g < R::NG; g++) {
			bx += djw[g] * fx * M.rho * phis[g][j] * rg[g];
			by += djw[g] * fy * M.rho * phis[g][j] * rg[g];
		}
		double2* bp = reinterpret_cast<double2*>(b) + n[j];
		double2 v = *bp;
		v.x += bx;
		v.y += by;
		*bp = v;
	}

#pragma unroll
	for (int j = 0; j < KIND; j++) {
#pragma unroll
		for (int k = j; k < KIND; k++) {
			double s11 = 0, s12 = 0, s21 = 0, s22 = 0;
#pragma unroll
			for (int g = 0; g < R::NG; g++) {
				double const xj = gxs[g][j], yj = gys[g][j], xk = gxs[g][k], yk = gys[g][k];
				double const pj = phis[g][j], pk = phis[g][k], r = rg[g];
				s11 += djw[g] * (a * xj * xk * r + c * yj * yk * r + pj * (bb * xk + a * pk / r) + xj * bb * pk);
				s12 += djw[g] * (bb * xj * yk * r + c * yj * xk * r + pj * bb * yk);
				s21 += djw[g] * (bb * yj * xk * r + c * xj * yk * r + yj * bb * pk);
				s22 += djw[g] * (a * yj * yk * r + c * xj * xk);
			}
			if (j == k) {
				s21 = s12;
			}
			add_block(val, slot(j, k, n), n[j] & 63, s11, s12, s21, s22);
			if (k != j) {
				add_block(val, slot(k, j, n), n[k] & 63, s11, s21, s12, s22);
			}
		}
	}
}

// Element e with the material's formulation.
template <int KIND, typename Slot>
__device__ __forceinline__ void assemble_any(int64_t e, const int32_t* __restrict__ el, const double* __restrict__ xy,
                                             const Material& M, Slot slot, double* __restrict__ val,
                                             double* __restrict__ b) {
	if (M.stress == 2) {
		assemble_element_axi<KIND>(e, el, xy, M, slot, val, b);
	} else {
		assemble_element<KIND>(e, el, xy, M, slot, val, b);
	}
}

// Dirichlet X or Y on node i: b[d] = v; the first value of each DOF is kept
// for the symmetric elimination.
__device__ __forceinline__ void dirichlet_xy(int32_t i, int shift, double v, uint8_t* __restrict__ fixed,
                                             double* __restrict__ vfirst, double* __restrict__ b) {
#pragma clang fp contract(on)
	int64_t const d = 2 * static_cast<int64_t>(i) + shift;
	if (!fixed[d]) {
		fixed[d] = 1;
		vfirst[d] = v;
	}
	b[d] = v;
}

// Dirichlet normal/tangent on node i (system.c:388-425): tangent = half the
// sum of unit vectors from each adjacent boundary-edge neighbour to the node.
template <typename IP>
__device__ __forceinline__ void dirichlet_nt(int32_t i, int tangent, double v, const IP* __restrict__ inc_ptr,
                                             const int32_t* __restrict__ inc, const int32_t* __restrict__ edge_nodes,
                                             const uint8_t* __restrict__ edge_bnd, const double* __restrict__ xy,
                                             uint8_t* __restrict__ fixed, double* __restrict__ vfirst,
                                             double* __restrict__ b) {
#pragma clang fp contract(on)
	double const xi = xy[2 * i], yi = xy[2 * i + 1];
	double tx = 0, ty = 0;
	for (int64_t q = inc_ptr[i]; q < inc_ptr[i + 1]; q++) {
		int32_t const e = inc[q];
		if (q > inc_ptr[i] && inc[q - 1] == e) {
			continue;   // degenerate edge listed twice; the reference visits it once
		}
		if (!edge_bnd[e]) {
			continue;
		}
		int32_t const n2 = edge_nodes[2 * e] == i ? edge_nodes[2 * e + 1] : edge_nodes[2 * e];
		double const dx = xi - xy[2 * n2], dy = yi - xy[2 * n2 + 1];
		double const length = sqrt(dx * dx + dy * dy);
		tx += dx / length / 2;
		ty += dy / length / 2;
	}
	double const val[2] = {v * (tangent ? tx : -ty), v * (tangent ? ty : tx)};
	for (int k = 0; k < 2; k++) {
		int64_t const d = 2 * static_cast<int64_t>(i) + k;
		if (!fixed[d]) {
			fixed[d] = 1;
			vfirst[d] = val[k];
		}
		b[d] = val[k];
	}
}

// Neumann loads of one condition on flagged node i (system.c:477-522): the
// incident flagged edges' loads in edge order. Axisymmetric (axi,
// system.c:584-617): X/Y loads carry the factor r1 + r2, each r formed as
// the reference forms it from one node's x and y; normal/tangent loads are
// not applied (the caller skips them).
template <typename IP>
__device__ __forceinline__ void neumann(int32_t i, int kind, double v, const uint8_t* __restrict__ flags,
                                        const IP* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                                        const int32_t* __restrict__ edge_nodes, const double* __restrict__ xy,
                                        bool axi, double* __restrict__ b) {
#pragma clang fp contract(on)
	double bx = b[2 * i], by = b[2 * i + 1];
	for (int64_t q = inc_ptr[i]; q < inc_ptr[i + 1]; q++) {
		int32_t const e = inc[q];
		int32_t const n1 = edge_nodes[2 * e], n2 = edge_nodes[2 * e + 1];
		if (!flags[n1] || !flags[n2]) {
			continue;
		}
		double const dx = xy[2 * n1] - xy[2 * n2];
		double const dy = xy[2 * n1 + 1] - xy[2 * n2 + 1];
		if (kind == BFMX_NEUMANN_X || kind == BFMX_NEUMANN_Y) {
			double const jac = sqrt(dx * dx + dy * dy) / 2;
			double load = jac * v;
			if (axi) {
				double const s3 = 1 / sqrt(3.0);
				double const r1 = xy[2 * n1] * (1 - s3) / 2 + xy[2 * n1 + 1] * (1 + s3) / 2;
				double const r2 = xy[2 * n2] * (1 - s3) / 2 + xy[2 * n2 + 1] * (1 + s3) / 2;
				load = (r1 + r2) * jac * v;
			}
			if (kind == BFMX_NEUMANN_X) {
				bx += load;
			} else {
				by += load;
			}
		} else {
			bool const tan = kind == BFMX_NEUMANN_TANGENT;
			bx += 0.5 * v * (tan ? dx : -dy);
			by += 0.5 * v * (tan ? dy : dx);
		}
	}
	b[2 * i] = bx;
	b[2 * i + 1] = by;
}

// Symmetric elimination of the constrained DOFs on row i: constrained rows
// become unit rows, constrained columns of free rows move v_first to b.
__device__ __forceinline__ void eliminate_row(int32_t i, const int32_t* __restrict__ slice_ptr,
                                              const int32_t* __restrict__ row_len, const int32_t* __restrict__ col,
                                              const uint8_t* __restrict__ fixed, const double* __restrict__ vfirst,
                                              double* __restrict__ val, double* __restrict__ b) {
#pragma clang fp contract(on)
	int const r = i & 63;
	int32_t const base = slice_ptr[i >> 6] + r;
	int32_t const len = row_len[i];
	bool const fr[2] = {fixed[2 * i] != 0, fixed[2 * i + 1] != 0};
	double bl[2] = {b[2 * i], b[2 * i + 1]};
	for (int k = 0; k < len; k++) {
		int32_t const blk = base + k * kSlice;
		int32_t const j = col[blk];
		double* top = val + 4 * static_cast<int64_t>(blk) - 2 * r;
		double* rowp[2] = {top, top + 2 * kSlice};
		bool const fc[2] = {fixed[2 * j] != 0, fixed[2 * j + 1] != 0};
		for (int rr = 0; rr < 2; rr++) {
			for (int cc = 0; cc < 2; cc++) {
				double& a = rowp[rr][cc];
				if (fr[rr]) {
					a = (j == i && rr == cc) ? 1.0 : 0.0;
				} else if (fc[cc]) {
					bl[rr] -= vfirst[2 * j + cc] * a;
					a = 0.0;
				}
			}
		}
	}
	for (int rr = 0; rr < 2; rr++) {
		if (!fr[rr]) {
			b[2 * i + rr] = bl[rr];
		}
	}
}

}  // namespace dev
}  // namespace bfmx
