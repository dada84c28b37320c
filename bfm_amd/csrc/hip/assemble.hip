// Element assembly and boundary conditions on the device.
//
// Assembly: one lane per element, launched once per element colour. The
// lane gathers its nodes, evaluates the Gauss rule (shape-function
// gradients, |det J|, B^T D B per node pair, body load) in registers and adds
// its 2x2 blocks straight into the sliced-ELL values; elements of one colour
// share no node, so there are no atomics and the summation order is fixed.
// Arithmetic follows fill_elasticity_elem (libbfm/src/system.c:136-224 in the
// reference tree); the Gauss-point contributions are summed in registers
// before the single add, and the element matrix is made exactly symmetric.
//
// Conditions: applied in call order with the semantics of create_planar's
// condition loop (system.c:358-526):
//   Dirichlet (X, Y, normal, tangent): b[d] = v; the first value of each DOF
//     is remembered for the symmetric elimination, done once at the end
//     (rows of never-constrained DOFs subtract v_first(d) K[i][d], constrained
//     rows and columns are zeroed, unit diagonal).
//   Neumann: each node adds its incident flagged edges' loads in edge order,
//     the order in which the reference adds them, so b matches it bit for bit
//     given the same body load.

#include <cmath>

#include "internal.h"

namespace bfmx {

namespace {

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

__device__ __forceinline__ void add_block(double* __restrict__ val, int blk, int r, double a00, double a01,
                                          double a10, double a11) {
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

// slots[pos * K*K + a*K + b] = block index of (node a, node b) of element pos.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_slots(int64_t ne, const int32_t* __restrict__ el,
                                                  const int32_t* __restrict__ slice_ptr,
                                                  const int32_t* __restrict__ row_len,
                                                  const int32_t* __restrict__ col, int32_t* __restrict__ slots) {
	int64_t const t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (t >= ne * KIND) {
		return;
	}
	int64_t const e = t / KIND;
	int const a = static_cast<int>(t % KIND);
	int32_t const row = el[e * KIND + a];
	int32_t const base = slice_ptr[row >> 6] + (row & 63);
	int32_t const len = row_len[row];
	for (int b = 0; b < KIND; b++) {
		int32_t const target = el[e * KIND + b];
		int32_t found = -1;
		for (int k = 0; k < len; k++) {
			if (col[base + k * kSlice] == target) {
				found = base + k * kSlice;
				break;
			}
		}
		slots[e * KIND * KIND + a * KIND + b] = found;
	}
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_assemble(int64_t e0, int64_t e1, const int32_t* __restrict__ el,
                                                     const int32_t* __restrict__ slots,
                                                     const double* __restrict__ xy,
                                                     const double* __restrict__ elemE, double E0, double nu,
                                                     int stress, double rho, const double* __restrict__ nforce,
                                                     double gfx, double gfy, double* __restrict__ val,
                                                     double* __restrict__ b) {
	using R = Rule<KIND>;
	int64_t const e = e0 + static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (e >= e1) {
		return;
	}
	int n[KIND];
	double x[KIND], y[KIND];
#pragma unroll
	for (int j = 0; j < KIND; j++) {
		n[j] = el[e * KIND + j];
		double2 const p = reinterpret_cast<const double2*>(xy)[n[j]];
		x[j] = p.x;
		y[j] = p.y;
	}
	double const E = elemE ? elemE[e] : E0;
	// D-matrix constants (system.c:454-458)
	double const a = !stress ? E * (1 - nu) / (1 + nu) / (1 - 2 * nu) : E / (1 - nu * nu);
	double const bb = !stress ? E * nu / (1 + nu) / (1 - 2 * nu) : E * nu / (1 - nu * nu);
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
		double fx = gfx, fy = gfy;
		if (nforce) {
			fx = nforce[2 * n[j]];
			fy = nforce[2 * n[j] + 1];
		}
		double bx = 0, by = 0;
#pragma unroll
		for (int g = 0; g < R::NG; g++) {
			bx += djw[g] * fx * rho * phis[g][j];
			by += djw[g] * fy * rho * phis[g][j];
		}
		double2* bp = reinterpret_cast<double2*>(b) + n[j];
		double2 v = *bp;
		v.x += bx;
		v.y += by;
		*bp = v;
	}

	// stiffness blocks (system.c:207-223); upper triangle, mirrored
	const int32_t* sl = slots + e * KIND * KIND;
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
			add_block(val, sl[j * KIND + k], n[j] & 63, s11, s12, s21, s22);
			if (k != j) {
				add_block(val, sl[k * KIND + j], n[k] & 63, s11, s21, s12, s22);
			}
		}
	}
}

__global__ __launch_bounds__(kBlock) void k_dirichlet_xy(int cnt, const int32_t* __restrict__ nodes, int shift,
                                                         double v, uint8_t* __restrict__ fixed,
                                                         double* __restrict__ vfirst, double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t >= cnt) {
		return;
	}
	int64_t const d = 2 * static_cast<int64_t>(nodes[t]) + shift;
	if (!fixed[d]) {
		fixed[d] = 1;
		vfirst[d] = v;
	}
	b[d] = v;
}

// Dirichlet normal/tangent (system.c:388-425): tangent = half the sum of unit
// vectors from each adjacent boundary-edge neighbour to the node.
__global__ __launch_bounds__(kBlock) void k_dirichlet_nt(int cnt, const int32_t* __restrict__ nodes, int tangent,
                                                         double v, const int64_t* __restrict__ inc_ptr,
                                                         const int32_t* __restrict__ inc,
                                                         const int32_t* __restrict__ edge_nodes,
                                                         const uint8_t* __restrict__ edge_bnd,
                                                         const double* __restrict__ xy, uint8_t* __restrict__ fixed,
                                                         double* __restrict__ vfirst, double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t >= cnt) {
		return;
	}
	int32_t const i = nodes[t];
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

// Neumann loads of one condition (system.c:477-522), per flagged node.
__global__ __launch_bounds__(kBlock) void k_neumann(int cnt, const int32_t* __restrict__ nodes, int kind, double v,
                                                    const uint8_t* __restrict__ flags,
                                                    const int64_t* __restrict__ inc_ptr,
                                                    const int32_t* __restrict__ inc,
                                                    const int32_t* __restrict__ edge_nodes,
                                                    const double* __restrict__ xy, double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t >= cnt) {
		return;
	}
	int32_t const i = nodes[t];
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
			if (kind == BFMX_NEUMANN_X) {
				bx += jac * v;
			} else {
				by += jac * v;
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

// Symmetric elimination of all constrained DOFs on the rows that hold them or
// reference them.
__global__ __launch_bounds__(kBlock) void k_eliminate(int cnt, const int32_t* __restrict__ rows,
                                                      const int32_t* __restrict__ slice_ptr,
                                                      const int32_t* __restrict__ row_len,
                                                      const int32_t* __restrict__ col,
                                                      const uint8_t* __restrict__ fixed,
                                                      const double* __restrict__ vfirst, double* __restrict__ val,
                                                      double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t >= cnt) {
		return;
	}
	int32_t const i = rows[t];
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

int grid(int64_t n) {
	return static_cast<int>((n + kBlock - 1) / kBlock);
}

}  // namespace

int compute_slots(bfmx_system* s) {
	int64_t const work = s->n_elems * s->kind;
	if (work == 0) {
		return BFMX_OK;
	}
	if (s->kind == 3) {
		k_slots<3><<<grid(work), kBlock, 0, s->stream>>>(s->n_elems, s->d_elem, s->d_slice_ptr, s->d_row_len, s->d_col,
		                                                s->d_slots);
	} else {
		k_slots<4><<<grid(work), kBlock, 0, s->stream>>>(s->n_elems, s->d_elem, s->d_slice_ptr, s->d_row_len, s->d_col,
		                                                s->d_slots);
	}
	BFMX_HIP(hipGetLastError());
	return BFMX_OK;
}

int assemble(bfmx_system* s, int stress, double E, double nu, double rho, const double* elemE,
             const double* node_force, double gx, double gy) {
	hipStream_t const st = s->stream;
	int64_t const n = s->n_nodes;
	double* d_force = nullptr;
	if (elemE) {
		std::vector<double> eE(s->n_elems);
		for (int64_t p = 0; p < s->n_elems; p++) {
			eE[p] = elemE[s->elem_order[p]];
		}
		if (!s->d_elemE) {
			BFMX_CHECK(dalloc(&s->d_elemE, s->n_elems));
		}
		BFMX_HIP(hipMemcpyAsync(s->d_elemE, eE.data(), eE.size() * sizeof(double), hipMemcpyHostToDevice, st));
		BFMX_HIP(hipStreamSynchronize(st));
	}
	if (node_force) {
		std::vector<double> f(2 * n);
		for (int64_t i = 0; i < n; i++) {
			f[2 * s->perm[i]] = node_force[2 * i];
			f[2 * s->perm[i] + 1] = node_force[2 * i + 1];
		}
		BFMX_CHECK(dalloc(&d_force, 2 * n));
		BFMX_HIP(hipMemcpyAsync(d_force, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, st));
		BFMX_HIP(hipStreamSynchronize(st));
	}
	BFMX_HIP(hipMemsetAsync(s->d_val, 0, 4 * s->nnzb_pad * sizeof(double), st));
	BFMX_HIP(hipMemsetAsync(s->d_b, 0, 2 * n * sizeof(double), st));
	BFMX_HIP(hipMemsetAsync(s->d_fixed, 0, 2 * n, st));
	BFMX_HIP(hipMemsetAsync(s->d_vfirst, 0, 2 * n * sizeof(double), st));
	const double* eE = elemE ? s->d_elemE : nullptr;
	for (int c = 0; c < s->n_colors; c++) {
		int64_t const e0 = s->color_ptr[c], e1 = s->color_ptr[c + 1];
		if (e1 == e0) {
			continue;
		}
		if (s->kind == 3) {
			k_assemble<3><<<grid(e1 - e0), kBlock, 0, st>>>(e0, e1, s->d_elem, s->d_slots, s->d_coords, eE, E, nu,
			                                               stress, rho, d_force, gx, gy, s->d_val, s->d_b);
		} else {
			k_assemble<4><<<grid(e1 - e0), kBlock, 0, st>>>(e0, e1, s->d_elem, s->d_slots, s->d_coords, eE, E, nu,
			                                               stress, rho, d_force, gx, gy, s->d_val, s->d_b);
		}
		BFMX_HIP(hipGetLastError());
	}
	if (d_force) {
		BFMX_HIP(hipStreamSynchronize(st));
		dfree(d_force);
	}
	s->stress = stress;
	s->assembled = true;
	s->finalized = false;
	s->has_u = false;
	s->dirichlet_nodes.assign(n, 0);
	return BFMX_OK;
}

int apply_condition(bfmx_system* s, int kind, double value, const uint8_t* flags) {
	if (!s->assembled) {
		set_error("apply_condition before assemble");
		return BFMX_E_ARG;
	}
	if (kind < 0 || kind > 7) {
		set_error("unknown condition kind %d", kind);
		return BFMX_E_ARG;
	}
	bool const neumann = kind >= BFMX_NEUMANN_X && kind <= BFMX_NEUMANN_TANGENT;
	bool const nt = kind == BFMX_DIRICHLET_NORMAL || kind == BFMX_DIRICHLET_TANGENT;
	if ((neumann || nt) && s->inc_ptr.empty()) {
		// no edge list: Neumann adds nothing, tangents are zero
		if (neumann) {
			return BFMX_OK;
		}
		std::vector<size_t> none;
		BFMX_CHECK(set_edges(s, 0, none.data(), nullptr));
	}
	int64_t const n = s->n_nodes;
	std::vector<int32_t> list;
	std::vector<uint8_t> fl(n, 0);
	for (int64_t i = 0; i < n; i++) {
		if (flags[i]) {
			fl[s->perm[i]] = 1;
		}
	}
	for (int64_t i = 0; i < n; i++) {
		if (fl[i]) {
			list.push_back(static_cast<int32_t>(i));
		}
	}
	s->finalized = false;
	s->has_u = false;
	if (list.empty()) {
		return BFMX_OK;
	}
	hipStream_t const st = s->stream;
	BFMX_HIP(hipMemcpyAsync(s->d_list, list.data(), list.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
	int const cnt = static_cast<int>(list.size());
	if (kind == BFMX_DIRICHLET_X || kind == BFMX_DIRICHLET_Y) {
		k_dirichlet_xy<<<grid(cnt), kBlock, 0, st>>>(cnt, s->d_list, kind, value, s->d_fixed, s->d_vfirst, s->d_b);
	} else if (nt) {
		k_dirichlet_nt<<<grid(cnt), kBlock, 0, st>>>(cnt, s->d_list, kind == BFMX_DIRICHLET_TANGENT, value,
		                                             s->d_inc_ptr, s->d_inc, s->d_edge_nodes, s->d_edge_bnd,
		                                             s->d_coords, s->d_fixed, s->d_vfirst, s->d_b);
	} else {
		BFMX_HIP(hipMemcpyAsync(s->d_flags, fl.data(), n, hipMemcpyHostToDevice, st));
		k_neumann<<<grid(cnt), kBlock, 0, st>>>(cnt, s->d_list, kind, value, s->d_flags, s->d_inc_ptr, s->d_inc,
		                                        s->d_edge_nodes, s->d_coords, s->d_b);
	}
	BFMX_HIP(hipGetLastError());
	if (!neumann) {
		for (int32_t i : list) {
			s->dirichlet_nodes[i] = 1;
		}
	}
	// list and fl are pageable and local: wait for the copies
	BFMX_HIP(hipStreamSynchronize(st));
	return BFMX_OK;
}

int finalize_conditions(bfmx_system* s) {
	if (s->finalized) {
		return BFMX_OK;
	}
	int64_t const n = s->n_nodes;
	std::vector<uint8_t> touched(n, 0);
	for (int64_t i = 0; i < n; i++) {
		if (s->dirichlet_nodes[i]) {
			for (int64_t t = s->adj_ptr[i]; t < s->adj_ptr[i + 1]; t++) {
				touched[s->adj[t]] = 1;
			}
		}
	}
	std::vector<int32_t> rows;
	for (int64_t i = 0; i < n; i++) {
		if (touched[i]) {
			rows.push_back(static_cast<int32_t>(i));
		}
	}
	if (!rows.empty()) {
		hipStream_t const st = s->stream;
		BFMX_HIP(hipMemcpyAsync(s->d_rows, rows.data(), rows.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
		int const cnt = static_cast<int>(rows.size());
		k_eliminate<<<grid(cnt), kBlock, 0, st>>>(cnt, s->d_rows, s->d_slice_ptr, s->d_row_len, s->d_col, s->d_fixed,
		                                          s->d_vfirst, s->d_val, s->d_b);
		BFMX_HIP(hipGetLastError());
		BFMX_HIP(hipStreamSynchronize(st));
	}
	s->finalized = true;
	return BFMX_OK;
}

}  // namespace bfmx
