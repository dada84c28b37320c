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
//
// The per-element, per-node and per-row bodies are the device functions of
// assemble_dev.h, which the batched persistent kernel shares.

#include <cmath>

#include "assemble_dev.h"

namespace bfmx {

namespace {

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
	for (int b = 0; b < KIND; b++) {
		slots[e * KIND * KIND + a * KIND + b] = dev::find_slot(slice_ptr, row_len, col, row, el[e * KIND + b]);
	}
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_assemble(int64_t e0, int64_t e1, const int32_t* __restrict__ el,
                                                     const int32_t* __restrict__ slots,
                                                     const double* __restrict__ xy, dev::Material M,
                                                     double* __restrict__ val, double* __restrict__ b) {
	int64_t const e = e0 + static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (e >= e1) {
		return;
	}
	const int32_t* sl = slots + e * KIND * KIND;
	dev::assemble_any<KIND>(e, el, xy, M, [&](int j, int k, const int*) { return sl[j * KIND + k]; }, val, b);
}

__global__ __launch_bounds__(kBlock) void k_dirichlet_xy(int cnt, const int32_t* __restrict__ nodes, int shift,
                                                         double v, uint8_t* __restrict__ fixed,
                                                         double* __restrict__ vfirst, double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t < cnt) {
		dev::dirichlet_xy(nodes[t], shift, v, fixed, vfirst, b);
	}
}

__global__ __launch_bounds__(kBlock) void k_dirichlet_nt(int cnt, const int32_t* __restrict__ nodes, int tangent,
                                                         double v, const int64_t* __restrict__ inc_ptr,
                                                         const int32_t* __restrict__ inc,
                                                         const int32_t* __restrict__ edge_nodes,
                                                         const uint8_t* __restrict__ edge_bnd,
                                                         const double* __restrict__ xy, uint8_t* __restrict__ fixed,
                                                         double* __restrict__ vfirst, double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t < cnt) {
		dev::dirichlet_nt(nodes[t], tangent, v, inc_ptr, inc, edge_nodes, edge_bnd, xy, fixed, vfirst, b);
	}
}

__global__ __launch_bounds__(kBlock) void k_neumann(int cnt, const int32_t* __restrict__ nodes, int kind, double v,
                                                    const uint8_t* __restrict__ flags,
                                                    const int64_t* __restrict__ inc_ptr,
                                                    const int32_t* __restrict__ inc,
                                                    const int32_t* __restrict__ edge_nodes,
                                                    const double* __restrict__ xy, int axi, double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t < cnt) {
		dev::neumann(nodes[t], kind, v, flags, inc_ptr, inc, edge_nodes, xy, axi != 0, b);
	}
}

__global__ __launch_bounds__(kBlock) void k_eliminate(int cnt, const int32_t* __restrict__ rows,
                                                      const int32_t* __restrict__ slice_ptr,
                                                      const int32_t* __restrict__ row_len,
                                                      const int32_t* __restrict__ col,
                                                      const uint8_t* __restrict__ fixed,
                                                      const double* __restrict__ vfirst, double* __restrict__ val,
                                                      double* __restrict__ b) {
	int const t = blockIdx.x * kBlock + threadIdx.x;
	if (t < cnt) {
		dev::eliminate_row(rows[t], slice_ptr, row_len, col, fixed, vfirst, val, b);
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
	s->val_gen++;
	s->e_spread = 1.0;
	if (elemE) {
		std::vector<double> eE(s->n_elems);
		double lo = 0, hi = 0;
		for (int64_t p = 0; p < s->n_elems; p++) {
			eE[p] = elemE[s->elem_order[p]];
			double const a = std::fabs(eE[p]);
			lo = p == 0 ? a : std::min(lo, a);
			hi = p == 0 ? a : std::max(hi, a);
		}
		s->e_spread = lo > 0 ? hi / lo : std::numeric_limits<double>::infinity();
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
	dev::Material const M{elemE ? s->d_elemE : nullptr, E, nu, rho, stress, d_force, gx, gy};
	for (int c = 0; c < s->n_colors; c++) {
		int64_t const e0 = s->color_ptr[c], e1 = s->color_ptr[c + 1];
		if (e1 == e0) {
			continue;
		}
		if (s->kind == 3) {
			k_assemble<3><<<grid(e1 - e0), kBlock, 0, st>>>(e0, e1, s->d_elem, s->d_slots, s->d_coords, M, s->d_val,
			                                               s->d_b);
		} else {
			k_assemble<4><<<grid(e1 - e0), kBlock, 0, st>>>(e0, e1, s->d_elem, s->d_slots, s->d_coords, M, s->d_val,
			                                               s->d_b);
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
	s->val_gen++;
	bool const neumann = kind >= BFMX_NEUMANN_X && kind <= BFMX_NEUMANN_TANGENT;
	bool const nt = kind == BFMX_DIRICHLET_NORMAL || kind == BFMX_DIRICHLET_TANGENT;
	if (s->stress == 2 && (kind == BFMX_NEUMANN_NORMAL || kind == BFMX_NEUMANN_TANGENT)) {
		return BFMX_OK;   // the axisymmetric condition loop has no such branch (system.c:575-619)
	}
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
		                                        s->d_edge_nodes, s->d_coords, s->stress == 2, s->d_b);
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
	s->val_gen++;
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
