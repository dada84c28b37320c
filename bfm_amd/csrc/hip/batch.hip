// Batched solves: many independent planar systems of uneven sizes solved by
// one persistent kernel launch per size class, the launches captured in one
// HIP graph.
//
// One workgroup runs the whole Jacobi or block-Jacobi PCG of one system: its
// passes over the rows (SpMV with the direction update folded into the
// gathers, the vector updates, the residual recomputation) are separated by
// workgroup barriers, and its dot products are workgroup reductions in a
// fixed order, so an iteration costs no kernel launch and no device-scope
// synchronisation. Systems are dealt to workgroups largest first; the
// hardware dispatcher balances them over the CUs as workgroups finish.
// A system's result does not depend on the other systems of the batch: a
// batch of one with the same configuration gives bitwise the same solution.
//
// Tuning: the workgroup size and the rows each thread takes per step (the
// independent row chains in flight per thread) are template parameters. The
// autotuner times every configuration on a sample of each size class and
// records its picks, which bfmx_batch_solve replays.
//
// Workspace: one packed FP64 allocation, per system x, r, z, p0, p1, q (2 per
// node each) and the Jacobi inverses (4 per node).
//
// Multigrid mode (BFMX_PRECOND_AMG): the batch solved as one block-diagonal
// system by the multigrid flexible CG of pcg.hip / amg.hip, every system a
// convergence segment with its own CG scalars and tolerance check (below).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include <atomic>
#include <cstring>
#include <string>
#include <thread>

#include "assemble_dev.h"

namespace bfmx {

namespace {

typedef double dbl2 __attribute__((ext_vector_type(2)));

// One condition of a batch built from meshes: its node list (renumbered,
// ascending) and, for Neumann, the per-node flags.
struct CondRec {
	int32_t kind;
	int32_t count;
	int64_t list_off;    // into the batch's lists
	int64_t flags_off;   // into the batch's Neumann flags (n bytes), Neumann only
	double value;
};

struct BatchInst {
	const int32_t* slice_ptr;
	const int32_t* row_len;
	const int32_t* col;
	double* val;
	const int32_t* diag_blk;
	double* b;
	double* x;
	double* r;
	double* z;
	double* p0;
	double* p1;
	double* q;
	double* dinv;
	int64_t n;
	// assembly inputs (batches built from meshes)
	int64_t nnzb_pad;
	const double* coords;
	const int32_t* elem;        // kind per element, colour order
	const int32_t* color_ptr;   // n_colors + 1
	int32_t n_colors;
	int32_t kind;
	dev::Material M;
	const CondRec* conds;
	int32_t n_cond;
	int32_t n_elim;
	const int32_t* lists;
	const uint8_t* nflags;
	const int32_t* elim;        // rows touched by the Dirichlet elimination
	const int64_t* inc_ptr;
	const int32_t* inc;
	const int32_t* edge_nodes;
	const uint8_t* edge_bnd;
	uint8_t* fixed;
	double* vfirst;
};

struct BatchOut {
	int64_t iter;
	double rr_true;
	double bb;
	int32_t status;   // 1 converged, 2 breakdown, 3 stagnated, 0 iteration limit
	int32_t checks;
};

// Configurations: workgroup size x rows per thread per step.
constexpr int kConfigs = 5;
constexpr int kCfgBS[kConfigs] = {256, 256, 512, 512, 1024};
constexpr int kCfgRPT[kConfigs] = {1, 2, 1, 2, 1};
// Size classes by nodes: [0, 2048), [2048, 8192), [8192, inf).
constexpr int kClasses = 3;
constexpr int64_t kClassEdge[kClasses - 1] = {2048, 8192};
constexpr int kDefaultPick[kClasses] = {0, 2, 4};
constexpr int kStallChecks = 8;
constexpr int kCheckEvery = 16;   // iterations between true-residual checks once one has failed

int size_class(int64_t n) {
	return n < kClassEdge[0] ? 0 : n < kClassEdge[1] ? 1 : 2;
}

__device__ __forceinline__ dbl2 ld2(const double* p) {
	return *reinterpret_cast<const dbl2*>(p);
}
__device__ __forceinline__ void st2(double* p, dbl2 v) {
	*reinterpret_cast<dbl2*>(p) = v;
}

// Workgroup sum of NV values (fixed order), returned to every thread.
template <int BS, int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double (*lds)[BS / 64]) {
#pragma unroll
	for (int k = 0; k < NV; k++) {
		double a = v[k];
#pragma unroll
		for (int off = 32; off > 0; off >>= 1) {
			a += __shfl_xor(a, off, 64);
		}
		v[k] = a;
	}
	__syncthreads();   // the previous reduction's reads of lds are done
	if ((threadIdx.x & 63) == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			lds[k][threadIdx.x >> 6] = v[k];
		}
	}
	__syncthreads();
#pragma unroll
	for (int k = 0; k < NV; k++) {
		double s = 0;
#pragma unroll
		for (int w = 0; w < BS / 64; w++) {
			s += lds[k][w];
		}
		v[k] = s;
	}
}

template <int PRE>
__device__ __forceinline__ dbl2 precond(const double* dinv, int64_t i, dbl2 r) {
	if (PRE == BFMX_PRECOND_JACOBI) {
		dbl2 const d = ld2(dinv + 2 * i);
		return dbl2{d.x * r.x, d.y * r.y};
	} else {
		dbl2 const d0 = ld2(dinv + 4 * i), d1 = ld2(dinv + 4 * i + 2);
		return dbl2{d0.x * r.x + d0.y * r.y, d1.x * r.x + d1.y * r.y};
	}
}

// y = A v for row i, v(c) the value at column c.
template <typename LoadV>
__device__ __forceinline__ dbl2 row_mul(const BatchInst& S, int64_t i, LoadV ldv) {
	int const r = static_cast<int>(i & 63);
	int32_t const beg = S.slice_ptr[i >> 6] + r;
	int const len = S.row_len[i];
	const double* v = S.val + 4 * static_cast<int64_t>(beg) - 2 * r;
	double y0 = 0, y1 = 0;
	// unrolled so that the column, value and gather loads of several blocks
	// are in flight together (a row holds 3-9 blocks)
#pragma unroll 4
	for (int k = 0; k < len; k++, v += 4 * kSlice) {
		int32_t const c = S.col[beg + k * kSlice];
		dbl2 const t = ld2(v), u = ld2(v + 2 * kSlice);
		dbl2 const pv = ldv(c);
		y0 += t.x * pv.x + t.y * pv.y;
		y1 += u.x * pv.x + u.y * pv.y;
	}
	return dbl2{y0, y1};
}

// Visits this thread's rows: RPT independent rows per step of BS * RPT.
template <int BS, int RPT, typename F>
__device__ __forceinline__ void for_rows(int64_t n, F f) {
	for (int64_t base = 0; base < n; base += BS * RPT) {
#pragma unroll
		for (int k = 0; k < RPT; k++) {
			int64_t const i = base + k * BS + threadIdx.x;
			if (i < n) {
				f(i);
			}
		}
	}
}

// Assembly of one system by one workgroup, in the order of the streaming
// kernels: zeroed values and load, the element colours one after another
// (no two elements of a colour share a node), the conditions in call order,
// then the symmetric elimination.
template <int BS, int KIND>
__device__ void assemble_elements(const BatchInst& S) {
	for (int c = 0; c < S.n_colors; c++) {
		int32_t const e1 = S.color_ptr[c + 1];
		for (int64_t e = S.color_ptr[c] + threadIdx.x; e < e1; e += BS) {
			dev::assemble_any<KIND>(
			    e, S.elem, S.coords, S.M,
			    [&](int j, int k, const int* nd) { return dev::find_slot(S.slice_ptr, S.row_len, S.col, nd[j], nd[k]); },
			    S.val, S.b);
		}
		__syncthreads();
	}
}

template <int BS>
__device__ void assemble_system(const BatchInst& S) {
	int64_t const n = S.n;
	for (int64_t k = threadIdx.x; k < 2 * S.nnzb_pad; k += BS) {
		st2(S.val + 2 * k, dbl2{0, 0});
	}
	for (int64_t i = threadIdx.x; i < n; i += BS) {
		st2(S.b + 2 * i, dbl2{0, 0});
		st2(S.vfirst + 2 * i, dbl2{0, 0});
		S.fixed[2 * i] = 0;
		S.fixed[2 * i + 1] = 0;
	}
	__syncthreads();
	if (S.kind == 3) {
		assemble_elements<BS, 3>(S);
	} else {
		assemble_elements<BS, 4>(S);
	}
	for (int q = 0; q < S.n_cond; q++) {
		CondRec const C = S.conds[q];
		for (int t = threadIdx.x; t < C.count; t += BS) {
			int32_t const i = S.lists[C.list_off + t];
			if (C.kind == BFMX_DIRICHLET_X || C.kind == BFMX_DIRICHLET_Y) {
				dev::dirichlet_xy(i, C.kind, C.value, S.fixed, S.vfirst, S.b);
			} else if (C.kind == BFMX_DIRICHLET_NORMAL || C.kind == BFMX_DIRICHLET_TANGENT) {
				dev::dirichlet_nt(i, C.kind == BFMX_DIRICHLET_TANGENT, C.value, S.inc_ptr, S.inc, S.edge_nodes,
				                  S.edge_bnd, S.coords, S.fixed, S.vfirst, S.b);
			} else {
				dev::neumann(i, C.kind, C.value, S.nflags + C.flags_off, S.inc_ptr, S.inc, S.edge_nodes, S.coords,
				             S.M.stress == 2,
				             S.b);
			}
		}
		__syncthreads();
	}
	for (int t = threadIdx.x; t < S.n_elim; t += BS) {
		dev::eliminate_row(S.elim[t], S.slice_ptr, S.row_len, S.col, S.fixed, S.vfirst, S.val, S.b);
	}
	__syncthreads();
}

// Assembly of every system of a batch built from meshes, one workgroup per
// system. A launch of its own: the element arithmetic holds the Gauss-point
// gradients in registers (about 200 VGPRs for Q1), which inside the solve
// kernel would cut its occupancy from 6-8 waves per SIMD to 2.
constexpr int kAsmBlock = 256;

__global__ __launch_bounds__(kAsmBlock) void k_batch_assemble(const BatchInst* __restrict__ insts,
                                                              const int32_t* __restrict__ order) {
	BatchInst const S = insts[order[blockIdx.x]];
	assemble_system<kAsmBlock>(S);
}

template <int BS, int RPT, int PRE>
__global__ __launch_bounds__(BS) void k_batch_pcg(const BatchInst* __restrict__ insts,
                                                  const int32_t* __restrict__ order, double rtol2, int64_t max_iter,
                                                  BatchOut* __restrict__ out) {
	__shared__ double lds[3][BS / 64];
	int32_t const id = order[blockIdx.x];
	BatchInst const S = insts[id];
	int64_t const n = S.n;
	// Jacobi inverses from the diagonal blocks
	for_rows<BS, RPT>(n, [&](int64_t i) {
		int const r = static_cast<int>(i & 63);
		const double* top = S.val + 4 * static_cast<int64_t>(S.diag_blk[i]) - 2 * r;
		double const a00 = top[0], a01 = top[1], a10 = top[2 * kSlice], a11 = top[2 * kSlice + 1];
		if (PRE == BFMX_PRECOND_JACOBI) {
			S.dinv[2 * i] = 1.0 / a00;
			S.dinv[2 * i + 1] = 1.0 / a11;
		} else {
			double const det = a00 * a11 - a01 * a10;
			S.dinv[4 * i] = a11 / det;
			S.dinv[4 * i + 1] = -a01 / det;
			S.dinv[4 * i + 2] = -a10 / det;
			S.dinv[4 * i + 3] = a00 / det;
		}
		st2(S.x + 2 * i, dbl2{0, 0});
		st2(S.p0 + 2 * i, dbl2{0, 0});
	});
	__syncthreads();
	// r = b - A x (x = 0 here, later the residual recomputation), z = M^-1 r
	auto residual = [&](double& rz, double& rr, double& bb) {
		double v[3] = {0, 0, 0};
		for_rows<BS, RPT>(n, [&](int64_t i) {
			dbl2 const ax = row_mul(S, i, [&](int32_t c) { return ld2(S.x + 2 * static_cast<int64_t>(c)); });
			dbl2 const bv = ld2(S.b + 2 * i);
			dbl2 const rv = bv - ax;
			dbl2 const zv = precond<PRE>(S.dinv, i, rv);
			st2(S.r + 2 * i, rv);
			st2(S.z + 2 * i, zv);
			v[0] += rv.x * zv.x + rv.y * zv.y;
			v[1] += rv.x * rv.x + rv.y * rv.y;
			v[2] += bv.x * bv.x + bv.y * bv.y;
		});
		wg_sum<BS, 3>(v, lds);
		rz = v[0];
		rr = v[1];
		bb = v[2];
	};
	double rz, rr, bb;
	residual(rz, rr, bb);
	double const tol2 = rtol2 * bb;
	double rr_true = rr, best = 0;
	int status = rr <= tol2 ? 1 : 0;
	int stall = 0, checks = 0;
	bool checking = false;
	int64_t it = 0;
	double beta = 0;
	double* pold = S.p0;
	double* pnew = S.p1;
	while (status == 0 && it < max_iter) {
		__syncthreads();   // z, r of the previous pass are written
		// p_new = z + beta p_old, q = A p_new, p_new . q
		double v1[1] = {0};
		for_rows<BS, RPT>(n, [&](int64_t i) {
			dbl2 const y = row_mul(S, i, [&](int32_t c) {
				int64_t const cc = 2 * static_cast<int64_t>(c);
				return ld2(S.z + cc) + beta * ld2(pold + cc);
			});
			dbl2 const pn = ld2(S.z + 2 * i) + beta * ld2(pold + 2 * i);
			st2(pnew + 2 * i, pn);
			st2(S.q + 2 * i, y);
			v1[0] += pn.x * y.x + pn.y * y.y;
		});
		wg_sum<BS, 1>(v1, lds);
		double const pq = v1[0];
		if (!(pq > 0) || !isfinite(pq)) {
			status = 2;
			break;
		}
		double const alpha = rz / pq;
		__syncthreads();   // every gather of z and p_old is done before z changes
		double v2[2] = {0, 0};
		for_rows<BS, RPT>(n, [&](int64_t i) {
			dbl2 const xv = ld2(S.x + 2 * i) + alpha * ld2(pnew + 2 * i);
			dbl2 const rv = ld2(S.r + 2 * i) - alpha * ld2(S.q + 2 * i);
			dbl2 const zv = precond<PRE>(S.dinv, i, rv);
			st2(S.x + 2 * i, xv);
			st2(S.r + 2 * i, rv);
			st2(S.z + 2 * i, zv);
			v2[0] += rv.x * zv.x + rv.y * zv.y;
			v2[1] += rv.x * rv.x + rv.y * rv.y;
		});
		wg_sum<BS, 2>(v2, lds);
		it++;
		double const rz_prev = rz;
		beta = v2[0] / rz_prev;
		rz = v2[0];
		rr = v2[1];
		double* t = pold;
		pold = pnew;
		pnew = t;
		if (rr <= tol2 || (checking && it % kCheckEvery == 0)) {
			// the true residual decides; a failed check keeps the directions
			__syncthreads();
			double rz_t, bb_t;
			residual(rz_t, rr_true, bb_t);
			checks++;
			if (rr_true <= tol2) {
				status = 1;
			} else {
				checking = true;
				beta = rz_prev != 0 ? rz_t / rz_prev : 0.0;
				rz = rz_t;
				rr = rr_true;
				if (best == 0 || rr_true < 0.25 * best) {
					best = rr_true;
					stall = 0;
				} else if (++stall >= kStallChecks) {
					status = 3;
				}
			}
		}
	}
	if (status != 1 && status != 2 && checks == 0) {
		__syncthreads();
		double rz_t, bb_t;
		residual(rz_t, rr_true, bb_t);   // final true residual for the report
	}
	if (threadIdx.x == 0) {
		out[id].iter = it;
		out[id].rr_true = rr_true;
		out[id].bb = bb;
		out[id].status = status;
		out[id].checks = checks;
	}
}

template <int PRE>
void launch_cfg(int cfg, int nblocks, hipStream_t st, const BatchInst* insts, const int32_t* order, double rtol2,
                int64_t max_iter, BatchOut* out) {
	switch (cfg) {
		case 0: k_batch_pcg<256, 1, PRE><<<nblocks, 256, 0, st>>>(insts, order, rtol2, max_iter, out); break;
		case 1: k_batch_pcg<256, 2, PRE><<<nblocks, 256, 0, st>>>(insts, order, rtol2, max_iter, out); break;
		case 2: k_batch_pcg<512, 1, PRE><<<nblocks, 512, 0, st>>>(insts, order, rtol2, max_iter, out); break;
		case 3: k_batch_pcg<512, 2, PRE><<<nblocks, 512, 0, st>>>(insts, order, rtol2, max_iter, out); break;
		default: k_batch_pcg<1024, 1, PRE><<<nblocks, 1024, 0, st>>>(insts, order, rtol2, max_iter, out); break;
	}
}

void launch(int pre, int cfg, int nblocks, hipStream_t st, const BatchInst* insts, const int32_t* order,
            double rtol2, int64_t max_iter, BatchOut* out) {
	if (pre == BFMX_PRECOND_JACOBI) {
		launch_cfg<BFMX_PRECOND_JACOBI>(cfg, nblocks, st, insts, order, rtol2, max_iter, out);
	} else {
		launch_cfg<BFMX_PRECOND_BLOCK_JACOBI>(cfg, nblocks, st, insts, order, rtol2, max_iter, out);
	}
}

// ------------------------------------------------------------ multigrid mode
// The batch as one block-diagonal system: the instances' rows one after
// another (each in its own renumbered order), solved by flexible CG with the
// multigrid V-cycle. No entry couples two instances, so no aggregate spans
// two and the coarsest level is inverted instance by instance. Every
// instance's load is scaled to unit norm, and the instances are the union's
// convergence segments: each runs its own CG scalars (when each has at least
// kBlock rows) and the union's tolerance checks test every instance's own
// true residual, so the run ends when each meets rtol.
constexpr int kUniBlock = 256;

// Per instance: b . b (fixed summation order).
__global__ __launch_bounds__(kUniBlock) void k_batch_bnorm(const BatchInst* __restrict__ insts,
                                                           double* __restrict__ bb) {
	__shared__ double lds[1][kUniBlock / 64];
	BatchInst const S = insts[blockIdx.x];
	double v[1] = {0};
	for (int64_t i = threadIdx.x; i < 2 * S.n; i += kUniBlock) {
		v[0] += S.b[i] * S.b[i];
	}
	wg_sum<kUniBlock, 1>(v, lds);
	if (threadIdx.x == 0) {
		bb[blockIdx.x] = v[0];
	}
}

// Instance of union row R: off[k] <= R < off[k + 1].
__device__ __forceinline__ int instance_of(int64_t R, int m, const int64_t* __restrict__ off) {
	int lo = 0, hi = m;
	while (hi - lo > 1) {
		int const mid = (lo + hi) >> 1;
		if (off[mid] <= R) {
			lo = mid;
		} else {
			hi = mid;
		}
	}
	return lo;
}

// Union row R from row i of its instance: the row's blocks (same column
// order, union slots at the union's slice offsets; padding slots stay
// zero), the load scaled by 1 / ||b_k|| and the fixed flags.
__global__ __launch_bounds__(kBlock) void k_union_gather(int64_t n, int m, const int64_t* __restrict__ off,
                                                         const BatchInst* __restrict__ insts,
                                                         const double* __restrict__ bb,
                                                         const int32_t* __restrict__ uslice, double* __restrict__ uval,
                                                         double* __restrict__ ub, uint8_t* __restrict__ ufixed) {
	int64_t const R = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
	if (R >= n) {
		return;
	}
	int const k = instance_of(R, m, off);
	const BatchInst& S = insts[k];
	int64_t const i = R - off[k];
	const double* src = S.val + 4 * static_cast<int64_t>(S.slice_ptr[i >> 6]) + 2 * (i & 63);
	double* dst = uval + 4 * static_cast<int64_t>(uslice[R >> 6]) + 2 * (R & 63);
	int const len = S.row_len[i];
	for (int t = 0; t < len; t++, src += 4 * kSlice, dst += 4 * kSlice) {
		st2(dst, ld2(src));
		st2(dst + 2 * kSlice, ld2(src + 2 * kSlice));
	}
	double const bk = bb[k];
	double const sc = bk > 0 ? 1.0 / sqrt(bk) : 0.0;
	st2(ub + 2 * R, sc * ld2(S.b + 2 * i));
	ufixed[2 * R] = S.fixed[2 * i];
	ufixed[2 * R + 1] = S.fixed[2 * i + 1];
}

// x of every instance from the union's solution, unscaled by ||b_k||.
__global__ __launch_bounds__(kBlock) void k_union_scatter(int64_t n, int m, const int64_t* __restrict__ off,
                                                          const BatchInst* __restrict__ insts,
                                                          const double* __restrict__ bb,
                                                          const double* __restrict__ ux) {
	int64_t const R = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
	if (R >= n) {
		return;
	}
	int const k = instance_of(R, m, off);
	int64_t const i = R - off[k];
	st2(insts[k].x + 2 * i, sqrt(bb[k]) * ld2(ux + 2 * R));
}

// Per instance: the true residual ||b - A x||^2 of its own system, and the
// report of the union's solve.
__global__ __launch_bounds__(kUniBlock) void k_batch_resid(const BatchInst* __restrict__ insts,
                                                           const double* __restrict__ bb, int64_t iters,
                                                           double rtol2, BatchOut* __restrict__ out) {
	__shared__ double lds[1][kUniBlock / 64];
	BatchInst const S = insts[blockIdx.x];
	double v[1] = {0};
	for (int64_t i = threadIdx.x; i < S.n; i += kUniBlock) {
		dbl2 const ax = row_mul(S, i, [&](int32_t c) { return ld2(S.x + 2 * static_cast<int64_t>(c)); });
		dbl2 const rv = ld2(S.b + 2 * i) - ax;
		v[0] += rv.x * rv.x + rv.y * rv.y;
	}
	wg_sum<kUniBlock, 1>(v, lds);
	if (threadIdx.x == 0) {
		BatchOut& o = out[blockIdx.x];
		o.iter = iters;
		o.rr_true = v[0];
		o.bb = bb[blockIdx.x];
		o.status = v[0] <= rtol2 * bb[blockIdx.x] ? 1 : 0;
		o.checks = 0;
	}
}

}  // namespace
}  // namespace bfmx

struct bfmx_batch {
	int device = 0;
	hipStream_t stream = nullptr;
	std::vector<bfmx_system*> systems;
	bfmx::BatchInst* d_inst = nullptr;
	bfmx::BatchOut* d_out = nullptr;
	int32_t* d_order = nullptr;             // per class: systems largest first
	std::vector<int32_t> class_off;         // kClasses + 1 offsets into d_order
	int32_t* d_order_all = nullptr;         // all systems largest first (one launch for equal picks)
	double* d_work = nullptr;
	std::vector<bfmx::BatchOut> h_out;
	struct Exec {
		hipGraphExec_t exec;
		int pre;
		double rtol2;
		int64_t max_iter;
		int picks[bfmx::kClasses];
	};
	std::vector<Exec> graphs;
	hipEvent_t ev[2] = {};
	// batches built from meshes: the systems are the batch's own (host
	// topology plus device arrays inside the arenas below), assembled by every
	// solve launch
	bool planar = false;
	bool assembled = false;
	double* d_f64 = nullptr;
	int32_t* d_i32 = nullptr;
	int64_t* d_i64 = nullptr;
	uint8_t* d_u8 = nullptr;
	bfmx::CondRec* d_cond = nullptr;
	// multigrid mode: the batch as one block-diagonal system (built at the
	// first such solve; shares the batch's stream)
	bfmx_system* uni = nullptr;
	int64_t* d_uoff = nullptr;     // first union row of each instance (m + 1)
	double* d_ubb = nullptr;       // b . b of each instance
};

namespace bfmx {

namespace {

// Frees what a solo solve of a batch-owned system allocated; its arrays in
// the batch's arenas stay.
void release_borrowed(bfmx_system* s) {
	if (s->stream) {
		(void) hipStreamSynchronize(s->stream);
	}
	release_solver(s);
	if (s->stream) {
		(void) hipStreamDestroy(s->stream);
		s->stream = nullptr;
	}
}

}  // namespace

int batch_n_configs() {
	return kConfigs;
}

int batch_n_classes() {
	return kClasses;
}

// The union system of the multigrid mode: what union_build allocated.
void union_release(bfmx_batch* b) {
	bfmx_system* U = b->uni;
	if (!U) {
		return;
	}
	release_solver(U);
	dfree(U->d_coords);
	dfree(U->d_slice_ptr);
	dfree(U->d_row_len);
	dfree(U->d_diag_blk);
	dfree(U->d_col);
	dfree(U->d_val);
	dfree(U->d_b);
	dfree(U->d_bsolve);
	dfree(U->d_fixed);
	dfree(U->d_conv_part);
	U->stream = nullptr;   // the batch's
	delete U;
	b->uni = nullptr;
	dfree(b->d_uoff);
	dfree(b->d_ubb);
}

void batch_destroy(bfmx_batch* b) {
	if (!b) {
		return;
	}
	(void) hipSetDevice(b->device);
	if (b->stream) {
		(void) hipStreamSynchronize(b->stream);
	}
	union_release(b);
	for (auto& g : b->graphs) {
		(void) hipGraphExecDestroy(g.exec);
	}
	dfree(b->d_inst);
	dfree(b->d_out);
	dfree(b->d_order);
	dfree(b->d_order_all);
	dfree(b->d_work);
	if (b->planar) {
		for (auto* s : b->systems) {
			if (s) {
				release_borrowed(s);
				delete s;
			}
		}
		dfree(b->d_f64);
		dfree(b->d_i32);
		dfree(b->d_i64);
		dfree(b->d_u8);
		dfree(b->d_cond);
	}
	for (auto& e : b->ev) {
		if (e) {
			(void) hipEventDestroy(e);
		}
	}
	if (b->stream) {
		(void) hipStreamDestroy(b->stream);
	}
	delete b;
}

namespace {

// Orders the systems by class, largest first, and uploads the instance
// records.
int finish_create(bfmx_batch* b, const std::vector<BatchInst>& h) {
	size_t const n = h.size();
	// per class, largest first (ties by index)
	std::vector<int32_t> order(n);
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
		int const ka = size_class(h[a].n), kc = size_class(h[c].n);
		return ka != kc ? ka < kc : h[a].n > h[c].n;
	});
	b->class_off.assign(kClasses + 1, 0);
	for (size_t k = 0; k < n; k++) {
		b->class_off[size_class(h[k].n) + 1]++;
	}
	for (int c = 0; c < kClasses; c++) {
		b->class_off[c + 1] += b->class_off[c];
	}
	std::vector<int32_t> order_all(n);
	std::iota(order_all.begin(), order_all.end(), 0);
	std::stable_sort(order_all.begin(), order_all.end(), [&](int32_t a, int32_t c) { return h[a].n > h[c].n; });
	BFMX_CHECK(dalloc(&b->d_inst, n));
	BFMX_CHECK(dalloc(&b->d_out, n));
	BFMX_CHECK(dalloc(&b->d_order, n));
	BFMX_CHECK(dalloc(&b->d_order_all, n));
	BFMX_HIP(hipMemcpyAsync(b->d_inst, h.data(), n * sizeof(BatchInst), hipMemcpyHostToDevice, b->stream));
	BFMX_HIP(hipMemcpyAsync(b->d_order, order.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
	BFMX_HIP(hipMemcpyAsync(b->d_order_all, order_all.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
	                        b->stream));
	BFMX_HIP(hipStreamSynchronize(b->stream));
	b->h_out.resize(n);
	return BFMX_OK;
}

}  // namespace

int batch_create(bfmx_batch** out, size_t n, bfmx_system* const* systems) {
	auto* b = new bfmx_batch();
	int const rc = [&]() -> int {
		BFMX_HIP(hipGetDevice(&b->device));
		BFMX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
		for (auto& e : b->ev) {
			BFMX_HIP(hipEventCreate(&e));
		}
		b->systems.assign(systems, systems + n);
		size_t words = 0;
		for (auto* s : b->systems) {
			if (!s || !s->assembled) {
				set_error("batch of unassembled systems");
				return BFMX_E_ARG;
			}
			if (s->device != b->device) {
				set_error("batch systems must live on the current device");
				return BFMX_E_ARG;
			}
			BFMX_CHECK(finalize_conditions(s));
			words += 16 * static_cast<size_t>(s->n_nodes);
		}
		BFMX_CHECK(dalloc(&b->d_work, words));
		std::vector<BatchInst> h(n);
		double* w = b->d_work;
		for (size_t k = 0; k < n; k++) {
			bfmx_system* s = b->systems[k];
			int64_t const nn = s->n_nodes;
			BatchInst& I = h[k];
			I.slice_ptr = s->d_slice_ptr;
			I.row_len = s->d_row_len;
			I.col = s->d_col;
			I.val = s->d_val;
			I.diag_blk = s->d_diag_blk;
			I.b = s->d_b;
			I.x = w;
			I.r = w + 2 * nn;
			I.z = w + 4 * nn;
			I.p0 = w + 6 * nn;
			I.p1 = w + 8 * nn;
			I.q = w + 10 * nn;
			I.dinv = w + 12 * nn;
			I.n = nn;
			I.fixed = s->d_fixed;   // read by the multigrid mode
			w += 16 * nn;
		}
		for (auto* s : b->systems) {
			BFMX_HIP(hipStreamSynchronize(s->stream));   // assembly done
		}
		BFMX_CHECK(finish_create(b, h));
		return BFMX_OK;
	}();
	if (rc != BFMX_OK) {
		batch_destroy(b);
		return rc;
	}
	*out = b;
	return BFMX_OK;
}

namespace {

// Assembly of a batch built from meshes (largest systems first).
void launch_assemble(bfmx_batch* b) {
	if (b->planar) {
		for (auto* s : b->systems) {
			s->val_gen++;
		}
		k_batch_assemble<<<static_cast<int>(b->systems.size()), kAsmBlock, 0, b->stream>>>(b->d_inst,
		                                                                                   b->d_order_all);
	}
}

// After a completed launch: the systems of a batch built from meshes hold
// their matrices and loads.
void mark_assembled(bfmx_batch* b) {
	if (b->planar && !b->assembled) {
		for (auto* s : b->systems) {
			s->assembled = true;
			s->finalized = true;
		}
		b->assembled = true;
	}
}

void launch_all(bfmx_batch* b, int pre, double rtol2, int64_t max_iter, const int* picks) {
	launch_assemble(b);
	if (std::all_of(picks, picks + kClasses, [&](int p) { return p == picks[0]; })) {
		// one configuration for all: one launch, largest systems first, so no
		// class waits for the previous one to drain
		launch(pre, picks[0], static_cast<int>(b->systems.size()), b->stream, b->d_inst, b->d_order_all, rtol2,
		       max_iter, b->d_out);
		return;
	}
	for (int c = kClasses - 1; c >= 0; c--) {   // largest systems first
		int const cnt = b->class_off[c + 1] - b->class_off[c];
		if (cnt > 0) {
			launch(pre, picks[c], cnt, b->stream, b->d_inst, b->d_order + b->class_off[c], rtol2, max_iter, b->d_out);
		}
	}
}

}  // namespace

namespace {

int blocks_of(int64_t n) {
	return static_cast<int>((n + kBlock - 1) / kBlock);
}

// The union system of the multigrid mode from the instances' host topology
// (renumbered adjacency and coordinates), its rows the instances' rows one
// after another. On failure nothing of it stays.
int union_build_parts(bfmx_batch* b);

int union_build(bfmx_batch* b) {
	if (b->uni) {
		return BFMX_OK;
	}
	int const rc = union_build_parts(b);
	if (rc != BFMX_OK) {
		(void) hipStreamSynchronize(b->stream);
		union_release(b);
	}
	return rc;
}

int union_build_parts(bfmx_batch* b) {
	size_t const m = b->systems.size();
	std::vector<int64_t> off(m + 1, 0);
	int64_t nnz = 0;
	for (size_t k = 0; k < m; k++) {
		bfmx_system const* s = b->systems[k];
		if (s->stress == BFMX_AXISYMMETRIC) {
			set_error("the multigrid mode needs SPD systems (axisymmetric instances are indefinite)");
			return BFMX_E_UNSUPPORTED;
		}
		off[k + 1] = off[k] + s->n_nodes;
		nnz += s->adj_ptr[s->n_nodes];
	}
	int64_t const n = off[m];
	if (n >= (int64_t(1) << 30)) {
		set_error("batch too large for the multigrid mode (%lld nodes)", (long long) n);
		return BFMX_E_ARG;
	}
	auto* U = new bfmx_system();
	b->uni = U;   // released by union_release, also on failure
	U->device = b->device;
	U->stream = b->stream;
	U->kind = b->systems[0]->kind;
	U->n_nodes = n;
	hipStream_t const st = b->stream;
	// the coordinates on the device (the multigrid's rigid-body modes are
	// formed there): a batch built from meshes keeps its instances' on the
	// device only
	BFMX_CHECK(dalloc(&U->d_coords, 2 * n));
	for (size_t k = 0; k < m; k++) {
		bfmx_system const* s = b->systems[k];
		size_t const bytes = 2 * static_cast<size_t>(s->n_nodes) * sizeof(double);
		if (s->d_coords) {
			BFMX_HIP(hipMemcpyAsync(U->d_coords + 2 * off[k], s->d_coords, bytes, hipMemcpyDeviceToDevice, st));
		} else {
			BFMX_HIP(hipMemcpyAsync(U->d_coords + 2 * off[k], s->coords.data(), bytes, hipMemcpyHostToDevice, st));
		}
	}
	// the instances' graphs one after another (instances spread over threads)
	std::vector<int64_t> eoff(m + 1, 0);
	for (size_t k = 0; k < m; k++) {
		eoff[k + 1] = eoff[k] + b->systems[k]->adj_ptr[b->systems[k]->n_nodes];
	}
	U->adj_ptr.assign(n + 1, 0);
	U->adj.resize(nnz);
	{
		unsigned const hw = std::max(1u, std::thread::hardware_concurrency());
		size_t const nthreads = std::min<size_t>({m, 16, hw});
		std::atomic<size_t> next{0};
		auto work = [&]() {
			for (size_t k; (k = next.fetch_add(1)) < m;) {
				bfmx_system const* s = b->systems[k];
				int64_t const o = off[k], e0 = eoff[k];
				for (int64_t i = 0; i < s->n_nodes; i++) {
					U->adj_ptr[o + i + 1] = e0 + s->adj_ptr[i + 1];
				}
				for (int64_t e = 0; e < s->adj_ptr[s->n_nodes]; e++) {
					U->adj[e0 + e] = static_cast<int32_t>(s->adj[e] + o);
				}
			}
		};
		std::vector<std::thread> pool;
		for (size_t t = 1; t < nthreads; t++) {
			pool.emplace_back(work);
		}
		work();
		for (auto& t : pool) {
			t.join();
		}
	}
	std::vector<int32_t> col;
	BFMX_CHECK(sell_layout(U, col));
	BFMX_CHECK(dalloc(&U->d_slice_ptr, U->n_slices + 1));
	BFMX_CHECK(dalloc(&U->d_row_len, n));
	BFMX_CHECK(dalloc(&U->d_diag_blk, n));
	BFMX_CHECK(dalloc(&U->d_col, U->nnzb_pad));
	BFMX_CHECK(dalloc(&U->d_val, 4 * U->nnzb_pad));
	BFMX_CHECK(dalloc(&U->d_b, 2 * n));
	BFMX_CHECK(dalloc(&U->d_bsolve, 2 * n));
	BFMX_CHECK(dalloc(&U->d_fixed, 2 * n));
	BFMX_CHECK(dalloc(&b->d_uoff, m + 1));
	BFMX_CHECK(dalloc(&b->d_ubb, m));
	BFMX_CHECK(dalloc(&U->d_conv_part, 2 * m));
	U->n_conv_seg = static_cast<int>(m);   // every instance meets the tolerance on its own
	U->d_conv_seg = b->d_uoff;
	BFMX_HIP(hipMemcpyAsync(U->d_slice_ptr, U->slice_ptr.data(), U->slice_ptr.size() * sizeof(int32_t),
	                        hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(U->d_row_len, U->row_len.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(U->d_diag_blk, U->diag_blk.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(U->d_col, col.data(), col.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(b->d_uoff, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
	// padding slots keep zero values; the gather writes every stored block
	BFMX_HIP(hipMemsetAsync(U->d_val, 0, 4 * static_cast<size_t>(U->nnzb_pad) * sizeof(double), st));
	BFMX_HIP(hipStreamSynchronize(st));
	U->assembled = true;
	U->finalized = true;
	return BFMX_OK;
}

// Multigrid mode of bfmx_batch_solve: assembly (batches built from meshes),
// load norms, gather into the union, flexible CG + V-cycle on the union,
// scatter of the solutions, per-instance true residuals.
int batch_solve_amg(bfmx_batch* b, double rtol, int64_t max_iter, bfmx_stats_t* stats, double* total_ms) {
	BFMX_CHECK(union_build(b));
	bfmx_system* U = b->uni;
	hipStream_t const st = b->stream;
	int const m = static_cast<int>(b->systems.size());
	int64_t const n = U->n_nodes;
	BFMX_HIP(hipEventRecord(b->ev[0], st));
	launch_assemble(b);
	k_batch_bnorm<<<m, kUniBlock, 0, st>>>(b->d_inst, b->d_ubb);
	U->val_gen++;
	k_union_gather<<<blocks_of(n), kBlock, 0, st>>>(n, m, b->d_uoff, b->d_inst, b->d_ubb, U->d_slice_ptr, U->d_val,
	                                                U->d_b, U->d_fixed);
	BFMX_HIP(hipGetLastError());
	std::vector<double> bb(m);
	BFMX_HIP(hipMemcpyAsync(bb.data(), b->d_ubb, m * sizeof(double), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipStreamSynchronize(st));
	int const nz = static_cast<int>(std::count_if(bb.begin(), bb.end(), [](double v) { return v > 0; }));
	bfmx_stats_t us{};
	int rc = BFMX_OK;
	if (nz > 0) {
		// the union's residual is checked instance by instance (its convergence
		// segments); with unit loads the checks start once ||r||^2 <= nz rtol^2
		rc = solve(U, nullptr, BFMX_PRECOND_AMG, BFMX_PRECISION_FP64, rtol, static_cast<size_t>(max_iter), &us);
		if (rc != BFMX_OK && rc != BFMX_E_BREAKDOWN && rc != BFMX_E_STAGNATION && rc != BFMX_E_MAXITER) {
			return rc;
		}
	} else {
		BFMX_CHECK(ensure_workspace_public(U));
		BFMX_HIP(hipMemsetAsync(U->d_x, 0, 2 * n * sizeof(double), st));
	}
	k_union_scatter<<<blocks_of(n), kBlock, 0, st>>>(n, m, b->d_uoff, b->d_inst, b->d_ubb, U->d_x);
	k_batch_resid<<<m, kUniBlock, 0, st>>>(b->d_inst, b->d_ubb, us.iterations, rtol * rtol, b->d_out);
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipEventRecord(b->ev[1], st));
	BFMX_HIP(hipMemcpyAsync(b->h_out.data(), b->d_out, b->h_out.size() * sizeof(BatchOut), hipMemcpyDeviceToHost,
	                        st));
	BFMX_HIP(hipStreamSynchronize(st));
	mark_assembled(b);
	float ms = 0;
	BFMX_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
	if (total_ms) {
		*total_ms = ms;
	}
	int worst = BFMX_OK;
	for (int k = 0; k < m; k++) {
		BatchOut const& o = b->h_out[k];
		int const sk = o.status == 1 ? BFMX_OK : rc != BFMX_OK ? rc : BFMX_E_STAGNATION;
		if (sk != BFMX_OK) {
			worst = sk;
		}
		if (stats) {
			bfmx_stats_t& S = stats[k];
			S = bfmx_stats_t{};
			S.status = sk;
			S.precond = BFMX_PRECOND_AMG;
			S.precision = BFMX_PRECISION_FP64;
			S.restarts = us.restarts;
			S.iterations = o.iter;
			S.n_dof = 2 * b->systems[k]->n_nodes;
			S.rtol = rtol;
			S.rel_residual = o.bb > 0 ? std::sqrt(o.rr_true / o.bb) : 0.0;
			S.solve_ms = ms;
			S.b_norm = std::sqrt(o.bb);
		}
	}
	if (worst != BFMX_OK) {
		set_error("batched multigrid solve: some systems did not reach rtol (per-system status in the stats)");
	}
	return worst;
}

}  // namespace

int batch_solve(bfmx_batch* b, int pre, double rtol, int64_t max_iter, const int32_t* picks_in, bfmx_stats_t* stats,
                double* total_ms) {
	if (pre == BFMX_PRECOND_AMG) {
		return batch_solve_amg(b, rtol, max_iter, stats, total_ms);
	}
	if (pre != BFMX_PRECOND_JACOBI && pre != BFMX_PRECOND_BLOCK_JACOBI) {
		set_error("batched solves use Jacobi or block-Jacobi PCG, or the multigrid mode");
		return BFMX_E_UNSUPPORTED;
	}
	int picks[kClasses];
	for (int c = 0; c < kClasses; c++) {
		picks[c] = picks_in ? picks_in[c] : kDefaultPick[c];
		if (picks[c] < 0 || picks[c] >= kConfigs) {
			set_error("configuration %d out of range", picks[c]);
			return BFMX_E_ARG;
		}
	}
	if (max_iter == 0) {
		for (auto* s : b->systems) {
			max_iter = std::max<int64_t>(max_iter, 20 * 2 * s->n_nodes);
		}
	}
	double const rtol2 = rtol * rtol;
	hipGraphExec_t exec = nullptr;
	bool const use_graph = graphs_enabled();
	for (auto const& g : b->graphs) {
		if (!use_graph) {
			break;
		}
		if (g.pre == pre && g.rtol2 == rtol2 && g.max_iter == max_iter && std::equal(picks, picks + kClasses, g.picks)) {
			exec = g.exec;
		}
	}
	if (!exec && use_graph) {
		hipGraph_t g = nullptr;
		BFMX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
		launch_all(b, pre, rtol2, max_iter, picks);
		BFMX_HIP(hipStreamEndCapture(b->stream, &g));
		hipError_t const e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
		(void) hipGraphDestroy(g);
		if (e != hipSuccess) {
			set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
			return BFMX_E_HIP;
		}
		bfmx_batch::Exec ex{exec, pre, rtol2, max_iter, {}};
		std::copy(picks, picks + kClasses, ex.picks);
		b->graphs.push_back(ex);
	}
	BFMX_HIP(hipEventRecord(b->ev[0], b->stream));
	if (use_graph) {
		BFMX_HIP(hipGraphLaunch(exec, b->stream));
	} else {
		launch_all(b, pre, rtol2, max_iter, picks);
		BFMX_HIP(hipGetLastError());
	}
	BFMX_HIP(hipEventRecord(b->ev[1], b->stream));
	BFMX_HIP(hipMemcpyAsync(b->h_out.data(), b->d_out, b->h_out.size() * sizeof(BatchOut), hipMemcpyDeviceToHost,
	                        b->stream));
	BFMX_HIP(hipStreamSynchronize(b->stream));
	mark_assembled(b);
	float ms = 0;
	BFMX_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
	if (total_ms) {
		*total_ms = ms;
	}
	int worst = BFMX_OK;
	for (size_t k = 0; k < b->systems.size(); k++) {
		BatchOut const& o = b->h_out[k];
		int const st = o.status == 1 ? BFMX_OK
		               : o.status == 2 ? BFMX_E_BREAKDOWN
		               : o.status == 3 ? BFMX_E_STAGNATION
		                               : BFMX_E_MAXITER;
		if (st != BFMX_OK && (worst == BFMX_OK || worst == BFMX_E_STAGNATION)) {
			worst = st;
		}
		if (stats) {
			bfmx_stats_t& S = stats[k];
			S = bfmx_stats_t{};
			S.status = st;
			S.precond = pre;
			S.precision = BFMX_PRECISION_FP64;
			S.restarts = o.checks;
			S.iterations = o.iter;
			S.n_dof = 2 * b->systems[k]->n_nodes;
			S.rtol = rtol;
			S.rel_residual = o.bb > 0 ? std::sqrt(o.rr_true / o.bb) : 0.0;
			S.solve_ms = ms;
			S.b_norm = std::sqrt(o.bb);
		}
	}
	if (worst != BFMX_OK) {
		set_error("batched solve: some systems did not converge (per-system status in the stats)");
	}
	return worst;
}

int batch_autotune(bfmx_batch* b, int pre, int sample, int iters, int32_t* picks, double* times_ms) {
	if (pre != BFMX_PRECOND_JACOBI && pre != BFMX_PRECOND_BLOCK_JACOBI) {
		set_error("batched solves use Jacobi or block-Jacobi PCG");
		return BFMX_E_UNSUPPORTED;
	}
	launch_assemble(b);
	BFMX_HIP(hipGetLastError());
	for (int c = 0; c < kClasses; c++) {
		int const all = b->class_off[c + 1] - b->class_off[c];
		int const cnt = sample > 0 ? std::min(sample, all) : all;
		int best = kDefaultPick[c];
		float best_ms = 0;
		for (int cfg = 0; cfg < kConfigs; cfg++) {
			float ms = 0;
			if (cnt > 0) {
				// tiny rtol: every sampled system runs exactly `iters` iterations
				launch(pre, cfg, cnt, b->stream, b->d_inst, b->d_order + b->class_off[c], 1e-300, iters, b->d_out);
				BFMX_HIP(hipEventRecord(b->ev[0], b->stream));
				launch(pre, cfg, cnt, b->stream, b->d_inst, b->d_order + b->class_off[c], 1e-300, iters, b->d_out);
				BFMX_HIP(hipEventRecord(b->ev[1], b->stream));
				BFMX_HIP(hipEventSynchronize(b->ev[1]));
				BFMX_HIP(hipGetLastError());
				BFMX_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
				if (cfg == 0 || ms < best_ms) {
					best_ms = ms;
					best = cfg;
				}
			}
			if (times_ms) {
				times_ms[c * kConfigs + cfg] = ms;
			}
		}
		picks[c] = best;
	}
	BFMX_HIP(hipStreamSynchronize(b->stream));
	mark_assembled(b);
	return BFMX_OK;
}

namespace {

// Host preparation of one instance of a batch built from meshes, as
// bfmx_system_create + set_edges + apply_condition + finalize_conditions do
// it, without device work.
struct Prepared {
	bfmx_system* s = nullptr;
	std::vector<int32_t> col, el_col, lists, elim, color_ptr;
	std::vector<uint8_t> nflags;
	std::vector<CondRec> conds;
	std::vector<double> elemE;   // colour order
	int rc = BFMX_OK;
	std::string err;
};

int prepare_instance(const bfmx_instance_t& I, int renumber, int device, Prepared& P) {
	if (I.kind != 3 && I.kind != 4) {
		set_error("element kind %d not supported (3 = P1 triangles, 4 = Q1 quads)", I.kind);
		return BFMX_E_UNSUPPORTED;
	}
	if (!I.coords || (I.n_elems && !I.elems) || (I.n_edges && !I.edge_nodes) ||
	    (I.n_conditions && (!I.cond_kind || !I.cond_value || !I.cond_flags))) {
		set_error("null array in an instance");
		return BFMX_E_ARG;
	}
	auto* s = new bfmx_system();
	P.s = s;
	s->device = device;
	s->borrowed = true;
	s->kind = I.kind;
	s->n_nodes = static_cast<int64_t>(I.n_nodes);
	s->n_elems = static_cast<int64_t>(I.n_elems);
	BFMX_CHECK(prepare_topology(s, I.coords, I.elems, renumber, P.col, P.el_col));
	BFMX_CHECK(prepare_edges(s, I.n_edges, I.edge_nodes, I.edge_boundary));
	int64_t const n = s->n_nodes;
	P.color_ptr.assign(s->color_ptr.begin(), s->color_ptr.end());
	if (I.elem_E) {
		P.elemE.resize(s->n_elems);
		for (int64_t p = 0; p < s->n_elems; p++) {
			P.elemE[p] = I.elem_E[s->elem_order[p]];
		}
	}
	if (I.stress < 0 || I.stress > 2) {
		set_error("formulation %d unknown (0 plane strain, 1 plane stress, 2 axisymmetric)", I.stress);
		return BFMX_E_ARG;
	}
	s->stress = I.stress;
	s->dirichlet_nodes.assign(n, 0);
	std::vector<uint8_t> fl(n);
	for (size_t q = 0; q < I.n_conditions; q++) {
		int const kind = I.cond_kind[q];
		if (kind < 0 || kind > 7) {
			set_error("unknown condition kind %d", kind);
			return BFMX_E_ARG;
		}
		bool const neumann = kind >= BFMX_NEUMANN_X && kind <= BFMX_NEUMANN_TANGENT;
		if (s->stress == 2 && (kind == BFMX_NEUMANN_NORMAL || kind == BFMX_NEUMANN_TANGENT)) {
			continue;   // the axisymmetric condition loop has no such branch
		}
		const uint8_t* flags = I.cond_flags + q * I.n_nodes;
		std::fill(fl.begin(), fl.end(), 0);
		for (int64_t i = 0; i < n; i++) {
			if (flags[i]) {
				fl[s->perm[i]] = 1;
			}
		}
		CondRec C{};
		C.kind = kind;
		C.value = I.cond_value[q];
		C.list_off = static_cast<int64_t>(P.lists.size());
		for (int64_t i = 0; i < n; i++) {
			if (fl[i]) {
				P.lists.push_back(static_cast<int32_t>(i));
				if (!neumann) {
					s->dirichlet_nodes[i] = 1;
				}
			}
		}
		C.count = static_cast<int32_t>(P.lists.size() - C.list_off);
		if (neumann) {
			C.flags_off = static_cast<int64_t>(P.nflags.size());
			P.nflags.insert(P.nflags.end(), fl.begin(), fl.end());
		}
		if (C.count > 0) {
			P.conds.push_back(C);
		}
	}
	std::vector<uint8_t> touched(n, 0);
	for (int64_t i = 0; i < n; i++) {
		if (s->dirichlet_nodes[i]) {
			for (int64_t t = s->adj_ptr[i]; t < s->adj_ptr[i + 1]; t++) {
				touched[s->adj[t]] = 1;
			}
		}
	}
	for (int64_t i = 0; i < n; i++) {
		if (touched[i]) {
			P.elim.push_back(static_cast<int32_t>(i));
		}
	}
	return BFMX_OK;
}

// Sub-array starts in the arenas are multiples of 32 elements (at least
// 128 bytes), so vector loads stay aligned.
int64_t take(int64_t& top, int64_t count) {
	int64_t const at = top;
	top += (count + 31) / 32 * 32;
	return at;
}

}  // namespace

int batch_create_planar(bfmx_batch** out, size_t n, const bfmx_instance_t* inst, int renumber) {
	auto* b = new bfmx_batch();
	b->planar = true;
	std::vector<Prepared> prep(n);
	int const rc = [&]() -> int {
		BFMX_HIP(hipGetDevice(&b->device));
		BFMX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
		for (auto& e : b->ev) {
			BFMX_HIP(hipEventCreate(&e));
		}
		// host preparation, instances spread over threads
		unsigned const hw = std::max(1u, std::thread::hardware_concurrency());
		size_t const nthreads = std::min<size_t>({n, 16, hw});
		std::atomic<size_t> next{0};
		auto work = [&]() {
			PrepThreads const one(1);   // the instances are the parallelism here
			for (size_t k; (k = next.fetch_add(1)) < n;) {
				Prepared& P = prep[k];
				try {
					P.rc = prepare_instance(inst[k], renumber, b->device, P);
					if (P.rc != BFMX_OK) {
						P.err = last_error_text();
					}
				} catch (const std::bad_alloc&) {
					P.rc = BFMX_E_NOMEM;
					P.err = "host allocation failed";
				}
			}
		};
		std::vector<std::thread> pool;
		for (size_t t = 1; t < nthreads; t++) {
			pool.emplace_back(work);
		}
		work();
		for (auto& t : pool) {
			t.join();
		}
		b->systems.resize(n);
		for (size_t k = 0; k < n; k++) {
			b->systems[k] = prep[k].s;
		}
		for (size_t k = 0; k < n; k++) {
			if (prep[k].rc != BFMX_OK) {
				set_error("instance %zu: %s", k, prep[k].err.c_str());
				return prep[k].rc;
			}
		}
		// arena layout
		struct Off {
			int64_t coords, elemE, val, b, bsolve, vfirst, work;                            // f64
			int64_t slice_ptr, row_len, diag_blk, col, elem, color_ptr, lists, elim, edges, inc;   // i32
			int64_t inc_ptr;                                                                // i64
			int64_t fixed, nflags, edge_bnd;                                               // u8
			int64_t cond;
		};
		std::vector<Off> off(n);
		int64_t f64 = 0, i32 = 0, i64 = 0, u8 = 0, nc = 0;
		for (size_t k = 0; k < n; k++) {
			bfmx_system* s = prep[k].s;
			int64_t const nn = s->n_nodes;
			Off& o = off[k];
			o.coords = take(f64, 2 * nn);
			o.elemE = prep[k].elemE.empty() ? -1 : take(f64, s->n_elems);
			o.val = take(f64, 4 * s->nnzb_pad);
			o.b = take(f64, 2 * nn);
			o.bsolve = take(f64, 2 * nn);
			o.vfirst = take(f64, 2 * nn);
			o.work = take(f64, 16 * nn);
			o.slice_ptr = take(i32, s->n_slices + 1);
			o.row_len = take(i32, nn);
			o.diag_blk = take(i32, nn);
			o.col = take(i32, s->nnzb_pad);
			o.elem = take(i32, s->n_elems * s->kind);
			o.color_ptr = take(i32, s->n_colors + 1);
			o.lists = take(i32, prep[k].lists.size());
			o.elim = take(i32, prep[k].elim.size());
			o.edges = take(i32, 2 * s->n_edges);
			o.inc = take(i32, s->inc.size());
			o.inc_ptr = take(i64, nn + 1);
			o.fixed = take(u8, 2 * nn);
			o.nflags = take(u8, prep[k].nflags.size());
			o.edge_bnd = take(u8, s->n_edges);
			o.cond = nc;
			nc += static_cast<int64_t>(prep[k].conds.size());
		}
		BFMX_CHECK(dalloc(&b->d_f64, f64));
		BFMX_CHECK(dalloc(&b->d_i32, i32));
		BFMX_CHECK(dalloc(&b->d_i64, i64));
		BFMX_CHECK(dalloc(&b->d_u8, u8));
		BFMX_CHECK(dalloc(&b->d_cond, std::max<int64_t>(nc, 1)));
		// staged uploads of everything the device reads (the rest is written
		// by the solve launches)
		std::vector<int32_t> h32(i32, 0);
		std::vector<int64_t> h64(i64, 0);
		std::vector<uint8_t> h8(u8, 0);
		std::vector<CondRec> hc(std::max<int64_t>(nc, 1));
		std::vector<std::pair<int64_t, const std::vector<double>*>> f64_parts;
		for (size_t k = 0; k < n; k++) {
			Prepared& P = prep[k];
			bfmx_system* s = P.s;
			Off const& o = off[k];
			auto put32 = [&](int64_t at, const std::vector<int32_t>& v) {
				std::copy(v.begin(), v.end(), h32.begin() + at);
			};
			put32(o.slice_ptr, s->slice_ptr);
			put32(o.row_len, s->row_len);
			put32(o.diag_blk, s->diag_blk);
			put32(o.col, P.col);
			put32(o.elem, P.el_col);
			put32(o.color_ptr, P.color_ptr);
			put32(o.lists, P.lists);
			put32(o.elim, P.elim);
			put32(o.edges, s->edge_nodes);
			put32(o.inc, s->inc);
			std::copy(s->inc_ptr.begin(), s->inc_ptr.end(), h64.begin() + o.inc_ptr);
			std::copy(P.nflags.begin(), P.nflags.end(), h8.begin() + o.nflags);
			std::copy(s->edge_bnd.begin(), s->edge_bnd.end(), h8.begin() + o.edge_bnd);
			std::copy(P.conds.begin(), P.conds.end(), hc.begin() + o.cond);
		}
		BFMX_HIP(hipMemcpy(b->d_i32, h32.data(), h32.size() * sizeof(int32_t), hipMemcpyHostToDevice));
		BFMX_HIP(hipMemcpy(b->d_i64, h64.data(), h64.size() * sizeof(int64_t), hipMemcpyHostToDevice));
		BFMX_HIP(hipMemcpy(b->d_u8, h8.data(), h8.size(), hipMemcpyHostToDevice));
		BFMX_HIP(hipMemcpy(b->d_cond, hc.data(), hc.size() * sizeof(CondRec), hipMemcpyHostToDevice));
		for (size_t k = 0; k < n; k++) {
			bfmx_system* s = prep[k].s;
			BFMX_HIP(hipMemcpy(b->d_f64 + off[k].coords, s->coords.data(), s->coords.size() * sizeof(double),
			                   hipMemcpyHostToDevice));
			if (off[k].elemE >= 0) {
				BFMX_HIP(hipMemcpy(b->d_f64 + off[k].elemE, prep[k].elemE.data(), prep[k].elemE.size() * sizeof(double),
				                   hipMemcpyHostToDevice));
			}
		}
		// the systems' device arrays and the kernels' instance records
		std::vector<BatchInst> h(n);
		for (size_t k = 0; k < n; k++) {
			Prepared& P = prep[k];
			bfmx_system* s = P.s;
			Off const& o = off[k];
			int64_t const nn = s->n_nodes;
			s->d_coords = b->d_f64 + o.coords;
			s->d_elemE = o.elemE >= 0 ? b->d_f64 + o.elemE : nullptr;
			s->d_val = b->d_f64 + o.val;
			s->d_b = b->d_f64 + o.b;
			s->d_bsolve = b->d_f64 + o.bsolve;
			s->d_vfirst = b->d_f64 + o.vfirst;
			s->d_slice_ptr = b->d_i32 + o.slice_ptr;
			s->d_row_len = b->d_i32 + o.row_len;
			s->d_diag_blk = b->d_i32 + o.diag_blk;
			s->d_col = b->d_i32 + o.col;
			s->d_elem = b->d_i32 + o.elem;
			s->d_edge_nodes = b->d_i32 + o.edges;
			s->d_inc = b->d_i32 + o.inc;
			s->d_inc_ptr = b->d_i64 + o.inc_ptr;
			s->d_fixed = b->d_u8 + o.fixed;
			s->d_edge_bnd = b->d_u8 + o.edge_bnd;
			BatchInst& I = h[k];
			I = BatchInst{};
			double* w = b->d_f64 + o.work;
			I.slice_ptr = s->d_slice_ptr;
			I.row_len = s->d_row_len;
			I.col = s->d_col;
			I.val = s->d_val;
			I.diag_blk = s->d_diag_blk;
			I.b = s->d_b;
			I.x = w;
			I.r = w + 2 * nn;
			I.z = w + 4 * nn;
			I.p0 = w + 6 * nn;
			I.p1 = w + 8 * nn;
			I.q = w + 10 * nn;
			I.dinv = w + 12 * nn;
			I.n = nn;
			I.nnzb_pad = s->nnzb_pad;
			I.coords = s->d_coords;
			I.elem = s->d_elem;
			I.color_ptr = b->d_i32 + o.color_ptr;
			I.n_colors = s->n_colors;
			I.kind = s->kind;
			I.M = dev::Material{s->d_elemE, inst[k].E, inst[k].nu, inst[k].rho, s->stress, nullptr, inst[k].gx,
			                    inst[k].gy};
			I.conds = b->d_cond + o.cond;
			I.n_cond = static_cast<int32_t>(P.conds.size());
			I.n_elim = static_cast<int32_t>(P.elim.size());
			I.lists = b->d_i32 + o.lists;
			I.nflags = b->d_u8 + o.nflags;
			I.elim = b->d_i32 + o.elim;
			I.inc_ptr = s->d_inc_ptr;
			I.inc = s->d_inc;
			I.edge_nodes = s->d_edge_nodes;
			I.edge_bnd = s->d_edge_bnd;
			I.fixed = s->d_fixed;
			I.vfirst = s->d_vfirst;
		}
		// host arrays only the device needed
		for (auto& P : prep) {
			bfmx_system* s = P.s;
			std::vector<double>().swap(s->coords);
			std::vector<int32_t>().swap(s->elem_order);
			std::vector<int32_t>().swap(s->inc);
			std::vector<int64_t>().swap(s->inc_ptr);
			std::vector<int32_t>().swap(s->edge_nodes);
			std::vector<uint8_t>().swap(s->edge_bnd);
		}
		BFMX_CHECK(finish_create(b, h));
		return BFMX_OK;
	}();
	if (rc != BFMX_OK) {
		batch_destroy(b);   // also the prepared systems (b->systems)
		return rc;
	}
	*out = b;
	return BFMX_OK;
}

int batch_system(bfmx_batch* b, size_t k, bfmx_system** out) {
	if (k >= b->systems.size()) {
		set_error("no system %zu in the batch", k);
		return BFMX_E_ARG;
	}
	bfmx_system* s = b->systems[k];
	if (!s->assembled) {
		set_error("the batch has not been solved yet: its systems are assembled by the solve launch");
		return BFMX_E_ARG;
	}
	BFMX_HIP(hipSetDevice(s->device));
	if (b->planar && !s->stream) {
		BFMX_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
	}
	*out = s;
	return BFMX_OK;
}

int batch_get_u(bfmx_batch* b, size_t k, double* u) {
	if (k >= b->systems.size()) {
		set_error("no system %zu in the batch", k);
		return BFMX_E_ARG;
	}
	bfmx_system* s = b->systems[k];
	int64_t const n = s->n_nodes;
	std::vector<BatchInst> h(1);
	BFMX_HIP(hipMemcpy(h.data(), b->d_inst + k, sizeof(BatchInst), hipMemcpyDeviceToHost));
	std::vector<double> v(2 * n);
	BFMX_HIP(hipMemcpy(v.data(), h[0].x, v.size() * sizeof(double), hipMemcpyDeviceToHost));
	for (int64_t i = 0; i < n; i++) {
		u[2 * i] = v[2 * s->perm[i]];
		u[2 * i + 1] = v[2 * s->perm[i] + 1];
	}
	return BFMX_OK;
}

}  // namespace bfmx
