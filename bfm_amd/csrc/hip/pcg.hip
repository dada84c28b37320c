// Preconditioned conjugate gradients on the device.
//
// One CG iteration is four kernels:
//   k_cg_spmv    p_new = z + beta p_old and q = A p_new for its rows, gathering
//                z and p_old at the block columns (so the direction update
//                costs no separate pass), plus per-workgroup sums of p_new . q;
//   k_red_alpha  one workgroup: alpha = (r . z) / (p . q), or a breakdown;
//   k_cg_update  x += alpha p, r -= alpha q, z = M^-1 r and per-workgroup sums
//                of r . z and r . r;
//   k_red_beta   one workgroup: beta and the iteration count.
// The cross-workgroup sums are finished by the one-workgroup kernels, ordered
// by the kernel boundary: a "last workgroup reduces" scheme would need a
// device-scope fence per workgroup, which on a multi-XCD GPU writes the
// XCD's L2 back each time.
// A captured graph holds kChunk iterations followed by k_residual, which
// recomputes r = b - A x with the FP64 matrix when the recursive residual has
// met the tolerance (then the true residual decides convergence) or, on the
// mixed path, has dropped by 1/delta since the last recomputation (a reliable
// update; the directions are kept). Scalars live in device memory, every
// kernel returns at once after convergence or breakdown, and the host only
// reads the scalars between graph replays. All reductions are FP64 and in a
// fixed order (deterministic).
//
// Precision: FP64 uses the FP64 matrix. MIXED streams the FP32 copy of the
// matrix values (half the bytes; the 2M-DOF copy fits the MI355X's 256 MB
// Infinity Cache) with FP64 vectors, FP64 accumulation, FP64 dots and the
// FP64 true-residual recomputation above.
//
// Preconditioners: point Jacobi (1 / diag) or 2x2 block Jacobi.
//
// Multigrid (BFMX_PRECOND_AMG): flexible CG in FP64 (matrix, vectors, dots)
// preconditioned by one FP32 V-cycle of aggregation multigrid (amg.hip). The
// V-cycle is not an exactly symmetric operator in FP32, so the direction is
// orthogonalised against the previous one explicitly (Notay's flexible CG
// with one stored direction): beta = -(z . q_prev) / (p_prev . q_prev),
// alpha = (p . r) / (p . q). One iteration is the V-cycle (its last kernel
// also forms the partials of z . q_prev) and five kernels:
//   k_red_fcg_beta, k_fcg_spmv (p = z + beta p_old, q = A p, partials of p . q
//   and p . r), k_red_fcg_alpha, k_fcg_update (x, r and partials of r . r),
//   k_red_fcg_rr.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "internal.h"

namespace bfmx {

// Chunks are replayed from a captured HIP graph unless the environment sets
// BFMX_GRAPHS=0, which launches the same kernels straight on the stream (for
// tools that do not follow graph launches).
bool graphs_enabled() {
	const char* e = std::getenv("BFMX_GRAPHS");
	return !(e && e[0] == '0');
}

namespace {

typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt2 __attribute__((ext_vector_type(2)));

constexpr double kDelta2 = 1e-2;   // reliable update when ||r|| dropped 10x since the last one

__device__ __forceinline__ dbl2 ld2(const double* p) {
	return *reinterpret_cast<const dbl2*>(p);
}
__device__ __forceinline__ void st2(double* p, dbl2 v) {
	*reinterpret_cast<dbl2*>(p) = v;
}

// Two consecutive matrix values, streamed (non-temporal: the matrix is not
// re-read within a kernel, keeping L2 for the gathered vectors).
__device__ __forceinline__ dbl2 ldv(const double* p) {
	return __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(p));
}
__device__ __forceinline__ dbl2 ldv(const float* p) {
	flt2 const f = __builtin_nontemporal_load(reinterpret_cast<const flt2*>(p));
	return dbl2{f.x, f.y};
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v += __shfl_xor(v, off, 64);
	}
	return v;
}

// Sums NV values over the workgroup; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*lds)[kBlock / 64]) {
#pragma unroll
	for (int k = 0; k < NV; k++) {
		v[k] = wave_sum(v[k]);
	}
	int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			lds[k][w] = v[k];
		}
	}
	__syncthreads();
	if (threadIdx.x == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			double s = 0;
#pragma unroll
			for (int i = 0; i < kBlock / 64; i++) {
				s += lds[k][i];
			}
			v[k] = s;
		}
	}
}

// Writes this workgroup's NV sums to partial[k * gridDim.x + blockIdx.x].
// The totals are formed by a one-workgroup reduce kernel launched next: the
// kernel boundary orders the writes, so no device-scope fence or atomic is
// needed (on multi-XCD parts such a fence writes the XCD's L2 back).
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* __restrict__ partial) {
	__shared__ double lds[NV][kBlock / 64];
	block_sum<NV>(v, lds);
	if (threadIdx.x == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			partial[k * gridDim.x + blockIdx.x] = v[k];
		}
	}
}

constexpr int kRed = 1024;   // threads of the one-workgroup reduce kernels
constexpr int32_t kHold = 4;   // Scalars::done while a flexible-CG chunk waits for its tolerance check

// Sums partial[k * n + i], i < n, for k < NV in a fixed order; the totals are
// valid in thread 0.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ partial, int n, double (&tot)[NV],
                                                const double (&first)[NV]) {
	__shared__ double lds[NV][kRed / 64];
	// each thread sums its partials in index order; they are loaded kLoads at
	// a time, all in flight together (a rolled loop waits for each in turn).
	// Slots past the end add 0.0, which changes no sum.
	constexpr int kLoads = 4;
	double acc[NV];
#pragma unroll
	for (int k = 0; k < NV; k++) {
		acc[k] = first[k];
	}
	for (int base = threadIdx.x + kRed; base < n; base += kLoads * kRed) {
		double v[NV][kLoads];
#pragma unroll
		for (int k = 0; k < NV; k++) {
#pragma unroll
			for (int u = 0; u < kLoads; u++) {
				int const i = base + u * kRed;
				v[k][u] = i < n ? partial[k * n + i] : 0.0;
			}
		}
#pragma unroll
		for (int k = 0; k < NV; k++) {
#pragma unroll
			for (int u = 0; u < kLoads; u++) {
				acc[k] += v[k][u];
			}
		}
	}
#pragma unroll
	for (int k = 0; k < NV; k++) {
		acc[k] = wave_sum(acc[k]);
	}
	int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			lds[k][w] = acc[k];
		}
	}
	__syncthreads();
	if (threadIdx.x == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			double s = 0;
#pragma unroll
			for (int i = 0; i < kRed / 64; i++) {
				s += lds[k][i];
			}
			tot[k] = s;
		}
	}
}

// This thread's first partial sums (index threadIdx.x; 0 past the end).
template <int NV>
__device__ __forceinline__ void first_partials(const double* __restrict__ partial, int n, double (&first)[NV]) {
#pragma unroll
	for (int k = 0; k < NV; k++) {
		first[k] = static_cast<int>(threadIdx.x) < n ? partial[k * n + threadIdx.x] : 0.0;
	}
}

template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ partial, int n, double (&tot)[NV]) {
	double first[NV];
	first_partials<NV>(partial, n, first);
	reduce_partials<NV>(partial, n, tot, first);
}

// reduce_partials of a run that is not done, with this thread's first loads
// in flight together with the done flag (done_after); false when it is done.
template <int NV>
__device__ __forceinline__ bool reduce_live(const double* __restrict__ partial, int n, double (&tot)[NV],
                                            const Scalars* __restrict__ S) {
	double first[NV];
	first_partials<NV>(partial, n, first);
	int const done = done_after(S, first[0]);
#pragma unroll
	for (int k = 1; k < NV; k++) {
		issued_with(done, first[k]);
	}
	if (done) {
		return false;
	}
	reduce_partials<NV>(partial, n, tot, first);
	return true;
}

// z = M^-1 r for one node.
template <int PRE>
__device__ __forceinline__ dbl2 apply_precond(const double* __restrict__ dinv, int i, dbl2 r) {
	if (PRE == BFMX_PRECOND_JACOBI) {
		dbl2 const d = ld2(dinv + 2 * i);
		return dbl2{d.x * r.x, d.y * r.y};
	} else {
		dbl2 const d0 = ld2(dinv + 4 * i), d1 = ld2(dinv + 4 * i + 2);
		return dbl2{d0.x * r.x + d0.y * r.y, d1.x * r.x + d1.y * r.y};
	}
}

// y = A x for block row `row`; MT is the stored value type (double or float),
// the accumulation is FP64. ldx(c) returns x at node c. The block loop is
// unrolled kSpmvUnroll times, so the loads of that many blocks are in flight
// together.
#ifndef BFMX_SPMV_UNROLL
#define BFMX_SPMV_UNROLL 3
#endif
constexpr int kSpmvUnroll = BFMX_SPMV_UNROLL;

template <typename MT, typename LoadX>
__device__ __forceinline__ dbl2 row_spmv(const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ col,
                                         const MT* __restrict__ val, int row, LoadX ldx) {
	int const s = row >> 6, r = row & 63;
	int32_t const beg = slice_ptr[s], end = slice_ptr[s + 1];
	const MT* v = val + 4 * static_cast<int64_t>(beg) + 2 * r;
	double y0 = 0, y1 = 0;
#pragma unroll kSpmvUnroll
	for (int32_t blk = beg + r; blk < end; blk += kSlice, v += 4 * kSlice) {
		int32_t const c = __builtin_nontemporal_load(col + blk);
		dbl2 const t = ldv(v), u = ldv(v + 2 * kSlice);
		dbl2 const xv = ldx(c);
		y0 += t.x * xv.x + t.y * xv.y;
		y1 += u.x * xv.x + u.y * xv.y;
	}
	return dbl2{y0, y1};
}

// s - a x accumulated without losing its rounding errors: s carries the
// running sum and c the sum of the errors, so that s + c is the row's
// b - sum A_ij x_j as accurately as if it were summed in twice the working
// precision and then rounded (the product's error exactly from an FMA,
// Knuth's TwoSum for the difference; Ogita, Rump and Oishi, SIAM J. Sci.
// Comput. 26 (2005) 1955-1988, Algorithm Dot2). Contraction is off here: a
// fused s - a x would not be the rounded difference the error terms assume.
__device__ __forceinline__ void sub_exact(double& s, double& c, double a, double x) {
#pragma clang fp contract(off)
	double const p = a * x;
	double const pe = fma(a, x, -p);   // a x = p + pe exactly
	double const t = s - p;
	double const z = t - s;
	double const e = (s - (t - z)) + (-p - z);   // s - p = t + e exactly
	s = t;
	c += e - pe;
}

// b - A x for one block row, compensated (sub_exact).
template <typename MT>
__device__ __forceinline__ dbl2 row_residual(const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ col,
                                             const MT* __restrict__ val, int row, dbl2 bv,
                                             const double* __restrict__ x) {
	int const sl = row >> 6, r = row & 63;
	int32_t const beg = slice_ptr[sl], end = slice_ptr[sl + 1];
	const MT* v = val + 4 * static_cast<int64_t>(beg) + 2 * r;
	double s0 = bv.x, s1 = bv.y, c0 = 0, c1 = 0;
#pragma unroll kSpmvUnroll
	for (int32_t blk = beg + r; blk < end; blk += kSlice, v += 4 * kSlice) {
		int32_t const c = __builtin_nontemporal_load(col + blk);
		dbl2 const t = ldv(v), u = ldv(v + 2 * kSlice);
		dbl2 const xv = ld2(x + 2 * c);
		sub_exact(s0, c0, t.x, xv.x);
		sub_exact(s0, c0, t.y, xv.y);
		sub_exact(s1, c1, u.x, xv.x);
		sub_exact(s1, c1, u.y, xv.y);
	}
	return dbl2{s0 + c0, s1 + c1};
}

// p_new = z + beta p_old, q = A p_new, partial p_new . q.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(int nb, const int32_t* __restrict__ slice_ptr,
                                                    const int32_t* __restrict__ col, const MT* __restrict__ val,
                                                    const double* __restrict__ z, const double* __restrict__ pold,
                                                    double* __restrict__ pnew, double* __restrict__ q,
                                                    double* __restrict__ partial, const Scalars* __restrict__ S) {
	if (S->done) {
		return;
	}
	double const beta = S->beta;
	int const row = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	double v[1] = {0};
	if (row < nb) {
		dbl2 const y = row_spmv(slice_ptr, col, val, row, [&](int c) {
			return ld2(z + 2 * c) + beta * ld2(pold + 2 * c);
		});
		dbl2 const pn = ld2(z + 2 * row) + beta * ld2(pold + 2 * row);
		st2(pnew + 2 * row, pn);
		st2(q + 2 * row, y);
		v[0] = pn.x * y.x + pn.y * y.y;
	}
	write_partials<1>(v, partial);
}

// alpha = (r . z) / (p . q), or a breakdown.
__global__ __launch_bounds__(kRed) void k_red_alpha(int nwg, const double* __restrict__ partial,
                                                    Scalars* __restrict__ S) {
	double tot[1];
	if (!reduce_live<1>(partial, nwg, tot, S)) {
		return;
	}
	if (threadIdx.x == 0) {
		double const pq = tot[0];
		S->pq = pq;
		if (!(pq > 0) || !isfinite(pq)) {
			S->done = 2;
		} else {
			S->alpha = S->rz / pq;
		}
	}
}

// x += alpha p, r -= alpha q, z = M^-1 r, partials r . z and r . r.
template <int PRE>
__global__ __launch_bounds__(kBlock) void k_cg_update(int nb, const double* __restrict__ p,
                                                      const double* __restrict__ q, const double* __restrict__ dinv,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ z, double* __restrict__ partial,
                                                      const Scalars* __restrict__ S) {
	if (S->done) {
		return;
	}
	double const alpha = S->alpha;
	int const i = blockIdx.x * kBlock + threadIdx.x;
	double v[2] = {0, 0};
	if (i < nb) {
		dbl2 const xv = ld2(x + 2 * i) + alpha * ld2(p + 2 * i);
		dbl2 const rv = ld2(r + 2 * i) - alpha * ld2(q + 2 * i);
		dbl2 const zv = apply_precond<PRE>(dinv, i, rv);
		st2(x + 2 * i, xv);
		st2(r + 2 * i, rv);
		st2(z + 2 * i, zv);
		v[0] = rv.x * zv.x + rv.y * zv.y;
		v[1] = rv.x * rv.x + rv.y * rv.y;
	}
	write_partials<2>(v, partial);
}

// beta = r'.z' / r.z, iteration count.
__global__ __launch_bounds__(kRed) void k_red_beta(int nwg, const double* __restrict__ partial,
                                                   Scalars* __restrict__ S) {
	double tot[2];
	if (!reduce_live<2>(partial, nwg, tot, S)) {
		return;
	}
	if (threadIdx.x == 0) {
		S->beta = tot[0] / S->rz;
		S->rz_prev = S->rz;
		S->rz = tot[0];
		S->rr = tot[1];
		S->iter += 1;
	}
}

// ------------------------------------------------------------ flexible CG

// beta = -(z . q_prev) / (p_prev . q_prev), 0 before the first direction.
__global__ __launch_bounds__(kRed) void k_red_fcg_beta(int nwg, const double* __restrict__ partial,
                                                       Scalars* __restrict__ S) {
	double tot[1];
	if (!reduce_live<1>(partial, nwg, tot, S)) {
		return;
	}
	if (threadIdx.x == 0) {
		S->beta = S->pq > 0 ? -tot[0] / S->pq : 0.0;
	}
}

// p_new = z + beta p_old (z the preconditioner's output: FP32 from the
// V-cycle, FP64 from the learned preconditioner), q = A p_new with the FP64
// matrix, partials p_new . q and p_new . r.
// SEG: a segmented flexible CG (beta of the row's segment, partial sums kept
// apart per segment).
#ifdef BFMX_FCG_WAVES
#define BFMX_FCG_ATTR __attribute__((amdgpu_waves_per_eu(BFMX_FCG_WAVES, BFMX_FCG_WAVES)))
#else
#define BFMX_FCG_ATTR
#endif
template <typename ZT, bool SEG = false>
__global__ __launch_bounds__(kBlock) BFMX_FCG_ATTR void k_fcg_spmv(int nb, const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ val, const ZT* __restrict__ z,
                                                     double* __restrict__ p0, double* __restrict__ p1,
                                                     double* __restrict__ q, const double* __restrict__ r,
                                                     double* __restrict__ partial, const Scalars* __restrict__ S,
                                                     SegView sv = SegView{}, const double* __restrict__ seg_scal = nullptr) {
	if (S->done) {
		return;
	}
	// the two direction buffers alternate with the count of iterations run
	// (the iterations a hold skips do not switch them)
	bool const odd = S->iter & 1;
	const double* __restrict__ pold = odd ? p1 : p0;
	double* __restrict__ pnew = odd ? p0 : p1;
	int const blk = xcd_remap(blockIdx.x, gridDim.x);
	bool const second = SEG && static_cast<int>(threadIdx.x) >= sv.split[blk];
	double const beta = SEG ? seg_scal[3 * sv.seg[2 * blk + second] + 1] : S->beta;
	int const row = blk * kBlock + threadIdx.x;
	auto ldz = [&](int c) {
		if constexpr (sizeof(ZT) == 4) {
			flt2 const f = *reinterpret_cast<const flt2*>(z + 2 * c);
			return dbl2{f.x, f.y};
		} else {
			return ld2(reinterpret_cast<const double*>(z) + 2 * c);
		}
	};
	double v[2] = {0, 0};
	if (row < nb) {
		dbl2 const y = row_spmv(slice_ptr, col, val, row, [&](int c) { return ldz(c) + beta * ld2(pold + 2 * c); });
		dbl2 const pn = ldz(row) + beta * ld2(pold + 2 * row);
		dbl2 const rv = ld2(r + 2 * row);
		st2(pnew + 2 * row, pn);
		st2(q + 2 * row, y);
		v[0] = pn.x * y.x + pn.y * y.y;
		v[1] = pn.x * rv.x + pn.y * rv.y;
	}
	if (SEG) {
		double lo[2] = {second ? 0.0 : v[0], second ? 0.0 : v[1]};
		double hi[2] = {second ? v[0] : 0.0, second ? v[1] : 0.0};
		seg_block_partials<2>(lo, hi, blk, gridDim.x, partial);
	} else {
		write_partials<2>(v, partial);
	}
}

// Segmented flexible CG: per segment (one wavefront each), the sums of its
// blocks' partials in block order, then MODE 0: beta_k = -(z . q_old)_k /
// (p . q)_k of the previous step; MODE 1: (p . q)_k and alpha_k =
// (p . r)_k / (p . q)_k. A segment whose direction vanishes (no load, or
// solved) gets alpha_k = 0; p . q < 0 or not finite is a breakdown.
// seg_scal holds alpha, beta, p . q per segment.
template <int MODE>
__global__ __launch_bounds__(64) void k_seg_red_fcg(const int64_t* __restrict__ off, SegView sv, int64_t nblk,
                                                    const double* __restrict__ partial,
                                                    double* __restrict__ seg_scal, Scalars* __restrict__ S) {
	if (S->done) {
		return;
	}
	int const k = blockIdx.x;
	int64_t const b0 = off[k] / kBlock, b1 = (off[k + 1] - 1) / kBlock;
	double a = 0, c = 0;
	for (int64_t b = b0 + threadIdx.x; b <= b1; b += 64) {
		int const h = sv.seg[2 * b] == k ? 0 : 1;
		a += partial[2 * b + h];
		if (MODE == 1) {
			c += partial[2 * (nblk + b) + h];
		}
	}
	a = wave_sum(a);
	c = wave_sum(c);
	if (threadIdx.x == 0) {
		double* sc = seg_scal + 3 * k;
		if (MODE == 0) {
			sc[1] = sc[2] > 0 ? -a / sc[2] : 0.0;
		} else {
			sc[2] = a;
			if (a > 0 && isfinite(a)) {
				sc[0] = c / a;
			} else {
				sc[0] = 0;
				if (!(a == 0 && c == 0)) {
					S->done = 2;
				}
			}
		}
	}
}

// alpha = (p . r) / (p . q), or a breakdown.
__global__ __launch_bounds__(kRed) void k_red_fcg_alpha(int nwg, const double* __restrict__ partial,
                                                        Scalars* __restrict__ S) {
	double tot[2];
	if (!reduce_live<2>(partial, nwg, tot, S)) {
		return;
	}
	if (threadIdx.x == 0) {
		double const pq = tot[0];
		S->pq = pq;
		if (!(pq > 0) || !isfinite(pq)) {
			S->done = 2;
		} else {
			S->alpha = tot[1] / pq;
		}
	}
}

// x += alpha p, r -= alpha q, partials r . r.
template <bool SEG = false>
__global__ __launch_bounds__(kBlock) void k_fcg_update(int nb, const double* __restrict__ p0,
                                                       const double* __restrict__ p1, const double* __restrict__ q,
                                                       double* __restrict__ x, double* __restrict__ r,
                                                       double* __restrict__ partial, const Scalars* __restrict__ S,
                                                       SegView sv = SegView{},
                                                       const double* __restrict__ seg_scal = nullptr) {
	if (S->done) {
		return;
	}
	const double* __restrict__ p = (S->iter & 1) ? p0 : p1;   // the direction k_fcg_spmv wrote
	bool const second = SEG && static_cast<int>(threadIdx.x) >= sv.split[blockIdx.x];
	double const alpha = SEG ? seg_scal[3 * sv.seg[2 * blockIdx.x + second]] : S->alpha;
	int const i = blockIdx.x * kBlock + threadIdx.x;
	double v[1] = {0};
	if (i < nb) {
		dbl2 const xv = ld2(x + 2 * i) + alpha * ld2(p + 2 * i);
		dbl2 const rv = ld2(r + 2 * i) - alpha * ld2(q + 2 * i);
		st2(x + 2 * i, xv);
		st2(r + 2 * i, rv);
		v[0] = rv.x * rv.x + rv.y * rv.y;
	}
	write_partials<1>(v, partial);
}

// r . r and the iteration count. When the recursive residual meets the
// tolerance (before any check has failed) the run holds (done = 4): the
// chunk's remaining kernels return at once and its closing residual kernel
// decides (converged, or running on with a check at every chunk end).
__global__ __launch_bounds__(kRed) void k_red_fcg_rr(int nwg, const double* __restrict__ partial,
                                                     Scalars* __restrict__ S) {
	double tot[1];
	if (!reduce_live<1>(partial, nwg, tot, S)) {
		return;
	}
	if (threadIdx.x == 0) {
		S->rr = tot[0];
		S->iter += 1;
		if (tot[0] <= S->tol2 && !S->checking) {
			S->done = kHold;
		}
	}
}

// Whether the residual recomputation runs (mode 1): the recursive residual
// met the tolerance, or dropped below delta2 times its value at the last
// recomputation, or a true-residual check has already failed (from then on
// the true residual is checked at the end of every chunk).
__device__ __forceinline__ bool residual_due(const Scalars* S, double delta2) {
	return (S->done == 0 || S->done == kHold) && (S->rr <= S->tol2 || S->rr <= delta2 * S->rr_ref || S->checking);
}

// r = b - A x with the FP64 matrix, compensated (row_residual: the true
// residual of the FP64 iterate, free of the rounding of its own evaluation,
// which on a strongly heterogeneous matrix is of the order of 1e-8 of ||b||),
// z = M^-1 r; partials r.z, r.r, b.b. mode 0 starts a run (directions
// zeroed); mode 1 runs when residual_due.
template <int PRE>
__global__ __launch_bounds__(kBlock) void k_residual(int nb, int mode, double delta2,
                                                     const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ val, const double* __restrict__ b,
                                                     const double* __restrict__ dinv, const double* __restrict__ x,
                                                     double* __restrict__ r, double* __restrict__ z,
                                                     double* __restrict__ p0, double* __restrict__ p1,
                                                     double* __restrict__ partial, const Scalars* __restrict__ S) {
	if (mode == 1 && !residual_due(S, delta2)) {
		return;
	}
	int const row = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	double v[3] = {0, 0, 0};
	if (row < nb) {
		dbl2 const bv = ld2(b + 2 * row);
		dbl2 const rv = row_residual(slice_ptr, col, val, row, bv, x);
		dbl2 const zv = apply_precond<PRE>(dinv, row, rv);
		st2(r + 2 * row, rv);
		st2(z + 2 * row, zv);
		if (mode == 0) {
			st2(p0 + 2 * row, dbl2{0, 0});
			st2(p1 + 2 * row, dbl2{0, 0});
		}
		v[0] = rv.x * zv.x + rv.y * zv.y;
		v[1] = rv.x * rv.x + rv.y * rv.y;
		v[2] = bv.x * bv.x + bv.y * bv.y;
	}
	write_partials<3>(v, partial);
}

// Convergence segments (a system made of independent parts, each to meet
// the tolerance on its own): per segment of rows off[k] .. off[k+1] - 1, the
// true residual's r.r and b.b, one workgroup per segment, after k_residual
// and under its condition.
constexpr int kSegBlock = 256;

__global__ __launch_bounds__(kSegBlock) void k_seg_norms(int mode, double delta2, const int64_t* __restrict__ off,
                                                         const double* __restrict__ r,
                                                         const double* __restrict__ b, double* __restrict__ part,
                                                         const Scalars* __restrict__ S) {
	if (mode == 1 && !residual_due(S, delta2)) {
		return;
	}
	__shared__ double lds[2][kSegBlock / 64];
	int64_t const d0 = 2 * off[blockIdx.x], d1 = 2 * off[blockIdx.x + 1];
	double rr = 0, bb = 0;
	for (int64_t d = d0 + threadIdx.x; d < d1; d += kSegBlock) {
		rr += r[d] * r[d];
		bb += b[d] * b[d];
	}
	rr = wave_sum(rr);
	bb = wave_sum(bb);
	if ((threadIdx.x & 63) == 0) {
		lds[0][threadIdx.x >> 6] = rr;
		lds[1][threadIdx.x >> 6] = bb;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double a = 0, c = 0;
		for (int w = 0; w < kSegBlock / 64; w++) {
			a += lds[0][w];
			c += lds[1][w];
		}
		part[2 * blockIdx.x] = a;
		part[2 * blockIdx.x + 1] = c;
	}
}

// Completes k_residual. mode 0: starts the run. mode 1 (when it ran): the
// recomputed residual replaces the recursive one and the directions are kept
// (beta = r'.z' / previous r.z). When the true residual meets the tolerance
// (with convergence segments: every segment with a nonzero load meets it on
// its own, r_k.r_k <= rtol2 b_k.b_k) the run ends (done = 1). Once a tolerance
// check has failed, every chunk ends with a check; when the true residual has
// not halved (in norm) below its best for kStallChecks checks, it sits at its
// rounding floor and the run ends (done = 3).
constexpr int kStallChecks = 8;

__global__ __launch_bounds__(kRed) void k_red_residual(int nwg, int mode, double rtol2, double delta2,
                                                       const double* __restrict__ partial, int nseg,
                                                       const double* __restrict__ seg_part,
                                                       Scalars* __restrict__ S) {
	if (mode == 1 && !residual_due(S, delta2)) {
		return;
	}
	bool const check = mode == 1 && (S->rr <= S->tol2 || S->checking);
	int seg_fail = 0;
	for (int k = threadIdx.x; k < nseg; k += kRed) {
		double const bk = seg_part[2 * k + 1];
		if (bk > 0 && !(seg_part[2 * k] <= rtol2 * bk)) {
			seg_fail = 1;
		}
	}
	seg_fail = __syncthreads_or(seg_fail);
	double tot[3];
	reduce_partials<3>(partial, nwg, tot);
	if (threadIdx.x == 0) {
		double const rr = tot[1];
		S->bb = tot[2];
		S->tol2 = rtol2 * tot[2];
		S->rr = rr;
		S->rr_true = rr;
		S->rr_ref = rr;
		S->updates += 1;
		if (mode == 0) {
			S->beta = 0;
			S->iter = 0;
			S->stall = 0;
			S->checking = 0;
			S->rr_best = 0;
		} else {
			S->beta = S->rz_prev != 0 ? tot[0] / S->rz_prev : 0;
		}
		S->rz = tot[0];
		if (S->done == kHold) {
			S->done = 0;
		}
		if (rr <= S->tol2 && !seg_fail) {
			S->done = 1;
		} else if (check) {
			S->checking = 1;
			if (S->rr_best == 0 || rr < 0.25 * S->rr_best) {
				S->rr_best = rr;
				S->stall = 0;
			} else if (++S->stall >= kStallChecks) {
				S->done = 3;
			}
		}
	}
}

// Point or 2x2 block Jacobi from the diagonal blocks.
__global__ __launch_bounds__(kBlock) void k_precond(int nb, int kind, const int32_t* __restrict__ diag_blk,
                                                    const double* __restrict__ val, double* __restrict__ dinv) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= nb) {
		return;
	}
	int const r = i & 63;
	const double* top = val + 4 * static_cast<int64_t>(diag_blk[i]) - 2 * r;
	double const a00 = top[0], a01 = top[1], a10 = top[2 * kSlice], a11 = top[2 * kSlice + 1];
	if (kind == BFMX_PRECOND_JACOBI) {
		dinv[2 * i] = 1.0 / a00;
		dinv[2 * i + 1] = 1.0 / a11;
	} else {
		double const det = a00 * a11 - a01 * a10;
		dinv[4 * i + 0] = a11 / det;
		dinv[4 * i + 1] = -a01 / det;
		dinv[4 * i + 2] = -a10 / det;
		dinv[4 * i + 3] = a00 / det;
	}
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (i < n) {
		b[i] = static_cast<float>(a[i]);
	}
}

__global__ __launch_bounds__(kBlock) void k_fill(int64_t n, double v, double* __restrict__ a) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (i < n) {
		a[i] = v;
	}
}

int grid(int64_t n) {
	return static_cast<int>((n + kBlock - 1) / kBlock);
}

int ensure_workspace(bfmx_system* s) {
	int64_t const n = s->n_nodes;
	if (s->d_x) {
		return BFMX_OK;
	}
	// from the device block cache (a fresh system's first solve would
	// otherwise spend ~0.4 ms in hipMalloc and hipHostMalloc here)
	BFMX_CHECK(palloc(&s->d_x, 2 * n));
	BFMX_CHECK(palloc(&s->d_r, 2 * n));
	BFMX_CHECK(palloc(&s->d_z, 2 * n));
	BFMX_CHECK(palloc(&s->d_q, 2 * n));
	BFMX_CHECK(palloc(&s->d_p[0], 2 * n));
	BFMX_CHECK(palloc(&s->d_p[1], 2 * n));
	BFMX_CHECK(palloc(&s->d_dinv, 4 * n));
	// per-workgroup partial sums: three per 256-row workgroup (the residual
	// kernel), one per 64-row workgroup (the learned preconditioner's decoder)
	s->partial_cap = std::max<int64_t>(3 * static_cast<int64_t>(grid(n)), (n + 63) / 64);
	BFMX_CHECK(palloc(&s->d_partial, static_cast<size_t>(s->partial_cap)));
	BFMX_CHECK(palloc(&s->d_scal, 1));
	BFMX_CHECK(pinned_small_alloc(reinterpret_cast<void**>(&s->h_scal), 4 * sizeof(Scalars)));
	for (auto& e : s->ev) {
		BFMX_HIP(hipEventCreate(&e));
	}
	BFMX_HIP(hipMemsetAsync(s->d_scal, 0, sizeof(Scalars), s->stream));
	BFMX_HIP(hipMemsetAsync(s->d_p[0], 0, 2 * n * sizeof(double), s->stream));
	BFMX_HIP(hipMemsetAsync(s->d_p[1], 0, 2 * n * sizeof(double), s->stream));
	return BFMX_OK;
}

// FP32 copy of the matrix values, refreshed at every call (the values change
// with assembly and conditions; the buffer, and so every captured graph that
// reads it, stays).
int ensure_f32(bfmx_system* s) {
	if (!s->d_val32) {
		BFMX_CHECK(palloc(&s->d_val32, 4 * s->nnzb_pad));
	} else if (s->f32_gen == s->val_gen) {
		return BFMX_OK;   // made from the current values
	}
	k_to_f32<<<grid(4 * s->nnzb_pad), kBlock, 0, s->stream>>>(4 * s->nnzb_pad, s->d_val, s->d_val32);
	BFMX_HIP(hipGetLastError());
	s->f32_gen = s->val_gen;
	return BFMX_OK;
}

void launch_residual(bfmx_system* s, int pre, int mode, double rtol2, double delta2) {
	int const nb = static_cast<int>(s->n_nodes);
	int const g = grid(nb);
	if (pre == BFMX_PRECOND_JACOBI) {
		k_residual<BFMX_PRECOND_JACOBI><<<g, kBlock, 0, s->stream>>>(
		    nb, mode, delta2, s->d_slice_ptr, s->d_col, s->d_val, s->d_bsolve, s->d_dinv, s->d_x, s->d_r, s->d_z,
		    s->d_p[0], s->d_p[1], s->d_partial, s->d_scal);
	} else {
		k_residual<BFMX_PRECOND_BLOCK_JACOBI><<<g, kBlock, 0, s->stream>>>(
		    nb, mode, delta2, s->d_slice_ptr, s->d_col, s->d_val, s->d_bsolve, s->d_dinv, s->d_x, s->d_r, s->d_z,
		    s->d_p[0], s->d_p[1], s->d_partial, s->d_scal);
	}
	if (s->n_conv_seg > 0) {
		k_seg_norms<<<s->n_conv_seg, kSegBlock, 0, s->stream>>>(mode, delta2, s->d_conv_seg, s->d_r, s->d_bsolve,
		                                                        s->d_conv_part, s->d_scal);
	}
	k_red_residual<<<1, kRed, 0, s->stream>>>(g, mode, rtol2, delta2, s->d_partial, s->n_conv_seg, s->d_conv_part,
	                                          s->d_scal);
}

void launch_spmv(bfmx_system* s, bool mixed, const double* pold, double* pnew) {
	int const nb = static_cast<int>(s->n_nodes);
	int const g = grid(nb);
	if (mixed) {
		k_cg_spmv<float><<<g, kBlock, 0, s->stream>>>(nb, s->d_slice_ptr, s->d_col, s->d_val32, s->d_z, pold, pnew,
		                                              s->d_q, s->d_partial, s->d_scal);
	} else {
		k_cg_spmv<double><<<g, kBlock, 0, s->stream>>>(nb, s->d_slice_ptr, s->d_col, s->d_val, s->d_z, pold, pnew,
		                                               s->d_q, s->d_partial, s->d_scal);
	}
}

// One CG iteration; `it` (position in the chunk) selects the direction buffers.
void launch_iteration(bfmx_system* s, int pre, bool mixed, int it) {
	int const nb = static_cast<int>(s->n_nodes);
	int const g = grid(nb);
	int const a = it & 1, b = a ^ 1;
	launch_spmv(s, mixed, s->d_p[a], s->d_p[b]);
	k_red_alpha<<<1, kRed, 0, s->stream>>>(g, s->d_partial, s->d_scal);
	if (pre == BFMX_PRECOND_JACOBI) {
		k_cg_update<BFMX_PRECOND_JACOBI><<<g, kBlock, 0, s->stream>>>(nb, s->d_p[b], s->d_q, s->d_dinv, s->d_x,
		                                                             s->d_r, s->d_z, s->d_partial, s->d_scal);
	} else {
		k_cg_update<BFMX_PRECOND_BLOCK_JACOBI><<<g, kBlock, 0, s->stream>>>(nb, s->d_p[b], s->d_q, s->d_dinv,
		                                                                   s->d_x, s->d_r, s->d_z, s->d_partial,
		                                                                   s->d_scal);
	}
	k_red_beta<<<1, kRed, 0, s->stream>>>(g, s->d_partial, s->d_scal);
}

// One flexible-CG iteration with the multigrid V-cycle or the learned
// preconditioner.
void launch_fcg_iteration(bfmx_system* s, int pre) {
	int const nb = static_cast<int>(s->n_nodes);
	int const g = grid(nb);
	int nwz = 0;
	if (pre == BFMX_PRECOND_AMG && s->fcg_sv.seg) {
		// segmented: every convergence segment runs its own alpha and beta
		int const m = s->n_conv_seg;
		float* z = amg_vcycle(s, s->d_r, s->d_q, s->d_fcg_part, s->d_scal, &nwz);
		k_seg_red_fcg<0><<<m, 64, 0, s->stream>>>(s->d_conv_seg, s->fcg_sv, g, s->d_fcg_part, s->d_fcg_scal,
		                                          s->d_scal);
		k_fcg_spmv<float, true><<<g, kBlock, 0, s->stream>>>(nb, s->d_slice_ptr, s->d_col, s->d_val, z, s->d_p[0],
		                                                     s->d_p[1], s->d_q, s->d_r, s->d_fcg_part, s->d_scal,
		                                                     s->fcg_sv, s->d_fcg_scal);
		k_seg_red_fcg<1><<<m, 64, 0, s->stream>>>(s->d_conv_seg, s->fcg_sv, g, s->d_fcg_part, s->d_fcg_scal,
		                                          s->d_scal);
		k_fcg_update<true><<<g, kBlock, 0, s->stream>>>(nb, s->d_p[0], s->d_p[1], s->d_q, s->d_x, s->d_r,
		                                                s->d_partial, s->d_scal, s->fcg_sv, s->d_fcg_scal);
		k_red_fcg_rr<<<1, kRed, 0, s->stream>>>(g, s->d_partial, s->d_scal);
		return;
	}
	if (pre == BFMX_PRECOND_AMG) {
		float* z = amg_vcycle(s, s->d_r, s->d_q, s->d_partial, s->d_scal, &nwz);
		k_red_fcg_beta<<<1, kRed, 0, s->stream>>>(nwz, s->d_partial, s->d_scal);
		k_fcg_spmv<float><<<g, kBlock, 0, s->stream>>>(nb, s->d_slice_ptr, s->d_col, s->d_val, z, s->d_p[0],
		                                               s->d_p[1], s->d_q, s->d_r, s->d_partial, s->d_scal);
	} else {
		// the learned network's correction added to block Jacobi, or to the
		// multigrid V-cycle (its dot partials are not needed: z . q comes
		// from the decoder)
		int nwv = 0;
		const float* base = pre == BFMX_PRECOND_AMG_LEARNED
		                        ? amg_vcycle(s, s->d_r, nullptr, nullptr, s->d_scal, &nwv)
		                        : nullptr;
		double* z = mpnn_apply(s, s->d_r, s->d_q, s->d_partial, s->d_scal, &nwz, base);
		k_red_fcg_beta<<<1, kRed, 0, s->stream>>>(nwz, s->d_partial, s->d_scal);
		k_fcg_spmv<double><<<g, kBlock, 0, s->stream>>>(nb, s->d_slice_ptr, s->d_col, s->d_val, z, s->d_p[0],
		                                                s->d_p[1], s->d_q, s->d_r, s->d_partial, s->d_scal);
	}
	k_red_fcg_alpha<<<1, kRed, 0, s->stream>>>(g, s->d_partial, s->d_scal);
	k_fcg_update<false><<<g, kBlock, 0, s->stream>>>(nb, s->d_p[0], s->d_p[1], s->d_q, s->d_x, s->d_r, s->d_partial,
	                                                 s->d_scal);
	k_red_fcg_rr<<<1, kRed, 0, s->stream>>>(g, s->d_partial, s->d_scal);
}

constexpr int kFcgChunk = 4;   // flexible-CG iterations per captured graph

bool is_learned(int pre) {
	return pre == BFMX_PRECOND_LEARNED || pre == BFMX_PRECOND_AMG_LEARNED;
}

bool is_fcg(int pre) {
	return pre == BFMX_PRECOND_AMG || is_learned(pre);
}

int chunk_iters(int pre) {
	return is_fcg(pre) ? kFcgChunk : kChunk;
}

// One chunk of iterations and the conditional residual recomputation. The
// multigrid path recomputes the residual only for the tolerance checks (its
// recursive residual is FP64 throughout). Flexible CG runs n iterations
// (1 .. kFcgChunk); the other solvers run kChunk.
void launch_chunk(bfmx_system* s, int pre, bool mixed, double rtol2, int n) {
	if (is_fcg(pre)) {
		for (int it = 0; it < n; it++) {
			launch_fcg_iteration(s, pre);
		}
		launch_residual(s, is_learned(pre) ? BFMX_PRECOND_BLOCK_JACOBI : BFMX_PRECOND_JACOBI, 1, rtol2,
		                0.0);
		return;
	}
	for (int it = 0; it < kChunk; it++) {
		launch_iteration(s, pre, mixed, it);
	}
	launch_residual(s, pre, 1, rtol2, mixed ? kDelta2 : 0.0);
}


// The captured chunk for this configuration (captured on first use).
int capture_chunk(bfmx_system* s, int pre, bool mixed, double rtol2, int n, hipGraphExec_t* out) {
	int const prec = mixed ? 1 : 0;
	bool const learned = is_learned(pre) && s->model;
	uint64_t const gen = learned ? s->model->gen : 0;
	const float* params = learned ? s->model->d_params : nullptr;
	for (auto const& gr : s->graphs) {
		if (gr.precond == pre && gr.precision == prec && gr.rtol2 == rtol2 && gr.model_gen == gen &&
		    gr.model_params == params && gr.iters == n) {
			*out = gr.exec;
			return BFMX_OK;
		}
	}
	hipGraph_t g = nullptr;
	BFMX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
	launch_chunk(s, pre, mixed, rtol2, n);
	BFMX_HIP(hipStreamEndCapture(s->stream, &g));
	Graph gr;
	hipError_t const e = hipGraphInstantiate(&gr.exec, g, nullptr, nullptr, 0);
	(void) hipGraphDestroy(g);
	if (e != hipSuccess) {
		set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
		return BFMX_E_HIP;
	}
	gr.precond = pre;
	gr.precision = prec;
	gr.rtol2 = rtol2;
	gr.model_gen = gen;
	gr.model_params = params;
	gr.iters = n;
	s->graphs.push_back(gr);
	*out = gr.exec;
	return BFMX_OK;
}

// A flexible-CG run (learned or multigrid preconditioner) stalls when its
// best r.r has not halved for twice as many iterations as the run has taken
// (and at least kFcgStall): a window that grows with the run, as CG's
// residual plateaus do. Both preconditioners are SPD, so this is a safeguard
// (a diverging or NaN run ends instead of running to the iteration cap), not
// the expected end of a run.
constexpr int64_t kFcgStall = 4000;

// Runs chunks of one solver from the current state (warm: x kept, else x = 0)
// until the device flags an end, cap iterations are launched or, with
// detect_stall, the recursive residual stops falling. *out: final scalars.
int run_chunks(bfmx_system* s, int precond, bool mixed, double rtol2, int64_t cap, bool warm, bool detect_stall,
               Scalars* out) {
	int64_t const n = s->n_nodes;
	hipStream_t const st = s->stream;
	// Chunks go through captured graphs from a system's second solve with a
	// preconditioner on: capturing and instantiating them costs about as much
	// as a large system's whole launch overhead (2M-DOF beam, multigrid FCG:
	// first solve 19.4 ms launching directly, 21.2 ms capturing; later solves
	// 10.5 ms either way), while a small system's repeated solves gain from
	// them.
	uint64_t const bit = uint64_t(1) << (precond & 63);
	bool const use_graph = graphs_enabled() && (s->solved_once & bit);
	s->solved_once |= bit;
	// the residual kernel's z (unused by flexible CG); the learned decoder reads
	// the 2x2 block inverses
	int const pre_res = precond == BFMX_PRECOND_AMG      ? BFMX_PRECOND_JACOBI
	                    : is_learned(precond)             ? BFMX_PRECOND_BLOCK_JACOBI
	                                                      : precond;
	k_precond<<<grid(n), kBlock, 0, st>>>(static_cast<int>(n), pre_res, s->d_diag_blk, s->d_val, s->d_dinv);
	if (!warm) {
		BFMX_HIP(hipMemsetAsync(s->d_x, 0, 2 * n * sizeof(double), st));
	}
	BFMX_HIP(hipMemsetAsync(s->d_q, 0, 2 * n * sizeof(double), st));
	BFMX_HIP(hipMemsetAsync(s->d_scal, 0, sizeof(Scalars), st));
	if (s->d_fcg_scal) {
		BFMX_HIP(hipMemsetAsync(s->d_fcg_scal, 0, 3 * static_cast<size_t>(s->n_conv_seg) * sizeof(double), st));
	}
	launch_residual(s, pre_res, 0, rtol2, 0.0);
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipMemcpyAsync(&s->h_scal[0], s->d_scal, sizeof(Scalars), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipStreamSynchronize(st));
	Scalars last = s->h_scal[0];
	double mark = last.rr;
	int64_t mark_iter = 0;
	// Every chunk length the loop may launch is captured here, while the
	// stream is idle: a capture begun on the stream while an earlier chunk's
	// graph launch is still running on it is avoided (capture mode is
	// per-thread, but the runtime and tools that intercept launches then see
	// a capture interleaved with in-flight work on the same stream).
	hipGraphExec_t execs[kFcgChunk > kChunk ? kFcgChunk + 1 : kChunk + 1] = {};
	if (use_graph) {
		int const lo = is_fcg(precond) ? 1 : kChunk;
		for (int len = lo; len <= chunk_iters(precond); len++) {
			BFMX_CHECK(capture_chunk(s, precond, mixed, rtol2, len, &execs[len]));
		}
	}

	// Replay chunks, keeping one in flight while the one before it is
	// inspected. Launches after the end return at once but still cost a
	// dispatch each, so the fall of r.r per iteration over the chunk inspected
	// last (q) sizes the end: no chunk is launched ahead when the one in
	// flight is expected to reach the tolerance (it is waited for instead; a
	// wrong guess leaves the GPU idle for one wait on the host), and a
	// flexible-CG chunk runs only the iterations expected to be needed (at
	// least one, at most kFcgChunk).
	int64_t launched = 0, chunks = 0, seen = -1;
	int its[4] = {};          // iterations of the chunks in the ring
	double rr_before = -1;    // r.r and iterations at the inspection before `last` (when consecutive)
	int64_t iter_before = 0;
	while (last.done == 0) {
		bool const more = launched < cap;
		bool const in_flight = seen < chunks - 1;
		double q = -1;
		if (rr_before > 0 && last.iter > iter_before && last.rr > 0) {
			q = std::pow(last.rr / rr_before, 1.0 / static_cast<double>(last.iter - iter_before));
		}
		double const ahead = in_flight && q > 0 ? last.rr * std::pow(q, its[(chunks - 1) & 3]) : last.rr;
		bool const ending = in_flight && q > 0 && ahead <= last.tol2;
		bool const launch = more && !ending;
		if (launch) {
			int n = chunk_iters(precond);
			if (is_fcg(precond) && q > 0 && q < 1 && ahead > last.tol2) {
				double const need = std::ceil(std::log(last.tol2 / ahead) / std::log(q));
				n = static_cast<int>(std::min<double>(std::max(need, 1.0), kFcgChunk));
			}
			if (use_graph) {
				BFMX_HIP(hipGraphLaunch(execs[n], st));
			} else {
				launch_chunk(s, precond, mixed, rtol2, n);
				BFMX_HIP(hipGetLastError());
			}
			launched += n;
			its[chunks & 3] = n;
			int const slot = static_cast<int>(chunks & 3);
			BFMX_HIP(hipMemcpyAsync(&s->h_scal[slot], s->d_scal, sizeof(Scalars), hipMemcpyDeviceToHost, st));
			BFMX_HIP(hipEventRecord(s->ev[slot], st));
			chunks++;
		}
		int64_t const look = (launch && chunks >= 2) ? chunks - 2 : chunks - 1;
		BFMX_HIP(hipEventSynchronize(s->ev[look & 3]));
		rr_before = look == seen + 1 ? last.rr : -1;
		iter_before = last.iter;
		last = s->h_scal[look & 3];
		seen = look;
		if (!more && look == chunks - 1) {
			break;
		}
		if (detect_stall && last.done == 0) {
			if (last.rr < 0.5 * mark) {
				mark = last.rr;
				mark_iter = last.iter;
			} else if (last.iter - mark_iter >= std::max(kFcgStall, 2 * mark_iter)) {
				break;
			}
		}
	}
	BFMX_HIP(hipStreamSynchronize(st));
	BFMX_HIP(hipMemcpy(out, s->d_scal, sizeof(Scalars), hipMemcpyDeviceToHost));
	return BFMX_OK;
}

}  // namespace

void drop_graphs(bfmx_system* s, int precond) {
	(void) hipStreamSynchronize(s->stream);
	auto keep = s->graphs.begin();
	for (auto& gr : s->graphs) {
		if (precond < 0 || gr.precond == precond) {
			(void) hipGraphExecDestroy(gr.exec);
		} else {
			*keep++ = gr;
		}
	}
	s->graphs.erase(keep, s->graphs.end());
}

void release_solver(bfmx_system* s) {
	if (s->stream) {
		(void) hipStreamSynchronize(s->stream);   // the block cache takes back drained blocks only
	}
	amg_release(s);
	dfree(s->d_fcg_seg);
	dfree(s->d_fcg_split);
	dfree(s->d_fcg_scal);
	dfree(s->d_fcg_part);
	s->fcg_sv = SegView{};
	mpnn_release(s);
	for (auto& gr : s->graphs) {
		(void) hipGraphExecDestroy(gr.exec);
	}
	s->graphs.clear();
	pfree(s->d_x);
	pfree(s->d_r);
	pfree(s->d_z);
	pfree(s->d_q);
	pfree(s->d_p[0]);
	pfree(s->d_p[1]);
	pfree(s->d_dinv);
	pfree(s->d_partial);
	pfree(s->d_scal);
	pfree(s->d_val32);
	pinned_small_free(s->h_scal, 4 * sizeof(Scalars));
	s->h_scal = nullptr;
	for (auto& e : s->ev) {
		if (e) {
			(void) hipEventDestroy(e);
			e = nullptr;
		}
	}
}

namespace {

// ------------------------------------------------------------------ MINRES
// Preconditioned MINRES (Paige and Saunders, SIAM J. Numer. Anal. 12 (1975)
// 617-629) for symmetric indefinite systems, with M = |diag(A)|. Per
// iteration: one SpMV pass (v = y / beta, y' = A v - (beta / beta_old) r1,
// partial v . y'), one update pass (y' -= (alpha / beta) r2, z = M^-1 y',
// partial y' . z), the Givens recurrences on the host, and the direction and
// solution update. The sums are fixed-order per-workgroup partials.

__global__ __launch_bounds__(kBlock) void k_mr_diag(int nb, const int32_t* __restrict__ diag_blk,
                                                    const double* __restrict__ val, double* __restrict__ dinv) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= nb) {
		return;
	}
	const double* top = val + 4 * static_cast<int64_t>(diag_blk[i]) - 2 * (i & 63);
	double const a0 = fabs(top[0]), a1 = fabs(top[2 * kSlice + 1]);
	dinv[2 * i] = a0 > 0 ? 1.0 / a0 : 1.0;
	dinv[2 * i + 1] = a1 > 0 ? 1.0 / a1 : 1.0;
}

// r1 = b, r2 = b, y = M^-1 b, x = w = w2 = 0; partials b . y, b . b.
__global__ __launch_bounds__(kBlock) void k_mr_start(int nb, const double* __restrict__ b,
                                                     const double* __restrict__ dinv, double* __restrict__ r1,
                                                     double* __restrict__ r2, double* __restrict__ y,
                                                     double* __restrict__ x, double* __restrict__ w,
                                                     double* __restrict__ w2, double* __restrict__ partial) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	double v[2] = {0, 0};
	if (i < nb) {
		dbl2 const bv = ld2(b + 2 * i), d = ld2(dinv + 2 * i);
		dbl2 const yv = dbl2{d.x * bv.x, d.y * bv.y};
		st2(r1 + 2 * i, bv);
		st2(r2 + 2 * i, bv);
		st2(y + 2 * i, yv);
		st2(x + 2 * i, dbl2{0, 0});
		st2(w + 2 * i, dbl2{0, 0});
		st2(w2 + 2 * i, dbl2{0, 0});
		v[0] = bv.x * yv.x + bv.y * yv.y;
		v[1] = bv.x * bv.x + bv.y * bv.y;
	}
	write_partials<2>(v, partial);
}

// v = sy y, yn = sy A y - sr r1; partial v . yn.
__global__ __launch_bounds__(kBlock) void k_mr_spmv(int nb, const int32_t* __restrict__ slice_ptr,
                                                    const int32_t* __restrict__ col, const double* __restrict__ val,
                                                    const double* __restrict__ y, const double* __restrict__ r1,
                                                    double sy, double sr, double* __restrict__ v,
                                                    double* __restrict__ yn, double* __restrict__ partial) {
	int const i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	double d[1] = {0};
	if (i < nb) {
		dbl2 const ay = row_spmv<double>(slice_ptr, col, val, i,
		                                 [&](int32_t c) { return ld2(y + 2 * static_cast<int64_t>(c)); });
		dbl2 const yi = ld2(y + 2 * i), ri = ld2(r1 + 2 * i);
		dbl2 const vi = dbl2{sy * yi.x, sy * yi.y};
		dbl2 const ni = dbl2{sy * ay.x - sr * ri.x, sy * ay.y - sr * ri.y};
		st2(v + 2 * i, vi);
		st2(yn + 2 * i, ni);
		d[0] = vi.x * ni.x + vi.y * ni.y;
	}
	write_partials<1>(d, partial);
}

// yn -= sa r2, z = M^-1 yn; partial yn . z.
__global__ __launch_bounds__(kBlock) void k_mr_update(int nb, double* __restrict__ yn, const double* __restrict__ r2,
                                                      const double* __restrict__ dinv, double sa,
                                                      double* __restrict__ z, double* __restrict__ partial) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	double d[1] = {0};
	if (i < nb) {
		dbl2 const a = ld2(yn + 2 * i), r = ld2(r2 + 2 * i), m = ld2(dinv + 2 * i);
		dbl2 const t = dbl2{a.x - sa * r.x, a.y - sa * r.y};
		dbl2 const zi = dbl2{m.x * t.x, m.y * t.y};
		st2(yn + 2 * i, t);
		st2(z + 2 * i, zi);
		d[0] = t.x * zi.x + t.y * zi.y;
	}
	write_partials<1>(d, partial);
}

// w_new = (v - oldeps w_older - delta w_old) / gamma into wout, x += phi w_new.
__global__ __launch_bounds__(kBlock) void k_mr_x(int nb, const double* __restrict__ v,
                                                 const double* __restrict__ w_older, const double* __restrict__ w_old,
                                                 double oldeps, double delta, double rgamma, double phi,
                                                 double* __restrict__ wout, double* __restrict__ x) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= nb) {
		return;
	}
	dbl2 const vi = ld2(v + 2 * i), a = ld2(w_older + 2 * i), c = ld2(w_old + 2 * i), xi = ld2(x + 2 * i);
	dbl2 const w = dbl2{(vi.x - oldeps * a.x - delta * c.x) * rgamma, (vi.y - oldeps * a.y - delta * c.y) * rgamma};
	st2(wout + 2 * i, w);
	st2(x + 2 * i, dbl2{xi.x + phi * w.x, xi.y + phi * w.y});
}

// partial |b - A x|^2.
__global__ __launch_bounds__(kBlock) void k_mr_resid(int nb, const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ x, const double* __restrict__ b,
                                                     double* __restrict__ partial) {
	int const i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	double d[1] = {0};
	if (i < nb) {
		dbl2 const ax = row_spmv<double>(slice_ptr, col, val, i,
		                                 [&](int32_t c) { return ld2(x + 2 * static_cast<int64_t>(c)); });
		dbl2 const bi = ld2(b + 2 * i);
		double const e0 = bi.x - ax.x, e1 = bi.y - ax.y;
		d[0] = e0 * e0 + e1 * e1;
	}
	write_partials<1>(d, partial);
}

template <int NV>
__global__ __launch_bounds__(kRed) void k_mr_reduce(int nwg, const double* __restrict__ partial,
                                                    double* __restrict__ out) {
	double tot[NV];
	reduce_partials<NV>(partial, nwg, tot);
	if (threadIdx.x == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			out[k] = tot[k];
		}
	}
}

}  // namespace

bool phase_times_on() {
	static bool const on = [] {
		const char* e = std::getenv("BFMX_SOLVE_TIMES");
		return e && e[0] == '1';
	}();
	return on;
}

static double host_ms() {
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

PhaseTimer::PhaseTimer(hipStream_t stream, const char* what) : st(stream), scope(what), t(0) {
	if (phase_times_on()) {
		(void) hipStreamSynchronize(st);
		t = host_ms();
	}
}

void PhaseTimer::mark(const char* what) {
	if (!phase_times_on()) {
		return;
	}
	(void) hipStreamSynchronize(st);
	double const now = host_ms();
	std::fprintf(stderr, "%s phase %-32s %8.3f ms\n", scope, what, now - t);
	t = now;
}

namespace {

// Totals of the NV partial sums of the last kernel (host copy, synchronous).
template <int NV>
int mr_sums(hipStream_t st, int nwg, const double* partial, double* d_out, double (&h)[NV]) {
	k_mr_reduce<NV><<<1, kRed, 0, st>>>(nwg, partial, d_out);
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipMemcpyAsync(h, d_out, NV * sizeof(double), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipStreamSynchronize(st));
	return BFMX_OK;
}

struct MrResult {
	int status;
	int64_t iter;
	double rr, bb;
	int checks;
};

int solve_minres(bfmx_system* s, double rtol, int64_t cap, MrResult* res) {
	int64_t const n = s->n_nodes;
	int const nb = static_cast<int>(n);
	int const g = grid(n);
	hipStream_t const st = s->stream;
	double* wbuf[3] = {nullptr, nullptr, nullptr};
	double* d_sum = nullptr;
	auto release = [&]() {
		for (auto& p : wbuf) {
			dfree(p);
		}
		dfree(d_sum);
	};
	int const rc = [&]() -> int {
		for (auto& p : wbuf) {
			BFMX_CHECK(dalloc(&p, 2 * n));
		}
		BFMX_CHECK(dalloc(&d_sum, 4));
		double* const b = s->d_bsolve;
		double* const x = s->d_x;
		double* const dinv = s->d_dinv;
		double* rbuf[3] = {s->d_r, s->d_z, s->d_q};   // r1, r2, and the next r2
		double* const y = s->d_p[0];
		double* const v = s->d_p[1];
		double* w = wbuf[0];        // latest direction
		double* w2 = wbuf[2];       // the one before
		double* wfree = wbuf[1];    // scratch for the next
		k_mr_diag<<<g, kBlock, 0, st>>>(nb, s->d_diag_blk, s->d_val, dinv);
		k_mr_start<<<g, kBlock, 0, st>>>(nb, b, dinv, rbuf[0], rbuf[1], y, x, w, w2, s->d_partial);
		double t2[2];
		BFMX_CHECK(mr_sums<2>(st, g, s->d_partial, d_sum, t2));
		res->bb = t2[1];
		res->rr = t2[1];
		res->iter = 0;
		res->checks = 0;
		res->status = 0;
		if (!(t2[0] > 0)) {
			res->status = t2[1] == 0 ? 1 : 2;   // b = 0: x = 0; otherwise M not positive
			return BFMX_OK;
		}
		double const beta1 = std::sqrt(t2[0]);
		double const tol2 = rtol * rtol * res->bb;
		double oldb = 0, beta = beta1, dbar = 0, epsln = 0, phibar = beta1, cs = -1, sn = 0;
		int64_t next_check = 0;
		for (int64_t it = 1; it <= cap; it++) {
			double* const r1 = rbuf[0];
			double* const r2 = rbuf[1];
			double* const yn = rbuf[2];
			k_mr_spmv<<<g, kBlock, 0, st>>>(nb, s->d_slice_ptr, s->d_col, s->d_val, y, r1, 1.0 / beta,
			                                it >= 2 ? beta / oldb : 0.0, v, yn, s->d_partial);
			double t1[1];
			BFMX_CHECK(mr_sums<1>(st, g, s->d_partial, d_sum, t1));
			double const alfa = t1[0];
			k_mr_update<<<g, kBlock, 0, st>>>(nb, yn, r2, dinv, alfa / beta, y, s->d_partial);
			BFMX_CHECK(mr_sums<1>(st, g, s->d_partial, d_sum, t1));
			rbuf[0] = r2;
			rbuf[1] = yn;
			rbuf[2] = r1;
			if (!(t1[0] >= 0) || !std::isfinite(alfa)) {
				res->status = 2;   // the preconditioner is not positive definite, or overflow
				res->iter = it;
				return BFMX_OK;
			}
			oldb = beta;
			beta = std::sqrt(t1[0]);
			double const oldeps = epsln;
			double const delta = cs * dbar + sn * alfa;
			double const gbar = sn * dbar - cs * alfa;
			epsln = sn * beta;
			dbar = -cs * beta;
			double const gamma = std::max(std::hypot(gbar, beta), 1e-300);
			cs = gbar / gamma;
			sn = beta / gamma;
			double const phi = cs * phibar;
			phibar = sn * phibar;
			// w_new = (v - oldeps w2 - delta w) / gamma; then (w2, w) <- (w, w_new)
			k_mr_x<<<g, kBlock, 0, st>>>(nb, v, w2, w, oldeps, delta, 1.0 / gamma, phi, wfree, x);
			BFMX_HIP(hipGetLastError());
			double* const older = w2;
			w2 = w;
			w = wfree;
			wfree = older;
			res->iter = it;
			// the M^-1-norm estimate phibar triggers a true FP64 residual check
			if ((phibar <= rtol * beta1 && it >= next_check) || beta == 0 || it == cap) {
				k_mr_resid<<<g, kBlock, 0, st>>>(nb, s->d_slice_ptr, s->d_col, s->d_val, x, b, s->d_partial);
				BFMX_CHECK(mr_sums<1>(st, g, s->d_partial, d_sum, t1));
				res->rr = t1[0];
				res->checks++;
				if (t1[0] <= tol2) {
					res->status = 1;
					return BFMX_OK;
				}
				if (beta == 0) {
					res->status = 3;   // the Krylov space is exhausted above rtol
					return BFMX_OK;
				}
				next_check = it + 10;
			}
		}
		return BFMX_OK;
	}();
	release();
	return rc;
}

}  // namespace

namespace {

// Segmented flexible CG for a system with convergence segments of at least
// kBlock rows each and a multigrid hierarchy of two or more levels (else the
// segments share one alpha and beta): the per-block segment view, per-segment
// scalars and the two-slot partial sums.
int fcg_segments(bfmx_system* s) {
	if (s->fcg_sv.seg || s->n_conv_seg == 0 || s->amg.size() < 2) {
		return BFMX_OK;
	}
	int const m = s->n_conv_seg;
	int64_t const n = s->n_nodes, nblk = grid(n);
	std::vector<int64_t> off(m + 1);
	BFMX_HIP(hipMemcpy(off.data(), s->d_conv_seg, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
	for (int k = 0; k < m; k++) {
		if (off[k + 1] - off[k] < kBlock) {
			return BFMX_OK;
		}
	}
	std::vector<int32_t> seg(2 * nblk), split(nblk);
	int k = 0;
	for (int64_t b = 0; b < nblk; b++) {
		int64_t const r0 = b * kBlock, r1 = std::min(n, r0 + kBlock);
		while (off[k + 1] <= r0) {
			k++;
		}
		seg[2 * b] = k;
		if (off[k + 1] < r1) {   // the next segment starts inside the block (at most one does)
			seg[2 * b + 1] = k + 1;
			split[b] = static_cast<int32_t>(off[k + 1] - r0);
		} else {
			seg[2 * b + 1] = k;
			split[b] = kBlock;
		}
	}
	BFMX_CHECK(dalloc(&s->d_fcg_seg, seg.size()));
	BFMX_CHECK(dalloc(&s->d_fcg_split, split.size()));
	BFMX_CHECK(dalloc(&s->d_fcg_scal, 3 * static_cast<size_t>(m)));
	BFMX_CHECK(dalloc(&s->d_fcg_part, 4 * static_cast<size_t>(nblk)));
	BFMX_HIP(hipMemcpy(s->d_fcg_seg, seg.data(), seg.size() * sizeof(int32_t), hipMemcpyHostToDevice));
	BFMX_HIP(hipMemcpy(s->d_fcg_split, split.data(), split.size() * sizeof(int32_t), hipMemcpyHostToDevice));
	s->fcg_sv.seg = s->d_fcg_seg;
	s->fcg_sv.split = s->d_fcg_split;
	return BFMX_OK;
}

}  // namespace

int solve(bfmx_system* s, const double* b_in, int precond, int precision, double rtol, size_t max_iter,
          bfmx_stats_t* stats) {
	if (precond < BFMX_PRECOND_JACOBI || precond > BFMX_PRECOND_AMG_LEARNED) {
		set_error("unknown preconditioner %d", precond);
		return BFMX_E_UNSUPPORTED;
	}
	if (precision != BFMX_PRECISION_FP64 && precision != BFMX_PRECISION_MIXED) {
		set_error("unknown precision %d", precision);
		return BFMX_E_UNSUPPORTED;
	}
	if (!(rtol > 0)) {
		set_error("rtol must be positive");
		return BFMX_E_ARG;
	}
	bool const amg = precond == BFMX_PRECOND_AMG || precond == BFMX_PRECOND_AMG_LEARNED;
	bool const learned = is_learned(precond);
	if (precond == BFMX_PRECOND_AMG_LEARNED && s->n_conv_seg > 0) {
		set_error("the multigrid-based learned preconditioner does not take systems with convergence segments");
		return BFMX_E_UNSUPPORTED;
	}
	bool const mixed = !is_fcg(precond) && precision == BFMX_PRECISION_MIXED;
	int64_t const n = s->n_nodes;
	int64_t const cap = max_iter ? static_cast<int64_t>(max_iter) : 20 * 2 * n;
	hipStream_t const st = s->stream;
	PhaseTimer phase(st, "solve");
	BFMX_CHECK(ensure_workspace(s));
	phase.mark("workspace");
	if (mixed || amg || learned) {
		// the FP32 copy first: the multigrid build queues kernels that read it
		BFMX_CHECK(ensure_f32(s));
	}
	if (amg) {
		BFMX_CHECK(amg_build(s));   // topology, once per system
		BFMX_CHECK(fcg_segments(s));
		phase.mark("multigrid topology");
	}
	if (learned) {
		BFMX_CHECK(mpnn_prepare(s));
	}
	if (b_in) {
		std::vector<double> bp(2 * n);
		for (int64_t i = 0; i < n; i++) {
			bp[2 * s->perm[i]] = b_in[2 * i];
			bp[2 * s->perm[i] + 1] = b_in[2 * i + 1];
		}
		BFMX_HIP(hipMemcpyAsync(s->d_bsolve, bp.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice, st));
		BFMX_HIP(hipStreamSynchronize(st));
	} else {
		BFMX_HIP(hipMemcpyAsync(s->d_bsolve, s->d_b, 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
	}
	double const rtol2 = rtol * rtol;

	auto const t0 = std::chrono::steady_clock::now();
	if (precond == BFMX_PRECOND_MINRES) {
		MrResult mr{};
		BFMX_CHECK(solve_minres(s, rtol, cap, &mr));
		auto const t1 = std::chrono::steady_clock::now();
		int32_t const status = mr.status == 1   ? BFMX_OK
		                       : mr.status == 2 ? BFMX_E_BREAKDOWN
		                       : mr.status == 3 ? BFMX_E_STAGNATION
		                                        : BFMX_E_MAXITER;
		s->last_iters = mr.iter;
		s->has_u = true;
		double const rel = mr.bb > 0 ? std::sqrt(mr.rr / mr.bb) : 0.0;
		if (stats) {
			*stats = bfmx_stats_t{};
			stats->status = status;
			stats->precond = precond;
			stats->precision = BFMX_PRECISION_FP64;
			stats->restarts = mr.checks;
			stats->iterations = mr.iter;
			stats->n_dof = 2 * n;
			stats->rtol = rtol;
			stats->rel_residual = rel;
			stats->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
			stats->b_norm = std::sqrt(mr.bb);
		}
		if (status != BFMX_OK) {
			set_error("MINRES %s after %lld iterations (relative residual %g, rtol %g)",
			          status == BFMX_E_BREAKDOWN ? "broke down" : status == BFMX_E_STAGNATION ? "stagnated"
			                                                                                  : "did not converge",
			          (long long) mr.iter, rel, rtol);
		}
		return status;
	}
	phase.mark("right-hand side");
	if (amg) {
		BFMX_CHECK(amg_setup_numeric(s));
		phase.mark("multigrid numeric set-up");
	}
	if (learned) {
		BFMX_CHECK(mpnn_setup_numeric(s));
	}
	Scalars last{};
	// flexible CG stops when its residual stops halving for long (a
	// divergent or stalled preconditioner) rather than running to the cap
	BFMX_CHECK(run_chunks(s, precond, mixed, rtol2, cap, false, learned || amg, &last));
	phase.mark("iterations (with captures)");
	int64_t iters = last.iter;
	int32_t fallback = 0;
	int64_t fb_iters = 0;
	if (learned && (last.done == 0 || last.done == 2)) {
		// breakdown, iteration limit or stall of the learned run: block-Jacobi
		// PCG from the current iterate
		BFMX_CHECK(run_chunks(s, BFMX_PRECOND_BLOCK_JACOBI, false, rtol2, cap, true, false, &last));
		fallback = 1;
		fb_iters = last.iter;
		iters += last.iter;
	}
	auto const t1 = std::chrono::steady_clock::now();
	int32_t status = BFMX_OK;
	switch (last.done) {
		case 1: status = BFMX_OK; break;
		case 2: status = BFMX_E_BREAKDOWN; break;
		case 3: status = BFMX_E_STAGNATION; break;
		default: status = BFMX_E_MAXITER; break;
	}
	s->last_iters = iters;
	s->has_u = true;
	double const rel = last.bb > 0 ? std::sqrt(last.rr_true / last.bb) : 0.0;
	if (stats) {
		stats->status = status;
		stats->precond = precond;
		stats->precision = precision;
		stats->restarts = static_cast<int32_t>(last.updates);
		stats->iterations = iters;
		stats->n_dof = 2 * n;
		stats->rtol = rtol;
		stats->rel_residual = rel;
		stats->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
		stats->b_norm = std::sqrt(last.bb);
		stats->fallback = fallback;
		stats->fallback_iterations = static_cast<int32_t>(fb_iters);
	}
	if (status == BFMX_E_BREAKDOWN) {
		set_error("CG breakdown after %lld iterations (p.Ap = %g): matrix not SPD or singular",
		          (long long) last.iter, last.pq);
	} else if (status == BFMX_E_STAGNATION) {
		set_error("CG stagnated: true relative residual %g stays above rtol %g (rounding floor)", rel, rtol);
	} else if (status == BFMX_E_MAXITER) {
		set_error("CG did not reach rtol %g in %lld iterations (relative residual %g)", rtol, (long long) last.iter,
		          last.bb > 0 ? std::sqrt(last.rr / last.bb) : 0.0);
	}
	return status;
}

namespace {

__global__ __launch_bounds__(kBlock) void k_zero(int64_t n, double* __restrict__ p) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (i < n) {
		p[i] = 0;
	}
}

}  // namespace

int pcg_in_buffers(bfmx_system* s, PcgBuffers const& w, int precond, double rtol, int64_t cap, Scalars* out) {
	if (precond != BFMX_PRECOND_JACOBI && precond != BFMX_PRECOND_BLOCK_JACOBI) {
		set_error("this solve runs point or block Jacobi PCG only (precond %d)", precond);
		return BFMX_E_UNSUPPORTED;
	}
	if (s->n_conv_seg > 0) {
		set_error("this solve does not take systems with convergence segments");
		return BFMX_E_UNSUPPORTED;
	}
	int64_t const n = s->n_nodes;
	hipStream_t const st = s->stream;
	// the launch helpers read the system's solver buffers: point them at the
	// caller's for the duration of the call
	struct Saved {
		bfmx_system* s;
		double *x, *r, *z, *q, *p0, *p1, *dinv, *partial, *bsolve;
		int64_t cap;
		Scalars* scal;
		~Saved() {
			s->d_x = x;
			s->d_r = r;
			s->d_z = z;
			s->d_q = q;
			s->d_p[0] = p0;
			s->d_p[1] = p1;
			s->d_dinv = dinv;
			s->d_partial = partial;
			s->d_bsolve = bsolve;
			s->partial_cap = cap;
			s->d_scal = scal;
		}
	} const saved{s, s->d_x, s->d_r, s->d_z, s->d_q, s->d_p[0], s->d_p[1], s->d_dinv, s->d_partial, s->d_bsolve,
	              s->partial_cap, s->d_scal};
	s->d_x = w.x;
	s->d_r = w.r;
	s->d_z = w.z;
	s->d_q = w.q;
	s->d_p[0] = w.p0;
	s->d_p[1] = w.p1;
	s->d_dinv = w.dinv;
	s->d_partial = w.partial;
	s->d_bsolve = w.bsolve;
	s->partial_cap = w.partial_cap;
	s->d_scal = w.scal;
	double const rtol2 = rtol * rtol;
	k_precond<<<grid(n), kBlock, 0, st>>>(static_cast<int>(n), precond, s->d_diag_blk, s->d_val, s->d_dinv);
	for (double* v : {w.x, w.q, w.p0, w.p1}) {
		k_zero<<<grid(2 * n), kBlock, 0, st>>>(2 * n, v);
	}
	static_assert(sizeof(Scalars) % sizeof(double) == 0, "Scalars holds whole doubles");
	k_zero<<<1, kBlock, 0, st>>>(sizeof(Scalars) / sizeof(double), reinterpret_cast<double*>(w.scal));
	launch_residual(s, precond, 0, rtol2, 0.0);
	BFMX_HIP(hipGetLastError());
	Scalars last{};
	BFMX_HIP(hipMemcpyAsync(&last, w.scal, sizeof(Scalars), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipStreamSynchronize(st));
	// chunks of kChunk iterations and a residual check, launched directly (no
	// graph: capturing one allocates), each inspected before the next
	for (int64_t launched = 0; last.done == 0 && launched < cap; launched += kChunk) {
		for (int it = 0; it < kChunk; it++) {
			launch_iteration(s, precond, false, it);
		}
		launch_residual(s, precond, 1, rtol2, 0.0);
		BFMX_HIP(hipGetLastError());
		BFMX_HIP(hipMemcpyAsync(&last, w.scal, sizeof(Scalars), hipMemcpyDeviceToHost, st));
		BFMX_HIP(hipStreamSynchronize(st));
	}
	*out = last;
	return BFMX_OK;
}

int record_pcg(bfmx_system* s, int P, int every, const std::function<int(int)>& rec) {
	int64_t const n = s->n_nodes;
	hipStream_t const st = s->stream;
	BFMX_CHECK(ensure_workspace(s));
	k_precond<<<grid(n), kBlock, 0, st>>>(static_cast<int>(n), BFMX_PRECOND_JACOBI, s->d_diag_blk, s->d_val,
	                                      s->d_dinv);
	BFMX_HIP(hipMemsetAsync(s->d_x, 0, 2 * n * sizeof(double), st));
	BFMX_HIP(hipMemsetAsync(s->d_q, 0, 2 * n * sizeof(double), st));
	BFMX_HIP(hipMemsetAsync(s->d_scal, 0, sizeof(Scalars), st));
	BFMX_HIP(hipMemcpyAsync(s->d_bsolve, s->d_b, 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
	launch_residual(s, BFMX_PRECOND_JACOBI, 0, 0.0, 0.0);
	int it = 0;
	for (int p = 0; p < P; p++) {
		if (p > 0) {
			for (int k = 0; k < every; k++) {
				launch_iteration(s, BFMX_PRECOND_JACOBI, false, it++);
			}
		}
		BFMX_CHECK(rec(p));
	}
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipStreamSynchronize(st));
	return BFMX_OK;
}

void precond_block_jacobi(bfmx_system* s) {
	k_precond<<<grid(s->n_nodes), kBlock, 0, s->stream>>>(static_cast<int>(s->n_nodes), BFMX_PRECOND_BLOCK_JACOBI,
	                                                      s->d_diag_blk, s->d_val, s->d_dinv);
}

int ensure_workspace_public(bfmx_system* s) {
	return ensure_workspace(s);
}

int ensure_f32_public(bfmx_system* s) {
	return ensure_f32(s);
}

int snapshot_f32(bfmx_system* s, float* dst) {
	k_to_f32<<<grid(4 * s->nnzb_pad), kBlock, 0, s->stream>>>(4 * s->nnzb_pad, s->d_val, dst);
	BFMX_HIP(hipGetLastError());
	return BFMX_OK;
}

// Prepares the solver buffers and the multigrid hierarchy as a solve does.
int prepare_amg(bfmx_system* s) {
	BFMX_CHECK(ensure_f32(s));
	BFMX_CHECK(amg_build(s));
	return amg_setup_numeric(s);
}

int time_kernel(bfmx_system* s, int kernel, int reps, double* avg_ms, double* bytes) {
	if (kernel == BFMX_KERNEL_CG_SPMV64 || kernel == BFMX_KERNEL_CG_SPMV32) {
		return time_spmv(s, kernel == BFMX_KERNEL_CG_SPMV64 ? BFMX_PRECISION_FP64 : BFMX_PRECISION_MIXED, reps,
		                 avg_ms, bytes);
	}
	int64_t const n = s->n_nodes;
	hipStream_t const st = s->stream;
	BFMX_CHECK(finalize_conditions(s));
	BFMX_CHECK(ensure_workspace(s));
	BFMX_CHECK(prepare_amg(s));
	Scalars init{};
	BFMX_HIP(hipMemcpyAsync(s->d_scal, &init, sizeof(Scalars), hipMemcpyHostToDevice, st));
	k_fill<<<grid(2 * n), kBlock, 0, st>>>(2 * n, 1.0, s->d_r);
	k_fill<<<grid(2 * n), kBlock, 0, st>>>(2 * n, 0.0, s->d_p[0]);
	k_fill<<<grid(2 * n), kBlock, 0, st>>>(2 * n, 1.0, s->d_q);
	float ms = 0;
	if (kernel == BFMX_KERNEL_FCG_SPMV) {
		int const nb = static_cast<int>(n);
		int const g = grid(nb);
		int nwz = 0;
		float* z = amg_vcycle(s, s->d_r, s->d_q, s->d_partial, s->d_scal, &nwz);   // an FP32 z
		auto launch = [&] {
			k_fcg_spmv<float><<<g, kBlock, 0, st>>>(nb, s->d_slice_ptr, s->d_col, s->d_val, z, s->d_p[0], s->d_p[1],
			                                        s->d_q, s->d_r, s->d_partial, s->d_scal);
		};
		for (int i = 0; i < 3; i++) {
			launch();
		}
		BFMX_HIP(hipGetLastError());
		BFMX_HIP(hipEventRecord(s->ev[4], st));
		for (int i = 0; i < reps; i++) {
			launch();
		}
		BFMX_HIP(hipEventRecord(s->ev[5], st));
		if (bytes) {
			// matrix blocks + slice pointers, gathered z (FP32) and p_old, read r,
			// written p_new and q
			double const nd = 2.0 * static_cast<double>(n);
			*bytes = static_cast<double>(s->nnzb) * 36 + 4.0 * (s->n_slices + 1) + nd * (4 + 8 + 8 + 8 + 8);
		}
	} else {
		BFMX_CHECK(amg_time_smoother(s, reps, s->ev[4], s->ev[5], bytes));
	}
	BFMX_HIP(hipEventSynchronize(s->ev[5]));
	BFMX_HIP(hipEventElapsedTime(&ms, s->ev[4], s->ev[5]));
	*avg_ms = static_cast<double>(ms) / reps;
	return BFMX_OK;
}

int time_spmv(bfmx_system* s, int precision, int reps, double* avg_ms, double* bytes) {
	bool const mixed = precision == BFMX_PRECISION_MIXED;
	int64_t const n = s->n_nodes;
	hipStream_t const st = s->stream;
	BFMX_CHECK(finalize_conditions(s));
	BFMX_CHECK(ensure_workspace(s));
	if (mixed) {
		BFMX_CHECK(ensure_f32(s));
	}
	// z = 1, p = 0, identical for every repetition. Uses the solver's z/p/q
	// buffers; x (the last solution) is untouched.
	Scalars init{};
	init.rz = 1.0;
	BFMX_HIP(hipMemcpyAsync(s->d_scal, &init, sizeof(Scalars), hipMemcpyHostToDevice, st));
	k_fill<<<grid(2 * n), kBlock, 0, st>>>(2 * n, 1.0, s->d_z);
	k_fill<<<grid(2 * n), kBlock, 0, st>>>(2 * n, 0.0, s->d_p[0]);
	for (int i = 0; i < 3; i++) {
		launch_spmv(s, mixed, s->d_p[0], s->d_p[1]);
	}
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipEventRecord(s->ev[4], st));
	for (int i = 0; i < reps; i++) {
		launch_spmv(s, mixed, s->d_p[0], s->d_p[1]);
	}
	BFMX_HIP(hipEventRecord(s->ev[5], st));
	BFMX_HIP(hipEventSynchronize(s->ev[5]));
	float ms = 0;
	BFMX_HIP(hipEventElapsedTime(&ms, s->ev[4], s->ev[5]));
	*avg_ms = static_cast<double>(ms) / reps;
	if (bytes) {
		double const nd = 2.0 * static_cast<double>(n);
		double const vb = mixed ? 4.0 : 8.0;
		// matrix blocks (4 values + column index) + slice pointers, the gathered
		// z and p_old (read once each), the written p_new and q
		*bytes = static_cast<double>(s->nnzb) * (4 * vb + 4) + 4.0 * (s->n_slices + 1) + 4 * nd * 8.0;
	}
	return BFMX_OK;
}

}  // namespace bfmx
