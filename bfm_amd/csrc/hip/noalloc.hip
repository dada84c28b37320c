// The solve that allocates nothing (bfmx_system_solve_noalloc): point or
// block Jacobi PCG in FP64 whose solver buffers (the vectors, the inverse
// diagonal, the partial sums, the scalars, the node permutation and the
// staging of the right-hand side and the solution) are carved out of one
// fixed device array, `scratch`, instead of hipMalloc'd workspace. Each GPU
// holds its own copy of the array until the process ends; the host reaches
// it through the address hipGetSymbolAddress returns on that GPU, looked up
// once per device at its first call. Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU race.

#include <algorithm>
#include <cmath>
#include <mutex>

#include "internal.h"

namespace bfmx {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // per GPU
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_lock;
unsigned char* g_scratch[kMaxDevices] = {};   // per device: the address of its copy of scratch

// The scratch array of the current device (which must be dev).
int scratch_of(int dev, unsigned char** out) {
	if (dev < 0 || dev >= kMaxDevices) {
		set_error("device %d outside the %d devices the no-allocation solve tracks", dev, kMaxDevices);
		return BFMX_E_ARG;
	}
	std::lock_guard<std::mutex> lock(g_scratch_lock);
	if (!g_scratch[dev]) {
		void* p = nullptr;
		hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) {
			set_error("hipGetSymbolAddress of the solve scratch on device %d: %s", dev, hipGetErrorString(e));
			return BFMX_E_HIP;
		}
		g_scratch[dev] = static_cast<unsigned char*>(p);
	}
	*out = g_scratch[dev];
	return BFMX_OK;
}

size_t up(size_t b) {
	return (b + kAlign - 1) / kAlign * kAlign;
}

// Byte offsets of the buffers of an n-node system within the array.
struct Layout {
	size_t x, r, z, q, p0, p1, dinv, bsolve, partial, scal, perm, stage, end;
	int64_t partial_cap;
	explicit Layout(int64_t n) {
		size_t const v = up(2 * static_cast<size_t>(n) * sizeof(double));
		partial_cap = std::max<int64_t>(3 * ((n + kBlock - 1) / kBlock), (n + 63) / 64);
		size_t o = 0;
		for (size_t* f : {&x, &r, &z, &q, &p0, &p1}) {
			*f = o;
			o += v;
		}
		dinv = o;
		o += up(4 * static_cast<size_t>(n) * sizeof(double));
		bsolve = o;
		o += v;
		partial = o;
		o += up(static_cast<size_t>(partial_cap) * sizeof(double));
		scal = o;
		o += up(sizeof(Scalars));
		perm = o;
		o += up(static_cast<size_t>(n) * sizeof(int32_t));
		stage = o;   // the right-hand side in, then the solution out (original order)
		o += v;
		end = o;
	}
};

// bsolve[2 perm[i] + c] = stage[2 i + c] (perm[old] = new).
__global__ __launch_bounds__(kBlock) void k_perm_in(int64_t n, const int32_t* __restrict__ perm,
                                                    const double* __restrict__ stage, double* __restrict__ bsolve) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (i < n) {
		int64_t const k = perm[i];
		bsolve[2 * k] = stage[2 * i];
		bsolve[2 * k + 1] = stage[2 * i + 1];
	}
}

// stage[2 i + c] = x[2 perm[i] + c].
__global__ __launch_bounds__(kBlock) void k_perm_out(int64_t n, const int32_t* __restrict__ perm,
                                                     const double* __restrict__ x, double* __restrict__ stage) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (i < n) {
		int64_t const k = perm[i];
		stage[2 * i] = x[2 * k];
		stage[2 * i + 1] = x[2 * k + 1];
	}
}

}  // namespace

int64_t noalloc_max_nodes() {
	int64_t lo = 0, hi = int64_t(1) << 31;
	while (lo < hi) {
		int64_t const mid = (lo + hi + 1) / 2;
		if (Layout(mid).end <= kScratchBytes) {
			lo = mid;
		} else {
			hi = mid - 1;
		}
	}
	return lo;
}

int solve_noalloc(bfmx_system* s, const double* b_in, double* u_out, int precond, double rtol, size_t max_iter,
                  bfmx_stats_t* stats) {
	if (!(rtol > 0)) {
		set_error("rtol must be positive");
		return BFMX_E_ARG;
	}
	int64_t const n = s->n_nodes;
	Layout const L(n);
	if (L.end > kScratchBytes) {
		set_error("system of %lld nodes needs %zu bytes of solve scratch; the no-allocation solve holds %zu per GPU "
		          "(at most %lld nodes)",
		          (long long) n, L.end, kScratchBytes, (long long) noalloc_max_nodes());
		return BFMX_E_ARG;
	}
	unsigned char* base = nullptr;
	BFMX_CHECK(scratch_of(s->device, &base));
	BFMX_CHECK(finalize_conditions(s));
	auto at = [&](size_t off) { return reinterpret_cast<double*>(base + off); };
	PcgBuffers w{at(L.x), at(L.r), at(L.z), at(L.q), at(L.p0), at(L.p1), at(L.dinv), at(L.bsolve), at(L.partial),
	             L.partial_cap, reinterpret_cast<Scalars*>(base + L.scal)};
	int32_t* perm = reinterpret_cast<int32_t*>(base + L.perm);
	double* stage = at(L.stage);
	hipStream_t const st = s->stream;
	int const g = static_cast<int>((n + kBlock - 1) / kBlock);
	BFMX_HIP(hipMemcpyAsync(perm, s->perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
	if (b_in) {
		BFMX_HIP(hipMemcpyAsync(stage, b_in, 2 * n * sizeof(double), hipMemcpyHostToDevice, st));
		k_perm_in<<<g, kBlock, 0, st>>>(n, perm, stage, w.bsolve);
	} else {
		BFMX_HIP(hipMemcpyAsync(w.bsolve, s->d_b, 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
	}
	BFMX_HIP(hipGetLastError());
	int64_t const cap = max_iter ? static_cast<int64_t>(max_iter) : 20 * 2 * n;
	Scalars last{};
	BFMX_CHECK(pcg_in_buffers(s, w, precond, rtol, cap, &last));
	int32_t const status = last.done == 1   ? BFMX_OK
	                       : last.done == 2 ? BFMX_E_BREAKDOWN
	                       : last.done == 3 ? BFMX_E_STAGNATION
	                                        : BFMX_E_MAXITER;
	if (u_out && (status == BFMX_OK || status == BFMX_E_MAXITER)) {
		k_perm_out<<<g, kBlock, 0, st>>>(n, perm, w.x, stage);
		BFMX_HIP(hipGetLastError());
		BFMX_HIP(hipMemcpyAsync(u_out, stage, 2 * n * sizeof(double), hipMemcpyDeviceToHost, st));
		BFMX_HIP(hipStreamSynchronize(st));
	}
	double const rel = last.bb > 0 ? std::sqrt(last.rr_true / last.bb) : 0.0;
	if (stats) {
		*stats = bfmx_stats_t{};
		stats->status = status;
		stats->precond = precond;
		stats->precision = BFMX_PRECISION_FP64;
		stats->restarts = static_cast<int32_t>(last.updates);
		stats->iterations = last.iter;
		stats->n_dof = 2 * n;
		stats->rtol = rtol;
		stats->rel_residual = rel;
		stats->b_norm = std::sqrt(last.bb);
	}
	if (status == BFMX_E_BREAKDOWN) {
		set_error("CG breakdown after %lld iterations: matrix not SPD or singular", (long long) last.iter);
	} else if (status == BFMX_E_STAGNATION) {
		set_error("CG stagnated: true relative residual %g stays above rtol %g (rounding floor)", rel, rtol);
	} else if (status == BFMX_E_MAXITER) {
		set_error("CG did not reach rtol %g in %lld iterations", rtol, (long long) last.iter);
	}
	return status;
}

}  // namespace bfmx
