// Multigrid hierarchy topology built on the GPU, one level at a time, from
// the level's graph in sliced ELL (the same arrays the solver reads).
//
//   * aggregation: a distance-2 maximal independent set of roots (MIS-2,
//     Bell, Dalton and Olson, SIAM J. Sci. Comput. 34 (2012) C123-C152): in
//     rounds, every eligible undecided row whose (state, eligibility, front
//     score, hashed priority, index) tuple is the largest within two hops
//     becomes a root, and every undecided row with a root within two hops
//     drops out; roots grow as fronts from sparse seeds (k_mis_front), which
//     packs the aggregates nearly as densely as a sequential sweep (about 10
//     rows each on a quad mesh instead of 13 with plain hashed priorities);
//     each root's aggregate is the root
//     and its neighbours (roots are at least three hops apart, so these do not
//     overlap), and every other row joins the aggregate of its first
//     neighbour (in column order) that has one; aggregates are numbered in
//     the order of their roots;
//   * aggregate members, sorted; the prolongator pattern (row i couples to
//     the aggregates of its neighbours, sorted) and the entry of each block
//     slot's column aggregate within its row; P's columns (restriction
//     lists: the rows next to the aggregate, sorted, with their entries);
//   * the Galerkin product's symbolic phase in two steps, A P then P^T (A P)
//     (the numeric kernels in amg.hip follow the same two steps): the
//     pattern of A P (row i: the prolongator columns of i's neighbours) and
//     the coarse pattern (row I: the A P columns of P's column I), its sliced
//     ELL, and each coarse slot's mirror (the slot of its transpose).
// Every result is independent of scheduling: counts use integer atomics,
// lists filled through atomics are sorted afterwards, and the sets built by
// a wavefront's lanes together (LDS hash tables, compare-and-swap inserts)
// are sorted before they are written.
//
// Bounds: a fine row has at most kMaxRow neighbours, a row of A P at most
// kMaxAP entries, a coarse row at most kMaxCoarse columns and a prolongator
// column at most kSetCap rows; a level that exceeds them is reported
// (BFMX_E_UNSUPPORTED) and the caller builds the hierarchy on the host
// instead.

#include <algorithm>
#include <cstring>
#include <vector>

#include "amg_topo.h"

namespace bfmx {

namespace {

constexpr int kT = 256;          // threads of the streaming set-up kernels
constexpr int kTW = 64;          // threads of the kernels with per-thread LDS tables
constexpr int kHash = 256;       // hash-table entries per coarse row (LDS)
constexpr int kHashStride = kHash + 1;
constexpr int kMisMaxRounds = 200;
constexpr int kMisGroup = 4;     // rounds launched between two checks for undecided rows
constexpr int kStash = 32;       // entries per row a count pass keeps for its fill pass
// Per-thread LDS tables sized for the common rows: prolongator rows of at
// most kSmallRow neighbours and rows of A P of at most kStash entries go
// through launches whose small tables let many wavefronts share a CU (the
// full-size tables, kMaxRow + 1 or kMaxAP + 1 entries per thread, allow four
// per CU); the other rows are listed and taken by kBigBlocks workgroups with
// the full tables.
constexpr int kSmallRow = 32;
constexpr int kBigBlocks = 128;
constexpr int kScanItems = 8;    // items per thread of a scan block
constexpr int kScanBlock = kT * kScanItems;

__host__ __device__ inline int nblk(int64_t n, int t = kT) {
	return static_cast<int>((n + t - 1) / t);
}

__device__ __forceinline__ int32_t nbr_slot(const int32_t* __restrict__ slice_ptr, int i, int k) {
	return slice_ptr[i >> 6] + k * kSlice + (i & 63);
}

// ------------------------------------------------------------------ scans

// Exclusive prefix sums of in[0 .. n) into out[0 .. n] (out[n] = total), in
// three passes: per-block sums, one workgroup over the block sums, and the
// block offsets added; one pass when a single block holds all n (bsum null:
// the block writes the total to out[n]).
template <typename T, typename U>
__global__ __launch_bounds__(kT) void k_scan_block(int64_t n, const T* __restrict__ in, U* __restrict__ out,
                                                   U* __restrict__ bsum) {
	__shared__ U wsum[kT / 64];
	int64_t const base = static_cast<int64_t>(blockIdx.x) * kScanBlock + threadIdx.x * kScanItems;
	U v[kScanItems];
	U s = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		v[k] = base + k < n ? static_cast<U>(in[base + k]) : U(0);
		s += v[k];
	}
	// inclusive scan of the thread sums within the wavefront
	U inc = s;
	int const lane = threadIdx.x & 63;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1) {
		U const o = __shfl_up(inc, off, 64);
		if (lane >= off) {
			inc += o;
		}
	}
	if (lane == 63) {
		wsum[threadIdx.x >> 6] = inc;
	}
	__syncthreads();
	U wbase = 0;
	for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); w++) {
		wbase += wsum[w];
	}
	U run = wbase + inc - s;
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		if (base + k < n) {
			out[base + k] = run;
		}
		run += v[k];
	}
	if (threadIdx.x == kT - 1) {
		if (bsum) {
			bsum[blockIdx.x] = wbase + inc;
		} else {
			out[n] = wbase + inc;
		}
	}
}

// Exclusive scan of bsum[0 .. nb) in place, bsum[nb] = total: one workgroup,
// kT sums at a time with a carry.
template <typename U>
__global__ __launch_bounds__(kT) void k_scan_top(int nb, U* __restrict__ bsum) {
	__shared__ U wsum[kT / 64];
	__shared__ U carry;
	if (threadIdx.x == 0) {
		carry = 0;
	}
	__syncthreads();
	int const lane = threadIdx.x & 63;
	for (int c0 = 0; c0 < nb; c0 += kT) {
		int const i = c0 + static_cast<int>(threadIdx.x);
		U const s = i < nb ? bsum[i] : U(0);
		U inc = s;
#pragma unroll
		for (int off = 1; off < 64; off <<= 1) {
			U const o = __shfl_up(inc, off, 64);
			if (lane >= off) {
				inc += o;
			}
		}
		if (lane == 63) {
			wsum[threadIdx.x >> 6] = inc;
		}
		__syncthreads();
		U wbase = carry;
		for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); w++) {
			wbase += wsum[w];
		}
		if (i < nb) {
			bsum[i] = wbase + inc - s;
		}
		__syncthreads();
		if (threadIdx.x == kT - 1) {
			carry = wbase + inc;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		bsum[nb] = carry;
	}
}

template <typename U>
__global__ __launch_bounds__(kT) void k_scan_add(int64_t n, U* __restrict__ out, const U* __restrict__ bsum) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
	if (i < n) {
		out[i] += bsum[i / kScanBlock];
	}
	if (i == 0) {
		out[n] = bsum[nblk(n, kScanBlock)];
	}
}

template <typename T>
__global__ __launch_bounds__(kT) void k_fill(int64_t n, T* __restrict__ p, T v) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
	if (i < n) {
		p[i] = v;
	}
}

// ------------------------------------------------------------------ MIS-2

__device__ __forceinline__ uint32_t mis_prio(uint32_t i, uint32_t salt) {
	// a 32-bit integer hash (the "lowbias32" finaliser), 30 bits kept
	uint32_t x = i * 0x9E3779B9u + salt;
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x >> 2;
}

// Rows that drop out (state 0) next to a root's neighbourhood steer where the
// next roots go: an undecided row next to an out row is on the front of the
// aggregated region, and its score (out neighbours of its neighbours) is
// highest in the region's concave corners. Roots are taken on the fronts
// first, so aggregates pack against each other as in a sequential sweep;
// elsewhere a row may start a region only when its priority falls below a
// seeding threshold that grows each round (every row is eligible after a
// few dozen rounds, which ends the rounds).

// s1[i] = out neighbours of row i (at most 255).
__global__ __launch_bounds__(kT) void k_mis_front(int n, const int32_t* __restrict__ slice_ptr,
                                                  const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ row_len,
                                                  const uint8_t* __restrict__ state, uint8_t* __restrict__ s1) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i >= n) {
		return;
	}
	int const len = row_len[i];
	int c = 0;
	for (int k = 0; k < len; k++) {
		c += state[col[nbr_slot(slice_ptr, i, k)]] == 0 ? 1 : 0;
	}
	s1[i] = static_cast<uint8_t>(min(c, 255));
}

// Tuple of row i: state (2 bits), eligible (1), score (6), priority (23
// bits), index (32); out rows 0. Eligible: on a front (s1 > 0) or priority
// below the seeding threshold.
__global__ __launch_bounds__(kT) void k_mis_tuples(int n, const int32_t* __restrict__ slice_ptr,
                                                   const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ row_len,
                                                   const uint8_t* __restrict__ state, const uint8_t* __restrict__ s1,
                                                   uint32_t salt, uint32_t seed_thr, uint64_t* __restrict__ t) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i >= n) {
		return;
	}
	uint8_t const st = state[i];
	if (st == 0) {
		t[i] = 0;
		return;
	}
	int const len = row_len[i];
	int score = 0;
	for (int k = 0; k < len; k++) {
		score += s1[col[nbr_slot(slice_ptr, i, k)]];
	}
	uint32_t const prio = mis_prio(static_cast<uint32_t>(i), salt);
	uint64_t const elig = (s1[i] > 0 || prio < seed_thr) ? 1 : 0;
	t[i] = (static_cast<uint64_t>(st) << 62) | (elig << 61) | (static_cast<uint64_t>(min(score, 63)) << 55) |
	       (static_cast<uint64_t>(prio >> 7) << 32) | static_cast<uint32_t>(i);
}

// tout[i] = max of tin over row i (its neighbours and itself).
__global__ __launch_bounds__(kT) void k_mis_max(int n, const int32_t* __restrict__ slice_ptr,
                                                const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                                const uint64_t* __restrict__ tin, uint64_t* __restrict__ tout) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i >= n) {
		return;
	}
	int const len = row_len[i];
	uint64_t m = tin[i];
	for (int k = 0; k < len; k++) {
		uint64_t const v = tin[col[nbr_slot(slice_ptr, i, k)]];
		m = v > m ? v : m;
	}
	tout[i] = m;
}

// An eligible undecided row whose tuple is the largest within two hops
// becomes a root; an undecided row that sees a root within two hops drops
// out. *left is set while rows stay undecided.
__global__ __launch_bounds__(kT) void k_mis_update(int n, uint8_t* __restrict__ state,
                                                   const uint64_t* __restrict__ own,
                                                   const uint64_t* __restrict__ t2, int32_t* __restrict__ left) {
	int const i = blockIdx.x * kT + threadIdx.x;
	int still = 0;
	if (i < n && state[i] == 1) {
		uint64_t const m = t2[i], o = own[i];
		if (m == o && ((o >> 61) & 1)) {
			state[i] = 2;
		} else if ((m >> 62) == 2) {
			state[i] = 0;
		} else {
			still = 1;
		}
	}
	if (__any(still) && (threadIdx.x & 63) == 0) {
		*left = 1;
	}
}

__global__ __launch_bounds__(kT) void k_is_root(int n, const uint8_t* __restrict__ state, int32_t* __restrict__ f) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i < n) {
		f[i] = state[i] == 2 ? 1 : 0;
	}
}

// agg1: a root's own aggregate, and for its neighbours the root's (a row
// has at most one root neighbour); -1 for the rest.
__global__ __launch_bounds__(kT) void k_attach_roots(int n, const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ row_len,
                                                     const uint8_t* __restrict__ state,
                                                     const int32_t* __restrict__ rootid, int32_t* __restrict__ agg1) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i >= n) {
		return;
	}
	int32_t a = -1;
	if (state[i] == 2) {
		a = rootid[i];
	} else {
		int const len = row_len[i];
		for (int k = 0; k < len; k++) {
			int const j = col[nbr_slot(slice_ptr, i, k)];
			if (state[j] == 2) {
				a = rootid[j];
				break;
			}
		}
	}
	agg1[i] = a;
}

// Every other row joins the aggregate of its first neighbour that has one.
__global__ __launch_bounds__(kT) void k_attach_rest(int n, const int32_t* __restrict__ slice_ptr,
                                                    const int32_t* __restrict__ col,
                                                    const int32_t* __restrict__ row_len,
                                                    const int32_t* __restrict__ agg1, int32_t* __restrict__ agg,
                                                    int32_t* __restrict__ bad) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i >= n) {
		return;
	}
	int32_t a = agg1[i];
	if (a < 0) {
		int const len = row_len[i];
		for (int k = 0; k < len; k++) {
			int32_t const b = agg1[col[nbr_slot(slice_ptr, i, k)]];
			if (b >= 0) {
				a = b;
				break;
			}
		}
		if (a < 0) {
			atomicOr(bad, 1);   // not reachable for a maximal MIS-2
		}
	}
	agg[i] = a;
}

__global__ __launch_bounds__(kT) void k_count_keys(int n, const int32_t* __restrict__ key, int32_t* __restrict__ cnt) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i < n) {
		atomicAdd(&cnt[key[i]], 1);
	}
}

__global__ __launch_bounds__(kT) void k_place_members(int n, const int32_t* __restrict__ agg,
                                                      const int32_t* __restrict__ aptr, int32_t* __restrict__ fill,
                                                      int32_t* __restrict__ anodes) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i < n) {
		int const a = agg[i];
		anodes[aptr[a] + atomicAdd(&fill[a], 1)] = i;
	}
}

// Sorts each list [ptr[k], ptr[k + 1]) of key (ascending), carrying val
// along when given (insertion sort: the lists are short).
__global__ __launch_bounds__(kT) void k_sort_lists(int nl, const int32_t* __restrict__ ptr, int32_t* __restrict__ key,
                                                   int32_t* __restrict__ val) {
	int const l = blockIdx.x * kT + threadIdx.x;
	if (l >= nl) {
		return;
	}
	int32_t const b = ptr[l], e = ptr[l + 1];
	for (int32_t x = b + 1; x < e; x++) {
		int32_t const k = key[x];
		int32_t const v = val ? val[x] : 0;
		int32_t y = x - 1;
		while (y >= b && key[y] > k) {
			key[y + 1] = key[y];
			if (val) {
				val[y + 1] = val[y];
			}
			y--;
		}
		key[y + 1] = k;
		if (val) {
			val[y + 1] = v;
		}
	}
}

// ------------------------------------------------------------------ prolongator pattern

// Sorted distinct aggregates of row i's neighbours into u (LDS); returns
// their count (the own aggregate only when not smoothed).
__device__ __forceinline__ int row_aggs(int i, const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ col,
                                        const int32_t* __restrict__ row_len, const int32_t* __restrict__ agg,
                                        bool smooth, int32_t* u) {
	if (!smooth) {
		u[0] = agg[i];
		return 1;
	}
	int const len = row_len[i];
	int cnt = 0;
	for (int k = 0; k < len; k++) {
		int32_t const a = agg[col[nbr_slot(slice_ptr, i, k)]];
		int y = cnt - 1;
		while (y >= 0 && u[y] > a) {
			u[y + 1] = u[y];
			y--;
		}
		if (y >= 0 && u[y] == a) {   // already present: undo the shift
			for (int z = y + 1; z < cnt; z++) {
				u[z] = u[z + 1];
			}
			continue;
		}
		u[y + 1] = a;
		cnt++;
	}
	return cnt;
}

// Row i's count of prolongator entries (u: a table of CAP + 1 entries).
__device__ __forceinline__ void p_count_row(int i, const int32_t* __restrict__ slice_ptr,
                                            const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                            const int32_t* __restrict__ agg, bool smooth, int32_t* u,
                                            int32_t* __restrict__ pcnt, int32_t* __restrict__ bad) {
	int const c = row_aggs(i, slice_ptr, col, row_len, agg, smooth, u);
	if (c > 255) {
		atomicOr(bad, 2);
	}
	pcnt[i] = c;
}

// Rows of at most kSmallRow neighbours; the others are listed in big
// (*nbig of them) for k_p_count_big.
__global__ __launch_bounds__(kTW) void k_p_count(int n, const int32_t* __restrict__ slice_ptr,
                                                 const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                                 const int32_t* __restrict__ agg, int smooth,
                                                 int32_t* __restrict__ pcnt, int32_t* __restrict__ bad,
                                                 int32_t* __restrict__ big, int32_t* __restrict__ nbig) {
	__shared__ int32_t lds[kTW * (kSmallRow + 1)];
	int const i = blockIdx.x * kTW + threadIdx.x;
	if (i >= n) {
		return;
	}
	if (row_len[i] > kSmallRow) {
		big[atomicAdd(nbig, 1)] = i;
		return;
	}
	p_count_row(i, slice_ptr, col, row_len, agg, smooth != 0, lds + threadIdx.x * (kSmallRow + 1), pcnt, bad);
}

// The listed rows with full-size tables (a row above kMaxRow is flagged).
__global__ __launch_bounds__(kTW) void k_p_count_big(const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ row_len,
                                                     const int32_t* __restrict__ agg, int smooth,
                                                     int32_t* __restrict__ pcnt, int32_t* __restrict__ bad,
                                                     const int32_t* __restrict__ big, const int32_t* __restrict__ nbig) {
	__shared__ int32_t lds[kTW * (kMaxRow + 1)];
	int const m = *nbig;
	for (int t = blockIdx.x * kTW + threadIdx.x; t < m; t += gridDim.x * kTW) {
		int const i = big[t];
		if (row_len[i] > kMaxRow) {
			atomicOr(bad, 1);
			pcnt[i] = 0;
			continue;
		}
		p_count_row(i, slice_ptr, col, row_len, agg, smooth != 0, lds + threadIdx.x * (kMaxRow + 1), pcnt, bad);
	}
}

// P's columns of row i, and for every block slot of the row (padding
// included) the entry of its column's aggregate within the row (padding
// slots and the unsmoothed case: the row's own aggregate).
__device__ __forceinline__ void p_fill_row(int i, const int32_t* __restrict__ slice_ptr,
                                           const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                           const int32_t* __restrict__ agg, bool smooth, int32_t* u,
                                           const int32_t* __restrict__ pptr, int32_t* __restrict__ pcol,
                                           uint8_t* __restrict__ ppos) {
	int const c = row_aggs(i, slice_ptr, col, row_len, agg, smooth, u);
	int32_t const p0 = pptr[i];
	for (int k = 0; k < c; k++) {
		pcol[p0 + k] = u[k];
	}
	auto pos = [&](int32_t a) {
		int lo = 0, hi = c - 1;
		while (lo < hi) {
			int const mid = (lo + hi) >> 1;
			if (u[mid] < a) {
				lo = mid + 1;
			} else {
				hi = mid;
			}
		}
		return lo;
	};
	uint8_t const own = static_cast<uint8_t>(pos(agg[i]));
	int const len = row_len[i];
	int32_t const s = i >> 6;
	int const width = (slice_ptr[s + 1] - slice_ptr[s]) / kSlice;
	for (int k = 0; k < width; k++) {
		int32_t const sl = nbr_slot(slice_ptr, i, k);
		ppos[sl] = smooth && k < len ? static_cast<uint8_t>(pos(agg[col[sl]])) : own;
	}
}

// Rows of at most kSmallRow neighbours (the others: k_p_fill_big over the
// count pass's list).
__global__ __launch_bounds__(kTW) void k_p_fill(int n, const int32_t* __restrict__ slice_ptr,
                                                const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                                const int32_t* __restrict__ agg, int smooth,
                                                const int32_t* __restrict__ pptr, int32_t* __restrict__ pcol,
                                                uint8_t* __restrict__ ppos) {
	__shared__ int32_t lds[kTW * (kSmallRow + 1)];
	int const i = blockIdx.x * kTW + threadIdx.x;
	if (i >= n || row_len[i] > kSmallRow) {
		return;
	}
	p_fill_row(i, slice_ptr, col, row_len, agg, smooth != 0, lds + threadIdx.x * (kSmallRow + 1), pptr, pcol, ppos);
}

__global__ __launch_bounds__(kTW) void k_p_fill_big(const int32_t* __restrict__ slice_ptr,
                                                    const int32_t* __restrict__ col,
                                                    const int32_t* __restrict__ row_len,
                                                    const int32_t* __restrict__ agg, int smooth,
                                                    const int32_t* __restrict__ pptr, int32_t* __restrict__ pcol,
                                                    uint8_t* __restrict__ ppos, const int32_t* __restrict__ big,
                                                    const int32_t* __restrict__ nbig) {
	__shared__ int32_t lds[kTW * (kMaxRow + 1)];
	int const m = *nbig;
	for (int t = blockIdx.x * kTW + threadIdx.x; t < m; t += gridDim.x * kTW) {
		int const i = big[t];
		if (row_len[i] <= kMaxRow) {
			p_fill_row(i, slice_ptr, col, row_len, agg, smooth != 0, lds + threadIdx.x * (kMaxRow + 1), pptr, pcol,
			           ppos);
		}
	}
}

// ------------------------------------------------------------------ wavefront sets

// A set of distinct non-negative int32 values built by the 64 lanes of one
// workgroup (one wavefront) in an LDS hash table: inserts are atomic
// compare-and-swaps, so the set does not depend on the lanes' order; its
// sorted contents are produced afterwards. At most kSetCap values are kept:
// an insert that finds the set full flags it (the table, twice that size
// plus a wavefront, never fills).
constexpr int kSetBits = 9;
constexpr int kSet = 1 << kSetBits;   // table entries
constexpr int kSetCap = kMaxPCol;     // values kept
static_assert(kSetCap + 64 < kSet, "the set's table must never fill");

struct WaveSet {
	int32_t* h;      // kSet entries
	int32_t* cnt;    // values inserted
	int32_t* over;   // set when an insert found kSetCap values

	__device__ void clear() {
		for (int k = threadIdx.x; k < kSet; k += 64) {
			h[k] = -1;
		}
		if (threadIdx.x == 0) {
			*cnt = 0;
			*over = 0;
		}
		__syncthreads();
	}
	// false once the set is full (over is set whenever more than kSetCap
	// values went in)
	__device__ bool insert(int32_t v) {
		if (*static_cast<volatile int32_t*>(cnt) >= kSetCap) {
			*over = 1;
			return false;
		}
		uint32_t x = (static_cast<uint32_t>(v) * 2654435761u) >> (32 - kSetBits);
		for (;;) {
			int32_t const old = atomicCAS(&h[x], -1, v);
			if (old == -1) {
				if (atomicAdd(cnt, 1) >= kSetCap) {
					*over = 1;
				}
				return true;
			}
			if (old == v) {
				return true;
			}
			x = (x + 1) & (kSet - 1);
		}
	}
	// The values, sorted ascending, into out[0 .. size) (LDS or global);
	// returns the size. Call after a __syncthreads() that follows the inserts.
	// scratch: kSetCap + 64 entries of LDS.
	__device__ int sorted(int32_t* scratch, int32_t* out) {
		constexpr int per = kSet / 64;
		int32_t v[per];
		int mine = 0;
#pragma unroll
		for (int s = 0; s < per; s++) {
			int32_t const x = h[threadIdx.x * per + s];
			if (x != -1) {
				v[mine++] = x;
			}
		}
		// exclusive prefix of the lanes' counts
		int inc = mine;
#pragma unroll
		for (int off = 1; off < 64; off <<= 1) {
			int const o = __shfl_up(inc, off, 64);
			if (static_cast<int>(threadIdx.x) >= off) {
				inc += o;
			}
		}
		int const total = __shfl(inc, 63, 64);
		int const base = inc - mine;
		for (int s = 0; s < mine; s++) {
			scratch[base + s] = v[s];
		}
		__syncthreads();
		for (int e = threadIdx.x; e < total; e += 64) {
			int32_t const x = scratch[e];
			int rank = 0;
			for (int y = 0; y < total; y++) {
				rank += scratch[y] < x ? 1 : 0;
			}
			out[rank] = x;
		}
		__syncthreads();
		return total;
	}
};

#define WAVE_SET(name)                          \
	__shared__ int32_t name##_h[kSet];          \
	__shared__ int32_t name##_c[2];             \
	WaveSet name{name##_h, name##_c, name##_c + 1}

// The largest v over the wavefront's lanes.
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v = max(v, __shfl_xor(v, off, 64));
	}
	return v;
}

// Position of J in the sorted list a[0 .. len) (J must be present).
__device__ __forceinline__ int find_sorted(const int32_t* a, int len, int32_t J) {
	int lo = 0, hi = len - 1;
	while (lo < hi) {
		int const mid = (lo + hi) >> 1;
		if (a[mid] < J) {
			lo = mid + 1;
		} else {
			hi = mid;
		}
	}
	return lo;
}

// ------------------------------------------------------------------ restriction lists

// P's column J (one workgroup of one wavefront per aggregate): the rows with
// a neighbour in aggregate J (smoothed; unsmoothed, its members), sorted:
// the rows whose prolongator row holds J. COUNT: their number into rcnt[J];
// else the rows into rrow and their P entries (the position of J in the
// row's sorted columns) into rent, from rptr[J].
template <bool COUNT>
__global__ __launch_bounds__(64) void k_r_lists(int na, int smooth, const int32_t* __restrict__ slice_ptr,
                                                const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                                const int32_t* __restrict__ agg_ptr,
                                                const int32_t* __restrict__ agg_nodes,
                                                const int32_t* __restrict__ pptr, const int32_t* __restrict__ pcol,
                                                int32_t* __restrict__ rcnt, const int32_t* __restrict__ rptr,
                                                int32_t* __restrict__ rrow, int32_t* __restrict__ rent,
                                                int32_t* __restrict__ bad) {
	WAVE_SET(set);
	__shared__ int32_t scratch[kSetCap + 64];
	__shared__ int32_t rows[kSetCap];
	int const J = blockIdx.x;
	set.clear();
	int32_t const a0 = agg_ptr[J];
	int const nm = agg_ptr[J + 1] - a0;
	if (!smooth) {
		for (int t = threadIdx.x; t < nm; t += 64) {
			set.insert(agg_nodes[a0 + t]);
		}
	} else {
		// the (member, neighbour slot) pairs spread over the lanes, w slots
		// per member (w: the widest member row)
		int w = 0;
		for (int t = threadIdx.x; t < nm; t += 64) {
			w = max(w, row_len[agg_nodes[a0 + t]]);
		}
		w = wave_max(w);
		for (int p = threadIdx.x; p < nm * w; p += 64) {
			int const mi = p / w, k = p - mi * w;
			int const m = agg_nodes[a0 + mi];
			if (k < row_len[m] && !set.insert(col[nbr_slot(slice_ptr, m, k)])) {
				break;
			}
		}
	}
	__syncthreads();
	if (*set.over) {
		if (threadIdx.x == 0) {
			atomicOr(bad, 8);
			if (COUNT) {
				rcnt[J] = 0;
			}
		}
		return;
	}
	if (COUNT) {
		if (threadIdx.x == 0) {
			rcnt[J] = *set.cnt;
		}
		return;
	}
	int const m = set.sorted(scratch, rows);
	int32_t const r0 = rptr[J];
	if (m != rptr[J + 1] - r0) {   // the rows next to the aggregate must be P's column
		if (threadIdx.x == 0) {
			atomicOr(bad, 32);
		}
		return;
	}
	for (int e = threadIdx.x; e < m; e += 64) {
		int const i = rows[e];
		int32_t const p0 = pptr[i];
		rrow[r0 + e] = i;
		rent[r0 + e] = p0 + find_sorted(pcol + p0, pptr[i + 1] - p0, J);
	}
}

// ------------------------------------------------------------------ A P pattern

// Row i of A P into u (a table of CAP + 1 entries): the aggregates in the
// prolongator rows of i's neighbours, sorted and distinct; returns their
// count, or -1 when there are more than CAP.
template <int CAP>
__device__ __forceinline__ int ap_row(int i, const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ col,
                                      const int32_t* __restrict__ row_len, const int32_t* __restrict__ pptr,
                                      const int32_t* __restrict__ pcol, int32_t* u) {
	int const len = row_len[i];
	int c = 0;
	for (int k = 0; k < len; k++) {
		int const j = col[nbr_slot(slice_ptr, i, k)];
		for (int32_t f = pptr[j]; f < pptr[j + 1]; f++) {
			int32_t const a = pcol[f];
			int y = c - 1;
			while (y >= 0 && u[y] > a) {
				y--;
			}
			if (y >= 0 && u[y] == a) {
				continue;
			}
			if (c == CAP) {
				return -1;
			}
			for (int z = c; z > y + 1; z--) {
				u[z] = u[z - 1];
			}
			u[y + 1] = a;
			c++;
		}
	}
	return c;
}

// Count pass: a row of at most kStash entries has its count in cnt[i] and
// its sorted entries in stash (kStash per row); a longer one is listed in
// big (*nbig of them) for k_ap_count_big.
__global__ __launch_bounds__(kTW) void k_ap_count(int n, const int32_t* __restrict__ slice_ptr,
                                                  const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ row_len,
                                                  const int32_t* __restrict__ pptr,
                                                  const int32_t* __restrict__ pcol, int32_t* __restrict__ cnt,
                                                  int32_t* __restrict__ stash, int32_t* __restrict__ big,
                                                  int32_t* __restrict__ nbig) {
	__shared__ int32_t lds[kTW * (kStash + 1)];
	int const i = blockIdx.x * kTW + threadIdx.x;
	if (i >= n) {
		return;
	}
	int32_t* u = lds + threadIdx.x * (kStash + 1);
	int const c = ap_row<kStash>(i, slice_ptr, col, row_len, pptr, pcol, u);
	if (c < 0) {
		big[atomicAdd(nbig, 1)] = i;
		return;
	}
	cnt[i] = c;
	int32_t* dst = stash + static_cast<int64_t>(i) * kStash;
	for (int k = 0; k < c; k++) {
		dst[k] = u[k];
	}
}

// The listed rows with full-size tables (a row above kMaxAP is flagged).
__global__ __launch_bounds__(kTW) void k_ap_count_big(const int32_t* __restrict__ slice_ptr,
                                                      const int32_t* __restrict__ col,
                                                      const int32_t* __restrict__ row_len,
                                                      const int32_t* __restrict__ pptr,
                                                      const int32_t* __restrict__ pcol, int32_t* __restrict__ cnt,
                                                      const int32_t* __restrict__ big,
                                                      const int32_t* __restrict__ nbig, int32_t* __restrict__ bad) {
	__shared__ int32_t lds[kTW * (kMaxAP + 1)];
	int const m = *nbig;
	for (int t = blockIdx.x * kTW + threadIdx.x; t < m; t += gridDim.x * kTW) {
		int const i = big[t];
		int const c = ap_row<kMaxAP>(i, slice_ptr, col, row_len, pptr, pcol, lds + threadIdx.x * (kMaxAP + 1));
		if (c < 0) {
			atomicOr(bad, 16);
		}
		cnt[i] = c < 0 ? 0 : c;
	}
}

// Fill pass: the rows the count pass kept, copied from stash to apcol from
// apptr[i], kFillLanes lanes a row (consecutive lanes, consecutive entries).
constexpr int kFillLanes = 8;

__global__ __launch_bounds__(kT) void k_ap_fill(int n, const int32_t* __restrict__ apptr,
                                                const int32_t* __restrict__ stash, int32_t* __restrict__ apcol) {
	int64_t const t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
	int64_t const i = t / kFillLanes;
	if (i >= n) {
		return;
	}
	int32_t const a0 = apptr[i], c = apptr[i + 1] - a0;
	if (c > kStash) {
		return;   // a listed row: k_ap_fill_big
	}
	const int32_t* src = stash + i * kStash;
	for (int k = static_cast<int>(t % kFillLanes); k < c; k += kFillLanes) {
		apcol[a0 + k] = src[k];
	}
}

// The listed rows, formed again with full-size tables.
__global__ __launch_bounds__(kTW) void k_ap_fill_big(const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ row_len,
                                                     const int32_t* __restrict__ pptr,
                                                     const int32_t* __restrict__ pcol,
                                                     const int32_t* __restrict__ apptr, int32_t* __restrict__ apcol,
                                                     const int32_t* __restrict__ big, const int32_t* __restrict__ nbig) {
	__shared__ int32_t lds[kTW * (kMaxAP + 1)];
	int const m = *nbig;
	for (int t = blockIdx.x * kTW + threadIdx.x; t < m; t += gridDim.x * kTW) {
		int const i = big[t];
		int32_t* u = lds + threadIdx.x * (kMaxAP + 1);
		int const c = ap_row<kMaxAP>(i, slice_ptr, col, row_len, pptr, pcol, u);
		int32_t* out = apcol + apptr[i];
		for (int k = 0; k < c; k++) {
			out[k] = u[k];
		}
	}
}

// ------------------------------------------------------------------ coarse pattern

// Coarse row I of P^T (A P) (one workgroup of one wavefront per row): the
// columns of the A P rows of P's column I, sorted and distinct. COUNT: their
// number into ccnt[I] and the sorted row into stash (kMaxCoarse per row);
// else the stashed row copied to cadj from cptr[I].
template <bool COUNT>
__global__ __launch_bounds__(64) void k_c_pattern(int na, const int32_t* __restrict__ rptr,
                                                  const int32_t* __restrict__ rrow,
                                                  const int32_t* __restrict__ apptr,
                                                  const int32_t* __restrict__ apcol, int32_t* __restrict__ ccnt,
                                                  const int64_t* __restrict__ cptr, int32_t* __restrict__ cadj,
                                                  int32_t* __restrict__ stash, int32_t* __restrict__ bad) {
	int const I = blockIdx.x;
	if (!COUNT) {   // the count pass kept the sorted row
		int64_t const c0 = cptr[I];
		int const m = static_cast<int>(cptr[I + 1] - c0);
		for (int e = threadIdx.x; e < m; e += 64) {
			cadj[c0 + e] = stash[static_cast<int64_t>(I) * kMaxCoarse + e];
		}
		return;
	}
	WAVE_SET(set);
	__shared__ int32_t scratch[kSetCap + 64];
	set.clear();
	// the (row of P's column, entry of its A P row) pairs spread over the
	// lanes, w slots per row (w: the longest of those A P rows)
	int32_t const r0 = rptr[I];
	int const nr = rptr[I + 1] - r0;
	int w = 0;
	for (int t = threadIdx.x; t < nr; t += 64) {
		int const i = rrow[r0 + t];
		w = max(w, apptr[i + 1] - apptr[i]);
	}
	w = wave_max(w);
	for (int p = threadIdx.x; p < nr * w; p += 64) {
		int const ri = p / w, k = p - ri * w;
		int const i = rrow[r0 + ri];
		int32_t const e = apptr[i] + k;
		if (e < apptr[i + 1] && !set.insert(apcol[e])) {
			break;
		}
	}
	__syncthreads();
	int const m = *set.cnt;
	if (*set.over || m > kMaxCoarse) {
		if (threadIdx.x == 0) {
			atomicOr(bad, 4);
			if (COUNT) {
				ccnt[I] = 0;
			}
		}
		return;
	}
	if (threadIdx.x == 0) {
		ccnt[I] = m;
	}
	set.sorted(scratch, stash + static_cast<int64_t>(I) * kMaxCoarse);
}

// Widest row of each slice of 64 rows, times 64: the slice's blocks.
__global__ __launch_bounds__(kT) void k_slice_blocks(int64_t n_slices, int na, const int64_t* __restrict__ cptr,
                                                     int32_t* __restrict__ sblocks) {
	int64_t const s = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
	if (s >= n_slices) {
		return;
	}
	int64_t w = 0;
	for (int64_t r = s * kSlice; r < min(static_cast<int64_t>(na), (s + 1) * kSlice); r++) {
		w = max(w, cptr[r + 1] - cptr[r]);
	}
	sblocks[s] = static_cast<int32_t>(w * kSlice);
}

// The coarse sliced ELL: columns (padding repeats the row's own column; the
// rows past na of the last slice use column 0), the diagonal slot and the
// row length of every row.
__global__ __launch_bounds__(kT) void k_c_sell(int64_t nrows_pad, int na, const int64_t* __restrict__ cptr,
                                               const int32_t* __restrict__ cadj,
                                               const int32_t* __restrict__ slice_ptr, int32_t* __restrict__ col,
                                               int32_t* __restrict__ diag_blk, int32_t* __restrict__ row_len) {
	int64_t const I = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
	if (I >= nrows_pad) {
		return;
	}
	int64_t const s = I >> 6;
	int const width = (slice_ptr[s + 1] - slice_ptr[s]) / kSlice;
	int const r = static_cast<int>(I & 63);
	int const len = I < na ? static_cast<int>(cptr[I + 1] - cptr[I]) : 0;
	for (int k = 0; k < width; k++) {
		int32_t const sl = slice_ptr[s] + k * kSlice + r;
		int32_t c = I < na ? static_cast<int32_t>(I) : 0;
		if (k < len) {
			c = cadj[cptr[I] + k];
			if (c == I) {
				diag_blk[I] = sl;
			}
		}
		col[sl] = c;
	}
	if (I < na) {
		row_len[I] = len;
	}
}

// ------------------------------------------------------------------ mirrors

// mirror[slot of (I, J)] = the slot of (J, I) for J >= I (the product is
// symmetric: the numeric kernel writes each block and its transpose), -2 for
// J < I; padding slots keep -1.
__global__ __launch_bounds__(kT) void k_mirror(int na, const int64_t* __restrict__ cptr,
                                               const int32_t* __restrict__ cadj,
                                               const int32_t* __restrict__ c_slice_ptr,
                                               int32_t* __restrict__ mirror) {
	int const I = blockIdx.x * kT + threadIdx.x;
	if (I >= na) {
		return;
	}
	int64_t const c0 = cptr[I];
	int const len = static_cast<int>(cptr[I + 1] - c0);
	int32_t const sbase = c_slice_ptr[I >> 6] + (I & 63);
	for (int q = 0; q < len; q++) {
		int32_t const J = cadj[c0 + q];
		int32_t m = -2;
		if (J == I) {
			m = sbase + q * kSlice;
		} else if (J > I) {
			int64_t const cj = cptr[J];
			// row J is sorted and holds I (the pattern is symmetric)
			int const p = find_sorted(cadj + cj, static_cast<int>(cptr[J + 1] - cj), I);
			m = c_slice_ptr[J >> 6] + p * kSlice + (J & 63);
		}
		mirror[sbase + q * kSlice] = m;
	}
}

// ------------------------------------------------------------------ segments

// cover[g] (g = 0 .. n-1: the gap after row g) counts the entries (r, c),
// r <= g < c: +1 at r for each column c > r, -1 at c.
__global__ __launch_bounds__(kT) void k_seg_diff(int n, const int32_t* __restrict__ slice_ptr,
                                                 const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                                 int32_t* __restrict__ diff) {
	int const i = blockIdx.x * kT + threadIdx.x;
	if (i >= n) {
		return;
	}
	int const len = row_len[i];
	int up = 0;
	for (int k = 0; k < len; k++) {
		int const c = col[nbr_slot(slice_ptr, i, k)];
		if (c > i) {
			up++;
			atomicAdd(&diff[c], -1);
		}
	}
	if (up) {
		atomicAdd(&diff[i], up);
	}
}

// brk[g] = 1 where no entry crosses the gap after row g (cover = prefix sum
// of diff, inclusive).
__global__ __launch_bounds__(kT) void k_seg_breaks(int n, const int32_t* __restrict__ diff,
                                                   const int64_t* __restrict__ excl, int32_t* __restrict__ brk) {
	int const g = blockIdx.x * kT + threadIdx.x;
	if (g < n) {
		brk[g] = excl[g] + diff[g] == 0 ? 1 : 0;
	}
}

__global__ __launch_bounds__(kT) void k_seg_list(int n, const int32_t* __restrict__ brk,
                                                 const int32_t* __restrict__ at, int32_t* __restrict__ seg) {
	int const g = blockIdx.x * kT + threadIdx.x;
	if (g < n && brk[g]) {
		seg[1 + at[g]] = g + 1;
	}
	if (g == 0) {
		seg[0] = 0;
	}
}

__global__ __launch_bounds__(kT) void k_gather_i32(int m, const int32_t* __restrict__ idx,
                                                   const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
	int const k = blockIdx.x * kT + threadIdx.x;
	if (k < m) {
		dst[k] = src[idx[k]];
	}
}

// ------------------------------------------------------------------ host helpers

// Device buffers from the block cache, returned when the scope ends (after
// the stream has drained).
struct Temps {
	hipStream_t st;
	std::vector<void*> p;
	explicit Temps(hipStream_t s) : st(s) {}
	~Temps() {
		(void) hipStreamSynchronize(st);
		for (void* q : p) {
			pool_free_bytes(q);
		}
	}
	template <typename T>
	int get(T** out, size_t n) {
		BFMX_CHECK(palloc(out, n));
		p.push_back(*out);
		return BFMX_OK;
	}
};

template <typename T, typename U>
int scan(hipStream_t st, const T* in, U* out, int64_t n, Temps& tmp) {
	int const nb = std::max(1, nblk(n, kScanBlock));
	if (n > 0 && nb == 1) {
		k_scan_block<T, U><<<1, kT, 0, st>>>(n, in, out, nullptr);
		return BFMX_OK;
	}
	U* bsum = nullptr;
	BFMX_CHECK(tmp.get(&bsum, static_cast<size_t>(nb) + 1));
	if (n > 0) {
		k_scan_block<T, U><<<nb, kT, 0, st>>>(n, in, out, bsum);
	} else {
		k_fill<U><<<1, kT, 0, st>>>(1, bsum, U(0));
	}
	k_scan_top<U><<<1, kT, 0, st>>>(nb, bsum);
	k_scan_add<U><<<std::max(1, nblk(n)), kT, 0, st>>>(n, out, bsum);
	return BFMX_OK;
}

// A count (and the overflow flags) read back in one round trip: a kernel
// writes them into pinned host memory, then the stream is drained.
__global__ void k_read_back(const int64_t* __restrict__ c64, const int32_t* __restrict__ c32,
                            const int32_t* __restrict__ flags, int64_t* __restrict__ out) {
	if (threadIdx.x == 0) {
		out[0] = c64 ? c64[0] : c32[0];
		out[1] = flags ? flags[0] : 0;
	}
}

template <typename T>
int read_count_flags(hipStream_t st, const T* d, T* h, const int32_t* dflags, int32_t* hflags) {
	void* buf = nullptr;
	BFMX_CHECK(pinned_staging(&buf, 2 * sizeof(int64_t), 1));
	int64_t* const out = static_cast<int64_t*>(buf);
	if constexpr (sizeof(T) == 8) {
		k_read_back<<<1, 64, 0, st>>>(reinterpret_cast<const int64_t*>(d), nullptr, dflags, out);
	} else {
		k_read_back<<<1, 64, 0, st>>>(nullptr, reinterpret_cast<const int32_t*>(d), dflags, out);
	}
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipStreamSynchronize(st));
	*h = static_cast<T>(out[0]);
	if (hflags) {
		*hflags = static_cast<int32_t>(out[1]);
	}
	return BFMX_OK;
}

template <typename T>
int read_one(hipStream_t st, const T* d, T* h) {
	return read_count_flags<T>(st, d, h, nullptr, nullptr);
}

// out[i] = in[i] (64-bit prefix sums known to fit in 32 bits).
__global__ __launch_bounds__(kT) void k_narrow(int64_t n, const int64_t* __restrict__ in, int32_t* __restrict__ out) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
	if (i < n) {
		out[i] = static_cast<int32_t>(in[i]);
	}
}

template <typename T>
int zeros(hipStream_t st, T* p, size_t n) {
	if (n) {
		k_fill<T><<<nblk(static_cast<int64_t>(n)), kT, 0, st>>>(static_cast<int64_t>(n), p, T(0));
	}
	return BFMX_OK;
}

// Slice pointers, row lengths and block columns of a sliced ELL, one after
// the other into out (a host buffer the device writes).
__global__ __launch_bounds__(kT) void k_pack_sell(int64_t ns1, int64_t n, int64_t total,
                                                  const int32_t* __restrict__ slice, const int32_t* __restrict__ len,
                                                  const int32_t* __restrict__ col, int32_t* __restrict__ out) {
	int64_t const t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
	if (t < ns1) {
		out[t] = slice[t];
	} else if (t < ns1 + n) {
		out[t] = len[t - ns1];
	} else if (t < total) {
		out[t] = col[t - ns1 - n];
	}
}

// One copy item of topo_upload: `bytes` from offset `off` of the packed
// block to dst.
struct UnpackDesc {
	void* dst;
	int64_t off;
	int64_t bytes;
};

// Workgroup k copies item k (words, then the last bytes).
__global__ __launch_bounds__(kT) void k_unpack(const UnpackDesc* __restrict__ desc, const char* __restrict__ src) {
	UnpackDesc const d = desc[blockIdx.x];
	const char* s = src + d.off;
	char* o = static_cast<char*>(d.dst);
	int64_t const words = d.bytes / 4;
	for (int64_t w = threadIdx.x; w < words; w += kT) {
		reinterpret_cast<int32_t*>(o)[w] = reinterpret_cast<const int32_t*>(s)[w];
	}
	for (int64_t b = words * 4 + threadIdx.x; b < d.bytes; b += kT) {
		o[b] = s[b];
	}
}

}  // namespace

int topo_upload(hipStream_t st, std::vector<UploadItem> const& items) {
	if (items.empty()) {
		return BFMX_OK;
	}
	auto const align = [](size_t b) { return (b + 15) & ~size_t(15); };
	size_t const ndesc = items.size();
	size_t total = align(ndesc * sizeof(UnpackDesc));
	for (auto const& it : items) {
		total += align(it.bytes);
	}
	void* hbuf = nullptr;
	BFMX_CHECK(pinned_staging(&hbuf, total, 0));
	char* const h = static_cast<char*>(hbuf);
	UnpackDesc* desc = reinterpret_cast<UnpackDesc*>(h);
	size_t off = align(ndesc * sizeof(UnpackDesc));
	for (size_t k = 0; k < ndesc; k++) {
		desc[k] = UnpackDesc{items[k].dst, static_cast<int64_t>(off), static_cast<int64_t>(items[k].bytes)};
		if (items[k].bytes) {
			std::memcpy(h + off, items[k].src, items[k].bytes);
		}
		off += align(items[k].bytes);
	}
	char* dbuf = nullptr;
	BFMX_CHECK(palloc(&dbuf, total));
	hipError_t e = hipMemcpyAsync(dbuf, h, total, hipMemcpyHostToDevice, st);
	if (e == hipSuccess) {
		k_unpack<<<static_cast<unsigned>(ndesc), kT, 0, st>>>(reinterpret_cast<const UnpackDesc*>(dbuf), dbuf);
		e = hipGetLastError();
	}
	hipError_t const e2 = hipStreamSynchronize(st);
	pfree(dbuf);
	if (e != hipSuccess || e2 != hipSuccess) {
		set_error("upload of a level's transfer arrays: %s", hipGetErrorString(e != hipSuccess ? e : e2));
		return BFMX_E_HIP;
	}
	return BFMX_OK;
}

int topo_sell_to_host(hipStream_t st, TopoGraph const& G, const int32_t** slice, const int32_t** len,
                      const int32_t** col, std::vector<int32_t>& spill) {
	int64_t const ns1 = G.n_slices + 1, n = G.n, total = ns1 + n + G.nnzb_pad;
	int32_t* out = nullptr;
	if (static_cast<size_t>(total) * sizeof(int32_t) > kPinnedGraphBytes) {
		spill.resize(static_cast<size_t>(total));
		out = spill.data();
		BFMX_HIP(hipMemcpyAsync(out, G.slice_ptr, ns1 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
		BFMX_HIP(hipMemcpyAsync(out + ns1, G.row_len, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
		BFMX_HIP(hipMemcpyAsync(out + ns1 + n, G.col, G.nnzb_pad * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	} else {
		void* buf = nullptr;
		BFMX_CHECK(pinned_staging(&buf, static_cast<size_t>(total) * sizeof(int32_t)));
		out = static_cast<int32_t*>(buf);
		k_pack_sell<<<nblk(total), kT, 0, st>>>(ns1, n, total, G.slice_ptr, G.row_len, G.col, out);
		BFMX_HIP(hipGetLastError());
	}
	BFMX_HIP(hipStreamSynchronize(st));
	*slice = out;
	*len = out + ns1;
	*col = out + ns1 + n;
	return BFMX_OK;
}

int topo_segments(hipStream_t st, TopoGraph const& G, std::vector<int32_t>& seg) {
	Temps tmp(st);
	int const n = G.n;
	int32_t *diff, *brk, *at, *dseg;
	int64_t* excl;
	BFMX_CHECK(tmp.get(&diff, n + 1));
	BFMX_CHECK(tmp.get(&excl, n + 1));
	BFMX_CHECK(tmp.get(&brk, n + 1));
	BFMX_CHECK(tmp.get(&at, n + 1));
	BFMX_CHECK(zeros(st, diff, n + 1));
	k_seg_diff<<<nblk(n), kT, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, diff);
	BFMX_CHECK(scan(st, diff, excl, n, tmp));
	k_seg_breaks<<<nblk(n), kT, 0, st>>>(n, diff, excl, brk);
	BFMX_CHECK(scan(st, brk, at, n, tmp));
	int32_t nbrk = 0;
	BFMX_CHECK(read_one(st, at + n, &nbrk));
	BFMX_CHECK(tmp.get(&dseg, static_cast<size_t>(nbrk) + 1));
	k_seg_list<<<nblk(n), kT, 0, st>>>(n, brk, at, dseg);
	seg.assign(static_cast<size_t>(nbrk) + 1, 0);
	BFMX_HIP(hipMemcpyAsync(seg.data(), dseg, seg.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipStreamSynchronize(st));
	BFMX_HIP(hipGetLastError());
	if (seg.back() != n) {   // the last row always ends a segment
		set_error("multigrid: segment boundaries of the level graph are inconsistent");
		return BFMX_E_ARG;
	}
	return BFMX_OK;
}

int topo_aggregate(hipStream_t st, TopoGraph const& G, uint32_t salt, double seed_frac, double seed_growth,
                   std::vector<int32_t> const& fseg, int32_t** agg_out, int* na_out, std::vector<int32_t>& cseg) {
	Temps tmp(st);
	PhaseTimer phase(st, "aggregation");
	int const n = G.n;
	uint8_t *state, *s1;
	uint64_t *t0, *t1, *t2;
	int32_t *left, *flag, *rootid, *agg1, *agg, *bad;
	BFMX_CHECK(tmp.get(&state, n));
	BFMX_CHECK(tmp.get(&s1, n));
	BFMX_CHECK(tmp.get(&t0, n));
	BFMX_CHECK(tmp.get(&t1, n));
	BFMX_CHECK(tmp.get(&t2, n));
	BFMX_CHECK(tmp.get(&left, 2));
	BFMX_CHECK(tmp.get(&flag, n));
	BFMX_CHECK(tmp.get(&rootid, n + 1));
	BFMX_CHECK(tmp.get(&agg1, n));
	BFMX_CHECK(palloc(&agg, n));
	*agg_out = agg;   // owned by the caller from here on
	bad = left + 1;
	k_fill<uint8_t><<<nblk(n), kT, 0, st>>>(n, state, 1);
	BFMX_CHECK(zeros(st, left, 2));
	phase.mark("buffers");
	// rounds in groups of kMisGroup between checks (a round after the last
	// undecided row is gone changes nothing)
	double thr = seed_frac * 1073741824.0;
	for (int rounds = 0;;) {
		for (int g = 0; g < kMisGroup; g++, rounds++) {
			if (g == kMisGroup - 1) {
				BFMX_CHECK(zeros(st, left, 1));
			}
			uint32_t const seed = static_cast<uint32_t>(std::min(thr, 1073741824.0));
			thr *= seed_growth;
			k_mis_front<<<nblk(n), kT, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, state, s1);
			k_mis_tuples<<<nblk(n), kT, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, state, s1, salt, seed, t0);
			k_mis_max<<<nblk(n), kT, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, t0, t1);
			k_mis_max<<<nblk(n), kT, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, t1, t2);
			k_mis_update<<<nblk(n), kT, 0, st>>>(n, state, t0, t2, left);
		}
		int32_t l = 0;
		BFMX_CHECK(read_one(st, left, &l));
		if (l == 0) {
			if (phase_times_on()) {
				std::fprintf(stderr, "aggregation: %d rows, %d independent-set rounds\n", n, rounds);
			}
			break;
		}
		if (rounds >= kMisMaxRounds) {
			set_error("multigrid: the independent-set rounds did not finish");
			return BFMX_E_UNSUPPORTED;
		}
	}
	phase.mark("independent-set rounds");
	k_is_root<<<nblk(n), kT, 0, st>>>(n, state, flag);
	BFMX_CHECK(scan(st, flag, rootid, n, tmp));
	k_attach_roots<<<nblk(n), kT, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, state, rootid, agg1);
	k_attach_rest<<<nblk(n), kT, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, agg1, agg, bad);
	int32_t h[2] = {0, 0};
	BFMX_HIP(hipMemcpyAsync(h, left, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	int32_t na = 0;
	BFMX_HIP(hipMemcpyAsync(&na, rootid + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
	// the first aggregate of each segment: the roots before its first row
	int const ns = static_cast<int>(fseg.size());
	int32_t *dfs, *dcs;
	BFMX_CHECK(tmp.get(&dfs, ns));
	BFMX_CHECK(tmp.get(&dcs, ns));
	BFMX_HIP(hipMemcpyAsync(dfs, fseg.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, st));
	k_gather_i32<<<nblk(ns), kT, 0, st>>>(ns, dfs, rootid, dcs);
	cseg.assign(ns, 0);
	BFMX_HIP(hipMemcpyAsync(cseg.data(), dcs, ns * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipStreamSynchronize(st));
	BFMX_HIP(hipGetLastError());
	if (h[1]) {
		set_error("multigrid: a row found no aggregate");
		return BFMX_E_ARG;
	}
	phase.mark("attach");
	*na_out = na;
	return BFMX_OK;
}

int topo_transfer(hipStream_t st, TopoGraph const& G, const int32_t* agg, int na, bool smooth, TopoLevel& T) {
	Temps tmp(st);
	PhaseTimer phase(st, "transfer");
	int const n = G.n;
	int32_t *cnt, *fill, *flags;
	BFMX_CHECK(tmp.get(&cnt, std::max(n, na) + 1));
	BFMX_CHECK(tmp.get(&fill, std::max(n, na) + 1));
	BFMX_CHECK(tmp.get(&flags, 3));   // overflow bits, then the counts of the two row lists
	BFMX_CHECK(zeros(st, flags, 3));
	int32_t* big;   // rows for the full-table kernels
	BFMX_CHECK(tmp.get(&big, n));
	// aggregate members, sorted
	BFMX_CHECK(palloc(&T.agg_ptr, na + 1));
	BFMX_CHECK(palloc(&T.agg_nodes, n));
	BFMX_CHECK(zeros(st, cnt, na + 1));
	k_count_keys<<<nblk(n), kT, 0, st>>>(n, agg, cnt);
	BFMX_CHECK(scan(st, cnt, T.agg_ptr, na, tmp));
	BFMX_CHECK(zeros(st, fill, na + 1));
	k_place_members<<<nblk(n), kT, 0, st>>>(n, agg, T.agg_ptr, fill, T.agg_nodes);
	k_sort_lists<<<nblk(na), kT, 0, st>>>(na, T.agg_ptr, T.agg_nodes, nullptr);
	phase.mark("aggregate members");
	// prolongator pattern
	BFMX_CHECK(palloc(&T.pptr, n + 1));
	k_p_count<<<nblk(n, kTW), kTW, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, agg, smooth ? 1 : 0, cnt, flags, big,
	                                        flags + 1);
	k_p_count_big<<<kBigBlocks, kTW, 0, st>>>(G.slice_ptr, G.col, G.row_len, agg, smooth ? 1 : 0, cnt, flags, big,
	                                          flags + 1);
	BFMX_CHECK(scan(st, cnt, T.pptr, n, tmp));
	int32_t nnzp = 0, bad = 0;
	BFMX_CHECK(read_count_flags(st, T.pptr + n, &nnzp, flags, &bad));
	if (bad) {
		set_error(bad & 1 ? "multigrid: a level row has more than %d neighbours for the device build"
		                  : "multigrid: a row couples to more than 255 aggregates",
		          kMaxRow);
		return BFMX_E_UNSUPPORTED;
	}
	if (static_cast<int64_t>(nnzp) >= (int64_t(1) << 31) / 9) {
		set_error("multigrid prolongator too large for 32-bit indices");
		return BFMX_E_ARG;
	}
	T.nnzp = nnzp;
	BFMX_CHECK(palloc(&T.pcol, nnzp));
	BFMX_CHECK(palloc(&T.ppos, static_cast<size_t>(G.nnzb_pad)));
	BFMX_CHECK(zeros(st, T.ppos, static_cast<size_t>(G.nnzb_pad)));   // the padding rows' slots
	k_p_fill<<<nblk(n, kTW), kTW, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, agg, smooth ? 1 : 0, T.pptr, T.pcol,
	                                       T.ppos);
	k_p_fill_big<<<kBigBlocks, kTW, 0, st>>>(G.slice_ptr, G.col, G.row_len, agg, smooth ? 1 : 0, T.pptr, T.pcol,
	                                         T.ppos, big, flags + 1);
	phase.mark("prolongator pattern");
	// restriction lists: P's columns, rows sorted; a column's length is the
	// count of its aggregate among P's entries (the lists are checked against
	// it as they are written; a mismatch or an overflow is reported with the
	// flags read after the A P count below)
	BFMX_CHECK(palloc(&T.rptr, na + 1));
	BFMX_CHECK(zeros(st, cnt, na + 1));
	k_count_keys<<<nblk(nnzp), kT, 0, st>>>(nnzp, T.pcol, cnt);
	BFMX_CHECK(scan(st, cnt, T.rptr, na, tmp));
	BFMX_CHECK(palloc(&T.rent, nnzp));
	BFMX_CHECK(palloc(&T.rrow, nnzp));
	k_r_lists<false><<<na, 64, 0, st>>>(na, smooth ? 1 : 0, G.slice_ptr, G.col, G.row_len, T.agg_ptr, T.agg_nodes, T.pptr, T.pcol,
	                                    nullptr, T.rptr, T.rrow, T.rent, flags);
	phase.mark("restriction lists");
	// pattern of A P
	BFMX_CHECK(palloc(&T.apptr, n + 1));
	int32_t* stash;   // each row's sorted list, when short, from the count pass to the fill pass
	BFMX_CHECK(tmp.get(&stash, static_cast<size_t>(n) * kStash));
	k_ap_count<<<nblk(n, kTW), kTW, 0, st>>>(n, G.slice_ptr, G.col, G.row_len, T.pptr, T.pcol, cnt, stash, big,
	                                         flags + 2);
	k_ap_count_big<<<kBigBlocks, kTW, 0, st>>>(G.slice_ptr, G.col, G.row_len, T.pptr, T.pcol, cnt, big, flags + 2,
	                                           flags);
	int32_t nap = 0;
	{
		// the prefix sums in 64 bits: the 32-bit row pointers must hold the total
		int64_t* ap64;
		BFMX_CHECK(tmp.get(&ap64, static_cast<size_t>(n) + 1));
		BFMX_CHECK(scan(st, cnt, ap64, n, tmp));
		int64_t total = 0;
		BFMX_CHECK(read_count_flags(st, ap64 + n, &total, flags, &bad));
		if (bad & 32) {
			set_error("multigrid: restriction lists do not match the prolongator");
			return BFMX_E_ARG;
		}
		if (bad & 8) {
			set_error("multigrid: a prolongator column has more than %d rows for the device build", kSetCap);
			return BFMX_E_UNSUPPORTED;
		}
		if (bad) {
			set_error("multigrid: a row of A P has more than %d entries for the device build", kMaxAP);
			return BFMX_E_UNSUPPORTED;
		}
		if (total >= (int64_t(1) << 31)) {
			set_error("multigrid: A P too large for 32-bit indices (%lld entries)", static_cast<long long>(total));
			return BFMX_E_UNSUPPORTED;
		}
		k_narrow<<<nblk(static_cast<int64_t>(n) + 1), kT, 0, st>>>(static_cast<int64_t>(n) + 1, ap64, T.apptr);
		nap = static_cast<int32_t>(total);
	}
	T.nnz_ap = nap;
	BFMX_CHECK(palloc(&T.apcol, nap));
	k_ap_fill<<<nblk(static_cast<int64_t>(n) * kFillLanes), kT, 0, st>>>(n, T.apptr, stash, T.apcol);
	k_ap_fill_big<<<kBigBlocks, kTW, 0, st>>>(G.slice_ptr, G.col, G.row_len, T.pptr, T.pcol, T.apptr, T.apcol, big,
	                                          flags + 2);
	phase.mark("A P pattern");
	// coarse pattern
	int64_t* cptr;
	BFMX_CHECK(tmp.get(&cptr, na + 1));
	int32_t* cstash;   // each coarse row, sorted, from the count pass to the fill pass
	BFMX_CHECK(tmp.get(&cstash, static_cast<size_t>(na) * kMaxCoarse));
	k_c_pattern<true><<<na, 64, 0, st>>>(na, T.rptr, T.rrow, T.apptr, T.apcol, cnt, nullptr, nullptr, cstash, flags);
	BFMX_CHECK(scan(st, cnt, cptr, na, tmp));
	int64_t cnnz = 0;
	BFMX_CHECK(read_count_flags(st, cptr + na, &cnnz, flags, &bad));
	if (bad) {
		set_error("multigrid: a coarse row couples to more than %d coarse rows for the device build", kMaxCoarse);
		return BFMX_E_UNSUPPORTED;
	}
	int32_t* cadj;
	BFMX_CHECK(tmp.get(&cadj, static_cast<size_t>(cnnz)));
	k_c_pattern<false><<<na, 64, 0, st>>>(na, T.rptr, T.rrow, T.apptr, T.apcol, nullptr, cptr, cadj, cstash, flags);
	phase.mark("coarse pattern");
	// coarse sliced ELL
	T.c_slices = (na + kSlice - 1) / kSlice;
	T.c_nnzb = cnnz;
	int32_t* sblocks;
	BFMX_CHECK(tmp.get(&sblocks, static_cast<size_t>(T.c_slices) + 1));
	k_slice_blocks<<<nblk(T.c_slices), kT, 0, st>>>(T.c_slices, na, cptr, sblocks);
	int64_t* sp64;
	BFMX_CHECK(tmp.get(&sp64, static_cast<size_t>(T.c_slices) + 1));
	BFMX_CHECK(scan(st, sblocks, sp64, T.c_slices, tmp));
	int64_t pad = 0;
	BFMX_CHECK(read_one(st, sp64 + T.c_slices, &pad));
	if (pad * 9 >= (int64_t(1) << 31)) {
		set_error("multigrid level too large for 32-bit block indices");
		return BFMX_E_ARG;
	}
	T.c_nnzb_pad = pad;
	BFMX_CHECK(palloc(&T.c_slice_ptr, static_cast<size_t>(T.c_slices) + 1));
	BFMX_CHECK(scan(st, sblocks, T.c_slice_ptr, T.c_slices, tmp));
	BFMX_CHECK(palloc(&T.c_col, static_cast<size_t>(pad)));
	BFMX_CHECK(palloc(&T.c_diag_blk, na));
	BFMX_CHECK(palloc(&T.c_row_len, na));
	k_c_sell<<<nblk(T.c_slices * kSlice), kT, 0, st>>>(T.c_slices * kSlice, na, cptr, cadj, T.c_slice_ptr, T.c_col,
	                                                   T.c_diag_blk, T.c_row_len);
	// mirrors of the coarse slots
	BFMX_CHECK(palloc(&T.mirror, static_cast<size_t>(pad)));
	k_fill<int32_t><<<nblk(pad), kT, 0, st>>>(pad, T.mirror, -1);
	k_mirror<<<nblk(na), kT, 0, st>>>(na, cptr, cadj, T.c_slice_ptr, T.mirror);
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipStreamSynchronize(st));
	phase.mark("coarse sliced ELL");
	return BFMX_OK;
}

void topo_release(TopoLevel& T) {
	pfree(T.agg_ptr);
	pfree(T.agg_nodes);
	pfree(T.pptr);
	pfree(T.pcol);
	pfree(T.ppos);
	pfree(T.rptr);
	pfree(T.rent);
	pfree(T.rrow);
	pfree(T.c_slice_ptr);
	pfree(T.c_col);
	pfree(T.c_diag_blk);
	pfree(T.c_row_len);
	pfree(T.apptr);
	pfree(T.apcol);
	pfree(T.mirror);
}

}  // namespace bfmx
