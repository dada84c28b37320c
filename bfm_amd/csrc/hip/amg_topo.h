// Device build of the multigrid hierarchy's topology (amg_topo.hip).
#pragma once

#include <vector>

#include "internal.h"

namespace bfmx {

constexpr int kMaxRow = 128;     // neighbours of a level row the device build handles
constexpr int kMaxCoarse = 128;  // columns of a coarse row the device build handles
constexpr int kMaxAP = 128;      // entries of a row of A P the device build handles
constexpr int kMaxPCol = 256;    // rows of a prolongator column the device build handles

// A level's graph on the device, as its sliced ELL: row i's k-th neighbour
// (k < row_len[i], sorted, the row itself included) is
// col[slice_ptr[i >> 6] + 64 k + (i & 63)].
struct TopoGraph {
	int n = 0;
	int64_t n_slices = 0;
	int64_t nnzb_pad = 0;
	const int32_t* slice_ptr = nullptr;
	const int32_t* col = nullptr;
	const int32_t* row_len = nullptr;
};

// The transfer from a level to the next and the next level's graph, as the
// host build produces them (device arrays, owned by whoever takes them).
struct TopoLevel {
	int64_t nnzp = 0;
	int32_t* agg_ptr = nullptr;
	int32_t* agg_nodes = nullptr;
	int32_t* pptr = nullptr;
	int32_t* pcol = nullptr;
	uint8_t* ppos = nullptr;
	int32_t* rptr = nullptr;
	int32_t* rent = nullptr;
	int32_t* rrow = nullptr;
	int64_t c_slices = 0, c_nnzb = 0, c_nnzb_pad = 0;
	int32_t* c_slice_ptr = nullptr;
	int32_t* c_col = nullptr;
	int32_t* c_diag_blk = nullptr;
	int32_t* c_row_len = nullptr;
	int32_t* mirror = nullptr;     // per coarse slot: the slot of its transpose (-2 below the diagonal, -1 padding)
	int64_t nnz_ap = 0;            // A P: CSR pattern over the fine rows
	int32_t* apptr = nullptr;
	int32_t* apcol = nullptr;
};

// Row ranges of the graph that no entry couples, as offsets 0, ..., n.
int topo_segments(hipStream_t st, TopoGraph const& G, std::vector<int32_t>& seg);

// The graph's sliced ELL on the host: *slice (n_slices + 1 entries), *len
// (n), *col (nnzb_pad). Up to kPinnedGraphBytes it is packed by one kernel
// into the calling thread's pinned staging buffer (valid until the thread's
// next use of it); a larger graph is copied into spill (pageable, released
// with it), so the staging buffer kept for the process stays small.
constexpr size_t kPinnedGraphBytes = size_t(32) << 20;
int topo_sell_to_host(hipStream_t st, TopoGraph const& G, const int32_t** slice, const int32_t** len,
                      const int32_t** col, std::vector<int32_t>& spill);

// Aggregates of the graph (MIS-2 roots with their neighbours; the other rows
// join a neighbour's aggregate): *agg (device, n entries, allocated here and
// owned by the caller, also on failure once set), their count, and cseg[c]
// = the first aggregate of segment c (fseg: the level's segment offsets).
// Roots grow from seeds: a fraction seed_frac of the rows may start a region
// in the first round, that fraction growing by seed_growth per round.
int topo_aggregate(hipStream_t st, TopoGraph const& G, uint32_t salt, double seed_frac, double seed_growth,
                   std::vector<int32_t> const& fseg, int32_t** agg, int* na, std::vector<int32_t>& cseg);

// The transfer operators' patterns, the pattern of A P, the coarse graph and
// its mirrors for the aggregates agg (na of them); smooth: P = (I - w D^-1
// A) P_t (else P = P_t). On failure the arrays already allocated stay in T
// (topo_release). BFMX_E_UNSUPPORTED when a row exceeds the device build's
// bounds.
int topo_transfer(hipStream_t st, TopoGraph const& G, const int32_t* agg, int na, bool smooth, TopoLevel& T);

void topo_release(TopoLevel& T);

// Host arrays copied to device arrays with one copy and one kernel: the
// items are packed into the calling thread's pinned staging buffer, copied
// to one device block and scattered from there; returns once they are in
// place.
struct UploadItem {
	void* dst;          // device array of at least `bytes`
	const void* src;    // host data
	size_t bytes;
};
int topo_upload(hipStream_t st, std::vector<UploadItem> const& items);

}  // namespace bfmx
