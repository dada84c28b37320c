// Internal declarations of libbfmhip.so (not installed, not part of the ABI).
//
// Device data layout of one planar system (all in the renumbered node order):
//   coords   double[2 n]            node positions
//   matrix   sliced ELL of 2x2 FP64 blocks, one slice = 64 consecutive block
//            rows = one wavefront. Slice s holds K_s = max row length block
//            columns; block k of row r (r = row & 63) is entry
//            slice_ptr[s] + 64 k + r of col[], and its values sit at
//            val[4 (slice_ptr[s] + 64 k)] + {2r, 2r+1} (top row a00 a01) and
//            + {128 + 2r, 128 + 2r + 1} (bottom row a10 a11), so one lane's
//            16-byte loads of a block row are contiguous across the wavefront.
//            Padding blocks repeat the row's own column with zero values.
//   elements int32 node ids in colour order; a colour's elements share no
//            node, so one lane per element adds into K without atomics and
//            the sum order is fixed (deterministic).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "bfmx.h"

namespace bfmx {

constexpr int kSlice = 64;     // block rows per sliced-ELL slice (one wavefront)
constexpr int kBlock = 256;    // threads per workgroup for the streaming kernels
constexpr int kChunk = 16;     // CG iterations per captured graph

void set_error(const char* fmt, ...);

#define BFMX_HIP(expr)                                                                              \
	do {                                                                                            \
		hipError_t e_ = (expr);                                                                     \
		if (e_ != hipSuccess) {                                                                     \
			::bfmx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
			return BFMX_E_HIP;                                                                      \
		}                                                                                           \
	} while (0)

#define BFMX_CHECK(expr)          \
	do {                          \
		int rc_ = (expr);         \
		if (rc_ != BFMX_OK) {     \
			return rc_;           \
		}                         \
	} while (0)

// BFMX_SOLVE_TIMES=1: the set-up and solve phases on stderr. Each mark
// waits for the stream, so timed runs are slightly slower than plain ones.
bool phase_times_on();
struct PhaseTimer {
	hipStream_t st;
	const char* scope;
	double t;   // host clock at the last mark, ms
	PhaseTimer(hipStream_t stream, const char* what);
	void mark(const char* what);
};

// Device-resident scalars of one CG run. Written by the one-workgroup reduce
// kernels, read by later kernels (kernel boundaries order the accesses).
struct alignas(128) Scalars {
	double rz;       // r . z
	double alpha;
	double beta;
	double rr;       // r . r (recursive residual)
	double pq;       // p . A p
	double tol2;     // (rtol ||b||)^2
	double bb;       // b . b
	double rr_true;  // ||b - A x||^2 at the last recomputation
	double rz_prev;  // r . z of the previous iteration
	double rr_ref;   // r . r at the last recomputation (reliable updates)
	double rr_best;  // lowest true r . r seen by the tolerance checks
	int64_t iter;
	int32_t done;    // 0 running, 1 converged, 2 breakdown, 3 stagnated, 4 holding for a tolerance check
	int32_t stall;   // consecutive tolerance checks without halving the true residual
	int32_t checking;// a tolerance check has failed: check at every chunk end
	int32_t updates; // true-residual recomputations
};

// Segmented flexible CG: a system whose convergence segments (at least kBlock
// rows each) run their own Krylov scalars. Every kBlock-row block then holds
// rows of at most two segments: seg[2 b] from its first row on and
// seg[2 b + 1] from its row split[b] on (split[b] = kBlock when the block lies
// in one segment). A block's partial sums are kept apart for the two, value
// k of block b of nblk at partial[(k nblk + b) 2 + h], h = 0 (first) or 1.
struct SegView {
	const int32_t* seg = nullptr;     // 2 per block (nullptr: not segmented)
	const int32_t* split = nullptr;   // 1 per block
};

template <int NV>
__device__ __forceinline__ void seg_block_partials(double (&lo)[NV], double (&hi)[NV], int64_t blk, int64_t nblk,
                                                   double* __restrict__ partial) {
	__shared__ double lds[2 * NV][kBlock / 64];
#pragma unroll
	for (int k = 0; k < NV; k++) {
#pragma unroll
		for (int off = 32; off > 0; off >>= 1) {
			lo[k] += __shfl_xor(lo[k], off, 64);
			hi[k] += __shfl_xor(hi[k], off, 64);
		}
	}
	if ((threadIdx.x & 63) == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			lds[k][threadIdx.x >> 6] = lo[k];
			lds[NV + k][threadIdx.x >> 6] = hi[k];
		}
	}
	__syncthreads();
	if (threadIdx.x == 0) {
#pragma unroll
		for (int k = 0; k < NV; k++) {
			double a = 0, c = 0;
#pragma unroll
			for (int w = 0; w < kBlock / 64; w++) {
				a += lds[k][w];
				c += lds[NV + k][w];
			}
			partial[(k * nblk + blk) * 2] = a;
			partial[(k * nblk + blk) * 2 + 1] = c;
		}
	}
}

// One level of the aggregation multigrid used as the FP32 preconditioner of
// the flexible-CG path. Level 0 is the system's own matrix (2x2 blocks);
// coarser levels have three DOFs per node (the coefficients of the three
// rigid-body modes of the aggregate). Block storage is sliced ELL as on
// level 0; the 3x3 blocks of a chunk of 64 (one per row of a slice) are
// stored block row by block row, each lane's three entries of a row next to
// each other (one 12-byte load per block row, contiguous across the lanes).
struct AmgLevel {
	int nb = 0;                 // block rows
	int bs = 2;                 // DOFs per block row
	int64_t n_slices = 0, nnzb = 0, nnzb_pad = 0;
	int32_t* d_slice_ptr = nullptr;   // level 0: the system's arrays (not owned)
	int32_t* d_row_len = nullptr;     // coarse levels built on the device: stored blocks per row
	int32_t* d_col = nullptr;
	int32_t* d_diag_blk = nullptr;
	double* d_val64 = nullptr;        // level 0: the system's FP64 values (not owned)
	float* d_val32 = nullptr;         // level 0: the system's FP32 copy (not owned)
	float* d_dinv = nullptr;          // bs*bs per row: inverse diagonal blocks
	double* d_scal = nullptr;         // lambda_max, then the Chebyshev coefficients
	double* d_B = nullptr;            // bs x 3 per row: rigid-body modes on this level
	double* d_B0 = nullptr;           // level 0: the modes before fixed DOFs are zeroed
	// transfer to the next coarser level (absent on the coarsest)
	int32_t* d_agg = nullptr;         // next-level row (aggregate) of each row
	float* d_Pt = nullptr;            // tentative prolongator: bs x 3 block per row
	int32_t* d_agg_ptr = nullptr;     // members of each aggregate
	int32_t* d_agg_nodes = nullptr;
	// prolongator P = (I - w D^-1 A) P_t, CSR of bs x 3 blocks over this
	// level's rows (unsmoothed: one entry per row, P = P_t)
	bool sa = true;                   // smoothed (false: P = P_t)
	int64_t nnzp = 0;
	int32_t* d_pptr = nullptr;        // nb + 1
	int32_t* d_pcol = nullptr;        // next-level row of each entry
	float* d_P = nullptr;             // bs x 3 per entry
	uint8_t* d_ppos = nullptr;        // per block slot: entry (within its row) of its column's aggregate
	int32_t* d_rptr = nullptr;        // per next-level row J: range of P's column J in the order below
	int32_t* d_rent = nullptr;        //   P's entries column by column (entry index)
	int32_t* d_rrow = nullptr;        //   and their rows on this level
	float* d_R = nullptr;             // those entries' values in that order (R = P^T read row by row)
	int64_t* d_gal_ptr = nullptr;     // per next-level block slot: range of contributions (slots (I, J >= I))
	int32_t* d_gal_mirror = nullptr;  // per next-level block slot: the slot of its transpose (-2: filled by
	                                  // it, -1: padding); both build paths
	int32_t* d_gal_blk = nullptr;     // this level's block slot
	int32_t* d_gal_pi = nullptr;      // entry of P in the slot's row
	int32_t* d_gal_pj = nullptr;      // entry of P in the slot's column
	// levels built on the device form the Galerkin product in two steps
	// instead of from contribution lists: A P as CSR of bs x 3 FP32 blocks
	// over this level's rows (absent for levels built on the host)
	int64_t nnz_ap = 0;
	int32_t* d_apptr = nullptr;
	int32_t* d_apcol = nullptr;
	float* d_apval = nullptr;
	// segments: row ranges that no entry of the level couples (the connected
	// parts of a mesh numbered one after another; a single mesh has one).
	// The coarsest level, when small enough, holds the dense inverse of each
	// segment: (bs m) x (bs m) FP64 for a segment of m rows, row-major, the
	// inverses one after another.
	int n_seg = 1;
	double* d_dense = nullptr;
	int64_t n_dense = 0;              // its entries
	int32_t* d_seg_ptr = nullptr;     // n_seg + 1 row offsets
	int32_t* d_seg_row = nullptr;     // segment of each row
	int64_t* d_seg_off = nullptr;     // first entry of each segment's inverse
	// FP32 work vectors (bs*nb each)
	float* d_b = nullptr;
	float* d_x[2] = {nullptr, nullptr};
	float* d_d = nullptr;
	float* d_r = nullptr;
	double* d_partial = nullptr;      // per-workgroup maxima / sums
};

// Device view of a coarse multigrid level (three DOFs per row) for the
// one-workgroup coarse cycle.
struct AmgLevelDesc {
	int nb;             // block rows
	int deg;            // Chebyshev steps (the coarsest: its iteration count)
	int dense;          // the coarsest level, solved by its dense inverse
	int pad;
	const int32_t* slice_ptr;
	const int32_t* col;
	const float* val32;
	const float* dinv;
	const double* scal;
	float* b;
	float* x0;
	float* x1;
	float* d;
	float* r;
	const int32_t* rptr;   // transfer to the next level (absent on the coarsest)
	const float* R;
	const int32_t* rrow;
	const float* P;
	const int32_t* pptr;
	const int32_t* pcol;
	const double* inv;     // dense inverses of the coarsest's segments
	const int32_t* seg_ptr;
	const int32_t* seg_row;
	const int64_t* seg_off;
};

// Multigrid parameters of a hierarchy, fixed when it is built (defaults,
// or the environment's BFMX_AMG_DEG, BFMX_AMG_DEG_COARSE, BFMX_AMG_DEG_DEEP, BFMX_AMG_RATIO, BFMX_AMG_COARSEST,
// BFMX_AMG_OMEGA).
struct AmgParams {
	int deg = 1;            // Chebyshev steps before and after each coarse correction (level 0);
	                        // kDegSpread when E spreads over more than kSpreadHetero
	int deg_coarse = 2;     // ... on level 1
	int deg_deep = 2;       // ... on the levels below it
	double ratio = 30.0;    // Chebyshev interval [lambda_max / ratio, lambda_max]
	int coarsest = 64;      // coarsen until every segment of a level has at most this many block rows
	double omega = 1.0;     // prolongator damping: omega * 4 / (3 lambda_max)
	int power = 10;         // power-iteration steps estimating lambda_max(D^-1 A); 0: Gershgorin bound
	int power_coarse = 10;  // ... on the coarser levels
	double seed_frac = 0.002;   // device aggregation: rows that may seed a region in the first round
	double seed_growth = 1.25;  // ... and that fraction's growth per round
	// device build: levels of at most this many rows, or whose segments all
	// are, are aggregated by host sweeps (the finest always, from the host
	// graph; 0: MIS-2 on the GPU at every level)
	int host_agg_rows = 131072;
	// device build: levels of at most this many rows that are swept on the
	// host get their transfer topology from the host as well (one upload
	// instead of the device build's chain of launches and read-backs, ~0.3 ms
	// a level; 0: the device build everywhere)
	int host_transfer_rows = 4096;
	// first build: workgroups at most of the finest level's prolongator
	// smoothing, A P and P^T (A P) kernels, which run on the second stream
	// beside the coarser levels' topology; a quarter of what the GPU holds
	// leaves the topology's small kernels room (2M-DOF beam, first solve,
	// on one MI355X: 15.5-15.7 ms uncapped, 15.35 at 512, 15.6-15.8 at 1024,
	// 16.1 at 256; on others within the spread between runs; 0: the natural
	// grid)
	int numeric_wgs = 512;
	// device build: from a level of at most this many rows on, the host build
	// (off: the host's Galerkin lists grow with the wide stencils of the
	// coarse levels; 7112 rows of the 2M-DOF beam's level 2 took 11.9 ms
	// there against about 2 ms for the device build of levels 2 to 4)
	int host_level_rows = 0;
};

struct Graph {
	hipGraphExec_t exec = nullptr;
	int precond = -1;
	int precision = -1;
	double rtol2 = -1;
	// the learned preconditioner's model a chunk was captured with: its
	// creation number (never reused, unlike a freed model's address) and the
	// parameter buffer the captured kernels read
	uint64_t model_gen = 0;
	const float* model_params = nullptr;
	int iters = 0;    // iterations in the chunk
};

}  // namespace bfmx

struct bfmx_model;
struct bfmx_trainset;
struct bfmx_batch;

struct bfmx_system {
	int device = 0;
	hipStream_t stream = nullptr;
	int kind = 0;
	int64_t n_nodes = 0;
	int64_t n_elems = 0;

	// renumbering: perm[old] = new, iperm[new] = old
	std::vector<int32_t> perm, iperm;
	std::vector<double> coords;      // renumbered, host copy
	std::vector<int64_t> adj_ptr;    // node adjacency (renumbered, sorted, incl. self)
	std::vector<int32_t> adj;

	// sliced ELL
	int64_t n_slices = 0, nnzb = 0, nnzb_pad = 0;
	std::vector<int32_t> slice_ptr;  // n_slices + 1, in blocks
	std::vector<int32_t> row_len;    // per block row
	std::vector<int32_t> diag_blk;   // block index of each row's diagonal block

	// elements in colour order; elem_order[pos] = original element index
	int n_colors = 0;
	std::vector<int64_t> color_ptr;
	std::vector<int32_t> elem_order;

	// edges (renumbered) and node -> incident edge list in edge order
	int64_t n_edges = 0;
	std::vector<int32_t> edge_nodes;
	std::vector<uint8_t> edge_bnd;
	std::vector<int64_t> inc_ptr;
	std::vector<int32_t> inc;

	// device buffers
	double* d_coords = nullptr;
	int32_t* d_elem = nullptr;       // kind per element, colour order
	int32_t* d_slots = nullptr;      // kind*kind block indices per element
	double* d_elemE = nullptr;
	double e_spread = 1.0;        // max / min of the per-element E of the last assembly (1: one E)
	int32_t* d_slice_ptr = nullptr;
	int32_t* d_row_len = nullptr;
	int32_t* d_diag_blk = nullptr;
	int32_t* d_col = nullptr;
	double* d_val = nullptr;         // 4 * nnzb_pad
	float* d_val32 = nullptr;        // FP32 copy for mixed precision (on demand)
	double* d_b = nullptr;           // assembled + condition-applied load
	double* d_bsolve = nullptr;      // right-hand side of the current solve
	uint8_t* d_fixed = nullptr;      // per DOF: constrained
	double* d_vfirst = nullptr;      // per DOF: value of the first constraint (lift)
	uint8_t* d_flags = nullptr;      // per node scratch: condition flags
	int32_t* d_list = nullptr;       // per node scratch: node lists
	int32_t* d_rows = nullptr;       // rows touched by Dirichlet elimination
	int32_t* d_edge_nodes = nullptr;
	uint8_t* d_edge_bnd = nullptr;
	int64_t* d_inc_ptr = nullptr;
	int32_t* d_inc = nullptr;

	// solver workspace
	double *d_x = nullptr, *d_r = nullptr, *d_z = nullptr, *d_q = nullptr;
	double* d_p[2] = {nullptr, nullptr};
	double* d_dinv = nullptr;        // 4 per node (block inverse) or 2 per node (point)
	double* d_partial = nullptr;     // per-workgroup partial sums
	int64_t partial_cap = 0;         // their number
	bfmx::Scalars* d_scal = nullptr;
	bfmx::Scalars* h_scal = nullptr; // pinned host mirror ring (4 entries)
	hipEvent_t ev[6] = {};
	std::vector<bfmx::Graph> graphs;   // captured chunks, one per solver configuration
	uint64_t solved_once = 0;          // bit p: a solve with preconditioner p has run (its chunks are captured from the next)
	std::vector<bfmx::AmgLevel> amg;   // multigrid hierarchy (built on first use)
	bfmx::AmgParams amg_par;           // its parameters
	int amg_wg = -1;                   // first level run by the one-workgroup coarse cycle (-1: none)
	bool amg_head0 = false;            // the finest level's numeric head was queued by the build
	// the hierarchy's first build queues each level's numeric set-up on this
	// second stream as soon as the level's topology is complete, so the GPU
	// runs it while the host builds the next level's topology
	hipStream_t amg_stream = nullptr;
	hipEvent_t amg_ev = nullptr;
	// generations of the matrix values: bumped by every change of d_val;
	// the FP32 copy and the multigrid's numeric set-up remember the one they
	// were made from and are redone only when it moved
	uint64_t val_gen = 1;
	uint64_t f32_gen = 0;
	uint64_t amg_num_gen = 0;
	bfmx::AmgLevelDesc* d_amg_desc = nullptr;   // descriptors of the levels from amg_wg on
	uint64_t* d_amg_pf = nullptr;      // addresses of the 128-byte lines those levels' arrays span
	int amg_npf = 0;
	// convergence segments: node ranges d_conv_seg[k] .. d_conv_seg[k+1] - 1
	// (borrowed) that must each meet the tolerance on their own (the union
	// system of a batch); d_conv_part holds their r.r and b.b at a check
	int n_conv_seg = 0;
	const int64_t* d_conv_seg = nullptr;
	double* d_conv_part = nullptr;
	// segmented flexible CG over the convergence segments (when each has at
	// least kBlock rows): the per-block view, per segment alpha, beta, p.q,
	// and the two-slot partial sums (4 per block)
	bfmx::SegView fcg_sv;
	int32_t* d_fcg_seg = nullptr;
	int32_t* d_fcg_split = nullptr;
	double* d_fcg_scal = nullptr;
	double* d_fcg_part = nullptr;
	bfmx_model* model = nullptr;       // learned preconditioner (borrowed)
	float* d_mp_feat = nullptr;        // its node features (8 per node)
	float* d_mp_h[4] = {nullptr, nullptr, nullptr, nullptr};   // hidden features (FP32, 64 per node)
	double* d_mp_z = nullptr;          // its output (FP64, 2 per node)

	bool assembled = false;
	bool finalized = false;
	bool borrowed = false;   // belongs to a batch built from meshes: device arrays inside its arenas
	int stress = 0;
	std::vector<uint8_t> dirichlet_nodes;  // union of Dirichlet-flagged nodes (renumbered)
	int64_t last_iters = 0;
	bool has_u = false;
};

// Learned preconditioner: parameters, optimiser state and training workspace.
struct bfmx_model {
	int device = 0;
	uint64_t gen = 0;           // creation number, unique in the process (keys captured graphs)
	hipStream_t stream = nullptr;
	float* d_params = nullptr;
	float* d_grad = nullptr;
	float* d_m = nullptr;
	float* d_v = nullptr;
	uint16_t* d_W = nullptr;    // 3 x 128 x 64 BF16 (layer weights)
	uint16_t* d_Wt = nullptr;   // 3 x 64 x 128 BF16 (transposed)
	int64_t adam_t = 0;
	// training workspace (rows capacity): forward passes of x and of g
	// (activations h, messages m), the adjoint's buffers, gradients
	int64_t cap = 0;
	uint16_t* h[4] = {nullptr, nullptr, nullptr, nullptr};
	uint16_t* m[3] = {nullptr, nullptr, nullptr};
	uint16_t* hg[4] = {nullptr, nullptr, nullptr, nullptr};
	uint16_t* mg[3] = {nullptr, nullptr, nullptr};
	uint16_t* adj[4] = {nullptr, nullptr, nullptr, nullptr};
	float* x = nullptr;
	float* xg = nullptr;
	float* z = nullptr;
	float* dA = nullptr;
	float* dB = nullptr;
	float* dM = nullptr;
	float* up = nullptr;
	float* partial = nullptr;   // gradient partials
	double* lpart = nullptr;    // loss partials
	double* loss = nullptr;
};

// Recorded (residual, error) pairs of one or more systems.
struct bfmx_trainset {
	struct Entry {
		bfmx_system* sys = nullptr;
		int64_t n = 0;          // nodes
		int P = 0;              // pairs
		float* d_val32 = nullptr;   // FP32 matrix snapshot
		float* d_feat = nullptr;    // node features
		float* d_r = nullptr;       // P x 2n
		float* d_e = nullptr;
		float* d_s = nullptr;       // P
		float* d_den = nullptr;     // P
	};
	std::vector<Entry> entries;
	int64_t pairs = 0;
};

namespace bfmx {

// prep.hip (host)
// Host half of build_topology: renumbering, adjacency, the sliced-ELL
// pattern (col: nnzb_pad block columns) and the colour-ordered elements
// (el_col: kind per element), without device work.
int prepare_topology(bfmx_system* s, const double* coords, const size_t* elems, int renumber,
                     std::vector<int32_t>& col, std::vector<int32_t>& el_col);
int build_topology(bfmx_system* s, const double* coords, const size_t* elems, int renumber);
// While alive, the mesh preparation called on this thread uses at most
// `limit` host threads (callers that already run it on many threads at once).
struct PrepThreads {
	int prev;
	explicit PrepThreads(int limit);
	~PrepThreads();
};
// Sliced-ELL pattern of s->adj_ptr / s->adj (slice_ptr, row_len, diag_blk,
// nnzb, nnzb_pad; col: the block column of every slot).
int sell_layout(bfmx_system* s, std::vector<int32_t>& col);
// Host half of set_edges: renumbered edges and the node -> edge incidence.
int prepare_edges(bfmx_system* s, size_t n_edges, const size_t* edge_nodes, const uint8_t* edge_boundary);
int set_edges(bfmx_system* s, size_t n_edges, const size_t* edge_nodes, const uint8_t* edge_boundary);
// assemble.hip
int compute_slots(bfmx_system* s);
int assemble(bfmx_system* s, int stress, double E, double nu, double rho, const double* elemE,
             const double* node_force, double gx, double gy);
int apply_condition(bfmx_system* s, int kind, double value, const uint8_t* flags);
int finalize_conditions(bfmx_system* s);
// amg.hip
int amg_build(bfmx_system* s);
void amg_release(bfmx_system* s);
int amg_setup_numeric(bfmx_system* s);
// V-cycle on level 0 for the FP64 residual r; returns the FP32 correction z.
// The last kernel also writes per-workgroup partial sums of z . q into
// partial (*n_partials of them), the flexible-CG orthogonalisation. Every
// kernel returns at once when S->done.
float* amg_vcycle(bfmx_system* s, const double* r, const double* q, double* partial, const Scalars* S,
                  int* n_partials);
// mpnn.hip
int mpnn_n_params();
int model_create(bfmx_model** out, uint64_t seed);
void model_destroy(bfmx_model* m);
int model_get_params(bfmx_model* m, float* p);
int model_set_params(bfmx_model* m, const float* p);
int model_loss_grad(bfmx_model* m, bfmx_trainset* ts, int entry, int p0, int B);
int model_train(bfmx_model* m, bfmx_trainset* ts, int steps, int batch, float lr, uint64_t seed, double* losses);
void trainset_destroy(bfmx_trainset* ts);
int trainset_add(bfmx_trainset* ts, bfmx_system* s, int max_pairs, int every, double rtol);
int trainset_get(bfmx_trainset* ts, int entry, float* feat, float* r, float* e, float* s_den);
int mpnn_prepare(bfmx_system* s);
int mpnn_setup_numeric(bfmx_system* s);
// One application of the learned preconditioner to the FP64 residual r;
// returns z (FP64, 2 per node: the base step plus the network's FP32
// correction; the base is the FP64 block-Jacobi step of the solver's inverses
// in s->d_dinv, or, given, the FP32 vector base, the multigrid V-cycle's
// output); the last kernel writes *n_partials partial sums of z . q. Every
// kernel returns at once when S->done.
double* mpnn_apply(bfmx_system* s, const double* r, const double* q, double* partial, const Scalars* S,
                   int* n_partials, const float* base = nullptr);
// Rebases the pairs of every entry for the multigrid-based preconditioner
// (bfmx_trainset_rebase).
int trainset_rebase(bfmx_trainset* ts);
// The multigrid hierarchy of a system set up as a solve sets it up.
int prepare_amg(bfmx_system* s);
void mpnn_release(bfmx_system* s);
// batch.hip
int batch_n_configs();
int batch_n_classes();
int batch_create(bfmx_batch** out, size_t n, bfmx_system* const* systems);
void batch_destroy(bfmx_batch* b);
int batch_solve(bfmx_batch* b, int pre, double rtol, int64_t max_iter, const int32_t* picks, bfmx_stats_t* stats,
                double* total_ms);
int batch_autotune(bfmx_batch* b, int pre, int sample, int iters, int32_t* picks, double* times_ms);
int batch_get_u(bfmx_batch* b, size_t k, double* u);
int batch_create_planar(bfmx_batch** out, size_t n, const bfmx_instance_t* inst, int renumber);
int batch_system(bfmx_batch* b, size_t k, bfmx_system** s);
// api.hip: the calling thread's last error message
const char* last_error_text();
// pcg.hip
// Plain Jacobi PCG from x = 0 on the current right-hand side, calling rec(p)
// after iteration p * every (p = 0 .. P-1) with the iterate in s->d_x and
// its residual in s->d_r.
int record_pcg(bfmx_system* s, int P, int every, const std::function<int(int)>& rec);
// FP32 copy of the current matrix values into dst (4 * nnzb_pad floats).
int snapshot_f32(bfmx_system* s, float* dst);
int ensure_workspace_public(bfmx_system* s);
// 2x2 block-Jacobi inverses of the current matrix into s->d_dinv.
void precond_block_jacobi(bfmx_system* s);
int ensure_f32_public(bfmx_system* s);
int solve(bfmx_system* s, const double* b_in, int precond, int precision, double rtol, size_t max_iter,
          bfmx_stats_t* stats);
int time_spmv(bfmx_system* s, int precision, int reps, double* avg_ms, double* bytes);
int time_kernel(bfmx_system* s, int kernel, int reps, double* avg_ms, double* bytes);
// amg.hip: reps launches of one finest-level smoothing step (set-up done)
int amg_time_smoother(bfmx_system* s, int reps, hipEvent_t e0, hipEvent_t e1, double* bytes);
void release_solver(bfmx_system* s);
// Solver buffers of a PCG run held by the caller (2 n doubles each for
// x, r, z, q, p0, p1, bsolve; 4 n for dinv; partial_cap partial sums).
struct PcgBuffers {
	double *x, *r, *z, *q, *p0, *p1, *dinv, *bsolve, *partial;
	int64_t partial_cap;
	Scalars* scal;
};
// Point or block Jacobi PCG (FP64) from u = 0 for the right-hand side in
// w.bsolve, in the buffers w, with kernel launches only (no allocation, no
// graph capture); *out receives the final scalars, the solution stays in
// w.x (internal node order). The system's own solver buffers are untouched.
int pcg_in_buffers(bfmx_system* s, PcgBuffers const& w, int precond, double rtol, int64_t cap, Scalars* out);
// noalloc.hip: bfmx_system_solve_noalloc, and the largest system (nodes) its
// fixed scratch array holds.
int solve_noalloc(bfmx_system* s, const double* b_in, double* u_out, int precond, double rtol, size_t max_iter,
                  bfmx_stats_t* stats);
int64_t noalloc_max_nodes();
// Destroys the system's captured chunks of one preconditioner (all with -1),
// after the system's stream has drained.
void drop_graphs(bfmx_system* s, int precond);
// false when the environment sets BFMX_GRAPHS=0: launches go straight to the
// stream instead of through captured graphs (for tools that do not follow
// graph launches)
bool graphs_enabled();

template <typename T>
int dalloc(T** p, size_t n) {
	if (n == 0) {
		n = 1;
	}
	hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
	if (e != hipSuccess) {
		*p = nullptr;
		set_error("hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
		return e == hipErrorOutOfMemory ? BFMX_E_NOMEM : BFMX_E_HIP;
	}
	return BFMX_OK;
}

template <typename T>
void dfree(T*& p) {
	if (p) {
		(void) hipFree(p);
		p = nullptr;
	}
}

// Device blocks from a per-device cache (pool.hip): pool_free_bytes takes a
// block back once every stream that used it has drained; a pointer the pool
// did not hand out goes to hipFree.
int pool_alloc_bytes(void** p, size_t bytes);
void pool_free_bytes(void* p);

// A pinned (device-visible) host buffer of at least `bytes`, one per host
// thread and slot (0 .. kStagingSlots - 1), grown as needed and kept until
// the process ends; valid until the thread's next call for the same slot.
// Slots of the multigrid build: 0 a level's graph, 1 counts read back.
constexpr int kStagingSlots = 2;
int pinned_staging(void** p, size_t bytes, int slot = 0);

// Small pinned host blocks from a process-wide cache (pool.hip); a block goes
// back with the size it was asked for, once no pending copy uses it.
int pinned_small_alloc(void** p, size_t bytes);
void pinned_small_free(void* p, size_t bytes);

template <typename T>
int palloc(T** p, size_t n) {
	void* q = nullptr;
	int const rc = pool_alloc_bytes(&q, (n ? n : 1) * sizeof(T));
	*p = static_cast<T*>(q);
	return rc;
}

template <typename T>
void pfree(T*& p) {
	if (p) {
		pool_free_bytes(const_cast<void*>(static_cast<const void*>(p)));
		p = nullptr;
	}
}

// Workgroup index remapped so that each XCD (workgroups are dealt round-robin
// over the 8 XCDs) processes one contiguous range of rows; neighbouring rows
// then share the XCD's L2 for the gathered vector entries.
__device__ __forceinline__ int xcd_remap(int wg, int nwg) {
	constexpr int kXcd = 8;
	int const per = nwg / kXcd;
	int const rem = nwg % kXcd;
	int const xcd = wg % kXcd;
	int const local = wg / kXcd;
	// XCD x takes per + 1 workgroups when x < rem, per otherwise; its range
	// starts at x per + min(x, rem) (no branch, so the kernel's argument loads
	// are issued together).
	return xcd * per + min(xcd, rem) + local;
}

// The done flag, read so that the loads the caller issued before this call
// are in flight together with it, one memory round trip instead of two at the
// start of a latency-bound kernel. The flag is read by a vector load (its
// index is a zero the compiler cannot see through): a scalar load would be
// waited for together with the kernel arguments that those loads' addresses
// need. The empty asm consumes the flag and the values, so the compiler can
// neither sink the loads below the caller's early return nor issue the
// flag's load after waiting for them.
template <typename T>
__device__ __forceinline__ void issued_with(int done, T const& v) {
	asm volatile("" ::"v"(done), "v"(v));
}
template <typename... T>
__device__ __forceinline__ int done_after(const Scalars* __restrict__ S, T const&... v) {
	int zero;
	asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
	int const done = (&S->done)[zero];
	(issued_with(done, v), ...);
	return done;
}

}  // namespace bfmx
