// Smoothed aggregation multigrid: the FP32 preconditioner of the
// flexible-CG path.
//
// Hierarchy topology (built once per system on the host from the node graph,
// by up to kBuildThreads threads over chunks of whole segments):
//   * segments: row ranges that no entry couples (the parts of a mesh, or the
//     systems of a batch solved as one); aggregates never span two, so the
//     segments follow the hierarchy down, and coarsening stops when every
//     segment has at most AmgParams::coarsest rows (or it stalls);
//   * aggregates: a node whose neighbours are all free starts an aggregate
//     with them (about 3 x 3 nodes on a quad mesh); a leftover node joins the
//     aggregate of its first aggregated neighbour;
//   * the prolongator pattern: row i couples to the aggregates of its
//     neighbours (P = (I - w D^-1 A) P_t has the sparsity of A P_t), stored
//     as CSR of bs x 3 blocks, with its transpose's pattern for the
//     restriction (P's entries column by column, with their rows);
//   * the coarse block pattern (that of P^T A P) in sliced ELL, and for every
//     coarse block slot on or above the diagonal the list of its
//     contributions (fine block A_ij, entry P_iI, entry P_jJ), so the
//     Galerkin product is formed without atomics and in a fixed order, each
//     block also stored as the transpose of its mirror; the coarse pattern is
//     the next level's graph.
// Numeric setup (every solve, on the GPU, because the Dirichlet set enters),
// level by level:
//   * level-0 modes B: the three rigid-body motions (two translations and the
//     rotation, in coordinates centred on the mesh and scaled to its size),
//     with the rows of constrained DOFs zeroed;
//   * inverse diagonal blocks, the Gershgorin bound of D^-1 A (the upper end
//     of the level's Chebyshev interval) and a power-iteration estimate
//     lambda of its largest eigenvalue (which sets the prolongator damping);
//   * per aggregate, a Cholesky QR of its rows of B: P_t = B R^-1 (orthonormal
//     columns) and the coarse modes B_c = R. A column whose remainder is below
//     kDropTol of its norm (a mode that the constrained DOFs make dependent) is
//     dropped: its coarse DOF is decoupled;
//   * the smoothed prolongator P = (I - w D^-1 A) P_t, w = 4 / (3 lambda),
//     stored FP32 (Vanek, Mandel and Brezina, Computing 56 (1996) 179-196),
//     also in column order (R = P^T row by row), which the restriction reads
//     contiguously;
//   * the Galerkin operator P^T A P in FP64 with an FP32 copy, and the dense
//     Gauss-Jordan inverse of every segment of the coarsest level.
// BFMX_AMG_SMOOTH=0 in the environment (read when a hierarchy is built)
// keeps P = P_t: plain aggregation, for comparisons.
// V-cycle (FP32 vectors and matrices, FP32 arithmetic): Chebyshev smoothing
// with block Jacobi (AmgParams::deg steps before and after the coarse
// correction on the finest level, deg_coarse / deg_deep below), residual,
// restriction P^T r, recursion, prolongation; the levels of at most kWgRows
// rows run in one workgroup; the coarsest level applies its dense inverses.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>

#include "amg_topo.h"
#include "internal.h"

namespace bfmx {

namespace {

constexpr int kCoarseDeg = 16;  // Chebyshev steps on a coarsest level too large for the dense inverse
constexpr int kMaxDeg = 16;
constexpr int kMaxLevels = 12;
constexpr int kDenseMax = 768;  // largest coarsest segment (DOFs) inverted densely
constexpr int64_t kDenseTotal = int64_t(1) << 27;   // entries of all the coarsest's dense inverses (1 GiB)
constexpr double kCoarseRatio = 200.0;  // Chebyshev interval of an iterated coarsest level
// A level's scalars: [0] lambda_max of its Chebyshev interval (the
// Gershgorin bound), [1 .. 2 kMaxDeg] the Chebyshev coefficients, then the
// power iteration's current scale (kScalScale) and its lambda_max estimate
// (kScalLambda), which sets the prolongator damping.
constexpr int kScalScale = 2 + 2 * kMaxDeg;
constexpr int kScalLambda = kScalScale + 1;
constexpr int kScalSize = kScalLambda + 1;
constexpr double kDropTol = 1e-10;      // relative remainder below which a mode column is dropped
constexpr int kGj = 1024;               // threads of the Gauss-Jordan workgroup (one segment)
constexpr int kGjSeg = 256;             // ... (one workgroup per segment of many)
constexpr int kBuildThreads = 16;       // host threads building a level's transfer operators
constexpr double kSpreadHetero = 10.0;  // per-element E spread (max / min) above which ...
constexpr int kDegSpread = 3;           // ... the finest level takes this many Chebyshev steps per side
// Blocks of a row whose loads the smoothing and residual row loops keep in
// flight together, for 2x2 (finest level) and 3x3 blocks. Measured (2M-DOF
// Q1 beam solve / 4096-panel P1 batch): no unrolling 12.20 ms / 797 ms,
// 2 and 2: 11.82 / 773, 3 and 3: 11.52 / 799, 4 and 4: 11.72 / 783,
// 2 and 3: 11.63 / 782.
#ifndef BFMX_ROW_UNROLL2
#define BFMX_ROW_UNROLL2 3
#endif
#ifndef BFMX_ROW_UNROLL3
#define BFMX_ROW_UNROLL3 3
#endif
template <int BS>
constexpr int kRowUnroll = BS == 2 ? BFMX_ROW_UNROLL2 : BFMX_ROW_UNROLL3;

int grid(int64_t n) {
	return static_cast<int>((n + kBlock - 1) / kBlock);
}

// Index of entry (a, b) of block slot blk (lane r = row & 63 of its row).
template <int BS>
struct Blk;
template <>
struct Blk<2> {   // the system's split layout: top rows, then bottom rows of 64 blocks
	__device__ __forceinline__ static int64_t at(int64_t blk, int r, int a, int b) {
		return 4 * blk - 2 * r + a * 2 * kSlice + b;
	}
};
template <>
struct Blk<3> {   // per block row a, the three entries of the 64 blocks of a chunk lane by lane
	__device__ __forceinline__ static int64_t at(int64_t blk, int r, int a, int b) {
		return (blk - r) * 9 + a * 3 * kSlice + 3 * r + b;
	}
};

// Loads block slot blk into A (two 8-byte loads per row for 2x2 blocks).
template <int BS, typename T>
__device__ __forceinline__ void ld_block(const T* __restrict__ val, int64_t blk, int r, T (&A)[BS][BS]) {
	if constexpr (BS == 2) {
		typedef T t2 __attribute__((ext_vector_type(2)));
		t2 const u = *reinterpret_cast<const t2*>(val + Blk<2>::at(blk, r, 0, 0));
		t2 const v = *reinterpret_cast<const t2*>(val + Blk<2>::at(blk, r, 1, 0));
		A[0][0] = u.x;
		A[0][1] = u.y;
		A[1][0] = v.x;
		A[1][1] = v.y;
	} else {
#pragma unroll
		for (int a = 0; a < BS; a++) {
#pragma unroll
			for (int b = 0; b < BS; b++) {
				A[a][b] = val[Blk<BS>::at(blk, r, a, b)];
			}
		}
	}
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v = fmax(v, __shfl_xor(v, off, 64));
	}
	return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v += __shfl_xor(v, off, 64);
	}
	return v;
}

// Per-workgroup sum (SUM) or maximum of v, written to partial[blockIdx.x].
template <bool SUM>
__device__ __forceinline__ void wg_partial(double v, double* __restrict__ partial) {
	__shared__ double lds[kBlock / 64];
	v = SUM ? wave_sum(v) : wave_max(v);
	if ((threadIdx.x & 63) == 0) {
		lds[threadIdx.x >> 6] = v;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double m = lds[0];
		for (int w = 1; w < kBlock / 64; w++) {
			m = SUM ? m + lds[w] : fmax(m, lds[w]);
		}
		partial[blockIdx.x] = m;
	}
}

// ------------------------------------------------------------ numeric setup

// Level-0 modes with the rows of constrained DOFs zeroed.
__global__ __launch_bounds__(kBlock) void k_fine_modes(int nb, const double* __restrict__ B0,
                                                       const uint8_t* __restrict__ fixed, double* __restrict__ B) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= nb) {
		return;
	}
#pragma unroll
	for (int a = 0; a < 2; a++) {
		bool const f = fixed[2 * i + a] != 0;
#pragma unroll
		for (int c = 0; c < 3; c++) {
			B[6 * i + 3 * a + c] = f ? 0.0 : B0[6 * i + 3 * a + c];
		}
	}
}

// Per-workgroup extents of the node coordinates: part[4 b + {0, 1, 2, 3}] =
// min x, min y, max x, max y over the workgroup's grid-stride rows.
constexpr int kBboxGrid = 256;
__global__ __launch_bounds__(kBlock) void k_bbox(int64_t n, const double* __restrict__ xy, double* __restrict__ part) {
	__shared__ double lds[4][kBlock / 64];
	double v[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
	for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
	     i += static_cast<int64_t>(gridDim.x) * kBlock) {
		double const x = xy[2 * i], y = xy[2 * i + 1];
		v[0] = fmin(v[0], x);
		v[1] = fmin(v[1], y);
		v[2] = fmax(v[2], x);
		v[3] = fmax(v[3], y);
	}
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v[0] = fmin(v[0], __shfl_xor(v[0], off, 64));
		v[1] = fmin(v[1], __shfl_xor(v[1], off, 64));
		v[2] = fmax(v[2], __shfl_xor(v[2], off, 64));
		v[3] = fmax(v[3], __shfl_xor(v[3], off, 64));
	}
	if ((threadIdx.x & 63) == 0) {
#pragma unroll
		for (int k = 0; k < 4; k++) {
			lds[k][threadIdx.x >> 6] = v[k];
		}
	}
	__syncthreads();
	if (threadIdx.x < 4) {
		int const k = threadIdx.x;
		double m = lds[k][0];
		for (int w = 1; w < kBlock / 64; w++) {
			m = k < 2 ? fmin(m, lds[k][w]) : fmax(m, lds[k][w]);
		}
		part[4 * blockIdx.x + k] = m;
	}
}

// Level-0 rigid-body modes (two translations, the rotation) in coordinates
// centred on the mesh and scaled to its half-size, from the workgroup
// extents of k_bbox (each thread folds the kBboxGrid partials: they sit in
// L2). Rows: (1, 0, -y), (0, 1, x).
__global__ __launch_bounds__(kBlock) void k_modes0(int64_t n, const double* __restrict__ part,
                                                   const double* __restrict__ xy, double* __restrict__ B0) {
	__shared__ double ext[4];
	if (threadIdx.x < 4) {
		int const k = threadIdx.x;
		double m = part[k];
		for (int b = 1; b < kBboxGrid; b++) {
			m = k < 2 ? fmin(m, part[4 * b + k]) : fmax(m, part[4 * b + k]);
		}
		ext[k] = m;
	}
	__syncthreads();
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	double const cx = 0.5 * (ext[0] + ext[2]), cy = 0.5 * (ext[1] + ext[3]);
	double L = 0.5 * fmax(ext[2] - ext[0], ext[3] - ext[1]);
	if (!(L > 0)) {
		L = 1;
	}
	double const x = (xy[2 * i] - cx) / L, y = (xy[2 * i + 1] - cy) / L;
	double* b = B0 + 6 * i;
	b[0] = 1;
	b[1] = 0;
	b[2] = -y;
	b[3] = 0;
	b[4] = 1;
	b[5] = x;
}

// One aggregate K per thread: G = sum_i B_i^T B_i, Cholesky G = R^T R with
// dropped columns, P_i = B_i R^-1 on the kept columns (0 on dropped ones),
// B_c[K] = R.
template <int BSF>
__global__ __launch_bounds__(kBlock) void k_tentative(int nc, const int32_t* __restrict__ agg_ptr,
                                                      const int32_t* __restrict__ agg_nodes,
                                                      const double* __restrict__ B, float* __restrict__ P,
                                                      double* __restrict__ Bc) {
	int const K = blockIdx.x * kBlock + threadIdx.x;
	if (K >= nc) {
		return;
	}
	int32_t const beg = agg_ptr[K], end = agg_ptr[K + 1];
	double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
	for (int32_t t = beg; t < end; t++) {
		const double* b = B + BSF * 3 * static_cast<int64_t>(agg_nodes[t]);
#pragma unroll
		for (int a = 0; a < BSF; a++) {
#pragma unroll
			for (int c = 0; c < 3; c++) {
#pragma unroll
				for (int d = 0; d < 3; d++) {
					G[c][d] += b[3 * a + c] * b[3 * a + d];
				}
			}
		}
	}
	double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
	bool keep[3];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		double s = G[k][k];
#pragma unroll
		for (int m = 0; m < k; m++) {
			s -= R[m][k] * R[m][k];
		}
		keep[k] = G[k][k] > 0 && s > kDropTol * G[k][k];
		if (keep[k]) {
			double const rkk = sqrt(s);
			R[k][k] = rkk;
#pragma unroll
			for (int j = k + 1; j < 3; j++) {
				double t = G[k][j];
#pragma unroll
				for (int m = 0; m < k; m++) {
					t -= R[m][k] * R[m][j];
				}
				R[k][j] = t / rkk;
			}
		}
	}
	for (int32_t t = beg; t < end; t++) {
		int64_t const i = agg_nodes[t];
		const double* b = B + BSF * 3 * i;
#pragma unroll
		for (int a = 0; a < BSF; a++) {
			double q[3];
#pragma unroll
			for (int k = 0; k < 3; k++) {
				double v = b[3 * a + k];
#pragma unroll
				for (int m = 0; m < k; m++) {
					v -= q[m] * R[m][k];
				}
				q[k] = keep[k] ? v / R[k][k] : 0.0;
				P[BSF * 3 * i + 3 * a + k] = static_cast<float>(q[k]);
			}
		}
	}
#pragma unroll
	for (int k = 0; k < 3; k++) {
#pragma unroll
		for (int c = 0; c < 3; c++) {
			Bc[9 * static_cast<int64_t>(K) + 3 * k + c] = R[k][c];
		}
	}
}

// Inverse of the diagonal block of row i in FP64 (a zero diagonal entry,
// a decoupled DOF, is treated as 1).
template <int BS>
__device__ __forceinline__ void diag_inverse(const double* __restrict__ val, int32_t dblk, int r,
                                             double (&Di)[BS][BS]) {
	double D[BS][BS];
	ld_block<BS>(val, dblk, r, D);
#pragma unroll
	for (int a = 0; a < BS; a++) {
		if (D[a][a] == 0) {
			D[a][a] = 1;
		}
	}
	if constexpr (BS == 2) {
		double const inv = 1.0 / (D[0][0] * D[1][1] - D[0][1] * D[1][0]);
		Di[0][0] = D[1][1] * inv;
		Di[0][1] = -D[0][1] * inv;
		Di[1][0] = -D[1][0] * inv;
		Di[1][1] = D[0][0] * inv;
	} else {
		double const c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
		double const c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
		double const c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
		double const inv = 1.0 / (D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02);
		Di[0][0] = c00 * inv;
		Di[1][0] = c01 * inv;
		Di[2][0] = c02 * inv;
		Di[0][1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * inv;
		Di[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * inv;
		Di[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * inv;
		Di[0][2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * inv;
		Di[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * inv;
		Di[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * inv;
	}
}

// Smoothed prolongator, one fine row i per thread: entry k of row i (the
// aggregate J = pcol[k]) is [J = agg(i)] Pt_i - w D_i^-1 sum over the row's
// blocks A_ij with agg(j) = J of A_ij Pt_j, w = omega 4 / (3 lambda_max(D^-1 A))
// with lambda_max the level's power-iteration estimate (or Gershgorin bound);
// omega = 0 keeps the tentative prolongator. Padding blocks are zero and add
// nothing. The off-diagonal blocks are read from the FP32 copy (P is stored
// FP32), the diagonal inverse from the FP64 values.
constexpr int kSmoothRegs = 6;   // entries per row summed in registers in one pass
template <int BS>
__global__ __launch_bounds__(kBlock) void k_smooth_p(int nb, const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ val,
                                                     const float* __restrict__ val32,
                                                     const int32_t* __restrict__ diag_blk,
                                                     const float* __restrict__ Pt, const int32_t* __restrict__ pptr,
                                                     const uint8_t* __restrict__ ppos, const double* __restrict__ scal,
                                                     double omega, float* __restrict__ P) {
	auto row = [&](int const i) {
		int const r = i & 63;
		double const w = omega * 4.0 / (3.0 * scal[kScalLambda]);
		int32_t const dblk = diag_blk[i];
		double Di[BS][BS];
		diag_inverse<BS>(val, dblk, r, Di);
		int const self = ppos[dblk];
		int32_t const p0 = pptr[i], np = pptr[i + 1] - p0;
		int32_t const sbeg = slice_ptr[i >> 6] + r, send = slice_ptr[(i >> 6) + 1];
		// writes entry k of the row from its sum acc = sum A_ij Pt_j
		auto store = [&](int k, const double (&acc)[BS][3]) {
			float* out = P + BS * 3 * static_cast<int64_t>(p0 + k);
	#pragma unroll
			for (int a = 0; a < BS; a++) {
	#pragma unroll
				for (int c = 0; c < 3; c++) {
					double t = 0;
	#pragma unroll
					for (int m = 0; m < BS; m++) {
						t += Di[a][m] * acc[m][c];
					}
					double v = -w * t;
					if (k == self) {
						v += static_cast<double>(Pt[BS * 3 * static_cast<int64_t>(i) + 3 * a + c]);
					}
					out[3 * a + c] = static_cast<float>(v);
				}
			}
		};
		// A_ij Pt_j for block blk (FP32 values, FP64 products)
		auto term = [&](int32_t blk, double (&T)[BS][3]) {
			int64_t const j = col[blk];
			float A[BS][BS];
			ld_block<BS>(val32, blk, r, A);
	#pragma unroll
			for (int a = 0; a < BS; a++) {
	#pragma unroll
				for (int c = 0; c < 3; c++) {
					double t = 0;
	#pragma unroll
					for (int m = 0; m < BS; m++) {
						t += static_cast<double>(A[a][m]) * static_cast<double>(Pt[BS * 3 * j + 3 * m + c]);
					}
					T[a][c] = t;
				}
			}
		};
		if (np <= kSmoothRegs) {
			// one pass over the row, each block's term added to its entry's sum
			// (the entry is selected, not indexed, so the sums stay in registers)
			double acc[kSmoothRegs][BS][3] = {};
			if (w != 0) {
	#pragma unroll 2
				for (int32_t blk = sbeg; blk < send; blk += kSlice) {
					int const p = ppos[blk];
					double T[BS][3];
					term(blk, T);
	#pragma unroll
					for (int k = 0; k < kSmoothRegs; k++) {
	#pragma unroll
						for (int a = 0; a < BS; a++) {
	#pragma unroll
							for (int c = 0; c < 3; c++) {
								acc[k][a][c] += p == k ? T[a][c] : 0.0;
							}
						}
					}
				}
			}
	#pragma unroll
			for (int k = 0; k < kSmoothRegs; k++) {
				if (k < np) {
					store(k, acc[k]);
				}
			}
			return;
		}
		// rows coupled to more aggregates: one pass per entry
		for (int k = 0; k < np; k++) {
			double acc[BS][3] = {};
			if (w != 0) {
				for (int32_t blk = sbeg; blk < send; blk += kSlice) {
					if (ppos[blk] != k) {
						continue;
					}
					double T[BS][3];
					term(blk, T);
	#pragma unroll
					for (int a = 0; a < BS; a++) {
	#pragma unroll
						for (int c = 0; c < 3; c++) {
							acc[a][c] += T[a][c];
						}
					}
				}
			}
			store(k, acc);
		}
	};
	// a grid-stride loop over the rows: the grid may be capped, leaving room for other work
	for (int64_t vb = blockIdx.x; vb * kBlock < nb; vb += gridDim.x) {
		int64_t const i = vb * kBlock + threadIdx.x;
		if (i < nb) {
			row(static_cast<int>(i));
		}
	}
}

// R[t] = P[rent[t]]: P's entries in column order, so that the restriction
// reads a coarse row's entries contiguously (written here, contiguously,
// rather than by k_smooth_p, whose stores in that order would scatter).
template <int BS>
__global__ __launch_bounds__(kBlock) void k_transpose_p(int64_t nnzp, const int32_t* __restrict__ rent,
                                                        const float* __restrict__ P, float* __restrict__ R) {
	int64_t const t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	if (t >= nnzp) {
		return;
	}
	const float* src = P + BS * 3 * static_cast<int64_t>(rent[t]);
	float v[BS * 3];
#pragma unroll
	for (int k = 0; k < BS * 3; k++) {
		v[k] = src[k];
	}
#pragma unroll
	for (int k = 0; k < BS * 3; k++) {
		R[BS * 3 * t + k] = v[k];
	}
}

// Coarse block slot b = sum over its contributions (fine block A_ij, entries
// pi = P_iI and pj = P_jJ) of P_iI^T A_ij P_jJ, in FP64 (P as stored, FP32).
// kGalLanes lanes share a slot: lane l takes contributions l, l + kGalLanes,
// ..., and the lanes' sums are combined by a fixed butterfly. P^T A P is
// symmetric: only the slots (I, J) with J >= I have contributions, and each
// also writes its transpose into the slot of (J, I) (mirror[b]; a slot with
// J < I has mirror -2 and is left to its mirror; a padding slot has -1 and
// receives zeros).
// Lanes per coarse slot and unrolling of their term loop: 16 and none
// measured best (2M beam solve 11.15 ms; 8 lanes 11.50, 16 unrolled by 2
// 11.15, 32 lanes 11.23, 32 unrolled by 2 11.35).
#ifndef BFMX_GAL_LANES
#define BFMX_GAL_LANES 16
#endif
#ifndef BFMX_GAL_UNROLL
#define BFMX_GAL_UNROLL 1
#endif
constexpr int kGalLanes = BFMX_GAL_LANES;
constexpr int kGalUnroll = BFMX_GAL_UNROLL;

template <int BSF, typename FT>
__global__ __launch_bounds__(kBlock) void k_galerkin(int64_t nslots, const int32_t* __restrict__ mirror,
                                                     const int64_t* __restrict__ gal_ptr,
                                                     const int32_t* __restrict__ gal_blk,
                                                     const int32_t* __restrict__ gal_pi,
                                                     const int32_t* __restrict__ gal_pj,
                                                     const FT* __restrict__ fval, const float* __restrict__ P,
                                                     double* __restrict__ cval64, float* __restrict__ cval32) {
	int64_t const b = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGalLanes;
	int const lane = threadIdx.x % kGalLanes;
	int32_t const mb = b < nslots ? mirror[b] : -2;
	bool const live = mb != -2;   // every lane stays for the butterfly
	double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
	int64_t const e0 = live ? gal_ptr[b] : 0, e1 = live ? gal_ptr[b + 1] : 0;
#pragma unroll kGalUnroll
	for (int64_t e = e0 + lane; e < e1; e += kGalLanes) {
		int32_t const blk = gal_blk[e];
		FT Af[BSF][BSF];
		ld_block<BSF>(fval, blk, blk & 63, Af);
		double A[BSF][BSF];
#pragma unroll
		for (int a = 0; a < BSF; a++) {
#pragma unroll
			for (int c = 0; c < BSF; c++) {
				A[a][c] = static_cast<double>(Af[a][c]);
			}
		}
		const float* pi = P + BSF * 3 * static_cast<int64_t>(gal_pi[e]);
		const float* pj = P + BSF * 3 * static_cast<int64_t>(gal_pj[e]);
		double T[BSF][3];   // A_ij P_jJ
#pragma unroll
		for (int a = 0; a < BSF; a++) {
#pragma unroll
			for (int c = 0; c < 3; c++) {
				double t = 0;
#pragma unroll
				for (int k = 0; k < BSF; k++) {
					t += A[a][k] * static_cast<double>(pj[3 * k + c]);
				}
				T[a][c] = t;
			}
		}
#pragma unroll
		for (int a = 0; a < 3; a++) {
#pragma unroll
			for (int c = 0; c < 3; c++) {
				double t = 0;
#pragma unroll
				for (int k = 0; k < BSF; k++) {
					t += static_cast<double>(pi[3 * k + a]) * T[k][c];
				}
				C[a][c] += t;
			}
		}
	}
#pragma unroll
	for (int a = 0; a < 3; a++) {
#pragma unroll
		for (int c = 0; c < 3; c++) {
			double v = C[a][c];
#pragma unroll
			for (int off = kGalLanes / 2; off > 0; off >>= 1) {
				v += __shfl_xor(v, off, 64);
			}
			C[a][c] = v;
		}
	}
	if (!live || lane != 0) {
		return;
	}
	int const r = static_cast<int>(b & 63);
	int const rm = mb & 63;
#pragma unroll
	for (int a = 0; a < 3; a++) {
#pragma unroll
		for (int c = 0; c < 3; c++) {
			int64_t const k = Blk<3>::at(b, r, a, c);
			cval64[k] = C[a][c];
			cval32[k] = static_cast<float>(C[a][c]);
			if (mb >= 0 && mb != b) {
				int64_t const km = Blk<3>::at(mb, rm, c, a);
				cval64[km] = C[a][c];
				cval32[km] = static_cast<float>(C[a][c]);
			}
		}
	}
}

// The Galerkin product in two steps (levels built on the device; the
// patterns come from amg_topo.hip). Step 1, A P: entry (i, J) of row i =
// sum over the row's blocks A_ij (FP32 copy) with J in P's row j of
// A_ij P_jJ, summed in FP64 and stored FP32. One thread per row takes the
// row's entries W at a time (their sums in registers): for each of
// the row's blocks it scans P's row j once (columns sorted) and adds the
// products of the columns that fall in the window, so a row's blocks and P
// rows are read once per window instead of once per entry.
template <int BS, int W>
__global__ __launch_bounds__(kBlock) void k_ap(int nb, const int32_t* __restrict__ slice_ptr,
                                               const int32_t* __restrict__ col, const int32_t* __restrict__ row_len,
                                               const float* __restrict__ val32, const int32_t* __restrict__ pptr,
                                               const int32_t* __restrict__ pcol, const float* __restrict__ P,
                                               const int32_t* __restrict__ apptr, const int32_t* __restrict__ apcol,
                                               float* __restrict__ apval) {
	auto row = [&](int const i) {
		int const r = i & 63;
		int32_t const sbeg = slice_ptr[i >> 6] + r;
		int const len = row_len[i];
		int32_t const q1 = apptr[i + 1];
		for (int32_t qa = apptr[i]; qa < q1; qa += W) {
			int const nw = min(W, q1 - qa);
			int32_t Jw[W];
	#pragma unroll
			for (int w = 0; w < W; w++) {
				Jw[w] = w < nw ? apcol[qa + w] : INT32_MAX;
			}
			int32_t const jlo = Jw[0], jhi = Jw[nw - 1];
			double acc[W][BS][3];
	#pragma unroll
			for (int w = 0; w < W; w++) {
	#pragma unroll
				for (int a = 0; a < BS; a++) {
	#pragma unroll
					for (int c = 0; c < 3; c++) {
						acc[w][a][c] = 0;
					}
				}
			}
			for (int k = 0; k < len; k++) {
				int32_t const blk = sbeg + k * kSlice;
				int const j = col[blk];
				int32_t const g1 = pptr[j + 1];
				bool have_a = false;
				float A[BS][BS];
				for (int32_t g = pptr[j]; g < g1; g++) {
					int32_t const cg = pcol[g];
					if (cg > jhi) {
						break;
					}
					if (cg < jlo) {
						continue;
					}
					if (!have_a) {
						ld_block<BS>(val32, blk, r, A);
						have_a = true;
					}
					const float* pj = P + BS * 3 * static_cast<int64_t>(g);
					float pv[BS * 3];
	#pragma unroll
					for (int e = 0; e < BS * 3; e++) {
						pv[e] = pj[e];
					}
					double tv[BS][3];
	#pragma unroll
					for (int a = 0; a < BS; a++) {
	#pragma unroll
						for (int c = 0; c < 3; c++) {
							double s = 0;
	#pragma unroll
							for (int m = 0; m < BS; m++) {
								s += static_cast<double>(A[a][m]) * static_cast<double>(pv[3 * m + c]);
							}
							tv[a][c] = s;
						}
					}
	#pragma unroll
					for (int w = 0; w < W; w++) {
						if (Jw[w] == cg) {
	#pragma unroll
							for (int a = 0; a < BS; a++) {
	#pragma unroll
								for (int c = 0; c < 3; c++) {
									acc[w][a][c] += tv[a][c];
								}
							}
						}
					}
				}
			}
	#pragma unroll
			for (int w = 0; w < W; w++) {
				if (w < nw) {
					float* out = apval + BS * 3 * static_cast<int64_t>(qa + w);
	#pragma unroll
					for (int a = 0; a < BS; a++) {
	#pragma unroll
						for (int c = 0; c < 3; c++) {
							out[3 * a + c] = static_cast<float>(acc[w][a][c]);
						}
					}
				}
			}
		}
	};
	// a grid-stride loop over the rows: the grid may be capped, leaving room for other work
	for (int64_t vb = blockIdx.x; vb * kBlock < nb; vb += gridDim.x) {
		int64_t const i = vb * kBlock + threadIdx.x;
		if (i < nb) {
			row(static_cast<int>(i));
		}
	}
}

// Step 1 with a lane per output entry: 16 lanes per fine row i (four rows
// per wavefront), lane l taking the row's entries l, l + 16, ...; the lanes
// of a row walk its blocks together (each block and P row read once per
// group, the same addresses across the lanes), each finding its own J in
// P's row j by a binary search and summing A_ij P_jJ in FP64, in block order.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_ap_lanes(int nb, const int32_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ row_len,
                                                     const float* __restrict__ val32, const int32_t* __restrict__ pptr,
                                                     const int32_t* __restrict__ pcol, const float* __restrict__ P,
                                                     const int32_t* __restrict__ apptr,
                                                     const int32_t* __restrict__ apcol, float* __restrict__ apval) {
	int64_t const i = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 4;
	int const l = threadIdx.x & 15;
	if (i >= nb) {
		return;
	}
	int const r = static_cast<int>(i & 63);
	int32_t const sbeg = slice_ptr[i >> 6] + r;
	int const len = row_len[i];
	int32_t const q1 = apptr[i + 1];
	for (int32_t qb = apptr[i]; qb < q1; qb += 16) {
		int32_t const q = qb + l;
		bool const act = q < q1;
		int32_t const J = act ? apcol[q] : -1;
		double acc[BS][3] = {};
		for (int k = 0; k < len; k++) {
			int32_t const blk = sbeg + k * kSlice;
			int const j = col[blk];
			int32_t lo = pptr[j], hi = pptr[j + 1] - 1;
			while (lo < hi) {
				int32_t const mid = (lo + hi) >> 1;
				if (pcol[mid] < J) {
					lo = mid + 1;
				} else {
					hi = mid;
				}
			}
			if (!act || lo > hi || pcol[lo] != J) {
				continue;
			}
			float A[BS][BS];
			ld_block<BS>(val32, blk, r, A);
			const float* pj = P + BS * 3 * static_cast<int64_t>(lo);
#pragma unroll
			for (int a = 0; a < BS; a++) {
#pragma unroll
				for (int c = 0; c < 3; c++) {
					double s = 0;
#pragma unroll
					for (int m = 0; m < BS; m++) {
						s += static_cast<double>(A[a][m]) * static_cast<double>(pj[3 * m + c]);
					}
					acc[a][c] += s;
				}
			}
		}
		if (act) {
			float* out = apval + BS * 3 * static_cast<int64_t>(q);
#pragma unroll
			for (int a = 0; a < BS; a++) {
#pragma unroll
				for (int c = 0; c < 3; c++) {
					out[3 * a + c] = static_cast<float>(acc[a][c]);
				}
			}
		}
	}
}

// Step 2 with a lane per coarse slot: 16 lanes per coarse row I, lane l
// taking the row's slots (I, J >= I) l, l + 16, ...; the lanes walk P's
// column I (R, rows in order) together, each finding its J in row i of A P
// by a binary search (within one or two cache lines of apcol) and summing
// P_iI^T (A P)_iJ in FP64 in row order: no cross-lane reduction. The block
// is stored with its transpose in the mirror slot; padding slots get zeros.
template <int BSF>
__global__ __launch_bounds__(kBlock) void k_rap_lanes(int64_t nrows_pad, int nc, const int32_t* __restrict__ c_slice_ptr,
                                                      const int32_t* __restrict__ c_col,
                                                      const int32_t* __restrict__ c_row_len,
                                                      const int32_t* __restrict__ c_diag_blk,
                                                      const int32_t* __restrict__ mirror,
                                                      const int32_t* __restrict__ rptr,
                                                      const int32_t* __restrict__ rrow, const float* __restrict__ R,
                                                      const int32_t* __restrict__ apptr,
                                                      const int32_t* __restrict__ apcol,
                                                      const float* __restrict__ apval, double* __restrict__ cval64,
                                                      float* __restrict__ cval32) {
	auto slot_lanes = [&](int64_t const I) {
		int const l = threadIdx.x & 15;
		int64_t const s = I >> 6;
		int const r = static_cast<int>(I & 63);
		int32_t const sp = c_slice_ptr[s];
		int const width = (c_slice_ptr[s + 1] - sp) / kSlice;
		int const len = I < nc ? c_row_len[I] : 0;
		for (int q = len + l; q < width; q += 16) {   // padding slots
			int64_t const slot = sp + static_cast<int64_t>(q) * kSlice + r;
	#pragma unroll
			for (int a = 0; a < 3; a++) {
	#pragma unroll
				for (int c = 0; c < 3; c++) {
					int64_t const k = Blk<3>::at(slot, r, a, c);
					cval64[k] = 0;
					cval32[k] = 0;
				}
			}
		}
		if (I >= nc) {
			return;
		}
		int const q0 = (c_diag_blk[I] - sp - r) / kSlice;
		int32_t const t0 = rptr[I], t1 = rptr[I + 1];
		for (int qb = q0; qb < len; qb += 16) {
			int const q = qb + l;
			bool const act = q < len;
			int64_t const slot = sp + static_cast<int64_t>(act ? q : q0) * kSlice + r;
			int32_t const J = act ? c_col[slot] : -1;
			double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
			for (int32_t t = t0; t < t1; t++) {
				int const i = rrow[t];
				int32_t lo = apptr[i], hi = apptr[i + 1] - 1;
				while (lo < hi) {
					int32_t const mid = (lo + hi) >> 1;
					if (apcol[mid] < J) {
						lo = mid + 1;
					} else {
						hi = mid;
					}
				}
				if (!act || lo > hi || apcol[lo] != J) {
					continue;
				}
				const float* rt = R + BSF * 3 * static_cast<int64_t>(t);
				const float* ap = apval + BSF * 3 * static_cast<int64_t>(lo);
	#pragma unroll
				for (int a = 0; a < 3; a++) {
	#pragma unroll
					for (int c = 0; c < 3; c++) {
						double v = 0;
	#pragma unroll
						for (int k = 0; k < BSF; k++) {
							v += static_cast<double>(rt[3 * k + a]) * static_cast<double>(ap[3 * k + c]);
						}
						C[a][c] += v;
					}
				}
			}
			if (act) {
				int32_t const mb = mirror[slot];
				int const rm = mb & 63;
	#pragma unroll
				for (int a = 0; a < 3; a++) {
	#pragma unroll
					for (int c = 0; c < 3; c++) {
						int64_t const k = Blk<3>::at(slot, r, a, c);
						cval64[k] = C[a][c];
						cval32[k] = static_cast<float>(C[a][c]);
						if (mb >= 0 && mb != slot) {
							int64_t const km = Blk<3>::at(mb, rm, c, a);
							cval64[km] = C[a][c];
							cval32[km] = static_cast<float>(C[a][c]);
						}
					}
				}
			}
		}
	};
	// a grid-stride loop over the rows: the grid may be capped, leaving room for other work
	for (int64_t vb = blockIdx.x; vb * kBlock < 16 * nrows_pad; vb += gridDim.x) {
		int64_t const I = (vb * kBlock + threadIdx.x) >> 4;
		if (I < nrows_pad) {
			slot_lanes(I);
		}
	}
}

// Step 2, P^T (A P): one wavefront per coarse row I (the rows past the
// level's end in its last slice included), in four groups of 16 lanes; a
// group takes the row's slots (I, J >= I) in turn, its lanes P's column I
// (R, entries in column order) with a stride of 16: C_IJ = sum over the
// column's rows i of P_iI^T (A P)_iJ (binary search of J in row i of A P),
// in FP64, combined by a fixed butterfly; the block is stored with its
// transpose in the mirror slot (FP64 and the FP32 copy). Padding slots are
// written with zeros.
template <int BSF>
__global__ __launch_bounds__(kBlock) void k_rap(int64_t nrows_pad, int nc, const int32_t* __restrict__ c_slice_ptr,
                                                const int32_t* __restrict__ c_col,
                                                const int32_t* __restrict__ c_row_len,
                                                const int32_t* __restrict__ c_diag_blk,
                                                const int32_t* __restrict__ mirror, const int32_t* __restrict__ rptr,
                                                const int32_t* __restrict__ rrow, const float* __restrict__ R,
                                                const int32_t* __restrict__ apptr, const int32_t* __restrict__ apcol,
                                                const float* __restrict__ apval, double* __restrict__ cval64,
                                                float* __restrict__ cval32) {
	int64_t const I = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 6;
	if (I >= nrows_pad) {
		return;
	}
	int const lane = threadIdx.x & 63;
	int const grp = lane >> 4, l = lane & 15;
	int64_t const s = I >> 6;
	int const r = static_cast<int>(I & 63);
	int32_t const sp = c_slice_ptr[s];
	int const width = (c_slice_ptr[s + 1] - sp) / kSlice;
	int const len = I < nc ? c_row_len[I] : 0;
	for (int q = len + lane; q < width; q += 64) {   // padding slots
		int64_t const slot = sp + static_cast<int64_t>(q) * kSlice + r;
#pragma unroll
		for (int a = 0; a < 3; a++) {
#pragma unroll
			for (int c = 0; c < 3; c++) {
				int64_t const k = Blk<3>::at(slot, r, a, c);
				cval64[k] = 0;
				cval32[k] = 0;
			}
		}
	}
	if (I >= nc) {
		return;
	}
	int const q0 = (c_diag_blk[I] - sp - r) / kSlice;
	int32_t const t0 = rptr[I], t1 = rptr[I + 1];
	for (int q = q0 + grp; q < len; q += 4) {
		int64_t const slot = sp + static_cast<int64_t>(q) * kSlice + r;
		int32_t const J = c_col[slot];
		double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
		for (int32_t t = t0 + l; t < t1; t += 16) {
			int const i = rrow[t];
			int32_t lo = apptr[i], hi = apptr[i + 1] - 1;
			while (lo < hi) {
				int32_t const mid = (lo + hi) >> 1;
				if (apcol[mid] < J) {
					lo = mid + 1;
				} else {
					hi = mid;
				}
			}
			if (lo > hi || apcol[lo] != J) {
				continue;
			}
			const float* rt = R + BSF * 3 * static_cast<int64_t>(t);
			const float* ap = apval + BSF * 3 * static_cast<int64_t>(lo);
#pragma unroll
			for (int a = 0; a < 3; a++) {
#pragma unroll
				for (int c = 0; c < 3; c++) {
					double v = 0;
#pragma unroll
					for (int k = 0; k < BSF; k++) {
						v += static_cast<double>(rt[3 * k + a]) * static_cast<double>(ap[3 * k + c]);
					}
					C[a][c] += v;
				}
			}
		}
#pragma unroll
		for (int a = 0; a < 3; a++) {
#pragma unroll
			for (int c = 0; c < 3; c++) {
				double v = C[a][c];
#pragma unroll
				for (int off = 8; off > 0; off >>= 1) {
					v += __shfl_xor(v, off, 64);
				}
				C[a][c] = v;
			}
		}
		if (l == 0) {
			int32_t const mb = mirror[slot];
			int const rm = mb & 63;
#pragma unroll
			for (int a = 0; a < 3; a++) {
#pragma unroll
				for (int c = 0; c < 3; c++) {
					int64_t const k = Blk<3>::at(slot, r, a, c);
					cval64[k] = C[a][c];
					cval32[k] = static_cast<float>(C[a][c]);
					if (mb >= 0 && mb != slot) {
						int64_t const km = Blk<3>::at(mb, rm, c, a);
						cval64[km] = C[a][c];
						cval32[km] = static_cast<float>(C[a][c]);
					}
				}
			}
		}
	}
}

// Inverse diagonal blocks (an exactly zero diagonal entry belongs to a
// decoupled DOF, whose row and column are zero: it is treated as 1) and the
// per-workgroup maxima of the Gershgorin bound max_k sum_j |(D^-1 A)_kj|.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_diag_gersh(int nb, const int32_t* __restrict__ slice_ptr,
                                                       const int32_t* __restrict__ diag_blk,
                                                       const double* __restrict__ val, float* __restrict__ dinv,
                                                       double* __restrict__ partial) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	double g = 0;
	if (i < nb) {
		int const r = i & 63;
		int32_t const dblk = diag_blk[i];
		double D[BS][BS], Di[BS][BS];
		ld_block<BS>(val, dblk, r, D);
#pragma unroll
		for (int a = 0; a < BS; a++) {
			if (D[a][a] == 0) {
				D[a][a] = 1;
			}
		}
		if constexpr (BS == 2) {
			double const inv = 1.0 / (D[0][0] * D[1][1] - D[0][1] * D[1][0]);
			Di[0][0] = D[1][1] * inv;
			Di[0][1] = -D[0][1] * inv;
			Di[1][0] = -D[1][0] * inv;
			Di[1][1] = D[0][0] * inv;
		} else {
			double const c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
			double const c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
			double const c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
			double const inv = 1.0 / (D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02);
			Di[0][0] = c00 * inv;
			Di[1][0] = c01 * inv;
			Di[2][0] = c02 * inv;
			Di[0][1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * inv;
			Di[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * inv;
			Di[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * inv;
			Di[0][2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * inv;
			Di[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * inv;
			Di[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * inv;
		}
#pragma unroll
		for (int a = 0; a < BS; a++) {
#pragma unroll
			for (int c = 0; c < BS; c++) {
				dinv[BS * BS * static_cast<int64_t>(i) + BS * a + c] = static_cast<float>(Di[a][c]);
			}
		}
		double rs[BS];
#pragma unroll
		for (int a = 0; a < BS; a++) {
			rs[a] = 0;
		}
#pragma unroll kRowUnroll<BS>
		for (int32_t blk = slice_ptr[i >> 6] + r; blk < slice_ptr[(i >> 6) + 1]; blk += kSlice) {
			double A[BS][BS];
			ld_block<BS>(val, blk, r, A);
			if (blk == dblk) {
#pragma unroll
				for (int a = 0; a < BS; a++) {
					A[a][a] = D[a][a];
				}
			}
#pragma unroll
			for (int a = 0; a < BS; a++) {
#pragma unroll
				for (int c = 0; c < BS; c++) {
					double t = 0;
#pragma unroll
					for (int k = 0; k < BS; k++) {
						t += Di[a][k] * A[k][c];
					}
					rs[a] += fabs(t);
				}
			}
		}
#pragma unroll
		for (int a = 0; a < BS; a++) {
			g = fmax(g, rs[a]);
		}
	}
	wg_partial<false>(g, partial);
}

// Power iteration for lambda_max(D^-1 A) on a level (FP32 matrix and
// inverse diagonal blocks): x_0 has entries 1 + (hash of the index) / 2^32
// (a fixed start with every component), then y = D^-1 A (s x) with s the
// previous step's 1 / ||y|| (scal[kScalScale]); the estimate after a step
// is ||D^-1 A x|| for the unit x (scal[kScalLambda]).
__global__ __launch_bounds__(kBlock) void k_power_init(int64_t m, float* __restrict__ x, double* __restrict__ partial) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
	double v = 0;
	if (i < m) {
		uint32_t h = static_cast<uint32_t>(i) * 2654435761u;
		h ^= h >> 15;
		float const xi = 1.0f + static_cast<float>(h) * 2.3283064e-10f;
		x[i] = xi;
		v = static_cast<double>(xi) * xi;
	}
	wg_partial<true>(v, partial);
}

template <int BS>
__global__ __launch_bounds__(kBlock) void k_power_step(int nb, const int32_t* __restrict__ slice_ptr,
                                                       const int32_t* __restrict__ col, const float* __restrict__ val,
                                                       const float* __restrict__ dinv, const float* __restrict__ x,
                                                       const double* __restrict__ scal, float* __restrict__ y,
                                                       double* __restrict__ partial) {
	int const i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	double v = 0;
	if (i < nb) {
		float const sc = static_cast<float>(scal[kScalScale]);
		int const r = i & 63;
		float ax[BS];
#pragma unroll
		for (int a = 0; a < BS; a++) {
			ax[a] = 0;
		}
#pragma unroll kRowUnroll<BS>
		for (int32_t blk = slice_ptr[i >> 6] + r; blk < slice_ptr[(i >> 6) + 1]; blk += kSlice) {
			int32_t const c = col[blk];
			float A[BS][BS];
			ld_block<BS>(val, blk, r, A);
#pragma unroll
			for (int a = 0; a < BS; a++) {
#pragma unroll
				for (int k = 0; k < BS; k++) {
					ax[a] += A[a][k] * x[BS * static_cast<int64_t>(c) + k];
				}
			}
		}
#pragma unroll
		for (int a = 0; a < BS; a++) {
			float z = 0;
#pragma unroll
			for (int k = 0; k < BS; k++) {
				z += dinv[BS * BS * static_cast<int64_t>(i) + BS * a + k] * ax[k];
			}
			z *= sc;
			y[BS * static_cast<int64_t>(i) + a] = z;
			v += static_cast<double>(z) * z;
		}
	}
	wg_partial<true>(v, partial);
}

// ||y||^2 = sum of the nwg partials: scal[kScalLambda] = ||y|| (the
// estimate, when y came from a unit x), scal[kScalScale] = 1 / ||y||.
__global__ __launch_bounds__(kBlock) void k_power_norm(int nwg, const double* __restrict__ partial,
                                                       double* __restrict__ scal) {
	__shared__ double lds[kBlock / 64];
	double m = 0;
	for (int i = threadIdx.x; i < nwg; i += kBlock) {
		m += partial[i];
	}
	m = wave_sum(m);
	if ((threadIdx.x & 63) == 0) {
		lds[threadIdx.x >> 6] = m;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double t = 0;
		for (int w = 0; w < kBlock / 64; w++) {
			t += lds[w];
		}
		double const nrm = sqrt(t);
		scal[kScalLambda] = nrm;
		scal[kScalScale] = nrm > 0 ? 1.0 / nrm : 0.0;
	}
}

// lambda_max = max of the nwg partials; Chebyshev coefficients for the
// interval [lambda_max / ratio, lambda_max]: scal[0] = lambda_max and step k
// updates d = scal[1 + 2k] d + scal[2 + 2k] D^-1 r, x += d (step 0 has no d
// term). This is the three-term Chebyshev recurrence for a polynomial
// smoother (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
// algorithm 12.1): sigma = theta / delta, rho_0 = 1 / sigma,
// rho_k = 1 / (2 sigma - rho_{k-1}), d_k = rho_k rho_{k-1} d_{k-1}
// + (2 rho_k / delta) D^-1 r_k, d_0 = D^-1 r_0 / theta.
__global__ __launch_bounds__(kBlock) void k_cheb_coef(int nwg, const double* __restrict__ partial, double ratio,
                                                      int degree, int use_power, double* __restrict__ scal) {
	__shared__ double lds[kBlock / 64];
	double m = 0;
	for (int i = threadIdx.x; i < nwg; i += kBlock) {
		m = fmax(m, partial[i]);
	}
	m = wave_max(m);
	if ((threadIdx.x & 63) == 0) {
		lds[threadIdx.x >> 6] = m;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double lmax = 0;
		for (int w = 0; w < kBlock / 64; w++) {
			lmax = fmax(lmax, lds[w]);
		}
		// the smoother's interval ends at the Gershgorin bound, a guaranteed
		// upper bound: below lambda_max the Chebyshev polynomial would amplify
		// the modes above its interval. The power estimate (a lower bound)
		// only sets the prolongator damping.
		if (!use_power) {
			scal[kScalLambda] = lmax;
		}
		double const lmin = lmax / ratio;
		double const theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
		double const sigma = theta / delta;
		double rho = 1.0 / sigma;
		scal[0] = lmax;
		scal[1] = 0;
		scal[2] = 1.0 / theta;
		for (int k = 1; k < degree; k++) {
			double const rn = 1.0 / (2.0 * sigma - rho);
			scal[1 + 2 * k] = rn * rho;
			scal[2 + 2 * k] = 2.0 * rn / delta;
			rho = rn;
		}
	}
}

// Dense copy of each segment of a level's operator (a segment of m block
// rows: BS m x BS m, row-major, at seg_off of its segment), one block row
// per thread; padding blocks carry zeros and add nothing.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_dense(int nb, const int32_t* __restrict__ slice_ptr,
                                                  const int32_t* __restrict__ col, const double* __restrict__ val,
                                                  const int32_t* __restrict__ seg_ptr,
                                                  const int32_t* __restrict__ seg_row,
                                                  const int64_t* __restrict__ seg_off, double* __restrict__ A) {
	int const i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= nb) {
		return;
	}
	int const sg = seg_row[i], base = seg_ptr[sg];
	int const n = BS * (seg_ptr[sg + 1] - base), r = i & 63, li = i - base;
	double* const As = A + seg_off[sg];
	for (int a = 0; a < BS; a++) {
		for (int j = 0; j < n; j++) {
			As[static_cast<int64_t>(BS * li + a) * n + j] = 0;
		}
	}
	for (int32_t blk = slice_ptr[i >> 6] + r; blk < slice_ptr[(i >> 6) + 1]; blk += kSlice) {
		int32_t const c = col[blk] - base;
		double B[BS][BS];
		ld_block<BS>(val, blk, r, B);
#pragma unroll
		for (int a = 0; a < BS; a++) {
#pragma unroll
			for (int b = 0; b < BS; b++) {
				As[static_cast<int64_t>(BS * li + a) * n + BS * c + b] += B[a][b];
			}
		}
	}
}

// In-place Gauss-Jordan inverse of the SPD n x n matrix of one segment
// (n = bs x its rows) by one workgroup per segment, without pivoting (the
// pivots of an SPD matrix are positive). A zero diagonal entry marks a
// decoupled DOF (zero row and column) and is replaced by 1.
// Step k: a_kk <- 1/p, a_kj <- a_kj/p, a_ik <- -a_ik/p,
// a_ij <- a_ij - a_ik a_kj / p (p = a_kk), with row and column k staged in LDS
// so every update reads the values of the step's start.
template <int NT>
__global__ __launch_bounds__(NT) void k_gauss_jordan(int bs, const int32_t* __restrict__ seg_ptr,
                                                     const int64_t* __restrict__ seg_off, double* __restrict__ A) {
	__shared__ double rowk[kDenseMax], colk[kDenseMax];
	int const n = bs * (seg_ptr[blockIdx.x + 1] - seg_ptr[blockIdx.x]);
	A += seg_off[blockIdx.x];
	for (int i = threadIdx.x; i < n; i += NT) {
		if (A[static_cast<int64_t>(i) * n + i] == 0) {
			A[static_cast<int64_t>(i) * n + i] = 1;
		}
	}
	__syncthreads();
	int64_t const nn = static_cast<int64_t>(n) * n;
	for (int k = 0; k < n; k++) {
		for (int j = threadIdx.x; j < n; j += NT) {
			rowk[j] = A[static_cast<int64_t>(k) * n + j];
			colk[j] = A[static_cast<int64_t>(j) * n + k];
		}
		__syncthreads();
		double const p = 1.0 / rowk[k];
		for (int64_t e = threadIdx.x; e < nn; e += NT) {
			int const i = static_cast<int>(e / n), j = static_cast<int>(e - static_cast<int64_t>(i) * n);
			double v;
			if (i == k) {
				v = j == k ? p : rowk[j] * p;
			} else if (j == k) {
				v = -colk[i] * p;
			} else {
				v = A[e] - colk[i] * rowk[j] * p;
			}
			A[e] = v;
		}
		__syncthreads();
	}
}

// ------------------------------------------------------------------ V-cycle

// One Chebyshev step on a level: r = b - A x_in (skipped when FIRST: x_in = 0),
// d = c1 d + c2 D^-1 r (c1 unused at step 0), x_out = x_in + d. With DOT, also
// the per-workgroup sums of x_out . q (q an FP64 vector over the same rows).
// The first loads of row i of a level: its slice's block range and b_i. A
// grid kernel issues them before it reads the done flag (done_after).
template <int BS>
struct RowHead {
	int32_t beg, end;
	float res[BS];
};

template <int BS, typename BT>
__device__ __forceinline__ RowHead<BS> row_head(int i, const int32_t* __restrict__ slice_ptr,
                                                const BT* __restrict__ b) {
	RowHead<BS> h;
	h.beg = slice_ptr[i >> 6];
	h.end = slice_ptr[(i >> 6) + 1];
#pragma unroll
	for (int a = 0; a < BS; a++) {
		h.res[a] = static_cast<float>(b[BS * static_cast<int64_t>(i) + a]);
	}
	return h;
}

// Row i of one Chebyshev step (k_cheb) from its head; xo receives the row's
// x_out.
template <int BS, bool FIRST>
__device__ __forceinline__ void cheb_row(int i, RowHead<BS> h, const int32_t* __restrict__ col,
                                         const float* __restrict__ val, const float* __restrict__ dinv,
                                         const float* __restrict__ xin, float* __restrict__ xout,
                                         float* __restrict__ d, float c1, float c2, int step, float (&xo)[BS]) {
	float (&res)[BS] = h.res;
	if (!FIRST) {
		int const r = i & 63;
#pragma unroll kRowUnroll<BS>
		for (int32_t blk = h.beg + r; blk < h.end; blk += kSlice) {
			int32_t const c = col[blk];
			float A[BS][BS];
			ld_block<BS>(val, blk, r, A);
			float xv[BS];
#pragma unroll
			for (int k = 0; k < BS; k++) {
				xv[k] = xin[BS * static_cast<int64_t>(c) + k];
			}
#pragma unroll
			for (int a = 0; a < BS; a++) {
#pragma unroll
				for (int k = 0; k < BS; k++) {
					res[a] -= A[a][k] * xv[k];
				}
			}
		}
	}
#pragma unroll
	for (int a = 0; a < BS; a++) {
		float z = 0;
#pragma unroll
		for (int k = 0; k < BS; k++) {
			z += dinv[BS * BS * static_cast<int64_t>(i) + BS * a + k] * res[k];
		}
		int64_t const o = BS * static_cast<int64_t>(i) + a;
		float const dn = (FIRST || step == 0) ? c2 * z : c1 * d[o] + c2 * z;
		float const xn = FIRST ? dn : xin[o] + dn;
		d[o] = dn;
		xout[o] = xn;
		xo[a] = xn;
	}
}

template <int BS, typename BT, bool FIRST, bool DOT>
__global__ __launch_bounds__(kBlock) void k_cheb(int nb, const int32_t* __restrict__ slice_ptr,
                                                 const int32_t* __restrict__ col, const float* __restrict__ val,
                                                 const BT* __restrict__ b, const float* __restrict__ dinv,
                                                 const float* __restrict__ xin, float* __restrict__ xout,
                                                 float* __restrict__ d, const double* __restrict__ scal, int step,
                                                 const double* __restrict__ q, double* __restrict__ partial,
                                                 SegView sv, const Scalars* __restrict__ S) {
	int const blk = xcd_remap(blockIdx.x, gridDim.x);
	int const i = blk * kBlock + threadIdx.x;
	RowHead<BS> const h = row_head<BS>(min(i, nb - 1), slice_ptr, b);
	if (FIRST ? done_after(S, h.res[0]) : done_after(S, h.beg, h.end, h.res[0])) {
		return;
	}
	float const c1 = static_cast<float>(scal[1 + 2 * step]);
	float const c2 = static_cast<float>(scal[2 + 2 * step]);
	double dot = 0;
	if (i < nb) {
		float xo[BS];
		cheb_row<BS, FIRST>(i, h, col, val, dinv, xin, xout, d, c1, c2, step, xo);
		if (DOT) {
#pragma unroll
			for (int a = 0; a < BS; a++) {
				dot += static_cast<double>(xo[a]) * q[BS * static_cast<int64_t>(i) + a];
			}
		}
	}
	if (DOT) {
		if (sv.seg) {
			// the two segments of the block apart (segmented flexible CG)
			bool const second = static_cast<int>(threadIdx.x) >= sv.split[blk];
			double lo[1] = {second ? 0.0 : dot}, hi[1] = {second ? dot : 0.0};
			seg_block_partials<1>(lo, hi, blk, gridDim.x, partial);
		} else {
			wg_partial<true>(dot, partial);
		}
	}
}

// Row i of r = b - A x from its head.
template <int BS>
__device__ __forceinline__ void resid_row(int i, RowHead<BS> h, const int32_t* __restrict__ col,
                                          const float* __restrict__ val, const float* __restrict__ x,
                                          float* __restrict__ rv) {
	int const r = i & 63;
	float (&res)[BS] = h.res;
#pragma unroll kRowUnroll<BS>
	for (int32_t blk = h.beg + r; blk < h.end; blk += kSlice) {
		int32_t const c = col[blk];
		float A[BS][BS];
		ld_block<BS>(val, blk, r, A);
#pragma unroll
		for (int a = 0; a < BS; a++) {
#pragma unroll
			for (int k = 0; k < BS; k++) {
				res[a] -= A[a][k] * x[BS * static_cast<int64_t>(c) + k];
			}
		}
	}
#pragma unroll
	for (int a = 0; a < BS; a++) {
		rv[BS * static_cast<int64_t>(i) + a] = res[a];
	}
}

// r = b - A x on a level.
template <int BS, typename BT>
__global__ __launch_bounds__(kBlock) void k_level_resid(int nb, const int32_t* __restrict__ slice_ptr,
                                                        const int32_t* __restrict__ col,
                                                        const float* __restrict__ val, const BT* __restrict__ b,
                                                        const float* __restrict__ x, float* __restrict__ rv,
                                                        const Scalars* __restrict__ S) {
	int const i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	RowHead<BS> const h = row_head<BS>(min(i, nb - 1), slice_ptr, b);
	if (done_after(S, h.beg, h.end, h.res[0])) {
		return;
	}
	if (i < nb) {
		resid_row<BS>(i, h, col, val, x, rv);
	}
}

// The first Chebyshev step and the residual after it in one pass, for a
// level smoothed by one step per side: x0 = c2 D^-1 b is row-local, so each
// neighbour's x0 is formed here from its b and inverse diagonal block instead
// of being written by one launch and read back by the next; each row writes
// its own x0 (d is not written: with one step per side no later step reads
// it), and r = b - A x0. The same FP32 operations as the first step of k_cheb.
template <int BS, typename BT>
__global__ __launch_bounds__(kBlock) void k_first_resid(int nb, const int32_t* __restrict__ slice_ptr,
                                                        const int32_t* __restrict__ col,
                                                        const float* __restrict__ val, const BT* __restrict__ b,
                                                        const float* __restrict__ dinv,
                                                        const double* __restrict__ scal, float* __restrict__ x0,
                                                        float* __restrict__ rv, const Scalars* __restrict__ S) {
	int const i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	RowHead<BS> h = row_head<BS>(min(i, nb - 1), slice_ptr, b);
	if (done_after(S, h.beg, h.end, h.res[0])) {
		return;
	}
	if (i >= nb) {
		return;
	}
	float const c2 = static_cast<float>(scal[2]);
	auto first_x = [&](int64_t j, float (&xv)[BS]) {
		float bj[BS];
#pragma unroll
		for (int k = 0; k < BS; k++) {
			bj[k] = static_cast<float>(b[BS * j + k]);
		}
#pragma unroll
		for (int a = 0; a < BS; a++) {
			float z = 0;
#pragma unroll
			for (int k = 0; k < BS; k++) {
				z += dinv[BS * BS * j + BS * a + k] * bj[k];
			}
			xv[a] = c2 * z;
		}
	};
	float xi[BS];
	first_x(i, xi);
#pragma unroll
	for (int a = 0; a < BS; a++) {
		x0[BS * static_cast<int64_t>(i) + a] = xi[a];
	}
	int const r = i & 63;
	float (&res)[BS] = h.res;
#pragma unroll kRowUnroll<BS>
	for (int32_t blk = h.beg + r; blk < h.end; blk += kSlice) {
		int32_t const c = col[blk];
		float A[BS][BS];
		ld_block<BS>(val, blk, r, A);
		float xv[BS];
		first_x(c, xv);
#pragma unroll
		for (int a = 0; a < BS; a++) {
#pragma unroll
			for (int k = 0; k < BS; k++) {
				res[a] -= A[a][k] * xv[k];
			}
		}
	}
#pragma unroll
	for (int a = 0; a < BS; a++) {
		rv[BS * static_cast<int64_t>(i) + a] = res[a];
	}
}

// b_c[J] = sum over the entries of column J of P (row i, entry e) of
// P_e^T r_i, the entries read from R (P's entries column by column, so a
// coarse row's are contiguous). kRestrictLanes lanes share a coarse row: lane
// l takes entries l, l + kRestrictLanes, ..., combined by a fixed butterfly.
// 16 measured best on the 2M beam's solve (10.62 ms; 8 lanes 10.66, 4 lanes
// 10.96).
#ifndef BFMX_RESTRICT_LANES
#define BFMX_RESTRICT_LANES 16
#endif
constexpr int kRestrictLanes = BFMX_RESTRICT_LANES;
static_assert(kRestrictLanes >= 1 && kRestrictLanes <= 64 && (kRestrictLanes & (kRestrictLanes - 1)) == 0,
              "a coarse row's lanes: a power of two within one wavefront (the butterfly)");

// Coarse row J (live) by the lane group of `lane`; every lane of the group
// must call it (the butterfly).
template <int BSF>
// [t0, t1) is the row's range of rptr (empty when not live).
__device__ __forceinline__ void restrict_row(int64_t J, bool live, int lane, int32_t t0, int32_t t1,
                                             const int32_t* __restrict__ rrow, const float* __restrict__ R,
                                             const float* __restrict__ r, float* __restrict__ bc) {
	float acc[3] = {0, 0, 0};
	for (int32_t t = t0 + lane; t < t1; t += kRestrictLanes) {
		int64_t const i = rrow[t];
		const float* p = R + BSF * 3 * static_cast<int64_t>(t);
#pragma unroll
		for (int k = 0; k < BSF; k++) {
			float const rk = r[BSF * i + k];
#pragma unroll
			for (int c = 0; c < 3; c++) {
				acc[c] += p[3 * k + c] * rk;
			}
		}
	}
#pragma unroll
	for (int c = 0; c < 3; c++) {
#pragma unroll
		for (int off = kRestrictLanes / 2; off > 0; off >>= 1) {
			acc[c] += __shfl_xor(acc[c], off, 64);
		}
	}
	if (live && lane == 0) {
#pragma unroll
		for (int c = 0; c < 3; c++) {
			bc[3 * J + c] = acc[c];
		}
	}
}

// With dinv: also the next level's first Chebyshev step on the restricted
// row (it reads only that row: x0 = d = c2 D^-1 b_c, c2 from cscal), which
// saves that step's launch.
template <int BSF>
__global__ __launch_bounds__(kBlock) void k_restrict(int nc, const int32_t* __restrict__ rptr,
                                                     const int32_t* __restrict__ rrow, const float* __restrict__ R,
                                                     const float* __restrict__ r, float* __restrict__ bc,
                                                     const float* __restrict__ dinv, const double* __restrict__ cscal,
                                                     float* __restrict__ x0, float* __restrict__ d,
                                                     const Scalars* __restrict__ S) {
	int64_t const J = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kRestrictLanes;
	int const lane = threadIdx.x % kRestrictLanes;
	bool const live = J < nc;
	int64_t const Jc = live ? J : nc - 1;
	int32_t const t0 = rptr[Jc], t1 = rptr[Jc + 1];
	if (done_after(S, t0, t1)) {
		return;
	}
	restrict_row<BSF>(J, live, lane, live ? t0 : 0, live ? t1 : 0, rrow, R, r, bc);
	if (dinv && J < nc && lane == 0) {
		float const c2 = static_cast<float>(cscal[2]);
		float const bj[3] = {bc[3 * J], bc[3 * J + 1], bc[3 * J + 2]};
#pragma unroll
		for (int a = 0; a < 3; a++) {
			float z = 0;
#pragma unroll
			for (int k = 0; k < 3; k++) {
				z += dinv[9 * J + 3 * a + k] * bj[k];
			}
			float const dn = c2 * z;
			d[3 * J + a] = dn;
			x0[3 * J + a] = dn;
		}
	}
}

// The first loads of row i of x += P x_c: x_i and the row's entry range.
template <int BSF>
struct ProlongHead {
	int32_t e0, e1;
	float acc[BSF];
};

template <int BSF>
__device__ __forceinline__ ProlongHead<BSF> prolong_head(int i, const int32_t* __restrict__ pptr,
                                                         const float* __restrict__ x) {
	ProlongHead<BSF> h;
	h.e0 = pptr[i];
	h.e1 = pptr[i + 1];
#pragma unroll
	for (int k = 0; k < BSF; k++) {
		h.acc[k] = x[BSF * static_cast<int64_t>(i) + k];
	}
	return h;
}

// Row i of x += P x_c from its head. The entries go in groups of
// kProlongGroup whose column indices, values and coarse values are all loaded
// before any is used (clamped to the row's last entry, the extra ones not
// added), so a row's few entries cost one round of dependent loads instead
// of one per entry; the sum keeps the entries' order. On the 2M-DOF beam's
// finest level (2.3 entries per row): 24.6 us against 27.4 us one by one
// (groups of 2: no gain).
constexpr int kProlongGroup = 4;

template <int BSF>
__device__ __forceinline__ void prolong_row(int i, ProlongHead<BSF> h, const int32_t* __restrict__ pcol,
                                            const float* __restrict__ P, const float* __restrict__ xc,
                                            float* __restrict__ x) {
	float (&acc)[BSF] = h.acc;
	for (int32_t e = h.e0; e < h.e1; e += kProlongGroup) {
		int64_t J[kProlongGroup];
		float pv[kProlongGroup][BSF * 3], xv[kProlongGroup][3];
#pragma unroll
		for (int u = 0; u < kProlongGroup; u++) {
			int32_t const eu = min(e + u, h.e1 - 1);
			J[u] = pcol[eu];
#pragma unroll
			for (int c = 0; c < BSF * 3; c++) {
				pv[u][c] = P[BSF * 3 * static_cast<int64_t>(eu) + c];
			}
		}
#pragma unroll
		for (int u = 0; u < kProlongGroup; u++) {
#pragma unroll
			for (int c = 0; c < 3; c++) {
				xv[u][c] = xc[3 * J[u] + c];
			}
		}
#pragma unroll
		for (int u = 0; u < kProlongGroup; u++) {
			if (e + u < h.e1) {
#pragma unroll
				for (int k = 0; k < BSF; k++) {
					acc[k] += pv[u][3 * k] * xv[u][0] + pv[u][3 * k + 1] * xv[u][1] + pv[u][3 * k + 2] * xv[u][2];
				}
			}
		}
	}
#pragma unroll
	for (int k = 0; k < BSF; k++) {
		x[BSF * static_cast<int64_t>(i) + k] = acc[k];
	}
}

// x_i += sum over the entries e of row i of P_e x_c[pcol[e]].
template <int BSF>
__global__ __launch_bounds__(kBlock) void k_prolong(int nb, const int32_t* __restrict__ pptr,
                                                    const int32_t* __restrict__ pcol, const float* __restrict__ P,
                                                    const float* __restrict__ xc, float* __restrict__ x,
                                                    const Scalars* __restrict__ S) {
	int const i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
	ProlongHead<BSF> const h = prolong_head<BSF>(min(i, nb - 1), pptr, x);
	if (done_after(S, h.e0, h.e1, h.acc[0])) {
		return;
	}
	if (i < nb) {
		prolong_row<BSF>(i, h, pcol, P, xc, x);
	}
}

// The prolongation and the first post-smoothing step of a level in one
// pass. r (the level's d_r) is the residual b - A x that the residual kernel
// left before the restriction, x_c the coarse correction. With x' = x + P x_c
// the step's residual is b - A x' = r - (A P) x_c, row-local given row i of P
// and of A P (the Galerkin set-up's A P, kept per level): d = c2 D^-1 (r -
// (A P) x_c), x_out = x' + d, as the first step of k_cheb takes it from x'.
// It saves the prolongation's launch and its write and re-read of x, and
// reads a row of A P (its coarse blocks) in place of the row of A.
// kProlongLanes lanes share a row: lane l takes entries l, l + kProlongLanes,
// ... of its rows of P and A P (consecutive lanes read consecutive entries),
// combined by a fixed butterfly.
constexpr int kProlongLanes = 8;

template <int BSF>
__global__ __launch_bounds__(kBlock) void k_prolong_cheb(
    int nb, const int32_t* __restrict__ pptr, const int32_t* __restrict__ pcol, const float* __restrict__ P,
    const int32_t* __restrict__ apptr, const int32_t* __restrict__ apcol, const float* __restrict__ apval,
    const float* __restrict__ xc, const float* __restrict__ r, const float* __restrict__ dinv,
    const double* __restrict__ scal, const float* __restrict__ xin, float* __restrict__ xout, float* __restrict__ d,
    const Scalars* __restrict__ S) {
	int64_t const t = static_cast<int64_t>(xcd_remap(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
	int64_t const i = t / kProlongLanes;
	int const lane = static_cast<int>(t % kProlongLanes);
	bool const live = i < nb;
	int64_t const ic = live ? i : nb - 1;
	int32_t const e0 = pptr[ic], e1 = pptr[ic + 1], f0 = apptr[ic], f1 = apptr[ic + 1];
	if (done_after(S, e0, e1, f0, f1)) {
		return;
	}
	float px[BSF] = {}, apx[BSF] = {};
	for (int32_t e = e0 + lane; e < e1; e += kProlongLanes) {
		int64_t const J = pcol[e];
		float const c0 = xc[3 * J], c1 = xc[3 * J + 1], c2 = xc[3 * J + 2];
		const float* p = P + BSF * 3 * static_cast<int64_t>(e);
#pragma unroll
		for (int k = 0; k < BSF; k++) {
			px[k] += p[3 * k] * c0 + p[3 * k + 1] * c1 + p[3 * k + 2] * c2;
		}
	}
	for (int32_t f = f0 + lane; f < f1; f += kProlongLanes) {
		int64_t const J = apcol[f];
		float const c0 = xc[3 * J], c1 = xc[3 * J + 1], c2 = xc[3 * J + 2];
		const float* q = apval + BSF * 3 * static_cast<int64_t>(f);
#pragma unroll
		for (int a = 0; a < BSF; a++) {
			apx[a] += q[3 * a] * c0 + q[3 * a + 1] * c1 + q[3 * a + 2] * c2;
		}
	}
#pragma unroll
	for (int a = 0; a < BSF; a++) {
#pragma unroll
		for (int off = kProlongLanes / 2; off > 0; off >>= 1) {
			px[a] += __shfl_xor(px[a], off, 64);
			apx[a] += __shfl_xor(apx[a], off, 64);
		}
	}
	if (!live || lane != 0) {
		return;
	}
	float res[BSF];
#pragma unroll
	for (int a = 0; a < BSF; a++) {
		res[a] = r[BSF * i + a] - apx[a];
	}
	float const c2 = static_cast<float>(scal[2]);
#pragma unroll
	for (int a = 0; a < BSF; a++) {
		float z = 0;
#pragma unroll
		for (int k = 0; k < BSF; k++) {
			z += dinv[BSF * BSF * i + BSF * a + k] * res[k];
		}
		int64_t const o = BSF * i + a;
		float const dn = c2 * z;
		d[o] = dn;
		xout[o] = xin[o] + px[a] + dn;
	}
}

// Row `row` (a DOF) of x = A^-1 b with the dense inverse of its segment, by
// one wavefront (lane `lane`); every lane returns the sum. The inverse is
// symmetric, so the row is read as a column: coalesced across the lanes.
template <int BS, typename BT>
__device__ __forceinline__ double dense_row(int row, int lane, const int32_t* __restrict__ seg_ptr,
                                            const int32_t* __restrict__ seg_row, const int64_t* __restrict__ seg_off,
                                            const double* __restrict__ Ainv, const BT* __restrict__ b) {
	int const sg = seg_row[row / BS];
	int const base = BS * seg_ptr[sg], n = BS * seg_ptr[sg + 1] - base;
	const double* const Ar = Ainv + seg_off[sg] + static_cast<int64_t>(row - base) * n;
	double acc = 0;
	for (int j = lane; j < n; j += 64) {
		acc += Ar[j] * static_cast<double>(b[base + j]);
	}
	return wave_sum(acc);
}

// x = A^-1 b with the dense inverses of the coarsest's segments, one
// wavefront per row.
constexpr int kDenseRowsPerBlock = kBlock / 64;

template <int BS, typename BT>
__global__ __launch_bounds__(kBlock) void k_dense_solve(int n, const int32_t* __restrict__ seg_ptr,
                                                        const int32_t* __restrict__ seg_row,
                                                        const int64_t* __restrict__ seg_off,
                                                        const double* __restrict__ Ainv, const BT* __restrict__ b,
                                                        float* __restrict__ x, const Scalars* __restrict__ S) {
	if (S->done) {
		return;
	}
	int const row = blockIdx.x * kDenseRowsPerBlock + (threadIdx.x >> 6);
	int const lane = threadIdx.x & 63;
	if (row >= n) {
		return;   // whole wavefronts leave together
	}
	double const acc = dense_row<BS, BT>(row, lane, seg_ptr, seg_row, seg_off, Ainv, b);
	if (lane == 0) {
		x[row] = static_cast<float>(acc);
	}
}

// Per-workgroup partial sums of x . q over n entries (x FP32, q FP64).
__global__ __launch_bounds__(kBlock) void k_dot_partials(int n, const float* __restrict__ x,
                                                         const double* __restrict__ q, double* __restrict__ partial,
                                                         const Scalars* __restrict__ S) {
	if (S->done) {
		return;
	}
	int const i = blockIdx.x * kBlock + threadIdx.x;
	double const v = i < n ? static_cast<double>(x[i]) * q[i] : 0.0;
	wg_partial<true>(v, partial);
}

// ------------------------------------------------- one-workgroup coarse cycle
// The small coarse levels (at most kWgRows block rows each) run their whole
// V-cycle in one workgroup: the same row bodies as the grid kernels,
// separated by workgroup barriers instead of kernel boundaries (a level of
// 1,400 rows takes a few microseconds per pass from L2, where each grid
// launch would cost its own launch latency). Descriptor 0 is the first such
// level; its correction ends in x1 (in x0 when it is a dense coarsest).
constexpr int kWg = 1024;
constexpr int kWgRows = 256;
constexpr int kPfLine = 128;                 // bytes per prefetched line
constexpr int kPfPerThread = 4;
constexpr size_t kPfMax = kWg * kPfPerThread;   // lines prefetched at most
static_assert(kPfPerThread == 4, "k_coarse_cycle passes its four line addresses to done_after");

__device__ __forceinline__ void wg_cheb(const AmgLevelDesc& L, bool first, int step, const float* xin, float* xout) {
	float const c1 = static_cast<float>(L.scal[1 + 2 * step]);
	float const c2 = static_cast<float>(L.scal[2 + 2 * step]);
	for (int i = threadIdx.x; i < L.nb; i += kWg) {
		float xo[3];
		if (first) {
			cheb_row<3, true>(i, row_head<3>(i, L.slice_ptr, L.b), L.col, L.val32, L.dinv, xin, xout, L.d, c1, c2, step, xo);
		} else {
			cheb_row<3, false>(i, row_head<3>(i, L.slice_ptr, L.b), L.col, L.val32, L.dinv, xin, xout, L.d, c1, c2, step, xo);
		}
	}
	__syncthreads();
}

// pf: addresses of the npf (at most kPfMax) lines of the arrays the cycle
// reads. They are all loaded at once at its start, so that its passes find
// them in this XCD's L2: between two cycles the finer levels' streams evict
// them, and each pass would otherwise wait on HBM for every step of its chain
// of dependent loads.
__global__ __launch_bounds__(kWg) void k_coarse_cycle(const AmgLevelDesc* __restrict__ lv, int nl,
                                                      const uint64_t* __restrict__ pf, int npf,
                                                      const Scalars* __restrict__ S) {
	uint64_t line[kPfPerThread];
#pragma unroll
	for (int j = 0; j < kPfPerThread; j++) {
		int const k = static_cast<int>(threadIdx.x) + j * kWg;
		line[j] = k < npf ? pf[k] : 0;
	}
	if (done_after(S, line[0], line[1], line[2], line[3])) {
		return;
	}
	uint32_t touched[kPfPerThread];
#pragma unroll
	for (int j = 0; j < kPfPerThread; j++) {
		// a global (not flat) load: the scalar loads' waits do not cover it
		touched[j] = line[j] ? *reinterpret_cast<const __attribute__((address_space(1))) uint32_t*>(line[j]) : 0;
	}
	int cur[kMaxLevels];
	for (int l = 0; l + 1 < nl; l++) {
		AmgLevelDesc const L = lv[l];
		wg_cheb(L, true, 0, nullptr, L.x0);
		int c = 0;
		for (int k = 1; k < L.deg; k++) {
			wg_cheb(L, false, k, c ? L.x1 : L.x0, c ? L.x0 : L.x1);
			c ^= 1;
		}
		cur[l] = c;
		float* const x = c ? L.x1 : L.x0;
		for (int i = threadIdx.x; i < L.nb; i += kWg) {
			resid_row<3>(i, row_head<3>(i, L.slice_ptr, L.b), L.col, L.val32, x, L.r);
		}
		__syncthreads();
		AmgLevelDesc const C = lv[l + 1];
		for (int64_t base = 0; base < C.nb; base += kWg / kRestrictLanes) {
			int64_t const J = base + threadIdx.x / kRestrictLanes;
			bool const live = J < C.nb;
			restrict_row<3>(J, live, threadIdx.x % kRestrictLanes, live ? L.rptr[J] : 0, live ? L.rptr[J + 1] : 0,
			                L.rrow, L.R, L.r, C.b);
		}
		__syncthreads();
	}
	AmgLevelDesc const Lc = lv[nl - 1];
	if (Lc.dense) {
		int const n = 3 * Lc.nb;
		int const lane = threadIdx.x & 63;
		for (int row = threadIdx.x >> 6; row < n; row += kWg / 64) {
			double const acc = dense_row<3, float>(row, lane, Lc.seg_ptr, Lc.seg_row, Lc.seg_off, Lc.inv, Lc.b);
			if (lane == 0) {
				Lc.x0[row] = static_cast<float>(acc);
			}
		}
		__syncthreads();
		cur[nl - 1] = 0;
	} else {
		wg_cheb(Lc, true, 0, nullptr, Lc.x0);
		int c = 0;
		for (int k = 1; k < Lc.deg; k++) {
			wg_cheb(Lc, false, k, c ? Lc.x1 : Lc.x0, c ? Lc.x0 : Lc.x1);
			c ^= 1;
		}
		cur[nl - 1] = c;
	}
	for (int l = nl - 2; l >= 0; l--) {
		AmgLevelDesc const L = lv[l];
		AmgLevelDesc const C = lv[l + 1];
		const float* xc = cur[l + 1] ? C.x1 : C.x0;
		int c = cur[l];
		float* const x = c ? L.x1 : L.x0;
		for (int i = threadIdx.x; i < L.nb; i += kWg) {
			prolong_row<3>(i, prolong_head<3>(i, L.pptr, x), L.pcol, L.P, xc, x);
		}
		__syncthreads();
		for (int k = 0; k < L.deg; k++) {
			wg_cheb(L, false, k, c ? L.x1 : L.x0, c ? L.x0 : L.x1);
			c ^= 1;
		}
		cur[l] = c;
	}
	// the prefetched values themselves are not needed: this keeps their loads
#pragma unroll
	for (int j = 0; j < kPfPerThread; j++) {
		asm volatile("" ::"v"(touched[j]));
	}
}

// ------------------------------------------------------------------ host topology

// A level's graph on the host (rows include themselves, sorted): as CSR
// (ptr, adj), or as the device's sliced ELL copied down (row i's k-th
// neighbour at col[slice[i >> 6] + 64 k + (i & 63)], k < len[i]), which the
// sweep reads as it is.
struct CsrRows {
	const int64_t* ptr;
	const int32_t* adj;
	static constexpr int kStride = 1;   // between a row's entries in row()
	int len(int i) const { return static_cast<int>(ptr[i + 1] - ptr[i]); }
	const int32_t* row(int i) const { return adj + ptr[i]; }
	int32_t at(int i, int k) const { return adj[ptr[i] + k]; }
};
struct SellRows {
	const int32_t* slice;
	const int32_t* col;
	const int32_t* rlen;
	static constexpr int kStride = kSlice;
	int len(int i) const { return rlen[i]; }
	const int32_t* row(int i) const { return col + slice[i >> 6] + (i & 63); }
	int32_t at(int i, int k) const { return col[slice[i >> 6] + k * kSlice + (i & 63)]; }
};

// Aggregates of rows r0 .. r1 - 1 of the graph G, whose entries stay in
// that range, numbered from 0 within the range; returns their count. Pass 1
// makes a row whose neighbourhood is still free the root of a new aggregate
// of that neighbourhood; pass 2 gives every other row the aggregate of its
// first neighbour aggregated by pass 1 (every such row has one). Pass 2
// writes aggregate a of a row as -2 - a, which its later rows still read as
// not aggregated (the same joins as writing them all after the pass):
// agg[i] < 0 holds -2 - agg_i on return, decoded by the caller. Pass 1
// notes the rows it passes over in a bit mask (one 64-bit word per 64 rows,
// written once per word), so pass 2 visits only those rows. A set *stop
// ends the sweep early (returns -1; checked every 4096 rows).
template <typename Rows>
int aggregate(int r0, int r1, Rows const& G, int32_t* agg, const std::atomic<bool>* stop = nullptr) {
	constexpr int S = Rows::kStride;
	std::fill(agg + r0, agg + r1, -1);
	int na = 0;
	int const nw = (r1 - r0 + 63) / 64;
	std::vector<uint64_t> passed(nw);
	for (int w = 0; w < nw; w++) {
		if (stop && (w & 63) == 0 && stop->load(std::memory_order_relaxed)) {
			return -1;
		}
		uint64_t bits = 0;
		int const i1 = std::min(r1, r0 + 64 * w + 64);
		for (int i = r0 + 64 * w; i < i1; i++) {
			if (agg[i] >= 0) {
				continue;
			}
			const int32_t* const p = G.row(i);
			int const m = G.len(i);
			bool free_nbhd = true;
			for (int k = 0; k < m && free_nbhd; k++) {
				free_nbhd = agg[p[k * S]] < 0;
			}
			if (free_nbhd) {
				for (int k = 0; k < m; k++) {
					agg[p[k * S]] = na;
				}
				agg[i] = na++;
			} else {
				bits |= uint64_t(1) << ((i - r0) & 63);
			}
		}
		passed[w] = bits;
	}
	for (int w = 0; w < nw; w++) {
		for (uint64_t bits = passed[w]; bits; bits &= bits - 1) {
			int const i = r0 + 64 * w + __builtin_ctzll(bits);
			if (agg[i] >= 0) {
				continue;   // taken by a later root
			}
			const int32_t* const p = G.row(i);
			int const m = G.len(i);
			int32_t a = -1;
			for (int k = 0; k < m && a < 0; k++) {
				a = std::max(agg[p[k * S]], -1);   // pass 2's own joins read as -1
			}
			if (a < 0) {
				a = na++;   // not reached: the sweep saw an aggregated neighbour
			}
			agg[i] = -2 - a;
		}
	}
	return na;
}

// Row offsets of up to kBuildThreads chunks of similar size (fewer when the
// machine has fewer hardware threads or BFMX_AMG_THREADS asks for fewer), cut
// at the boundaries of the segments seg (offsets 0 .. n); one chunk for one
// segment.
std::vector<int32_t> chunk_rows(const std::vector<int32_t>& seg, int n) {
	int hw = static_cast<int>(std::thread::hardware_concurrency());
	if (const char* e = std::getenv("BFMX_AMG_THREADS")) {
		hw = std::atoi(e);
	}
	int const T = std::max(1, std::min(kBuildThreads, hw));
	int64_t const target = (static_cast<int64_t>(n) + T - 1) / T;
	std::vector<int32_t> rb{0};
	for (size_t c = 1; c + 1 < seg.size(); c++) {
		if (seg[c] - rb.back() >= target) {
			rb.push_back(seg[c]);
		}
	}
	rb.push_back(n);
	return rb;
}

// f(c) for c = 0 .. T - 1, one thread each (T = 1: on this thread); an
// exception of any call is rethrown here after all have ended.
template <typename F>
void parallel_chunks(int T, const F& f) {
	if (T <= 1) {
		f(0);
		return;
	}
	std::vector<std::exception_ptr> err(T);
	std::vector<std::thread> th;
	th.reserve(T);
	for (int c = 0; c < T; c++) {
		th.emplace_back([&, c] {
			try {
				f(c);
			} catch (...) {
				err[c] = std::current_exception();
			}
		});
	}
	for (auto& t : th) {
		t.join();
	}
	for (auto& e : err) {
		if (e) {
			std::rethrow_exception(e);
		}
	}
}

// Aggregates of a level by aggregate() on the chunks of whole segments rb
// (chunk_rows), in parallel, numbered over the chunks in row order: agg (one
// entry per row, sized by the caller) and cseg[c] = the first aggregate of
// segment c (cseg.back() = their count); returns the first aggregate of each
// chunk. The first row of a segment is a root (its neighbours are all in the
// segment and not yet aggregated), and the sweeps of whole segments do not
// depend on how they are grouped.
template <typename Rows>
std::vector<int32_t> aggregate_chunks(Rows const& G, const std::vector<int32_t>& seg, const std::vector<int32_t>& rb,
                                      int32_t* agg, std::vector<int32_t>& cseg) {
	int const T = static_cast<int>(rb.size()) - 1;
	std::vector<int32_t> ab(T + 1, 0);   // first aggregate of each chunk
	parallel_chunks(T, [&](int c) { ab[c + 1] = aggregate(rb[c], rb[c + 1], G, agg); });
	for (int c = 0; c < T; c++) {
		ab[c + 1] += ab[c];
	}
	parallel_chunks(T, [&](int c) {
		for (int i = rb[c]; i < rb[c + 1]; i++) {
			int32_t const a = agg[i];
			agg[i] = (a < 0 ? -2 - a : a) + ab[c];   // pass 2's joins decoded
		}
	});
	cseg.assign(seg.size(), ab[T]);
	for (size_t c = 0; c + 1 < seg.size(); c++) {
		cseg[c] = agg[seg[c]];
	}
	return ab;
}

struct Sell {
	std::vector<int32_t> slice_ptr, col, diag_blk;
	int64_t n_slices = 0, nnzb = 0, nnzb_pad = 0;
	std::vector<int32_t> slot_of;   // block slot of each adjacency entry
};

// Sliced ELL of 3x3 blocks for the graph (ptr, adj); padding repeats the
// row's own column (with zero values).
int build_sell(int n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& adj, Sell& S) {
	S.n_slices = (n + kSlice - 1) / kSlice;
	S.slice_ptr.assign(S.n_slices + 1, 0);
	for (int64_t sl = 0; sl < S.n_slices; sl++) {
		int64_t kmax = 0;
		for (int64_t r = sl * kSlice; r < std::min<int64_t>(n, (sl + 1) * kSlice); r++) {
			kmax = std::max(kmax, ptr[r + 1] - ptr[r]);
		}
		int64_t const next = S.slice_ptr[sl] + kmax * kSlice;
		if (next * 9 >= (int64_t(1) << 31)) {
			set_error("multigrid level too large for 32-bit block indices");
			return BFMX_E_ARG;
		}
		S.slice_ptr[sl + 1] = static_cast<int32_t>(next);
	}
	S.nnzb = ptr[n];
	S.nnzb_pad = S.slice_ptr[S.n_slices];
	S.col.assign(S.nnzb_pad, 0);
	S.diag_blk.assign(n, 0);
	S.slot_of.assign(S.nnzb, 0);
	for (int64_t sl = 0; sl < S.n_slices; sl++) {
		int64_t const kmax = (S.slice_ptr[sl + 1] - S.slice_ptr[sl]) / kSlice;
		for (int r = 0; r < kSlice; r++) {
			int64_t const row = sl * kSlice + r;
			int64_t const len = row < n ? ptr[row + 1] - ptr[row] : 0;
			for (int64_t k = 0; k < kmax; k++) {
				int64_t const slot = S.slice_ptr[sl] + k * kSlice + r;
				if (k < len) {
					int32_t const c = adj[ptr[row] + k];
					S.col[slot] = c;
					S.slot_of[ptr[row] + k] = static_cast<int32_t>(slot);
					if (c == row) {
						S.diag_blk[row] = static_cast<int32_t>(slot);
					}
				} else {
					S.col[slot] = row < n ? static_cast<int32_t>(row) : 0;
				}
			}
		}
	}
	return BFMX_OK;
}

template <typename T>
int upload(T** d, const std::vector<T>& h, hipStream_t st) {
	BFMX_CHECK(palloc(d, h.size()));
	if (!h.empty()) {
		BFMX_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
	}
	return BFMX_OK;
}

void numeric_level_head(bfmx_system* s, size_t l, bool coarsest, hipStream_t st);
void numeric_level(bfmx_system* s, size_t l, hipStream_t st, bool skip_head);

// Smoothed aggregation unless the environment sets BFMX_AMG_SMOOTH=0
// (unsmoothed aggregation, P = P_t: for comparisons).
bool amg_smoothing() {
	const char* e = std::getenv("BFMX_AMG_SMOOTH");
	return !(e && e[0] == '0');
}

// Chebyshev steps per side on level l (not the coarsest).
int level_deg(const AmgParams& p, size_t l) {
	return l == 0 ? p.deg : l == 1 ? p.deg_coarse : p.deg_deep;
}

// Row ranges of the graph (ptr, adj) that no entry couples, as offsets (0,
// ..., n): a boundary follows row i when no row up to i has a neighbour
// after i (the adjacency is symmetric).
std::vector<int32_t> segments(int n, const int64_t* ptr, const int32_t* adj) {
	std::vector<int32_t> seg{0};
	int reach = 0;
	for (int i = 0; i < n; i++) {
		for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
			reach = std::max(reach, adj[e]);
		}
		if (reach <= i) {
			seg.push_back(i + 1);
		}
	}
	if (seg.back() != n) {
		seg.push_back(n);
	}
	return seg;
}

int seg_largest(const std::vector<int32_t>& seg) {
	int m = 0;
	for (size_t c = 0; c + 1 < seg.size(); c++) {
		m = std::max(m, seg[c + 1] - seg[c]);
	}
	return m;
}

// Entries of the coarsest's dense inverses (0: too large, Chebyshev instead).
int64_t dense_entries(int bs, const std::vector<int32_t>& seg) {
	int64_t total = 0;
	for (size_t c = 0; c + 1 < seg.size(); c++) {
		int64_t const n = static_cast<int64_t>(bs) * (seg[c + 1] - seg[c]);
		if (n > kDenseMax) {
			return 0;
		}
		total += n * n;
	}
	return total <= kDenseTotal ? total : 0;
}

}  // namespace

void amg_release(bfmx_system* s) {
	if (!s->amg.empty() || s->d_amg_desc || s->d_amg_pf) {
		// the pool takes blocks back only from drained streams
		(void) hipStreamSynchronize(s->stream);
	}
	if (s->amg_stream) {
		(void) hipStreamSynchronize(s->amg_stream);
		(void) hipStreamDestroy(s->amg_stream);
		s->amg_stream = nullptr;
	}
	if (s->amg_ev) {
		(void) hipEventDestroy(s->amg_ev);
		s->amg_ev = nullptr;
	}
	for (size_t l = 0; l < s->amg.size(); l++) {
		AmgLevel& L = s->amg[l];
		if (l > 0) {
			pfree(L.d_slice_ptr);
			pfree(L.d_row_len);
			pfree(L.d_col);
			pfree(L.d_diag_blk);
			pfree(L.d_val64);
			pfree(L.d_val32);
		}
		pfree(L.d_dinv);
		pfree(L.d_scal);
		pfree(L.d_B);
		pfree(L.d_B0);
		pfree(L.d_agg);
		pfree(L.d_Pt);
		pfree(L.d_agg_ptr);
		pfree(L.d_agg_nodes);
		pfree(L.d_pptr);
		pfree(L.d_pcol);
		pfree(L.d_P);
		pfree(L.d_ppos);
		pfree(L.d_rptr);
		pfree(L.d_rent);
		pfree(L.d_R);
		pfree(L.d_rrow);
		pfree(L.d_gal_ptr);
		pfree(L.d_gal_mirror);
		pfree(L.d_gal_blk);
		pfree(L.d_gal_pi);
		pfree(L.d_gal_pj);
		pfree(L.d_apptr);
		pfree(L.d_apcol);
		pfree(L.d_apval);
		pfree(L.d_dense);
		pfree(L.d_seg_ptr);
		pfree(L.d_seg_row);
		pfree(L.d_seg_off);
		pfree(L.d_b);
		pfree(L.d_x[0]);
		pfree(L.d_x[1]);
		pfree(L.d_d);
		pfree(L.d_r);
		pfree(L.d_partial);
	}
	s->amg.clear();
	pfree(s->d_amg_desc);
	pfree(s->d_amg_pf);
	s->amg_npf = 0;
	s->amg_wg = -1;
	s->amg_head0 = false;
	s->amg_num_gen = 0;
}

namespace {

// The host build's state for the level being coarsened: its graph (fptr,
// fadj: sorted rows with the diagonal), the block slot of each graph entry,
// the level's slice pointers and its segments.
struct HostLevelGraph {
	std::vector<int64_t> fptr;
	std::vector<int32_t> fadj;
	std::vector<int32_t> fslot;
	std::vector<int32_t> fslice;
};

int host_levels_from(bfmx_system* s, bool smooth, bool finest, HostLevelGraph& H, std::vector<int32_t>& fseg);

// A level's sliced ELL copied from the device: slice pointers, block
// columns and row lengths.
int level_sell_to_host(hipStream_t st, TopoGraph const& G, std::vector<int32_t>& slice, std::vector<int32_t>& col,
                       std::vector<int32_t>& len) {
	slice.resize(static_cast<size_t>(G.n_slices) + 1);
	col.resize(static_cast<size_t>(G.nnzb_pad));
	len.resize(static_cast<size_t>(G.n));
	BFMX_HIP(hipMemcpyAsync(slice.data(), G.slice_ptr, slice.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipMemcpyAsync(col.data(), G.col, col.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipMemcpyAsync(len.data(), G.row_len, len.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	BFMX_HIP(hipStreamSynchronize(st));
	return BFMX_OK;
}

// A level's graph (sliced ELL on the device) as the host build's state: its
// rows in CSR and the block slot of every entry (rows in parallel).
int level_graph_to_host(hipStream_t st, TopoGraph const& G, HostLevelGraph& H) {
	int const nf = G.n;
	std::vector<int32_t> len, col;
	BFMX_CHECK(level_sell_to_host(st, G, H.fslice, col, len));
	H.fptr.assign(static_cast<size_t>(nf) + 1, 0);
	for (int i = 0; i < nf; i++) {
		H.fptr[i + 1] = H.fptr[i] + len[i];
	}
	H.fadj.resize(H.fptr[nf]);
	H.fslot.resize(H.fptr[nf]);
	int const T = std::max(1, std::min(kBuildThreads, nf / 16384));
	parallel_chunks(T, [&](int c) {
		for (int i = static_cast<int>(static_cast<int64_t>(nf) * c / T);
		     i < static_cast<int>(static_cast<int64_t>(nf) * (c + 1) / T); i++) {
			for (int k = 0; k < len[i]; k++) {
				int32_t const slot = H.fslice[i >> 6] + k * kSlice + (i & 63);
				H.fadj[H.fptr[i] + k] = col[slot];
				H.fslot[H.fptr[i] + k] = slot;
			}
		}
	});
	return BFMX_OK;
}

// The levels below level 0, built on the host from the node adjacency
// (s->adj_ptr, s->adj). fseg receives the coarsest level's segments.
int build_levels_host(bfmx_system* s, bool smooth, std::vector<int32_t>& fseg) {
	int64_t const n0 = s->n_nodes;
	HostLevelGraph H;
	H.fslot.resize(s->nnzb);
	H.fslice = s->slice_ptr;
	for (int64_t i = 0; i < n0; i++) {
		for (int64_t k = 0; k < s->adj_ptr[i + 1] - s->adj_ptr[i]; k++) {
			H.fslot[s->adj_ptr[i] + k] = static_cast<int32_t>(s->slice_ptr[i >> 6] + k * kSlice + (i & 63));
		}
	}
	// segments of the current level: aggregates never span two, and the
	// first row of a segment is an aggregate root (its neighbours are all
	// in the segment and not yet aggregated), so the aggregates of a
	// segment are numbered consecutively and give the next level's segments
	fseg = segments(static_cast<int>(n0), s->adj_ptr.data(), s->adj.data());
	return host_levels_from(s, smooth, true, H, fseg);
}

// The levels below the current coarsest one (s->amg.back()), built on the
// host from its graph H (finest: the system's own graph s->adj_ptr / s->adj
// instead of H's), by up to kBuildThreads threads over chunks of whole
// segments; fseg: the level's segments, on return the coarsest level's.
int host_levels_from(bfmx_system* s, bool smooth, bool finest, HostLevelGraph& H, std::vector<int32_t>& fseg) {
	hipStream_t const st = s->stream;
	std::vector<AmgLevel>& lv = s->amg;
	AmgParams const& par = s->amg_par;
	std::vector<int64_t>& fptr = H.fptr;
	std::vector<int32_t>& fadj = H.fadj;
	std::vector<int32_t>& fslot = H.fslot;
	std::vector<int32_t>& fslice = H.fslice;
	int const lf0 = static_cast<int>(lv.size()) - 1;
	// coarsen to at most par.coarsest block rows (the one-workgroup
	// dense inverse of the coarsest operator costs O(n^3) per set-up:
	// ~0.1 ms at 192 DOFs, ~40 ms at 500)
	while (static_cast<int>(lv.size()) < kMaxLevels && seg_largest(fseg) > par.coarsest) {
		int const lf = static_cast<int>(lv.size()) - 1;
		bool const sys = finest && lf == lf0;
		const int64_t* ptr = sys ? s->adj_ptr.data() : fptr.data();
		const int32_t* adj = sys ? s->adj.data() : fadj.data();
		int const nf = lv[lf].nb;
		// the rows split at segment boundaries into chunks of similar size,
		// one thread each: the aggregates, prolongator rows, restriction
		// columns and coarse rows of a chunk stay inside it, so the result
		// does not depend on the split
		std::vector<int32_t> const rb = chunk_rows(fseg, nf);
		int const T = static_cast<int>(rb.size()) - 1;
		std::atomic<int> bad{0};
		std::vector<int32_t> agg(nf), cseg;
		std::vector<int32_t> const ab = aggregate_chunks(CsrRows{ptr, adj}, fseg, rb, agg.data(), cseg);
		int const na = cseg.back();
		if (!std::is_sorted(cseg.begin(), cseg.end()) || cseg[0] != 0) {
			set_error("multigrid: aggregates do not follow the level's segments");
			return BFMX_E_ARG;
		}
		// stalled when the segments still above the target size shrink by
		// less than a fifth
		int64_t big_f = 0, big_a = 0;
		for (size_t c = 0; c + 1 < fseg.size(); c++) {
			if (fseg[c + 1] - fseg[c] > par.coarsest) {
				big_f += fseg[c + 1] - fseg[c];
				big_a += cseg[c + 1] - cseg[c];
			}
		}
		if (big_a * 10 > big_f * 8) {
			break;   // coarsening stalled
		}
		// members of each aggregate (a chunk's members fill its own rows' range)
		std::vector<int32_t> aptr(na + 1, 0), anodes(nf);
		parallel_chunks(T, [&](int c) {
			std::vector<int32_t> fill(ab[c + 1] - ab[c], 0);
			for (int i = rb[c]; i < rb[c + 1]; i++) {
				fill[agg[i] - ab[c]]++;
			}
			int32_t run = rb[c];
			for (int I = ab[c]; I < ab[c + 1]; I++) {
				int32_t const cnt = fill[I - ab[c]];
				fill[I - ab[c]] = run;
				run += cnt;
				aptr[I + 1] = run;
			}
			for (int i = rb[c]; i < rb[c + 1]; i++) {
				anodes[fill[agg[i] - ab[c]]++] = i;
			}
		});
		// prolongator pattern: row i couples to the aggregates of its
		// neighbours (smoothed) or to its own aggregate (unsmoothed)
		std::vector<int32_t> pptr(nf + 1, 0);
		std::vector<std::vector<int32_t>> pcol_c(T);
		parallel_chunks(T, [&](int c) {
			std::vector<int32_t> row;
			std::vector<int32_t>& pc = pcol_c[c];
			pc.reserve(static_cast<size_t>(smooth ? 3 : 1) * (rb[c + 1] - rb[c]));
			for (int i = rb[c]; i < rb[c + 1]; i++) {
				row.clear();
				if (smooth) {
					for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
						row.push_back(agg[adj[e]]);
					}
					std::sort(row.begin(), row.end());
					row.erase(std::unique(row.begin(), row.end()), row.end());
				} else {
					row.push_back(agg[i]);
				}
				if (row.size() > 255) {
					bad = 1;
				}
				pc.insert(pc.end(), row.begin(), row.end());
				pptr[i + 1] = static_cast<int32_t>(pc.size());   // within the chunk, offset below
			}
		});
		if (bad) {
			set_error("multigrid: a row couples to more than 255 aggregates");
			return BFMX_E_UNSUPPORTED;
		}
		std::vector<int64_t> pb(T + 1, 0);   // first entry of each chunk
		for (int c = 0; c < T; c++) {
			pb[c + 1] = pb[c] + static_cast<int64_t>(pcol_c[c].size());
		}
		if (pb[T] >= (int64_t(1) << 31) / 9) {
			set_error("multigrid prolongator too large for 32-bit indices");
			return BFMX_E_ARG;
		}
		std::vector<int32_t> pcol(pb[T]);
		parallel_chunks(T, [&](int c) {
			for (int i = rb[c]; i < rb[c + 1]; i++) {
				pptr[i + 1] += static_cast<int32_t>(pb[c]);
			}
			std::copy(pcol_c[c].begin(), pcol_c[c].end(), pcol.begin() + pb[c]);
			std::vector<int32_t>().swap(pcol_c[c]);
		});
		auto entry_of = [&](int i, int32_t J) -> int32_t {   // entry of (i, J) in P; J must be in row i
			auto const beg = pcol.begin() + pptr[i], end = pcol.begin() + pptr[i + 1];
			return static_cast<int32_t>(std::lower_bound(beg, end, J) - pcol.begin());
		};
		// per fine block slot, the entry (within its row) of its column's
		// aggregate; padding slots and the unsmoothed case point at the
		// row's own aggregate
		std::vector<uint8_t> ppos(fslice.back(), 0);
		parallel_chunks(T, [&](int c) {
			for (int i = rb[c]; i < rb[c + 1]; i++) {
				auto const own = static_cast<uint8_t>(entry_of(i, agg[i]) - pptr[i]);
				for (int32_t slot = fslice[i >> 6] + (i & 63); slot < fslice[(i >> 6) + 1]; slot += kSlice) {
					ppos[slot] = own;
				}
				if (smooth) {
					for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
						ppos[fslot[e]] = static_cast<uint8_t>(entry_of(i, agg[adj[e]]) - pptr[i]);
					}
				}
			}
		});
		// columns of P (restriction), entries in row order
		std::vector<int32_t> rptr(na + 1, 0), rent(pcol.size()), rrow(pcol.size());
		parallel_chunks(T, [&](int c) {
			std::vector<int32_t> fill(ab[c + 1] - ab[c], 0);
			for (int32_t e = static_cast<int32_t>(pb[c]); e < pb[c + 1]; e++) {
				if (pcol[e] < ab[c] || pcol[e] >= ab[c + 1]) {
					bad = 1;
					return;
				}
				fill[pcol[e] - ab[c]]++;
			}
			int32_t run = static_cast<int32_t>(pb[c]);
			for (int I = ab[c]; I < ab[c + 1]; I++) {
				int32_t const cnt = fill[I - ab[c]];
				fill[I - ab[c]] = run;
				run += cnt;
				rptr[I + 1] = run;
			}
			for (int i = rb[c]; i < rb[c + 1]; i++) {
				for (int32_t e = pptr[i]; e < pptr[i + 1]; e++) {
					int32_t const at = fill[pcol[e] - ab[c]]++;
					rent[at] = e;
					rrow[at] = i;
				}
			}
		});
		if (bad) {
			set_error("multigrid: a prolongator row leaves its segment");
			return BFMX_E_ARG;
		}
		// coarse graph, the pattern of P^T A P (sorted rows, with the diagonal)
		std::vector<int64_t> cptr(na + 1, 0);
		std::vector<std::vector<int32_t>> cadj_c(T);
		parallel_chunks(T, [&](int c) {
			std::vector<int32_t> row;
			std::vector<int32_t>& ca = cadj_c[c];
			for (int I = ab[c]; I < ab[c + 1]; I++) {
				row.clear();
				for (int32_t t = rptr[I]; t < rptr[I + 1]; t++) {
					int const i = rrow[t];
					for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
						int const j = adj[e];
						row.insert(row.end(), pcol.begin() + pptr[j], pcol.begin() + pptr[j + 1]);
					}
				}
				std::sort(row.begin(), row.end());
				row.erase(std::unique(row.begin(), row.end()), row.end());
				ca.insert(ca.end(), row.begin(), row.end());
				cptr[I + 1] = static_cast<int64_t>(ca.size());   // within the chunk, offset below
			}
		});
		std::vector<int64_t> cb(T + 1, 0);
		for (int c = 0; c < T; c++) {
			cb[c + 1] = cb[c] + static_cast<int64_t>(cadj_c[c].size());
		}
		std::vector<int32_t> cadj(cb[T]);
		parallel_chunks(T, [&](int c) {
			for (int I = ab[c]; I < ab[c + 1]; I++) {
				cptr[I + 1] += cb[c];
			}
			std::copy(cadj_c[c].begin(), cadj_c[c].end(), cadj.begin() + cb[c]);
			std::vector<int32_t>().swap(cadj_c[c]);
		});
		Sell CS;
		BFMX_CHECK(build_sell(na, cptr, cadj, CS));
		// Galerkin gather lists: every (fine block A_ij, entry P_iI, entry
		// P_jJ) grouped by the coarse slot (I, J), in fine row order (the
		// slots of coarse row I belong to its chunk)
		auto coarse_slot = [&](int32_t I, int32_t J) -> int32_t {
			auto const beg = cadj.begin() + cptr[I], end = cadj.begin() + cptr[I + 1];
			return CS.slot_of[std::lower_bound(beg, end, J) - cadj.begin()];
		};
		std::vector<int64_t> gptr(CS.nnzb_pad + 1, 0);
		std::vector<int32_t> mirror(CS.nnzb_pad, -1);
		parallel_chunks(T, [&](int c) {
			for (int I = ab[c]; I < ab[c + 1]; I++) {
				for (int64_t e = cptr[I]; e < cptr[I + 1]; e++) {
					int32_t const J = cadj[e];
					mirror[CS.slot_of[e]] = J < I ? -2 : coarse_slot(J, I);
				}
				for (int32_t t = rptr[I]; t < rptr[I + 1]; t++) {
					int const i = rrow[t];
					for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
						int const j = adj[e];
						for (int32_t f = pptr[j]; f < pptr[j + 1]; f++) {
							if (pcol[f] >= I) {
								gptr[coarse_slot(I, pcol[f]) + 1]++;
							}
						}
					}
				}
			}
		});
		for (int64_t b = 0; b < CS.nnzb_pad; b++) {
			gptr[b + 1] += gptr[b];
		}
		int64_t const nterms = gptr[CS.nnzb_pad];
		std::vector<int32_t> gblk(nterms), gpi(nterms), gpj(nterms);
		{
			std::vector<int64_t> fill(gptr.begin(), gptr.end() - 1);
			parallel_chunks(T, [&](int c) {
				for (int I = ab[c]; I < ab[c + 1]; I++) {
					for (int32_t t = rptr[I]; t < rptr[I + 1]; t++) {
						int const i = rrow[t];
						int32_t const pi = rent[t];
						for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
							int const j = adj[e];
							for (int32_t f = pptr[j]; f < pptr[j + 1]; f++) {
								if (pcol[f] < I) {
									continue;
								}
								int64_t const at = fill[coarse_slot(I, pcol[f])]++;
								gblk[at] = fslot[e];
								gpi[at] = pi;
								gpj[at] = f;
							}
						}
					}
				}
			});
		}
		AmgLevel& F = lv[lf];
		F.sa = smooth;
		F.nnzp = static_cast<int64_t>(pcol.size());
		BFMX_CHECK(upload(&F.d_agg, agg, st));
		BFMX_CHECK(palloc(&F.d_Pt, static_cast<size_t>(F.bs) * 3 * nf));
		BFMX_CHECK(upload(&F.d_agg_ptr, aptr, st));
		BFMX_CHECK(upload(&F.d_agg_nodes, anodes, st));
		BFMX_CHECK(upload(&F.d_pptr, pptr, st));
		BFMX_CHECK(upload(&F.d_pcol, pcol, st));
		BFMX_CHECK(palloc(&F.d_P, static_cast<size_t>(F.bs) * 3 * pcol.size()));
		BFMX_CHECK(upload(&F.d_ppos, ppos, st));
		BFMX_CHECK(upload(&F.d_rptr, rptr, st));
		BFMX_CHECK(upload(&F.d_rent, rent, st));
		BFMX_CHECK(palloc(&F.d_R, static_cast<size_t>(F.bs) * 3 * pcol.size()));
		BFMX_CHECK(upload(&F.d_rrow, rrow, st));
		BFMX_CHECK(upload(&F.d_gal_ptr, gptr, st));
		BFMX_CHECK(upload(&F.d_gal_mirror, mirror, st));
		BFMX_CHECK(upload(&F.d_gal_blk, gblk, st));
		BFMX_CHECK(upload(&F.d_gal_pi, gpi, st));
		BFMX_CHECK(upload(&F.d_gal_pj, gpj, st));
		lv.emplace_back();   // its buffers are released with the hierarchy, also on failure
		AmgLevel& C = lv.back();
		C.nb = na;
		C.bs = 3;
		C.n_slices = CS.n_slices;
		C.nnzb = CS.nnzb;
		C.nnzb_pad = CS.nnzb_pad;
		BFMX_CHECK(upload(&C.d_slice_ptr, CS.slice_ptr, st));
		BFMX_CHECK(upload(&C.d_col, CS.col, st));
		BFMX_CHECK(upload(&C.d_diag_blk, CS.diag_blk, st));
		BFMX_CHECK(palloc(&C.d_val64, 9 * static_cast<size_t>(CS.nnzb_pad)));
		BFMX_CHECK(palloc(&C.d_val32, 9 * static_cast<size_t>(CS.nnzb_pad)));
		BFMX_HIP(hipStreamSynchronize(st));   // the host sources above are temporaries
		fptr.swap(cptr);
		fadj.swap(cadj);
		fslot.swap(CS.slot_of);
		fslice.swap(CS.slice_ptr);
		fseg.swap(cseg);
	}
	return BFMX_OK;
}

// The transfer topology of a small level, built on the host from its graph
// G (n rows, n_slices slices, as copied down) and its aggregates agg (na of
// them): the arrays topo_transfer builds on the GPU, with the same contents,
// order and conventions and the same bounds, uploaded into T's device arrays
// in one copy (topo_upload). BFMX_E_UNSUPPORTED where the device build
// reports a row beyond its bounds.
int host_transfer(hipStream_t st, SellRows const& G, int n, int64_t n_slices, const int32_t* agg, int na,
                  bool smooth, TopoLevel& T) {
	PhaseTimer phase(st, "host transfer");
	int64_t const nnzb_pad = G.slice[n_slices];
	// aggregate members in row order
	std::vector<int32_t> agg_ptr(na + 1, 0), agg_nodes(n);
	for (int i = 0; i < n; i++) {
		agg_ptr[agg[i] + 1]++;
	}
	for (int a = 0; a < na; a++) {
		agg_ptr[a + 1] += agg_ptr[a];
	}
	{
		std::vector<int32_t> fill(agg_ptr.begin(), agg_ptr.end() - 1);
		for (int i = 0; i < n; i++) {
			agg_nodes[fill[agg[i]]++] = i;
		}
	}
	// prolongator rows: the sorted distinct aggregates of the neighbours
	// (smoothed) or the own aggregate; per block slot the entry of its
	// column's aggregate (padding slots and the unsmoothed case: the own)
	std::vector<int32_t> pptr(n + 1, 0), pcol;
	std::vector<uint8_t> ppos(static_cast<size_t>(nnzb_pad), 0);
	std::vector<int32_t> row;
	for (int i = 0; i < n; i++) {
		int const len = G.len(i);
		if (len > kMaxRow) {
			set_error("multigrid: a level row has more than %d neighbours for the device build", kMaxRow);
			return BFMX_E_UNSUPPORTED;
		}
		row.clear();
		if (smooth) {
			for (int k = 0; k < len; k++) {
				row.push_back(agg[G.at(i, k)]);
			}
			std::sort(row.begin(), row.end());
			row.erase(std::unique(row.begin(), row.end()), row.end());
		} else {
			row.push_back(agg[i]);
		}
		if (row.size() > 255) {
			set_error("multigrid: a row couples to more than 255 aggregates");
			return BFMX_E_UNSUPPORTED;
		}
		auto const pos = [&](int32_t a) {
			return static_cast<uint8_t>(std::lower_bound(row.begin(), row.end(), a) - row.begin());
		};
		uint8_t const own = pos(agg[i]);
		int32_t const s0 = G.slice[i >> 6];
		int const width = (G.slice[(i >> 6) + 1] - s0) / kSlice;
		for (int k = 0; k < width; k++) {
			ppos[s0 + k * kSlice + (i & 63)] = smooth && k < len ? pos(agg[G.at(i, k)]) : own;
		}
		pcol.insert(pcol.end(), row.begin(), row.end());
		pptr[i + 1] = static_cast<int32_t>(pcol.size());
	}
	int64_t const nnzp = static_cast<int64_t>(pcol.size());
	if (nnzp >= (int64_t(1) << 31) / 9) {
		set_error("multigrid prolongator too large for 32-bit indices");
		return BFMX_E_ARG;
	}
	// P's columns: rows in order and their entries
	std::vector<int32_t> rptr(na + 1, 0), rent(nnzp), rrow(nnzp);
	for (int64_t e = 0; e < nnzp; e++) {
		rptr[pcol[e] + 1]++;
	}
	for (int a = 0; a < na; a++) {
		if (rptr[a + 1] > kMaxPCol) {
			set_error("multigrid: a prolongator column has more than %d rows for the device build", kMaxPCol);
			return BFMX_E_UNSUPPORTED;
		}
		rptr[a + 1] += rptr[a];
	}
	{
		std::vector<int32_t> fill(rptr.begin(), rptr.end() - 1);
		for (int i = 0; i < n; i++) {
			for (int32_t e = pptr[i]; e < pptr[i + 1]; e++) {
				int32_t const at = fill[pcol[e]]++;
				rrow[at] = i;
				rent[at] = e;
			}
		}
	}
	// A P: the sorted distinct aggregates of the neighbours' prolongator rows
	// (a stamp per aggregate keeps each once, so only the distinct ones are
	// sorted)
	std::vector<int32_t> apptr(n + 1, 0), apcol;
	std::vector<int32_t> stamp(na, -1);
	for (int i = 0; i < n; i++) {
		row.clear();
		int const len = G.len(i);
		for (int k = 0; k < len; k++) {
			int32_t const j = G.at(i, k);
			for (int32_t e = pptr[j]; e < pptr[j + 1]; e++) {
				if (stamp[pcol[e]] != i) {
					stamp[pcol[e]] = i;
					row.push_back(pcol[e]);
				}
			}
		}
		std::sort(row.begin(), row.end());
		if (row.size() > static_cast<size_t>(kMaxAP)) {
			set_error("multigrid: a row of A P has more than %d entries for the device build", kMaxAP);
			return BFMX_E_UNSUPPORTED;
		}
		apcol.insert(apcol.end(), row.begin(), row.end());
		apptr[i + 1] = static_cast<int32_t>(apcol.size());
	}
	// coarse rows: the sorted distinct columns of the A P rows of P's column
	std::vector<int64_t> cptr(na + 1, 0);
	std::vector<int32_t> cadj;
	std::fill(stamp.begin(), stamp.end(), -1);
	for (int I = 0; I < na; I++) {
		row.clear();
		for (int32_t t = rptr[I]; t < rptr[I + 1]; t++) {
			int32_t const i = rrow[t];
			for (int32_t e = apptr[i]; e < apptr[i + 1]; e++) {
				if (stamp[apcol[e]] != I) {
					stamp[apcol[e]] = I;
					row.push_back(apcol[e]);
				}
			}
		}
		std::sort(row.begin(), row.end());
		if (row.size() > static_cast<size_t>(kMaxCoarse)) {
			set_error("multigrid: a coarse row couples to more than %d coarse rows for the device build", kMaxCoarse);
			return BFMX_E_UNSUPPORTED;
		}
		cadj.insert(cadj.end(), row.begin(), row.end());
		cptr[I + 1] = static_cast<int64_t>(cadj.size());
	}
	// the coarse sliced ELL (padding repeats the row's own column, rows past
	// na use column 0) and the mirror of every coarse slot
	int64_t const cs = (na + kSlice - 1) / kSlice;
	std::vector<int32_t> c_slice(cs + 1, 0);
	for (int64_t sl = 0; sl < cs; sl++) {
		int64_t w = 0;
		for (int64_t r = sl * kSlice; r < std::min<int64_t>(na, (sl + 1) * kSlice); r++) {
			w = std::max(w, cptr[r + 1] - cptr[r]);
		}
		int64_t const next = c_slice[sl] + w * kSlice;
		if (next * 9 >= (int64_t(1) << 31)) {
			set_error("multigrid level too large for 32-bit block indices");
			return BFMX_E_ARG;
		}
		c_slice[sl + 1] = static_cast<int32_t>(next);
	}
	int64_t const pad = c_slice[cs];
	std::vector<int32_t> c_col(pad, 0), c_diag(na, 0), c_len(na, 0), mirror(pad, -1);
	for (int64_t I = 0; I < cs * kSlice; I++) {
		int64_t const sl = I >> 6;
		int const width = (c_slice[sl + 1] - c_slice[sl]) / kSlice;
		int const len = I < na ? static_cast<int>(cptr[I + 1] - cptr[I]) : 0;
		for (int k = 0; k < width; k++) {
			int32_t const slot = c_slice[sl] + k * kSlice + static_cast<int32_t>(I & 63);
			int32_t c = I < na ? static_cast<int32_t>(I) : 0;
			if (k < len) {
				c = cadj[cptr[I] + k];
				if (c == I) {
					c_diag[I] = slot;
				}
				int32_t m = -2;
				if (c == I) {
					m = slot;
				} else if (c > I) {
					auto const b = cadj.begin() + cptr[c], e = cadj.begin() + cptr[c + 1];
					int32_t const p = static_cast<int32_t>(std::lower_bound(b, e, static_cast<int32_t>(I)) - b);
					m = c_slice[c >> 6] + p * kSlice + (c & 63);
				}
				mirror[slot] = m;
			}
			c_col[slot] = c;
		}
		if (I < na) {
			c_len[I] = len;
		}
	}
	phase.mark("arrays");
	// device arrays, filled by one upload
	T.nnzp = nnzp;
	T.nnz_ap = static_cast<int64_t>(apcol.size());
	T.c_slices = cs;
	T.c_nnzb = cptr[na];
	T.c_nnzb_pad = pad;
	BFMX_CHECK(palloc(&T.agg_ptr, na + 1));
	BFMX_CHECK(palloc(&T.agg_nodes, n));
	BFMX_CHECK(palloc(&T.pptr, n + 1));
	BFMX_CHECK(palloc(&T.pcol, nnzp));
	BFMX_CHECK(palloc(&T.ppos, static_cast<size_t>(nnzb_pad)));
	BFMX_CHECK(palloc(&T.rptr, na + 1));
	BFMX_CHECK(palloc(&T.rent, nnzp));
	BFMX_CHECK(palloc(&T.rrow, nnzp));
	BFMX_CHECK(palloc(&T.apptr, n + 1));
	BFMX_CHECK(palloc(&T.apcol, apcol.size()));
	BFMX_CHECK(palloc(&T.c_slice_ptr, c_slice.size()));
	BFMX_CHECK(palloc(&T.c_col, c_col.size()));
	BFMX_CHECK(palloc(&T.c_diag_blk, na));
	BFMX_CHECK(palloc(&T.c_row_len, na));
	BFMX_CHECK(palloc(&T.mirror, mirror.size()));
	auto item = [](auto* d, auto const& v) {
		return UploadItem{d, v.data(), v.size() * sizeof(v[0])};
	};
	int const rc = topo_upload(st, {item(T.agg_ptr, agg_ptr), item(T.agg_nodes, agg_nodes), item(T.pptr, pptr),
	                                item(T.pcol, pcol), item(T.ppos, ppos), item(T.rptr, rptr), item(T.rent, rent),
	                                item(T.rrow, rrow), item(T.apptr, apptr), item(T.apcol, apcol),
	                                item(T.c_slice_ptr, c_slice), item(T.c_col, c_col), item(T.c_diag_blk, c_diag),
	                                item(T.c_row_len, c_len), item(T.mirror, mirror)});
	phase.mark("upload");
	return rc;
}

constexpr uint32_t kAggSalt = 0x2545F491u;   // seeds the aggregation's row priorities (per level)

// The levels below level 0, built on the GPU from the level graphs' sliced
// ELL (amg_topo.hip): the same arrays as build_levels_host, from MIS-2
// aggregates. before_finest() runs just before the finest level's
// aggregation, after_level(l) once level l's transfer topology and the next
// level's arrays exist. BFMX_E_UNSUPPORTED when a row exceeds the device
// build's bounds (the caller then builds on the host).
// The finest level's sweep over the host graph the system keeps, on a host
// thread: aggregates of all rows as one chunk (whole segments sweep alike
// however they are grouped), decoded; na = their count, or -1 when stopped.
// Destroying it stops and joins the thread.
struct EarlySweep {
	std::vector<int32_t> agg;
	int na = -1;
	std::atomic<bool> stop{false};
	std::thread th;
	void start(bfmx_system* s, int n) {
		th = std::thread([this, s, n] {
			try {
				agg.resize(n);   // here: its first touch off the build's thread
				int const a = aggregate(0, n, CsrRows{s->adj_ptr.data(), s->adj.data()}, agg.data(), &stop);
				for (int i = 0; a >= 0 && i < n; i++) {
					agg[i] = agg[i] < 0 ? -2 - agg[i] : agg[i];   // pass 2's joins
				}
				na = a;
			} catch (...) {
				na = -1;
			}
		});
	}
	void finish(bool cancel) {
		if (th.joinable()) {
			if (cancel) {
				stop = true;
			}
			th.join();
		}
	}
	~EarlySweep() { finish(true); }
};

int build_levels_device(bfmx_system* s, bool smooth, std::vector<int32_t>& fseg,
                        const std::function<int()>& before_finest, const std::function<int(int)>& after_level) {
	hipStream_t const st = s->stream;
	std::vector<AmgLevel>& lv = s->amg;
	AmgParams const& par = s->amg_par;
	TopoGraph G;
	G.n = static_cast<int>(s->n_nodes);
	G.n_slices = s->n_slices;
	G.nnzb_pad = s->nnzb_pad;
	G.slice_ptr = s->d_slice_ptr;
	G.col = s->d_col;
	G.row_len = s->d_row_len;
	PhaseTimer phase(st, "multigrid levels");
	// the finest level's sweep starts at once, beside the segments' kernels and
	// the finest level's numeric head; it is kept when the level's segments
	// make one chunk (else the chunks are swept in parallel below)
	EarlySweep early;
	if (par.host_agg_rows > 0 && G.n > par.host_level_rows && s->n_conv_seg <= 1 &&
	    s->adj_ptr.size() == static_cast<size_t>(G.n) + 1) {   // not a batch's union (many segments)
		early.start(s, G.n);
	}
	BFMX_CHECK(topo_segments(st, G, fseg));
	phase.mark("segments");
	while (static_cast<int>(lv.size()) < kMaxLevels && seg_largest(fseg) > par.coarsest) {
		int const lf = static_cast<int>(lv.size()) - 1;
		int const nf = lv[lf].nb;
		if (nf <= par.host_level_rows) {
			// a small level and those below it by the host build
			if (lf == 0 && s->adj_ptr.size() == static_cast<size_t>(nf) + 1) {
				return build_levels_host(s, smooth, fseg);
			}
			HostLevelGraph H;
			BFMX_CHECK(level_graph_to_host(st, G, H));
			phase.mark("level graph to the host");
			int const rc = host_levels_from(s, smooth, false, H, fseg);
			phase.mark("host build of the small levels");
			return rc;
		}
		int na = 0;
		std::vector<int32_t> cseg;
		int32_t* agg = nullptr;
		int ra = BFMX_OK;
		bool const sys_graph = lf == 0 && s->adj_ptr.size() == static_cast<size_t>(nf) + 1;
		bool const sweeps = par.host_agg_rows > 0 &&
		                    (sys_graph || nf <= par.host_agg_rows || seg_largest(fseg) <= par.host_agg_rows);
		// a small level swept on the host gets its transfer topology there too
		bool const host_xfer = sweeps && !sys_graph && nf <= par.host_transfer_rows;
		const int32_t *gslice = nullptr, *grlen = nullptr, *gcol = nullptr;   // the level's graph, copied down
		std::vector<int32_t> gspill;   // ... when too large for the pinned staging buffer
		std::vector<int32_t> hagg_buf;
		int32_t* hagg = nullptr;
		if (lf == 0) {
			BFMX_CHECK(before_finest());
		}
		if (sweeps) {
			// the host build's sequential sweeps, in parallel over chunks of
			// whole segments: on the finest level over the host graph the system
			// keeps, below it over the level's graph copied down, when the level
			// is small or made of small segments (a batch's union). Sweeps pack
			// the aggregates as a regular lattice on structured parts of a mesh,
			// which the FCG iterations show (21 against 26 to 35 for sweeps on
			// chunks cut inside a segment or MIS-2 aggregates, on the 2M-DOF Q1
			// beam)
			if (sys_graph) {
				std::vector<int32_t> const rb = chunk_rows(fseg, nf);
				early.finish(rb.size() > 2);
				if (rb.size() == 2 && early.na >= 0 && static_cast<int>(early.agg.size()) == nf) {
					hagg = early.agg.data();
					cseg.assign(fseg.size(), early.na);
					for (size_t c = 0; c + 1 < fseg.size(); c++) {
						cseg[c] = hagg[fseg[c]];
					}
				} else {
					hagg_buf.resize(nf);
					hagg = hagg_buf.data();
					aggregate_chunks(CsrRows{s->adj_ptr.data(), s->adj.data()}, fseg, rb, hagg, cseg);
				}
			} else {
				hagg_buf.resize(nf);
				hagg = hagg_buf.data();
				// the level's sliced ELL as it is (no CSR needed for the sweep),
				// packed by one kernel into pinned host memory: one launch and
				// one wait instead of three pageable copies
				BFMX_CHECK(topo_sell_to_host(st, G, &gslice, &grlen, &gcol, gspill));
				phase.mark("level graph to the host");
				aggregate_chunks(SellRows{gslice, gcol, grlen}, fseg, chunk_rows(fseg, nf), hagg, cseg);
			}
			phase.mark("sweeps (host)");
			na = cseg.back();
			ra = host_xfer ? BFMX_OK : palloc(&agg, nf);   // the host transfer needs no device copy
			if (ra == BFMX_OK && !host_xfer) {
				hipError_t e = hipMemcpyAsync(agg, hagg, nf * sizeof(int32_t), hipMemcpyHostToDevice, st);
				if (e == hipSuccess) {
					e = hipStreamSynchronize(st);
				}
				if (e != hipSuccess) {
					set_error("upload of the aggregates: %s", hipGetErrorString(e));
					ra = BFMX_E_HIP;
				}
			}
		} else {
			ra = topo_aggregate(st, G, kAggSalt + static_cast<uint32_t>(lf), par.seed_frac, par.seed_growth, fseg,
			                    &agg, &na, cseg);
		}
		lv[lf].d_agg = agg;   // released with the hierarchy
		BFMX_CHECK(ra);
		phase.mark("aggregation");
		if (!std::is_sorted(cseg.begin(), cseg.end()) || cseg[0] != 0 || cseg.back() != na) {
			set_error("multigrid: aggregates do not follow the level's segments");
			return BFMX_E_ARG;
		}
		// stalled when the segments still above the target size shrink by
		// less than a fifth
		int64_t big_f = 0, big_a = 0;
		for (size_t c = 0; c + 1 < fseg.size(); c++) {
			if (fseg[c + 1] - fseg[c] > par.coarsest) {
				big_f += fseg[c + 1] - fseg[c];
				big_a += cseg[c + 1] - cseg[c];
			}
		}
		if (big_a * 10 > big_f * 8) {
			pfree(lv[lf].d_agg);
			break;   // coarsening stalled
		}
		TopoLevel T;
		int const rt = host_xfer ? host_transfer(st, SellRows{gslice, gcol, grlen}, nf, G.n_slices, hagg, na, smooth, T)
		                         : topo_transfer(st, G, agg, na, smooth, T);
		if (rt != BFMX_OK) {
			topo_release(T);
			return rt;
		}
		phase.mark("transfer patterns");
		{
			AmgLevel& F = lv[lf];
			F.sa = smooth;
			F.nnzp = T.nnzp;
			F.d_agg_ptr = T.agg_ptr;
			F.d_agg_nodes = T.agg_nodes;
			F.d_pptr = T.pptr;
			F.d_pcol = T.pcol;
			F.d_ppos = T.ppos;
			F.d_rptr = T.rptr;
			F.d_rent = T.rent;
			F.d_rrow = T.rrow;
			F.d_gal_mirror = T.mirror;
			F.nnz_ap = T.nnz_ap;
			F.d_apptr = T.apptr;
			F.d_apcol = T.apcol;
		}
		lv.emplace_back();   // its buffers are released with the hierarchy, also on failure
		{
			AmgLevel& C = lv.back();
			C.nb = na;
			C.bs = 3;
			C.n_slices = T.c_slices;
			C.nnzb = T.c_nnzb;
			C.nnzb_pad = T.c_nnzb_pad;
			C.d_slice_ptr = T.c_slice_ptr;
			C.d_col = T.c_col;
			C.d_diag_blk = T.c_diag_blk;
			C.d_row_len = T.c_row_len;
			BFMX_CHECK(palloc(&C.d_val64, 9 * static_cast<size_t>(C.nnzb_pad)));
			BFMX_CHECK(palloc(&C.d_val32, 9 * static_cast<size_t>(C.nnzb_pad)));
			G.n = na;
			G.n_slices = C.n_slices;
			G.nnzb_pad = C.nnzb_pad;
			G.slice_ptr = C.d_slice_ptr;
			G.col = C.d_col;
			G.row_len = C.d_row_len;
		}
		AmgLevel& F = lv[lf];
		BFMX_CHECK(palloc(&F.d_Pt, static_cast<size_t>(F.bs) * 3 * nf));
		BFMX_CHECK(palloc(&F.d_P, static_cast<size_t>(F.bs) * 3 * F.nnzp));
		BFMX_CHECK(palloc(&F.d_R, static_cast<size_t>(F.bs) * 3 * F.nnzp));
		BFMX_CHECK(palloc(&F.d_apval, static_cast<size_t>(F.bs) * 3 * F.nnz_ap));
		phase.mark("level values");
		fseg.swap(cseg);
		BFMX_CHECK(after_level(lf));
	}
	return BFMX_OK;
}

}  // namespace

// Needs the system's FP32 matrix buffer (s->d_val32) allocated: level 0
// reads it. device: the levels' topology is built on the GPU (amg_topo.hip),
// else on the host.
int amg_build_impl(bfmx_system* s, bool device) {
	if (!s->amg.empty()) {
		return BFMX_OK;
	}
	if (!s->d_val32) {
		set_error("multigrid set-up before the FP32 matrix buffer");
		return BFMX_E_ARG;
	}
	if (!s->d_coords && s->coords.size() != static_cast<size_t>(2 * s->n_nodes)) {
		set_error("multigrid set-up needs the node coordinates");
		return BFMX_E_ARG;
	}
	int const rc = [&]() -> int {
		hipStream_t const st = s->stream;
		PhaseTimer phase(st, "multigrid build");
		std::vector<AmgLevel>& lv = s->amg;
		int64_t const n0 = s->n_nodes;
		lv.emplace_back();
		{
			AmgLevel& L0 = lv[0];
			L0.nb = static_cast<int>(n0);
			L0.bs = 2;
			L0.n_slices = s->n_slices;
			L0.nnzb = s->nnzb;
			L0.nnzb_pad = s->nnzb_pad;
			L0.d_slice_ptr = s->d_slice_ptr;
			L0.d_row_len = s->d_row_len;
			L0.d_col = s->d_col;
			L0.d_diag_blk = s->d_diag_blk;
			L0.d_val64 = s->d_val;
			L0.d_val32 = s->d_val32;
			// rigid-body modes in coordinates centred on the mesh, scaled to its
			// half-size: on the device from its coordinates, else (the union of a
			// batch, whose coordinates are gathered on the host) on the host
			if (s->d_coords) {
				double* part = nullptr;
				BFMX_CHECK(palloc(&L0.d_B0, 6 * n0));
				BFMX_CHECK(palloc(&part, 4 * kBboxGrid));
				k_bbox<<<kBboxGrid, kBlock, 0, st>>>(n0, s->d_coords, part);
				k_modes0<<<grid(n0), kBlock, 0, st>>>(n0, part, s->d_coords, L0.d_B0);
				hipError_t const e = hipStreamSynchronize(st);
				pfree(part);
				BFMX_HIP(e);
				BFMX_HIP(hipGetLastError());
			} else {
				double lo[2] = {s->coords[0], s->coords[1]}, hi[2] = {s->coords[0], s->coords[1]};
				for (int64_t i = 0; i < n0; i++) {
					for (int a = 0; a < 2; a++) {
						lo[a] = std::min(lo[a], s->coords[2 * i + a]);
						hi[a] = std::max(hi[a], s->coords[2 * i + a]);
					}
				}
				double const cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]);
				double L = 0.5 * std::max(hi[0] - lo[0], hi[1] - lo[1]);
				if (!(L > 0)) {
					L = 1;
				}
				std::vector<double> B0(6 * n0);
				for (int64_t i = 0; i < n0; i++) {
					double const x = (s->coords[2 * i] - cx) / L, y = (s->coords[2 * i + 1] - cy) / L;
					double const m[6] = {1, 0, -y, 0, 1, x};
					std::copy(m, m + 6, &B0[6 * i]);
				}
				BFMX_CHECK(upload(&L0.d_B0, B0, st));
				BFMX_HIP(hipStreamSynchronize(st));
			}
		}
		phase.mark("rigid-body modes");
		bool const smooth = amg_smoothing();
		AmgParams& par = s->amg_par;
		par = AmgParams{};
		// per-element E spread over more than a decade: three finest-level
		// steps (2M-DOF beam, E log-uniform over two decades: 51 / 32 / 26
		// FCG iterations and a first solve of 31.5 / 27.1 / 25.3 ms with one /
		// two / three steps; one step stays best with a single E, 20.0 ms
		// against 21.0 / 22.1)
		if (s->e_spread > kSpreadHetero) {
			par.deg = kDegSpread;
		}
		if (const char* e = std::getenv("BFMX_AMG_DEG")) {
			par.deg = std::min(kMaxDeg, std::max(1, std::atoi(e)));
		}
		if (const char* e = std::getenv("BFMX_AMG_DEG_COARSE")) {
			par.deg_coarse = std::min(kMaxDeg, std::max(1, std::atoi(e)));
		}
		if (const char* e = std::getenv("BFMX_AMG_DEG_DEEP")) {
			par.deg_deep = std::min(kMaxDeg, std::max(1, std::atoi(e)));
		}
		if (const char* e = std::getenv("BFMX_AMG_RATIO")) {
			par.ratio = std::max(1.5, std::atof(e));
		}
		if (const char* e = std::getenv("BFMX_AMG_COARSEST")) {
			par.coarsest = std::max(1, std::atoi(e));
		}
		if (const char* e = std::getenv("BFMX_AMG_OMEGA")) {
			par.omega = std::atof(e);
		}
		if (const char* e = std::getenv("BFMX_AMG_SEED")) {
			par.seed_frac = std::min(1.0, std::max(1e-6, std::atof(e)));
		}
		if (const char* e = std::getenv("BFMX_AMG_SEED_GROWTH")) {
			par.seed_growth = std::max(1.01, std::atof(e));
		}
		if (const char* e = std::getenv("BFMX_AMG_HOST_AGG")) {
			par.host_agg_rows = std::max(0, std::atoi(e));
		}
		if (const char* e = std::getenv("BFMX_AMG_HOST_TRANSFER")) {
			par.host_transfer_rows = std::max(0, std::atoi(e));
		}
		if (const char* e = std::getenv("BFMX_AMG_NUMERIC_WGS")) {
			par.numeric_wgs = std::max(0, std::atoi(e));
		}
		if (const char* e = std::getenv("BFMX_AMG_HOST_LEVELS")) {
			par.host_level_rows = std::max(0, std::atoi(e));
		}
		if (const char* e = std::getenv("BFMX_AMG_POWER")) {
			par.power = std::max(0, std::atoi(e));
		}
		par.power_coarse = par.power;
		if (const char* e = std::getenv("BFMX_AMG_POWER_COARSE")) {
			par.power_coarse = std::max(0, std::atoi(e));
		}
		auto vectors = [&](AmgLevel& L) -> int {
			if (L.d_b) {
				return BFMX_OK;
			}
			size_t const m = static_cast<size_t>(L.bs) * L.nb;
			BFMX_CHECK(palloc(&L.d_b, m));
			BFMX_CHECK(palloc(&L.d_x[0], m));
			BFMX_CHECK(palloc(&L.d_x[1], m));
			BFMX_CHECK(palloc(&L.d_d, m));
			BFMX_CHECK(palloc(&L.d_r, m));
			BFMX_CHECK(palloc(&L.d_dinv, static_cast<size_t>(L.bs) * L.bs * L.nb));
			BFMX_CHECK(palloc(&L.d_scal, kScalSize));
			BFMX_CHECK(palloc(&L.d_partial, grid(L.nb) + grid(static_cast<int64_t>(L.bs) * L.nb)));
			BFMX_CHECK(palloc(&L.d_B, 3 * m));
			return BFMX_OK;
		};
		// the finest level's own part of the numeric set-up (inverse
		// diagonal, power iteration, modes) needs no hierarchy: the device
		// build queues it just before the host aggregates the finest level, so
		// the GPU runs it meanwhile (the FP32 copy of the matrix it reads is
		// queued before the build)
		auto head0 = [&]() -> int {
			if (n0 > 2 * par.coarsest) {
				BFMX_CHECK(vectors(lv[0]));
				numeric_level_head(s, 0, false, st);
				BFMX_HIP(hipGetLastError());
				s->amg_head0 = true;
			}
			return BFMX_OK;
		};
		// a level's numeric set-up (its prolongator and the next level's
		// Galerkin operator) needs only its own topology and the levels above:
		// it goes on a second stream as soon as that topology is complete, so
		// the GPU runs it while the host builds the levels below (the set-up
		// kernels of all levels are the same as on one stream; BFMX_AMG_OVERLAP=0
		// queues them after the build, on the system's stream)
		bool overlap = true;
		if (const char* e = std::getenv("BFMX_AMG_OVERLAP"); e && e[0] == '0') {
			overlap = false;
		}
		if (overlap && !s->amg_stream) {
			BFMX_HIP(hipStreamCreateWithFlags(&s->amg_stream, hipStreamNonBlocking));
		}
		if (overlap && !s->amg_ev) {
			BFMX_HIP(hipEventCreateWithFlags(&s->amg_ev, hipEventDisableTiming));
		}
		size_t queued = 0;   // levels whose numeric set-up is on the second stream
		auto queue_numeric = [&](size_t upto) -> int {
			BFMX_HIP(hipEventRecord(s->amg_ev, st));
			BFMX_HIP(hipStreamWaitEvent(s->amg_stream, s->amg_ev, 0));
			for (; queued < upto; queued++) {
				numeric_level(s, queued, s->amg_stream, queued == 0 && s->amg_head0);
			}
			BFMX_HIP(hipGetLastError());
			return BFMX_OK;
		};
		auto after_level = [&](int lf) -> int {
			if (overlap) {
				BFMX_CHECK(vectors(lv[lf]));
				BFMX_CHECK(vectors(lv[lf + 1]));
				BFMX_CHECK(queue_numeric(static_cast<size_t>(lf) + 1));
			}
			return BFMX_OK;
		};
		std::vector<int32_t> fseg;
		BFMX_CHECK(device ? build_levels_device(s, smooth, fseg, head0, after_level)
		                  : build_levels_host(s, smooth, fseg));
		phase.mark("levels");
		for (auto& L : lv) {
			BFMX_CHECK(vectors(L));
		}
		{
			AmgLevel& Lc = lv.back();
			Lc.n_seg = static_cast<int>(fseg.size()) - 1;
			int64_t const entries = dense_entries(Lc.bs, fseg);
			if (entries > 0) {
				std::vector<int32_t> seg_row(Lc.nb);
				std::vector<int64_t> seg_off(Lc.n_seg);
				int64_t off = 0;
				for (int c = 0; c < Lc.n_seg; c++) {
					std::fill(seg_row.begin() + fseg[c], seg_row.begin() + fseg[c + 1], c);
					seg_off[c] = off;
					int64_t const n = static_cast<int64_t>(Lc.bs) * (fseg[c + 1] - fseg[c]);
					off += n * n;
				}
				BFMX_CHECK(palloc(&Lc.d_dense, static_cast<size_t>(entries)));
				Lc.n_dense = entries;
				BFMX_CHECK(upload(&Lc.d_seg_ptr, fseg, st));
				BFMX_CHECK(upload(&Lc.d_seg_row, seg_row, st));
				BFMX_CHECK(upload(&Lc.d_seg_off, seg_off, st));
				BFMX_HIP(hipStreamSynchronize(st));
			}
		}
		phase.mark("level vectors, coarsest segments");
		if (const char* e = std::getenv("BFMX_AMG_VERBOSE"); e && e[0] == '1') {
			for (size_t l = 0; l < lv.size(); l++) {
				std::fprintf(stderr, "multigrid level %zu: %d rows of %d DOFs, %lld blocks (%lld with the slice padding)\n",
				             l, lv[l].nb, lv[l].bs, static_cast<long long>(lv[l].nnzb),
				             static_cast<long long>(lv[l].nnzb_pad));
			}
			AmgLevel const& Lc = lv.back();
			std::fprintf(stderr, "multigrid coarsest: %d segments, largest %d rows, %s\n", Lc.n_seg, seg_largest(fseg),
			             Lc.d_dense ? "dense inverses" : "Chebyshev iteration");
		}
		// the small coarse levels run in one workgroup
		s->amg_wg = -1;
		for (size_t l = 1; l < lv.size(); l++) {
			if (lv[l].nb <= kWgRows) {
				s->amg_wg = static_cast<int>(l);
				break;
			}
		}
		if (s->amg_wg > 0) {
			std::vector<AmgLevelDesc> desc;
			for (size_t l = s->amg_wg; l < lv.size(); l++) {
				AmgLevel const& L = lv[l];
				bool const last = l + 1 == lv.size();
				AmgLevelDesc D{};
				D.nb = L.nb;
				D.deg = last ? kCoarseDeg : level_deg(s->amg_par, l);
				D.dense = last && L.d_dense ? 1 : 0;
				D.slice_ptr = L.d_slice_ptr;
				D.col = L.d_col;
				D.val32 = L.d_val32;
				D.dinv = L.d_dinv;
				D.scal = L.d_scal;
				D.b = L.d_b;
				D.x0 = L.d_x[0];
				D.x1 = L.d_x[1];
				D.d = L.d_d;
				D.r = L.d_r;
				D.rptr = L.d_rptr;
				D.R = L.d_R;
				D.rrow = L.d_rrow;
				D.P = L.d_P;
				D.pptr = L.d_pptr;
				D.pcol = L.d_pcol;
				D.inv = L.d_dense;
				D.seg_ptr = L.d_seg_ptr;
				D.seg_row = L.d_seg_row;
				D.seg_off = L.d_seg_off;
				desc.push_back(D);
			}
			BFMX_CHECK(upload(&s->d_amg_desc, desc, st));
			// the lines of the arrays the cycle reads before writing (its
			// scratch vectors aside), prefetched at its start
			std::vector<uint64_t> pf;
			auto lines = [&](const void* p, size_t bytes) {
				for (size_t off = 0; off < bytes && pf.size() < kPfMax; off += kPfLine) {
					pf.push_back(reinterpret_cast<uint64_t>(p) + off);
				}
			};
			for (size_t l = s->amg_wg; l < lv.size(); l++) {
				AmgLevel const& L = lv[l];
				size_t const f4 = sizeof(float), i4 = sizeof(int32_t), nb = L.nb, nnzp = L.nnzp;
				size_t const pad = L.nnzb_pad;
				lines(L.d_slice_ptr, (L.n_slices + 1) * i4);
				lines(L.d_col, pad * i4);
				lines(L.d_val32, 9 * pad * f4);
				lines(L.d_dinv, 9 * nb * f4);
				lines(L.d_scal, kScalSize * sizeof(double));
				if (static_cast<int>(l) == s->amg_wg) {
					lines(L.d_b, 3 * nb * f4);
				}
				if (l + 1 < lv.size()) {
					lines(L.d_pptr, (nb + 1) * i4);
					lines(L.d_pcol, nnzp * i4);
					lines(L.d_P, 9 * nnzp * f4);
					lines(L.d_rptr, (static_cast<size_t>(lv[l + 1].nb) + 1) * i4);
					lines(L.d_R, 9 * nnzp * f4);
					lines(L.d_rrow, nnzp * i4);
				} else if (L.d_dense) {
					lines(L.d_dense, static_cast<size_t>(L.n_dense) * sizeof(double));
					lines(L.d_seg_ptr, (static_cast<size_t>(L.n_seg) + 1) * i4);
					lines(L.d_seg_row, nb * i4);
					lines(L.d_seg_off, static_cast<size_t>(L.n_seg) * sizeof(int64_t));
				}
			}
			BFMX_CHECK(upload(&s->d_amg_pf, pf, st));
			s->amg_npf = static_cast<int>(pf.size());
			BFMX_HIP(hipStreamSynchronize(st));
		}
		if (overlap) {
			// the levels left (the coarsest, levels built on the host), then
			// the system's stream waits for the second one
			BFMX_CHECK(queue_numeric(lv.size()));
			BFMX_HIP(hipEventRecord(s->amg_ev, s->amg_stream));
			BFMX_HIP(hipStreamWaitEvent(st, s->amg_ev, 0));
			s->amg_head0 = false;
			s->amg_num_gen = s->val_gen;
		}
		return BFMX_OK;
	}();
	if (rc != BFMX_OK) {
		amg_release(s);
	}
	return rc;
}

// The hierarchy's topology is built on the GPU unless the environment sets
// BFMX_AMG_BUILD=host, or a level exceeds the device build's bounds (then on
// the host, from the node adjacency).
int amg_build(bfmx_system* s) {
	if (!s->amg.empty()) {
		return BFMX_OK;
	}
	const char* e = std::getenv("BFMX_AMG_BUILD");
	bool const device = !(e && std::string(e) == "host") && s->d_row_len;
	int rc = amg_build_impl(s, device);
	if (rc == BFMX_E_UNSUPPORTED && device) {
		rc = amg_build_impl(s, false);
	}
	return rc;
}

// Modes, prolongators, Galerkin operators, smoother data and the coarsest
// inverse, from the current FP64 matrix and Dirichlet set, level by level:
// a level's smoother data (inverse diagonal blocks, lambda_max) come first,
// since its prolongator smoothing uses them.
namespace {

// The part of a level's numeric set-up that needs only its own matrix: the
// inverse diagonal blocks and Gershgorin bounds, the power iteration for
// lambda_max(D^-1 A) and the Chebyshev coefficients (on the finest level
// also its near-null-space modes with the fixed DOFs removed).
void numeric_level_head(bfmx_system* s, size_t l, bool coarsest, hipStream_t st) {
	AmgLevel& L = s->amg[l];
	if (l == 0) {
		k_fine_modes<<<grid(L.nb), kBlock, 0, st>>>(L.nb, L.d_B0, s->d_fixed, L.d_B);
	}
	int const g = grid(L.nb);
	if (L.bs == 2) {
		k_diag_gersh<2><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_diag_blk, L.d_val64, L.d_dinv, L.d_partial);
	} else {
		k_diag_gersh<3><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_diag_blk, L.d_val64, L.d_dinv, L.d_partial);
	}
	int const power = coarsest ? 0 : l == 0 ? s->amg_par.power : s->amg_par.power_coarse;
	if (power > 0) {
		// the Gershgorin maxima stay in d_partial's first g entries; the
		// power sums use the entries after them
		double* const pp = L.d_partial + g;
		int64_t const m = static_cast<int64_t>(L.bs) * L.nb;
		k_power_init<<<grid(m), kBlock, 0, st>>>(m, L.d_x[0], pp);
		k_power_norm<<<1, kBlock, 0, st>>>(grid(m), pp, L.d_scal);
		// the iterates are normalised only for the last step, whose input is
		// then a unit vector: lambda_max(D^-1 A) lies between 1 and a few
		// units, so ten unnormalised FP32 steps neither overflow nor underflow
		for (int k = 0; k < power; k++) {
			float* const xin = L.d_x[k & 1];
			float* const yout = L.d_x[(k & 1) ^ 1];
			if (L.bs == 2) {
				k_power_step<2><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val32, L.d_dinv, xin,
				                                      L.d_scal, yout, pp);
			} else {
				k_power_step<3><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val32, L.d_dinv, xin,
				                                      L.d_scal, yout, pp);
			}
			if (k + 2 >= power) {
				k_power_norm<<<1, kBlock, 0, st>>>(g, pp, L.d_scal);
			}
		}
	}
	k_cheb_coef<<<1, kBlock, 0, st>>>(g, L.d_partial, coarsest ? kCoarseRatio : s->amg_par.ratio,
	                                  coarsest ? kCoarseDeg : level_deg(s->amg_par, l), power > 0, L.d_scal);
}

// Level l's numeric set-up on stream st: its head (unless skip_head: the
// build queued it already) and, below the coarsest level, its prolongator and
// the next level's Galerkin operator; on a coarsest level with dense inverses,
// those inverses instead.
void numeric_level(bfmx_system* s, size_t l, hipStream_t st, bool skip_head) {
	std::vector<AmgLevel>& lv = s->amg;
	AmgLevel& L = lv[l];
	bool const coarsest = l + 1 == lv.size();
	if (coarsest && L.d_dense) {
		if (L.bs == 2) {
			k_dense<2><<<grid(L.nb), kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val64, L.d_seg_ptr,
			                                          L.d_seg_row, L.d_seg_off, L.d_dense);
		} else {
			k_dense<3><<<grid(L.nb), kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val64, L.d_seg_ptr,
			                                          L.d_seg_row, L.d_seg_off, L.d_dense);
		}
		if (L.n_seg == 1) {
			k_gauss_jordan<kGj><<<1, kGj, 0, st>>>(L.bs, L.d_seg_ptr, L.d_seg_off, L.d_dense);
		} else {
			k_gauss_jordan<kGjSeg><<<L.n_seg, kGjSeg, 0, st>>>(L.bs, L.d_seg_ptr, L.d_seg_off, L.d_dense);
		}
		return;
	}
	int const g = grid(L.nb);
	// on the second stream of a first build, the largest kernels' grids capped
	int const cap = st == s->amg_stream ? s->amg_par.numeric_wgs : 0;
	auto capped = [cap](int natural) { return cap > 0 ? std::min(natural, cap) : natural; };
	if (!skip_head) {
		numeric_level_head(s, l, coarsest, st);
	}
	if (coarsest) {
		return;
	}
	AmgLevel& C = lv[l + 1];
	int64_t const ns = C.nnzb_pad;
	if (L.bs == 2) {
		k_tentative<2><<<grid(C.nb), kBlock, 0, st>>>(C.nb, L.d_agg_ptr, L.d_agg_nodes, L.d_B, L.d_Pt, C.d_B);
		k_smooth_p<2><<<capped(g), kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val64, L.d_val32, L.d_diag_blk,
		                                    L.d_Pt, L.d_pptr, L.d_ppos, L.d_scal, L.sa ? s->amg_par.omega : 0.0,
		                                    L.d_P);
		k_transpose_p<2><<<grid(L.nnzp), kBlock, 0, st>>>(L.nnzp, L.d_rent, L.d_P, L.d_R);
		if (L.d_apptr) {
			// finest level: windows of 8 entries per thread for A P (332 us
			// on the 2M-DOF beam against 402 / 383 us with 4 / 16 and 738 us
			// with a lane per entry), a lane per coarse slot for P^T (A P)
			// (351 us against 382 us with 16 lanes per slot)
			k_ap<2, 8><<<capped(g), kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_row_len, L.d_val32, L.d_pptr,
			                                 L.d_pcol, L.d_P, L.d_apptr, L.d_apcol, L.d_apval);
			k_rap_lanes<2><<<capped(grid(C.n_slices * kSlice * 16)), kBlock, 0, st>>>(
			    C.n_slices * kSlice, C.nb, C.d_slice_ptr, C.d_col, C.d_row_len, C.d_diag_blk, L.d_gal_mirror,
			    L.d_rptr, L.d_rrow, L.d_R, L.d_apptr, L.d_apcol, L.d_apval, C.d_val64, C.d_val32);
		} else {
			k_galerkin<2, float><<<grid(ns * kGalLanes), kBlock, 0, st>>>(ns, L.d_gal_mirror, L.d_gal_ptr,
			                                                               L.d_gal_blk, L.d_gal_pi, L.d_gal_pj,
			                                                               L.d_val32, L.d_P, C.d_val64,
			                                                               C.d_val32);
		}
	} else {
		k_tentative<3><<<grid(C.nb), kBlock, 0, st>>>(C.nb, L.d_agg_ptr, L.d_agg_nodes, L.d_B, L.d_Pt, C.d_B);
		k_smooth_p<3><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val64, L.d_val32, L.d_diag_blk,
		                                    L.d_Pt, L.d_pptr, L.d_ppos, L.d_scal, L.sa ? s->amg_par.omega : 0.0,
		                                    L.d_P);
		k_transpose_p<3><<<grid(L.nnzp), kBlock, 0, st>>>(L.nnzp, L.d_rent, L.d_P, L.d_R);
		if (L.d_apptr) {
			// coarser levels (wider rows, longer prolongator columns): a lane
			// per entry for A P (39 us per level against 51-69 us with
			// windows), 16 lanes per coarse slot for P^T (A P) (23 us against
			// 52 us with a lane per slot)
			k_ap_lanes<3><<<grid(static_cast<int64_t>(L.nb) * 16), kBlock, 0, st>>>(
			    L.nb, L.d_slice_ptr, L.d_col, L.d_row_len, L.d_val32, L.d_pptr, L.d_pcol, L.d_P, L.d_apptr,
			    L.d_apcol, L.d_apval);
			k_rap<3><<<grid(C.n_slices * kSlice * 64), kBlock, 0, st>>>(
			    C.n_slices * kSlice, C.nb, C.d_slice_ptr, C.d_col, C.d_row_len, C.d_diag_blk, L.d_gal_mirror,
			    L.d_rptr, L.d_rrow, L.d_R, L.d_apptr, L.d_apcol, L.d_apval, C.d_val64, C.d_val32);
		} else {
			k_galerkin<3, float><<<grid(ns * kGalLanes), kBlock, 0, st>>>(ns, L.d_gal_mirror, L.d_gal_ptr,
			                                                               L.d_gal_blk, L.d_gal_pi, L.d_gal_pj,
			                                                               L.d_val32, L.d_P, C.d_val64,
			                                                               C.d_val32);
		}
	}
}

}  // namespace

int amg_setup_numeric(bfmx_system* s) {
	if (!s->amg_head0 && s->amg_num_gen == s->val_gen) {
		return BFMX_OK;   // made from the current matrix values
	}
	bool const head0 = s->amg_head0;   // launched by the build, ahead of the finest level's aggregation
	s->amg_head0 = false;
	for (size_t l = 0; l < s->amg.size(); l++) {
		numeric_level(s, l, s->stream, l == 0 && head0);
	}
	BFMX_HIP(hipGetLastError());
	s->amg_num_gen = s->val_gen;
	return BFMX_OK;
}

namespace {

template <int BS, typename BT>
void cheb(hipStream_t st, AmgLevel& L, const BT* b, bool first, int step, const float* xin, float* xout,
          const double* q, double* partial, const Scalars* S, SegView sv = SegView{}) {
	int const g = grid(L.nb);
	if (first) {
		k_cheb<BS, BT, true, false><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val32, b, L.d_dinv, xin,
		                                                  xout, L.d_d, L.d_scal, step, nullptr, nullptr, sv, S);
	} else if (q) {
		k_cheb<BS, BT, false, true><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val32, b, L.d_dinv, xin,
		                                                  xout, L.d_d, L.d_scal, step, q, partial, sv, S);
	} else {
		k_cheb<BS, BT, false, false><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val32, b, L.d_dinv,
		                                                   xin, xout, L.d_d, L.d_scal, step, nullptr, nullptr, sv,
		                                                   S);
	}
}

// V-cycle on level l for the right-hand side b (FP64 on level 0, FP32
// below); returns the buffer that holds the level's correction. With q, the
// last kernel also writes the per-workgroup partial sums of x . q.
template <int BS, typename BT>
float* vcycle(bfmx_system* s, size_t l, const BT* b, const double* q, double* partial, const Scalars* S,
              bool first_done = false) {
	hipStream_t const st = s->stream;
	AmgLevel& L = s->amg[l];
	bool const coarsest = l + 1 == s->amg.size();
	if (BS == 3 && static_cast<int>(l) == s->amg_wg) {
		// b is this level's own d_b (restricted into it); the workgroup cycle
		// runs every level from here down
		int const nl = static_cast<int>(s->amg.size() - l);
		k_coarse_cycle<<<1, kWg, 0, st>>>(s->d_amg_desc, nl, s->d_amg_pf, s->amg_npf, S);
		return nl == 1 && L.d_dense ? L.d_x[0] : L.d_x[1];
	}
	if (coarsest && L.d_dense) {
		int const n = BS * L.nb;
		int const g = (n + kDenseRowsPerBlock - 1) / kDenseRowsPerBlock;
		k_dense_solve<BS, BT><<<g, kBlock, 0, st>>>(n, L.d_seg_ptr, L.d_seg_row, L.d_seg_off, L.d_dense, b, L.d_x[0],
		                                            S);
		if (q) {
			k_dot_partials<<<grid(n), kBlock, 0, st>>>(n, L.d_x[0], q, partial, S);
		}
		return L.d_x[0];
	}
	int const deg = coarsest ? kCoarseDeg : level_deg(s->amg_par, l);
	int cur = 0;
	// one step per side: the first step goes with the residual (k_first_resid)
	bool const first_resid = !coarsest && !first_done && deg == 1;
	if (!first_done && !first_resid) {
		cheb<BS, BT>(st, L, b, true, 0, nullptr, L.d_x[0], nullptr, nullptr, S);
	}
	if (coarsest) {
		for (int k = 1; k < deg; k++) {
			bool const last = k + 1 == deg;
			cheb<BS, BT>(st, L, b, false, k, L.d_x[cur], L.d_x[cur ^ 1], last ? q : nullptr, last ? partial : nullptr,
			             S);
			cur ^= 1;
		}
		return L.d_x[cur];
	}
	for (int k = 1; k < deg; k++) {
		cheb<BS, BT>(st, L, b, false, k, L.d_x[cur], L.d_x[cur ^ 1], nullptr, nullptr, S);
		cur ^= 1;
	}
	AmgLevel& C = s->amg[l + 1];
	int const g = grid(L.nb);
	if (first_resid) {
		k_first_resid<BS, BT><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val32, b, L.d_dinv, L.d_scal,
		                                            L.d_x[0], L.d_r, S);
	} else {
		k_level_resid<BS, BT><<<g, kBlock, 0, st>>>(L.nb, L.d_slice_ptr, L.d_col, L.d_val32, b, L.d_x[cur], L.d_r,
		                                            S);
	}
	// the restriction also takes the next level's first Chebyshev step when
	// that level starts with one on the grid (not the one-workgroup cycle, not
	// a dense coarsest)
	bool const fuse = static_cast<int>(l + 1) != s->amg_wg && !(l + 2 == s->amg.size() && C.d_dense);
	k_restrict<BS><<<grid(static_cast<int64_t>(C.nb) * kRestrictLanes), kBlock, 0, st>>>(
	    C.nb, L.d_rptr, L.d_rrow, L.d_R, L.d_r, C.d_b, fuse ? C.d_dinv : nullptr, C.d_scal, C.d_x[0],
	    C.d_d, S);
	float* xc = vcycle<3, float>(s, l + 1, C.d_b, nullptr, nullptr, S, fuse);
	int k0 = 0;
	if (BS == 3 && L.d_apptr && !q) {
		// a coarse level: the prolongation with the first post-smoothing step
		// (k_prolong_cheb; on the finest level a row of A P costs more than the
		// row of A it replaces, and the x . q sums go with the last step)
		k_prolong_cheb<BS><<<grid(static_cast<int64_t>(L.nb) * kProlongLanes), kBlock, 0, st>>>(
		    L.nb, L.d_pptr, L.d_pcol, L.d_P, L.d_apptr, L.d_apcol, L.d_apval, xc, L.d_r, L.d_dinv, L.d_scal,
		    L.d_x[cur], L.d_x[cur ^ 1], L.d_d, S);
		cur ^= 1;
		k0 = 1;
	} else {
		k_prolong<BS><<<g, kBlock, 0, st>>>(L.nb, L.d_pptr, L.d_pcol, L.d_P, xc, L.d_x[cur], S);
	}
	// the finest level's last step also writes the partial sums of x . q,
	// split by segment in a segmented flexible CG
	SegView const sv = l == 0 ? s->fcg_sv : SegView{};
	for (int k = k0; k < deg; k++) {
		bool const last = k + 1 == deg;
		cheb<BS, BT>(st, L, b, false, k, L.d_x[cur], L.d_x[cur ^ 1], last ? q : nullptr, last ? partial : nullptr, S,
		             last ? sv : SegView{});
		cur ^= 1;
	}
	return L.d_x[cur];
}

}  // namespace

// reps launches of the finest level's last post-smoothing step as the
// V-cycle runs it (Chebyshev step deg - 1 with its SpMV, fused with the
// partial sums of x . q), bracketed by events e0 and e1 on the system's
// stream; b is the solver's residual buffer and q its q buffer. Algorithmic
// bytes per launch: FP32 blocks and column indices, slice pointers, and per
// node the FP64 b, the FP32 inverse diagonal block, x_in (gathered once), d
// read and written, x_out written and the FP64 q.
int amg_time_smoother(bfmx_system* s, int reps, hipEvent_t e0, hipEvent_t e1, double* bytes) {
	hipStream_t const st = s->stream;
	AmgLevel& L = s->amg[0];
	if (s->amg.size() == 1 && L.d_dense) {
		set_error("the finest level is solved densely (no smoother to time)");
		return BFMX_E_UNSUPPORTED;
	}
	const Scalars* S = s->d_scal;
	int const step = s->amg_par.deg - 1;
	cheb<2, double>(st, L, s->d_r, true, 0, nullptr, L.d_x[0], nullptr, nullptr, S);
	for (int i = 0; i < 3; i++) {
		cheb<2, double>(st, L, s->d_r, false, step, L.d_x[0], L.d_x[1], s->d_q, s->d_partial, S);
	}
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipEventRecord(e0, st));
	for (int i = 0; i < reps; i++) {
		cheb<2, double>(st, L, s->d_r, false, step, L.d_x[0], L.d_x[1], s->d_q, s->d_partial, S);
	}
	BFMX_HIP(hipEventRecord(e1, st));
	BFMX_HIP(hipGetLastError());
	if (bytes) {
		*bytes = static_cast<double>(L.nnzb) * 20 + 4.0 * (L.n_slices + 1) + 80.0 * L.nb;
	}
	return BFMX_OK;
}

float* amg_vcycle(bfmx_system* s, const double* r, const double* q, double* partial, const Scalars* S,
                  int* n_partials) {
	AmgLevel const& L0 = s->amg[0];
	*n_partials = s->amg.size() == 1 && L0.d_dense ? grid(2 * static_cast<int64_t>(L0.nb)) : grid(L0.nb);
	return vcycle<2, double>(s, 0, r, q, partial, S);
}

}  // namespace bfmx
