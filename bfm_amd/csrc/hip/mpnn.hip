// Learned preconditioner: a message-passing network over the mesh's node
// graph, applied inside flexible CG, and its training on the GPU.
//
// Preconditioner (SPD by construction, so CG stays stable whatever the
// weights):
//   z = D^-1 r + P Delta^-1/2 N^T N Delta^-1/2 P r
// D^-1 the 2x2 block-Jacobi inverse, Delta the mean diagonal of each node's
// block, P the projection onto the unconstrained DOFs, and N a linear
// message-passing network over the node graph (hidden width H = 64):
//   h0 = x W0                                  (x = Delta^-1/2 P r, 2 -> 64)
//   m_i = sum_j a_ij[g(c)] h_j[c],  a_ij = Delta_i^-1/2 A_ij Delta_j^-1/2
//         (the channels are split in four groups of 16, group g = 2a + b
//          weighted by entry (a, b) of the 2x2 stiffness block)
//   h' = [h, m] W_l                            (3 layers, 128 -> 64, MFMA)
// N^T is applied by the adjoint layers: [s, t] = a W_l^T, a_prev = s +
// agg^T(t); the decoder forms a0 = s0 + agg^T(t0), u = a0 W0^T and
// z = D^-1 r + P Delta^-1/2 u. With N = 0 this is exactly block Jacobi.
// Layer weights are BF16 and MFMA accumulates in FP32. Training stores the
// activations in BF16. A solve keeps them in FP32 and splits every MFMA input
// in two BF16 parts (hi + lo, about 16 significant bits): with BF16 rounding
// of the activations the operator would drift from symmetric and linear by
// ~1e-3, enough to slow flexible CG down; the block-Jacobi part is FP64. A new model has a small random
// encoder: block Jacobi plus a small PSD term.
//
// Training: pairs (r, e) recorded from plain PCG runs (e = x* - x_k,
// r = b - A x_k); loss = mean over a batch of ||z(r) - e||_A^2 / ||e||_A^2 on
// the FP32 matrix. With g = P Delta^-1/2 dL/dz, the gradient of u = N^T N x
// is the weight gradient of <N g, N(theta) x> plus that of <N x, N(theta) g>:
// two forward passes (x and g) and two backward passes through them.
//
// Layer GEMM tiles: 64 rows x 64 (forward) or 128 (adjoint) outputs per
// workgroup, four waves of v_mfma_f32_16x16x32_bf16 (16 rows each); operands
// are staged in LDS with padded rows (no bank conflicts for the 16-byte
// operand reads). Each lane's eight A and eight B values occupy the same
// eight positions of K, so the product does not depend on how the hardware
// orders K inside a lane group.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <random>

#include "internal.h"

namespace bfmx {

namespace {

typedef uint16_t u16;
typedef u16 u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kH = 64;        // hidden channels
constexpr int kT = 64;        // rows per GEMM tile
constexpr int kP128 = 136;    // LDS row stride (u16) of 128-wide tiles
constexpr int kP64 = 72;      // LDS row stride (u16) of 64-wide tiles
constexpr int kWG = 256;

// parameter vector layout (FP32 master copy)
constexpr int kOffW0 = 0;                    // [2][64]
constexpr int kLayer = 2 * kH * kH;          // W_l [128][64]
constexpr int kOffL = kOffW0 + 2 * kH;       // layer l at kOffL + l * kLayer
constexpr int kNParams = kOffL + 3 * kLayer;
constexpr int kGradParts = 512;              // workgroups of the gradient kernels (partials)
constexpr float kEncInit = 0.015f;           // initial encoder weights: uniform in +-kEncInit

__device__ __forceinline__ u16 f2bf(float f) {   // round to nearest even
	uint32_t const u = __float_as_uint(f);
	if ((u & 0x7fffffffu) > 0x7f800000u) {
		return 0x7fc0;
	}
	return static_cast<u16>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) {
	return __uint_as_float(static_cast<uint32_t>(h) << 16);
}
__device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
	                                               0, 0);
}

// Entry (a, b) of the system's 2x2 block slot blk (lane r of its row).
__device__ __forceinline__ int64_t blk_at(int64_t blk, int r, int a, int b) {
	return 4 * blk - 2 * r + a * 2 * kSlice + b;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v += __shfl_xor(v, off, 64);
	}
	return v;
}

__device__ __forceinline__ void wg_sum_partial(double v, double* __restrict__ partial) {
	__shared__ double lds[kWG / 64];
	v = wsum(v);
	if ((threadIdx.x & 63) == 0) {
		lds[threadIdx.x >> 6] = v;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double s = 0;
		for (int w = 0; w < kWG / 64; w++) {
			s += lds[w];
		}
		partial[blockIdx.x] = s;
	}
}

// 16 activations of a row as float: BF16 storage (training) or FP32 (a solve).
__device__ __forceinline__ void ld16(const u16* p, float (&v)[16]) {
	u16x8 const a = *reinterpret_cast<const u16x8*>(p);
	u16x8 const b = *reinterpret_cast<const u16x8*>(p + 8);
#pragma unroll
	for (int c = 0; c < 8; c++) {
		v[c] = bf2f(a[c]);
		v[8 + c] = bf2f(b[c]);
	}
}
__device__ __forceinline__ void ld16(const float* p, float (&v)[16]) {
#pragma unroll
	for (int k = 0; k < 4; k++) {
		f32x4 const a = *reinterpret_cast<const f32x4*>(p + 4 * k);
		v[4 * k] = a[0];
		v[4 * k + 1] = a[1];
		v[4 * k + 2] = a[2];
		v[4 * k + 3] = a[3];
	}
}

// The BF16 rounding of 16 values (8-byte-aligned LDS or global row).
__device__ __forceinline__ void st16(u16* p, const float (&v)[16]) {
	u16x8 a, b;
#pragma unroll
	for (int c = 0; c < 8; c++) {
		a[c] = f2bf(v[c]);
		b[c] = f2bf(v[8 + c]);
	}
	*reinterpret_cast<u16x8*>(p) = a;
	*reinterpret_cast<u16x8*>(p + 8) = b;
}

// v = hi + lo with hi, lo BF16 (hi the rounding of v, lo of the remainder):
// two BF16 MFMAs on (hi, lo) give the product to about 16 significant bits.
__device__ __forceinline__ void split16(u16* hi, u16* lo, const float (&v)[16]) {
	u16x8 h0, h1, l0, l1;
#pragma unroll
	for (int c = 0; c < 8; c++) {
		h0[c] = f2bf(v[c]);
		h1[c] = f2bf(v[8 + c]);
		l0[c] = f2bf(v[c] - bf2f(h0[c]));
		l1[c] = f2bf(v[8 + c] - bf2f(h1[c]));
	}
	*reinterpret_cast<u16x8*>(hi) = h0;
	*reinterpret_cast<u16x8*>(hi + 8) = h1;
	*reinterpret_cast<u16x8*>(lo) = l0;
	*reinterpret_cast<u16x8*>(lo + 8) = l1;
}

// acc[c] += sum over the neighbours k of node i (row slot order) of
// w_ik T[(k B + b) H + 16 g + c], w_ik = entry (a, b') of block (i, k) times
// isd_i isd_k with (a, b') = (g / 2, g % 2), or (g % 2, g / 2) when TR (the
// adjoint: for the symmetric A, entry (a, b') of block (k, i) is entry
// (b', a) of block (i, k)).
template <bool TR, typename AT>
__device__ __forceinline__ void gather16(int64_t i, int64_t b, int B, int g, const int32_t* __restrict__ slice_ptr,
                                         const int32_t* __restrict__ row_len, const int32_t* __restrict__ col,
                                         const float* __restrict__ val32, const float* __restrict__ feat,
                                         const AT* __restrict__ T, float (&acc)[16]) {
	int const r = static_cast<int>(i & 63);
	int const a = TR ? (g & 1) : (g >> 1), bb = TR ? (g >> 1) : (g & 1);
	float const si = feat[8 * i + 5];
	int32_t const beg = slice_ptr[i >> 6] + r;
	int const len = row_len[i];
	// three neighbours' loads in flight together (a learned solve on the
	// 2M-DOF beam: 141 -> 126 ms; two: 136 ms, four: 134 ms)
#pragma unroll 3
	for (int k = 0; k < len; k++) {
		int64_t const blk = beg + k * kSlice;
		int64_t const j = col[blk];
		float const w = val32[blk_at(blk, r, a, bb)] * si * feat[8 * j + 5];
		float v[16];
		ld16(T + (j * B + b) * kH + 16 * g, v);
#pragma unroll
		for (int c = 0; c < 16; c++) {
			acc[c] += w * v[c];
		}
	}
}

// --------------------------------------------------------------- features

// Per node: [dinv00, dinv01, dinv10, dinv11, delta, 1/sqrt(delta), c, fixed
// bits (1 = x, 2 = y)] from the FP64 diagonal block; c is filled by k_feat2.
__global__ __launch_bounds__(kWG) void k_feat1(int nb, const int32_t* __restrict__ diag_blk,
                                               const double* __restrict__ val, const uint8_t* __restrict__ fixed,
                                               float* __restrict__ feat) {
	int const i = blockIdx.x * kWG + threadIdx.x;
	if (i >= nb) {
		return;
	}
	int const r = i & 63;
	int64_t const d = diag_blk[i];
	double const a00 = val[blk_at(d, r, 0, 0)], a01 = val[blk_at(d, r, 0, 1)];
	double const a10 = val[blk_at(d, r, 1, 0)], a11 = val[blk_at(d, r, 1, 1)];
	double const inv = 1.0 / (a00 * a11 - a01 * a10);
	double const delta = 0.5 * (a00 + a11);
	float* f = feat + 8 * static_cast<int64_t>(i);
	f[0] = static_cast<float>(a11 * inv);
	f[1] = static_cast<float>(-a01 * inv);
	f[2] = static_cast<float>(-a10 * inv);
	f[3] = static_cast<float>(a00 * inv);
	f[4] = static_cast<float>(delta);
	f[5] = static_cast<float>(1.0 / sqrt(delta));
	f[6] = 0.f;
	f[7] = static_cast<float>((fixed[2 * i] ? 1 : 0) + (fixed[2 * i + 1] ? 2 : 0));
}

// c_i = log delta_i - mean over the row's neighbours (itself included) of
// log delta_j (a material-contrast indicator kept with the features).
__global__ __launch_bounds__(kWG) void k_feat2(int nb, const int32_t* __restrict__ slice_ptr,
                                               const int32_t* __restrict__ row_len,
                                               const int32_t* __restrict__ col, float* __restrict__ feat) {
	int const i = blockIdx.x * kWG + threadIdx.x;
	if (i >= nb) {
		return;
	}
	int const r = i & 63;
	int32_t const beg = slice_ptr[i >> 6] + r;
	int const len = row_len[i];
	double acc = 0;
	for (int k = 0; k < len; k++) {
		acc += log(static_cast<double>(feat[8 * static_cast<int64_t>(col[beg + k * kSlice]) + 4]));
	}
	float* f = feat + 8 * static_cast<int64_t>(i);
	f[6] = static_cast<float>(log(static_cast<double>(f[4])) - acc / len);
}

// ------------------------------------------------------------------ forward

// h0 = x W0 (AT: BF16 when training, FP32 in a solve), row rg = node * B + b.
// RESID: x = Delta^-1/2 P r with r of pair b at r + b * pstride (FP64 in a
// solve, FP32 when training); otherwise x is given (2 per row, FP32). X (may
// be null) receives x.
template <bool RESID, typename RT, typename AT>
__global__ __launch_bounds__(kWG) void k_encode(int64_t R, int B, const RT* __restrict__ r, int64_t pstride,
                                                const float* __restrict__ feat, const float* __restrict__ params,
                                                AT* __restrict__ H0, float* __restrict__ X,
                                                const Scalars* __restrict__ S) {
	if (S && S->done) {
		return;
	}
	__shared__ float W0[2 * kH];
	for (int k = threadIdx.x; k < 2 * kH; k += kWG) {
		W0[k] = params[kOffW0 + k];
	}
	__syncthreads();
	int64_t const rg = static_cast<int64_t>(blockIdx.x) * kWG + threadIdx.x;
	if (rg >= R) {
		return;
	}
	float x0, x1;
	if (RESID) {
		int64_t const i = rg / B;
		int64_t const b = rg - i * B;
		const float* f = feat + 8 * i;
		int const fx = static_cast<int>(f[7]);
		x0 = (fx & 1) ? 0.f : f[5] * static_cast<float>(r[b * pstride + 2 * i]);
		x1 = (fx & 2) ? 0.f : f[5] * static_cast<float>(r[b * pstride + 2 * i + 1]);
	} else {
		x0 = static_cast<float>(r[2 * rg]);
		x1 = static_cast<float>(r[2 * rg + 1]);
	}
	if (X) {
		X[2 * rg] = x0;
		X[2 * rg + 1] = x1;
	}
#pragma unroll
	for (int c16 = 0; c16 < kH / 16; c16++) {
		float o[16];
#pragma unroll
		for (int t = 0; t < 16; t++) {
			int const c = 16 * c16 + t;
			o[t] = x0 * W0[c] + x1 * W0[kH + c];
		}
		if constexpr (sizeof(AT) == 2) {
			st16(H0 + rg * kH + 16 * c16, o);
		} else {
#pragma unroll
			for (int k = 0; k < 4; k++) {
				*reinterpret_cast<f32x4*>(H0 + rg * kH + 16 * c16 + 4 * k) =
				    f32x4{o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]};
			}
		}
	}
}

// One forward layer over a 64-row tile: message m (FP32 sums of neighbour
// features), then h_out = [h_in, m] W on MFMA (Wt: the layer weight
// transposed, 64 x 128, BF16). AT = u16: BF16 activations (training; Mout,
// when given, receives m); AT = float: FP32 activations (a solve) and each
// MFMA input split in two BF16 parts (hi + lo).
template <typename AT>
__global__ __launch_bounds__(kWG) void k_layer(int64_t R, int B, const int32_t* __restrict__ slice_ptr,
                                               const int32_t* __restrict__ row_len,
                                               const int32_t* __restrict__ col, const float* __restrict__ val32,
                                               const float* __restrict__ feat, const AT* __restrict__ Hin,
                                               AT* __restrict__ Hout, u16* __restrict__ Mout,
                                               const u16* __restrict__ Wt, const Scalars* __restrict__ S) {
	constexpr bool kSplit = sizeof(AT) == 4;
	if (S && S->done) {
		return;
	}
	__shared__ u16 X[kT][kP128];
	__shared__ u16 Xl[kSplit ? kT : 1][kP128];
	__shared__ u16 Ws[kH][kP128];
	for (int e = threadIdx.x; e < kH * 16; e += kWG) {
		int const n = e >> 4, k8 = e & 15;
		*reinterpret_cast<u16x8*>(&Ws[n][8 * k8]) = *reinterpret_cast<const u16x8*>(Wt + n * 2 * kH + 8 * k8);
	}
	int const t = threadIdx.x, row = t >> 2, g = t & 3;
	int64_t const rg = static_cast<int64_t>(blockIdx.x) * kT + row;
	float self[16], acc[16];
#pragma unroll
	for (int c = 0; c < 16; c++) {
		self[c] = 0.f;
		acc[c] = 0.f;
	}
	if (rg < R) {
		int64_t const i = rg / B;
		int64_t const b = rg - i * B;
		ld16(Hin + rg * kH + 16 * g, self);
		gather16<false>(i, b, B, g, slice_ptr, row_len, col, val32, feat, Hin, acc);
	}
	if constexpr (kSplit) {
		split16(&X[row][16 * g], &Xl[row][16 * g], self);
		split16(&X[row][kH + 16 * g], &Xl[row][kH + 16 * g], acc);
	} else {
		st16(&X[row][16 * g], self);   // exact: self holds BF16 values
		st16(&X[row][kH + 16 * g], acc);
		if (Mout && rg < R) {
			st16(Mout + rg * kH + 16 * g, acc);
		}
	}
	__syncthreads();
	int const w = t >> 6, l = t & 63, lr = l & 15, lq = l >> 4;
	f32x4 d[4];
#pragma unroll
	for (int nb = 0; nb < 4; nb++) {
		d[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
	}
#pragma unroll
	for (int ks = 0; ks < 4; ks++) {
		u16x8 const av = *reinterpret_cast<const u16x8*>(&X[16 * w + lr][32 * ks + 8 * lq]);
#pragma unroll
		for (int nb = 0; nb < 4; nb++) {
			u16x8 const bv = *reinterpret_cast<const u16x8*>(&Ws[16 * nb + lr][32 * ks + 8 * lq]);
			d[nb] = mfma(av, bv, d[nb]);
			if constexpr (kSplit) {
				u16x8 const al = *reinterpret_cast<const u16x8*>(&Xl[16 * w + lr][32 * ks + 8 * lq]);
				d[nb] = mfma(al, bv, d[nb]);
			}
		}
	}
	if constexpr (kSplit) {
#pragma unroll
		for (int q = 0; q < 4; q++) {
			int64_t const ro = static_cast<int64_t>(blockIdx.x) * kT + 16 * w + 4 * lq + q;
			if (ro < R) {
#pragma unroll
				for (int nb = 0; nb < 4; nb++) {
					Hout[ro * kH + 16 * nb + lr] = d[nb][q];
				}
			}
		}
	} else {
		__syncthreads();   // X is reused for the output tile
#pragma unroll
		for (int nb = 0; nb < 4; nb++) {
#pragma unroll
			for (int q = 0; q < 4; q++) {
				X[16 * w + 4 * lq + q][16 * nb + lr] = f2bf(d[nb][q]);
			}
		}
		__syncthreads();
		if (rg < R) {
			*reinterpret_cast<u16x8*>(Hout + rg * kH + 16 * g) = *reinterpret_cast<const u16x8*>(&X[row][16 * g]);
			*reinterpret_cast<u16x8*>(Hout + rg * kH + 16 * g + 8) =
			    *reinterpret_cast<const u16x8*>(&X[row][16 * g + 8]);
		}
	}
}

// One adjoint layer over a 64-row tile: a = Sin (FIRST: the output v of N) or
// a = Sin + agg^T(Tin), then [Sout, Tout] = a W^T on MFMA (W: the layer
// weight, 128 x 64, BF16, natural layout). AT = u16: BF16 activations and a
// rounded to BF16 (training); AT = float: FP32 activations, a split in two
// BF16 parts (a solve).
template <bool FIRST, typename AT>
__global__ __launch_bounds__(kWG) void k_adjoint(int64_t R, int B, const int32_t* __restrict__ slice_ptr,
                                                 const int32_t* __restrict__ row_len,
                                                 const int32_t* __restrict__ col, const float* __restrict__ val32,
                                                 const float* __restrict__ feat, const AT* __restrict__ Sin,
                                                 const AT* __restrict__ Tin, AT* __restrict__ Sout,
                                                 AT* __restrict__ Tout, const u16* __restrict__ W,
                                                 const Scalars* __restrict__ S) {
	constexpr bool kSplit = sizeof(AT) == 4;
	if (S && S->done) {
		return;
	}
	__shared__ u16 A[kT][kP64];
	__shared__ u16 Al[kSplit ? kT : 1][kP64];
	__shared__ u16 Ws[2 * kH][kP64];
	__shared__ u16 O[kSplit ? 1 : kT][kP128];
	for (int e = threadIdx.x; e < 2 * kH * 8; e += kWG) {
		int const n = e >> 3, k8 = e & 7;
		*reinterpret_cast<u16x8*>(&Ws[n][8 * k8]) = *reinterpret_cast<const u16x8*>(W + n * kH + 8 * k8);
	}
	int const t = threadIdx.x, row = t >> 2, g = t & 3;
	int64_t const rg = static_cast<int64_t>(blockIdx.x) * kT + row;
	float a[16];
#pragma unroll
	for (int c = 0; c < 16; c++) {
		a[c] = 0.f;
	}
	if (rg < R) {
		ld16(Sin + rg * kH + 16 * g, a);
		if (!FIRST) {
			int64_t const i = rg / B;
			int64_t const b = rg - i * B;
			gather16<true>(i, b, B, g, slice_ptr, row_len, col, val32, feat, Tin, a);
		}
	}
	if constexpr (kSplit) {
		split16(&A[row][16 * g], &Al[row][16 * g], a);
	} else {
		st16(&A[row][16 * g], a);
	}
	__syncthreads();
	int const w = t >> 6, l = t & 63, lr = l & 15, lq = l >> 4;
	f32x4 d[8];
#pragma unroll
	for (int nb = 0; nb < 8; nb++) {
		d[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
	}
#pragma unroll
	for (int ks = 0; ks < 2; ks++) {
		u16x8 const av = *reinterpret_cast<const u16x8*>(&A[16 * w + lr][32 * ks + 8 * lq]);
#pragma unroll
		for (int nb = 0; nb < 8; nb++) {
			u16x8 const bv = *reinterpret_cast<const u16x8*>(&Ws[16 * nb + lr][32 * ks + 8 * lq]);
			d[nb] = mfma(av, bv, d[nb]);
			if constexpr (kSplit) {
				u16x8 const al = *reinterpret_cast<const u16x8*>(&Al[16 * w + lr][32 * ks + 8 * lq]);
				d[nb] = mfma(al, bv, d[nb]);
			}
		}
	}
	if constexpr (kSplit) {
#pragma unroll
		for (int q = 0; q < 4; q++) {
			int64_t const ro = static_cast<int64_t>(blockIdx.x) * kT + 16 * w + 4 * lq + q;
			if (ro < R) {
#pragma unroll
				for (int nb = 0; nb < 8; nb++) {
					int const c = 16 * nb + lr;
					if (c < kH) {
						Sout[ro * kH + c] = d[nb][q];
					} else {
						Tout[ro * kH + c - kH] = d[nb][q];
					}
				}
			}
		}
	} else {
#pragma unroll
		for (int nb = 0; nb < 8; nb++) {
#pragma unroll
			for (int q = 0; q < 4; q++) {
				O[16 * w + 4 * lq + q][16 * nb + lr] = f2bf(d[nb][q]);
			}
		}
		__syncthreads();
		if (rg < R) {
			u16* so = reinterpret_cast<u16*>(Sout) + rg * kH + 16 * g;
			u16* to = reinterpret_cast<u16*>(Tout) + rg * kH + 16 * g;
			*reinterpret_cast<u16x8*>(so) = *reinterpret_cast<const u16x8*>(&O[row][16 * g]);
			*reinterpret_cast<u16x8*>(so + 8) = *reinterpret_cast<const u16x8*>(&O[row][16 * g + 8]);
			*reinterpret_cast<u16x8*>(to) = *reinterpret_cast<const u16x8*>(&O[row][kH + 16 * g]);
			*reinterpret_cast<u16x8*>(to + 8) = *reinterpret_cast<const u16x8*>(&O[row][kH + 16 * g + 8]);
		}
	}
}

// Decoder (64 rows per workgroup, four threads per row): a0 = S0 + agg^T(T0)
// (FP32), u = a0 W0^T, z = y + P Delta^-1/2 u. In a solve (!TRAIN) y = D^-1 r
// with the solver's FP64 2x2 inverses (dinv64), or y = base when given (the
// multigrid V-cycle's output), and z is FP64; DOT adds
// per-workgroup partial sums of z . q. Training: y from the FP32 features,
// z FP32.
template <bool TRAIN, bool DOT, typename RT, typename ZT, typename AT>
__global__ __launch_bounds__(kWG) void k_decode(int64_t R, int B, const RT* __restrict__ r, int64_t pstride,
                                                const int32_t* __restrict__ slice_ptr,
                                                const int32_t* __restrict__ row_len,
                                                const int32_t* __restrict__ col, const float* __restrict__ val32,
                                                const float* __restrict__ feat, const float* __restrict__ params,
                                                const AT* __restrict__ S0, const AT* __restrict__ T0,
                                                const double* __restrict__ dinv64, const float* __restrict__ base,
                                                ZT* __restrict__ z, const double* __restrict__ q,
                                                double* __restrict__ partial, const Scalars* __restrict__ S) {
	if (S && S->done) {
		return;
	}
	__shared__ float W0[2 * kH];
	__shared__ float part[kWG][2];
	for (int k = threadIdx.x; k < 2 * kH; k += kWG) {
		W0[k] = params[kOffW0 + k];
	}
	__syncthreads();
	int const t = threadIdx.x, rr = t >> 2, g = t & 3;
	int64_t const rg = static_cast<int64_t>(blockIdx.x) * kT + rr;
	float u0 = 0.f, u1 = 0.f;
	if (rg < R) {
		int64_t const i = rg / B;
		int64_t const b = rg - i * B;
		float a[16];
		ld16(S0 + rg * kH + 16 * g, a);
		gather16<true>(i, b, B, g, slice_ptr, row_len, col, val32, feat, T0, a);
#pragma unroll
		for (int c = 0; c < 16; c++) {
			u0 += a[c] * W0[16 * g + c];
			u1 += a[c] * W0[kH + 16 * g + c];
		}
	}
	part[t][0] = u0;
	part[t][1] = u1;
	__syncthreads();
	double dot = 0;
	if (g == 0 && rg < R) {
		u0 = part[t][0] + part[t + 1][0] + part[t + 2][0] + part[t + 3][0];
		u1 = part[t][1] + part[t + 1][1] + part[t + 2][1] + part[t + 3][1];
		int64_t const i = rg / B;
		int64_t const b = rg - i * B;
		const float* f = feat + 8 * i;
		int const fx = static_cast<int>(f[7]);
		float const c0 = (fx & 1) ? 0.f : f[5] * u0;
		float const c1 = (fx & 2) ? 0.f : f[5] * u1;
		ZT z0, z1;
		if (TRAIN) {
			float const r0 = static_cast<float>(r[b * pstride + 2 * i]);
			float const r1 = static_cast<float>(r[b * pstride + 2 * i + 1]);
			z0 = static_cast<ZT>(f[0] * r0 + f[1] * r1 + c0);
			z1 = static_cast<ZT>(f[2] * r0 + f[3] * r1 + c1);
		} else if (base) {
			z0 = static_cast<ZT>(static_cast<double>(base[2 * i]) + static_cast<double>(c0));
			z1 = static_cast<ZT>(static_cast<double>(base[2 * i + 1]) + static_cast<double>(c1));
		} else {
			double const r0 = static_cast<double>(r[2 * i]), r1 = static_cast<double>(r[2 * i + 1]);
			const double* d = dinv64 + 4 * i;
			z0 = static_cast<ZT>(d[0] * r0 + d[1] * r1 + static_cast<double>(c0));
			z1 = static_cast<ZT>(d[2] * r0 + d[3] * r1 + static_cast<double>(c1));
		}
		z[2 * rg] = z0;
		z[2 * rg + 1] = z1;
		if (DOT) {
			dot = static_cast<double>(z0) * q[2 * rg] + static_cast<double>(z1) * q[2 * rg + 1];
		}
	}
	if (DOT) {
		wg_sum_partial(dot, partial);
	}
}

// ------------------------------------------------------------------ training

// d = z - e, q = A d (FP32 matrix); loss partial sum d . q / (den_b B) and
// g = P Delta^-1/2 dL/dz with dL/dz = 2 q / (den_b B).
__global__ __launch_bounds__(kWG) void k_loss(int64_t R, int B, const int32_t* __restrict__ slice_ptr,
                                              const int32_t* __restrict__ row_len, const int32_t* __restrict__ col,
                                              const float* __restrict__ val32, const float* __restrict__ feat,
                                              const float* __restrict__ z, const float* __restrict__ e,
                                              int64_t pstride, const float* __restrict__ den, float* __restrict__ gout,
                                              double* __restrict__ partial) {
	int64_t const rg = static_cast<int64_t>(blockIdx.x) * kWG + threadIdx.x;
	double lv = 0;
	if (rg < R) {
		int64_t const i = rg / B;
		int const b = static_cast<int>(rg - i * B);
		int const r = static_cast<int>(i & 63);
		int32_t const beg = slice_ptr[i >> 6] + r;
		int const len = row_len[i];
		const float* eb = e + b * pstride;
		float q0 = 0.f, q1 = 0.f;
		for (int k = 0; k < len; k++) {
			int64_t const blk = beg + k * kSlice;
			int64_t const j = col[blk];
			float const d0 = z[2 * (j * B + b)] - eb[2 * j];
			float const d1 = z[2 * (j * B + b) + 1] - eb[2 * j + 1];
			q0 += val32[blk_at(blk, r, 0, 0)] * d0 + val32[blk_at(blk, r, 0, 1)] * d1;
			q1 += val32[blk_at(blk, r, 1, 0)] * d0 + val32[blk_at(blk, r, 1, 1)] * d1;
		}
		float const sc = 1.f / (den[b] * static_cast<float>(B));
		float const d0 = z[2 * rg] - eb[2 * i], d1 = z[2 * rg + 1] - eb[2 * i + 1];
		const float* f = feat + 8 * i;
		int const fx = static_cast<int>(f[7]);
		gout[2 * rg] = (fx & 1) ? 0.f : f[5] * (2.f * q0 * sc);
		gout[2 * rg + 1] = (fx & 2) ? 0.f : f[5] * (2.f * q1 * sc);
		lv = (static_cast<double>(d0) * q0 + static_cast<double>(d1) * q1) * sc;
	}
	wg_sum_partial(lv, partial);
}

// Backward through one forward layer over 64-row tiles (persistent over the
// grid), with the upstream gradient dHout (FP32):
//   dpre = dHout                                  (rounded to BF16)
//   [dh_in, dm] = dpre W^T                        (MFMA; written FP32)
//   dW += [h_in, m]^T dpre                        (MFMA; per-workgroup partials)
// W is the layer weight (128 x 64, BF16, natural layout).
__global__ __launch_bounds__(kWG) void k_layer_bwd(int64_t R, const u16* __restrict__ Hin, const u16* __restrict__ M,
                                                   const float* __restrict__ dHout, const u16* __restrict__ W,
                                                   float* __restrict__ dHin, float* __restrict__ dM,
                                                   float* __restrict__ partial) {
	__shared__ u16 Ws[2 * kH][kP64];
	__shared__ u16 XT[2 * kH][kP64];
	__shared__ u16 dp[kT][kP64];
	__shared__ u16 dpT[kH][kP64];
	for (int e = threadIdx.x; e < 2 * kH * 8; e += kWG) {
		int const n = e >> 3, k8 = e & 7;
		*reinterpret_cast<u16x8*>(&Ws[n][8 * k8]) = *reinterpret_cast<const u16x8*>(W + n * kH + 8 * k8);
	}
	int const t = threadIdx.x, row = t >> 2, g = t & 3;
	int const w = t >> 6, l = t & 63, lr = l & 15, lq = l >> 4;
	f32x4 acc[8];
#pragma unroll
	for (int k = 0; k < 8; k++) {
		acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
	}
	int64_t const n_tiles = (R + kT - 1) / kT;
	for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
		int64_t const rg = tile * kT + row;
		__syncthreads();   // the previous tile's LDS reads are done
		if (rg < R) {
			u16x8 const hi0 = *reinterpret_cast<const u16x8*>(Hin + rg * kH + 16 * g);
			u16x8 const hi1 = *reinterpret_cast<const u16x8*>(Hin + rg * kH + 16 * g + 8);
			u16x8 const m0 = *reinterpret_cast<const u16x8*>(M + rg * kH + 16 * g);
			u16x8 const m1 = *reinterpret_cast<const u16x8*>(M + rg * kH + 16 * g + 8);
#pragma unroll
			for (int c = 0; c < 16; c++) {
				u16 const dh = f2bf(dHout[rg * kH + 16 * g + c]);
				dp[row][16 * g + c] = dh;
				dpT[16 * g + c][row] = dh;
				XT[16 * g + c][row] = c < 8 ? hi0[c] : hi1[c - 8];
				XT[kH + 16 * g + c][row] = c < 8 ? m0[c] : m1[c - 8];
			}
		} else {
#pragma unroll
			for (int c = 0; c < 16; c++) {
				dp[row][16 * g + c] = 0;
				dpT[16 * g + c][row] = 0;
				XT[16 * g + c][row] = 0;
				XT[kH + 16 * g + c][row] = 0;
			}
		}
		__syncthreads();
		// [dh_in, dm] = dpre W^T: rows 16 w .. 16 w + 15, 128 columns, K = 64
		{
			f32x4 dx[8];
#pragma unroll
			for (int nb = 0; nb < 8; nb++) {
				dx[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
			}
#pragma unroll
			for (int ks = 0; ks < 2; ks++) {
				u16x8 const av = *reinterpret_cast<const u16x8*>(&dp[16 * w + lr][32 * ks + 8 * lq]);
#pragma unroll
				for (int nb = 0; nb < 8; nb++) {
					u16x8 const bv = *reinterpret_cast<const u16x8*>(&Ws[16 * nb + lr][32 * ks + 8 * lq]);
					dx[nb] = mfma(av, bv, dx[nb]);
				}
			}
#pragma unroll
			for (int q = 0; q < 4; q++) {
				int64_t const ro = tile * kT + 16 * w + 4 * lq + q;
				if (ro < R) {
#pragma unroll
					for (int nb = 0; nb < 8; nb++) {
						int const c = 16 * nb + lr;
						if (c < kH) {
							dHin[ro * kH + c] = dx[nb][q];
						} else {
							dM[ro * kH + c - kH] = dx[nb][q];
						}
					}
				}
			}
		}
		// dW += X^T dpre: rows of W 32 w .. 32 w + 31, 64 columns, K = 64 tile rows
#pragma unroll
		for (int ks = 0; ks < 2; ks++) {
#pragma unroll
			for (int mb = 0; mb < 2; mb++) {
				u16x8 const av = *reinterpret_cast<const u16x8*>(&XT[32 * w + 16 * mb + lr][32 * ks + 8 * lq]);
#pragma unroll
				for (int nb = 0; nb < 4; nb++) {
					u16x8 const bv = *reinterpret_cast<const u16x8*>(&dpT[16 * nb + lr][32 * ks + 8 * lq]);
					acc[4 * mb + nb] = mfma(av, bv, acc[4 * mb + nb]);
				}
			}
		}
	}
	float* p = partial + static_cast<int64_t>(blockIdx.x) * kLayer;
#pragma unroll
	for (int mb = 0; mb < 2; mb++) {
#pragma unroll
		for (int nb = 0; nb < 4; nb++) {
#pragma unroll
			for (int q = 0; q < 4; q++) {
				p[(32 * w + 16 * mb + 4 * lq + q) * kH + 16 * nb + lr] = acc[4 * mb + nb][q];
			}
		}
	}
}

// dh_in[j] += agg^T(dm)[j] = sum over the neighbours i of j of
// a_ij[g(c)] dm_i[c] (read from row j: a_ij[(a, b)] = a_ji[(b, a)]). FP32.
__global__ __launch_bounds__(kWG) void k_agg_bwd(int64_t R, int B, const int32_t* __restrict__ slice_ptr,
                                                 const int32_t* __restrict__ row_len,
                                                 const int32_t* __restrict__ col, const float* __restrict__ val32,
                                                 const float* __restrict__ feat, const float* __restrict__ dM,
                                                 float* __restrict__ dHin) {
	int const t = threadIdx.x, row = t >> 2, g = t & 3;
	int64_t const rg = static_cast<int64_t>(blockIdx.x) * kT + row;
	if (rg >= R) {
		return;
	}
	int64_t const j = rg / B;
	int64_t const b = rg - j * B;
	int const r = static_cast<int>(j & 63), a = g & 1, bb = g >> 1;
	float const sj = feat[8 * j + 5];
	int32_t const beg = slice_ptr[j >> 6] + r;
	int const len = row_len[j];
	float acc[16];
#pragma unroll
	for (int c = 0; c < 16; c++) {
		acc[c] = dHin[rg * kH + 16 * g + c];
	}
	for (int k = 0; k < len; k++) {
		int64_t const blk = beg + k * kSlice;
		int64_t const i = col[blk];
		float const wgt = val32[blk_at(blk, r, a, bb)] * sj * feat[8 * i + 5];
		const float* dmi = dM + (i * B + b) * kH + 16 * g;
#pragma unroll
		for (int c = 0; c < 16; c++) {
			acc[c] += wgt * dmi[c];
		}
	}
#pragma unroll
	for (int c = 0; c < 16; c++) {
		dHin[rg * kH + 16 * g + c] = acc[c];
	}
}

// Encoder weight gradient: per-workgroup partial sums of dW0 = x^T dh0
// (2 x 64). Thread: column c = t % 64, row lane t / 64; rows strided.
__global__ __launch_bounds__(kWG) void k_encode_bwd(int64_t R, const float* __restrict__ X,
                                                    const float* __restrict__ dH0, float* __restrict__ partial) {
	__shared__ float red[4][2 * kH];
	int const t = threadIdx.x, c = t & 63, rs = t >> 6;
	float a0 = 0.f, a1 = 0.f;
	for (int64_t rg = static_cast<int64_t>(blockIdx.x) * 4 + rs; rg < R; rg += static_cast<int64_t>(gridDim.x) * 4) {
		float const d = dH0[rg * kH + c];
		a0 += X[2 * rg] * d;
		a1 += X[2 * rg + 1] * d;
	}
	red[rs][c] = a0;
	red[rs][kH + c] = a1;
	__syncthreads();
	if (t < 2 * kH) {
		partial[static_cast<int64_t>(blockIdx.x) * 2 * kH + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
	}
}

// out[k] (+)= sum over the G partial vectors (length len) of partial[wg * len + k].
__global__ __launch_bounds__(kWG) void k_reduce_partials(int G, int len, const float* __restrict__ partial,
                                                         float* __restrict__ out, int accumulate) {
	int const k = blockIdx.x * kWG + threadIdx.x;
	if (k >= len) {
		return;
	}
	float s = accumulate ? out[k] : 0.f;
	for (int wg = 0; wg < G; wg++) {
		s += partial[static_cast<int64_t>(wg) * len + k];
	}
	out[k] = s;
}

__global__ __launch_bounds__(kWG) void k_sum_loss(int n, const double* __restrict__ partial, double* __restrict__ out) {
	__shared__ double lds[kWG / 64];
	double s = 0;
	for (int i = threadIdx.x; i < n; i += kWG) {
		s += partial[i];
	}
	s = wsum(s);
	if ((threadIdx.x & 63) == 0) {
		lds[threadIdx.x >> 6] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double tot = 0;
		for (int w = 0; w < kWG / 64; w++) {
			tot += lds[w];
		}
		*out = tot;
	}
}

__global__ __launch_bounds__(kWG) void k_bf16_to_f32(int64_t n, const u16* __restrict__ a, float* __restrict__ b) {
	int64_t const k = static_cast<int64_t>(blockIdx.x) * kWG + threadIdx.x;
	if (k < n) {
		b[k] = bf2f(a[k]);
	}
}

__global__ __launch_bounds__(kWG) void k_adam(int n, float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
	int const k = blockIdx.x * kWG + threadIdx.x;
	if (k >= n) {
		return;
	}
	float const gk = g[k];
	float const mk = b1 * m[k] + (1.f - b1) * gk;
	float const vk = b2 * v[k] + (1.f - b2) * gk * gk;
	m[k] = mk;
	v[k] = vk;
	w[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

// BF16 copies of the three layer weights: natural (128 x 64) and transposed (64 x 128).
__global__ __launch_bounds__(kWG) void k_pack(const float* __restrict__ params, u16* __restrict__ W,
                                              u16* __restrict__ Wt) {
	int const k = blockIdx.x * kWG + threadIdx.x;
	if (k >= 3 * kLayer) {
		return;
	}
	int const l = k / kLayer, e = k - l * kLayer, kin = e / kH, o = e - kin * kH;
	u16 const h = f2bf(params[kOffL + k]);
	W[k] = h;
	Wt[l * kLayer + o * 2 * kH + kin] = h;
}

// Training pairs: r (FP32) and e = x* - x (FP32) of one recorded iterate.
__global__ __launch_bounds__(kWG) void k_record(int64_t n2, const double* __restrict__ r, const double* __restrict__ x,
                                                const double* __restrict__ xs, float* __restrict__ rp,
                                                float* __restrict__ ep) {
	int64_t const k = static_cast<int64_t>(blockIdx.x) * kWG + threadIdx.x;
	if (k < n2) {
		rp[k] = static_cast<float>(r[k]);
		ep[k] = static_cast<float>(xs[k] - x[k]);
	}
}

// Per pair: partial sums of r . r and e^T A e (FP64 over the FP32 pair data
// and the FP64 matrix).
__global__ __launch_bounds__(kWG) void k_pair_norms(int nb, const int32_t* __restrict__ slice_ptr,
                                                    const int32_t* __restrict__ row_len,
                                                    const int32_t* __restrict__ col, const double* __restrict__ val,
                                                    const float* __restrict__ rp, const float* __restrict__ ep,
                                                    double* __restrict__ partial) {
	int const i = blockIdx.x * kWG + threadIdx.x;
	double rr = 0, eae = 0;
	if (i < nb) {
		int const r = i & 63;
		int32_t const beg = slice_ptr[i >> 6] + r;
		int const len = row_len[i];
		double q0 = 0, q1 = 0;
		for (int k = 0; k < len; k++) {
			int64_t const blk = beg + k * kSlice;
			int64_t const j = col[blk];
			double const e0 = ep[2 * j], e1 = ep[2 * j + 1];
			q0 += val[blk_at(blk, r, 0, 0)] * e0 + val[blk_at(blk, r, 0, 1)] * e1;
			q1 += val[blk_at(blk, r, 1, 0)] * e0 + val[blk_at(blk, r, 1, 1)] * e1;
		}
		eae = q0 * ep[2 * i] + q1 * ep[2 * i + 1];
		rr = static_cast<double>(rp[2 * i]) * rp[2 * i] + static_cast<double>(rp[2 * i + 1]) * rp[2 * i + 1];
	}
	__shared__ double lds[2][kWG / 64];
	rr = wsum(rr);
	eae = wsum(eae);
	if ((threadIdx.x & 63) == 0) {
		lds[0][threadIdx.x >> 6] = rr;
		lds[1][threadIdx.x >> 6] = eae;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double a = 0, b = 0;
		for (int w = 0; w < kWG / 64; w++) {
			a += lds[0][w];
			b += lds[1][w];
		}
		partial[2 * blockIdx.x] = a;
		partial[2 * blockIdx.x + 1] = b;
	}
}

int grid_of(int64_t n, int per) {
	return static_cast<int>((n + per - 1) / per);
}

// Graph, matrix and features that the network kernels read.
struct Net {
	const int32_t* slice_ptr;
	const int32_t* row_len;
	const int32_t* col;
	const float* val32;
	const float* feat;
};

// Forward N over R = n B rows: h[1..3] from h[0] (and the messages when ms).
template <typename AT>
void net_forward(hipStream_t st, const Net& g, int64_t R, int B, const bfmx_model* m, AT* const* h, u16* const* ms,
                 const Scalars* S) {
	int const gt = grid_of(R, kT);
	for (int l = 0; l < 3; l++) {
		k_layer<AT><<<gt, kWG, 0, st>>>(R, B, g.slice_ptr, g.row_len, g.col, g.val32, g.feat, h[l], h[l + 1],
		                                ms ? ms[l] : nullptr, m->d_Wt + l * kLayer, S);
	}
}

// Adjoint N^T from v through the three layers; leaves (s0, t0) in (sa, ta).
template <typename AT>
void net_adjoint(hipStream_t st, const Net& g, int64_t R, int B, const bfmx_model* m, const AT* v, AT* sa, AT* ta,
                 AT* sb, AT* tb, const Scalars* S) {
	int const gt = grid_of(R, kT);
	k_adjoint<true, AT><<<gt, kWG, 0, st>>>(R, B, g.slice_ptr, g.row_len, g.col, g.val32, g.feat, v, nullptr, sa, ta,
	                                        m->d_W + 2 * kLayer, S);
	k_adjoint<false, AT><<<gt, kWG, 0, st>>>(R, B, g.slice_ptr, g.row_len, g.col, g.val32, g.feat, sa, ta, sb, tb,
	                                         m->d_W + 1 * kLayer, S);
	k_adjoint<false, AT><<<gt, kWG, 0, st>>>(R, B, g.slice_ptr, g.row_len, g.col, g.val32, g.feat, sb, tb, sa, ta,
	                                         m->d_W + 0 * kLayer, S);
}

void release_workspace(bfmx_model* m) {
	for (auto& p : m->h) {
		dfree(p);
	}
	for (auto& p : m->m) {
		dfree(p);
	}
	for (auto& p : m->hg) {
		dfree(p);
	}
	for (auto& p : m->mg) {
		dfree(p);
	}
	for (auto& p : m->adj) {
		dfree(p);
	}
	dfree(m->x);
	dfree(m->xg);
	dfree(m->z);
	dfree(m->dA);
	dfree(m->dB);
	dfree(m->dM);
	dfree(m->up);
	dfree(m->partial);
	dfree(m->lpart);
	m->cap = 0;
}

int ensure_train_ws(bfmx_model* m, int64_t rows) {
	if (rows <= m->cap) {
		return BFMX_OK;
	}
	release_workspace(m);
	size_t const hr = static_cast<size_t>(rows) * kH;
	for (auto& p : m->h) {
		BFMX_CHECK(dalloc(&p, hr));
	}
	for (auto& p : m->m) {
		BFMX_CHECK(dalloc(&p, hr));
	}
	for (auto& p : m->hg) {
		BFMX_CHECK(dalloc(&p, hr));
	}
	for (auto& p : m->mg) {
		BFMX_CHECK(dalloc(&p, hr));
	}
	for (auto& p : m->adj) {
		BFMX_CHECK(dalloc(&p, hr));
	}
	BFMX_CHECK(dalloc(&m->x, 2 * static_cast<size_t>(rows)));
	BFMX_CHECK(dalloc(&m->xg, 2 * static_cast<size_t>(rows)));
	BFMX_CHECK(dalloc(&m->z, 2 * static_cast<size_t>(rows)));
	BFMX_CHECK(dalloc(&m->dA, hr));
	BFMX_CHECK(dalloc(&m->dB, hr));
	BFMX_CHECK(dalloc(&m->dM, hr));
	BFMX_CHECK(dalloc(&m->up, hr));
	BFMX_CHECK(dalloc(&m->partial, static_cast<size_t>(kGradParts) * kLayer));
	BFMX_CHECK(dalloc(&m->lpart, static_cast<size_t>(grid_of(rows, kWG))));
	m->cap = rows;
	return BFMX_OK;
}

// Weight gradient of <up, N(theta) x> through the stored forward pass (h, ms)
// of input x (FP32, 2 per row), into (or, with accumulate, added to) m->d_grad.
void net_weight_grad(hipStream_t st, const Net& g, int64_t R, int B, bfmx_model* m, u16* const* h, u16* const* ms,
                     const float* x, const float* up, int accumulate) {
	int const gt = grid_of(R, kT);
	int const gp = std::min(kGradParts, gt);
	const float* dh = up;
	float* bufs[2] = {m->dA, m->dB};
	int which = 0;
	for (int l = 2; l >= 0; l--) {
		float* dprev = bufs[which];
		k_layer_bwd<<<gp, kWG, 0, st>>>(R, h[l], ms[l], dh, m->d_W + l * kLayer, dprev, m->dM, m->partial);
		k_reduce_partials<<<grid_of(kLayer, kWG), kWG, 0, st>>>(gp, kLayer, m->partial, m->d_grad + kOffL + l * kLayer,
		                                                       accumulate);
		k_agg_bwd<<<gt, kWG, 0, st>>>(R, B, g.slice_ptr, g.row_len, g.col, g.val32, g.feat, m->dM, dprev);
		dh = dprev;
		which ^= 1;
	}
	int const ge = std::min(kGradParts, grid_of(R, 4));
	k_encode_bwd<<<ge, kWG, 0, st>>>(R, x, dh, m->partial);
	k_reduce_partials<<<1, kWG, 0, st>>>(ge, 2 * kH, m->partial, m->d_grad + kOffW0, accumulate);
}

}  // namespace

int mpnn_n_params() {
	return kNParams;
}

int model_create(bfmx_model** out, uint64_t seed) {
	static std::atomic<uint64_t> created{0};
	auto* m = new bfmx_model();
	m->gen = ++created;
	int const rc = [&]() -> int {
		BFMX_HIP(hipGetDevice(&m->device));
		BFMX_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
		std::vector<float> p(kNParams, 0.f);
		std::mt19937_64 rng(seed);
		// Glorot-uniform layers and a small encoder: the untrained
		// preconditioner is block Jacobi plus a small PSD term (N = 0 exactly
		// would be a stationary point of N^T N: zero gradient)
		std::uniform_real_distribution<float> U(-std::sqrt(6.f / (3 * kH)), std::sqrt(6.f / (3 * kH)));
		std::uniform_real_distribution<float> U0(-kEncInit, kEncInit);
		for (int k = kOffW0; k < kOffL; k++) {
			p[k] = U0(rng);
		}
		for (int k = kOffL; k < kNParams; k++) {
			p[k] = U(rng);
		}
		BFMX_CHECK(dalloc(&m->d_params, kNParams));
		BFMX_CHECK(dalloc(&m->d_grad, kNParams));
		BFMX_CHECK(dalloc(&m->d_m, kNParams));
		BFMX_CHECK(dalloc(&m->d_v, kNParams));
		BFMX_CHECK(dalloc(&m->d_W, 3 * kLayer));
		BFMX_CHECK(dalloc(&m->d_Wt, 3 * kLayer));
		BFMX_CHECK(dalloc(&m->loss, 1));
		BFMX_HIP(hipMemcpyAsync(m->d_params, p.data(), kNParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
		BFMX_HIP(hipMemsetAsync(m->d_m, 0, kNParams * sizeof(float), m->stream));
		BFMX_HIP(hipMemsetAsync(m->d_v, 0, kNParams * sizeof(float), m->stream));
		k_pack<<<grid_of(3 * kLayer, kWG), kWG, 0, m->stream>>>(m->d_params, m->d_W, m->d_Wt);
		BFMX_HIP(hipGetLastError());
		BFMX_HIP(hipStreamSynchronize(m->stream));
		return BFMX_OK;
	}();
	if (rc != BFMX_OK) {
		model_destroy(m);
		return rc;
	}
	*out = m;
	return BFMX_OK;
}

void model_destroy(bfmx_model* m) {
	if (!m) {
		return;
	}
	(void) hipSetDevice(m->device);
	if (m->stream) {
		(void) hipStreamSynchronize(m->stream);
	}
	release_workspace(m);
	dfree(m->d_params);
	dfree(m->d_grad);
	dfree(m->d_m);
	dfree(m->d_v);
	dfree(m->d_W);
	dfree(m->d_Wt);
	dfree(m->loss);
	if (m->stream) {
		(void) hipStreamDestroy(m->stream);
	}
	delete m;
}

int model_get_params(bfmx_model* m, float* p) {
	BFMX_HIP(hipStreamSynchronize(m->stream));
	BFMX_HIP(hipMemcpy(p, m->d_params, kNParams * sizeof(float), hipMemcpyDeviceToHost));
	return BFMX_OK;
}

int model_set_params(bfmx_model* m, const float* p) {
	BFMX_HIP(hipStreamSynchronize(m->stream));
	BFMX_HIP(hipMemcpy(m->d_params, p, kNParams * sizeof(float), hipMemcpyHostToDevice));
	BFMX_HIP(hipMemsetAsync(m->d_m, 0, kNParams * sizeof(float), m->stream));
	BFMX_HIP(hipMemsetAsync(m->d_v, 0, kNParams * sizeof(float), m->stream));
	m->adam_t = 0;
	k_pack<<<grid_of(3 * kLayer, kWG), kWG, 0, m->stream>>>(m->d_params, m->d_W, m->d_Wt);
	BFMX_HIP(hipGetLastError());
	BFMX_HIP(hipStreamSynchronize(m->stream));
	return BFMX_OK;
}

// Features of a system's current matrix and Dirichlet set into feat (8 per node).
int mpnn_features(bfmx_system* s, float* feat, hipStream_t st) {
	int const nb = static_cast<int>(s->n_nodes);
	k_feat1<<<grid_of(nb, kWG), kWG, 0, st>>>(nb, s->d_diag_blk, s->d_val, s->d_fixed, feat);
	k_feat2<<<grid_of(nb, kWG), kWG, 0, st>>>(nb, s->d_slice_ptr, s->d_row_len, s->d_col, feat);
	BFMX_HIP(hipGetLastError());
	return BFMX_OK;
}

// One forward and backward pass over pairs p0 .. p0 + B - 1 of entry e:
// the loss into m->loss, the gradient into m->d_grad.
int model_loss_grad(bfmx_model* m, bfmx_trainset* ts, int entry, int p0, int B) {
	auto const& E = ts->entries[entry];
	bfmx_system* s = E.sys;
	hipStream_t const st = m->stream;
	int64_t const R = E.n * B;
	int64_t const ps = 2 * E.n;
	BFMX_CHECK(ensure_train_ws(m, R));
	float const* r = E.d_r + static_cast<int64_t>(p0) * ps;
	float const* e = E.d_e + static_cast<int64_t>(p0) * ps;
	Net const g{s->d_slice_ptr, s->d_row_len, s->d_col, E.d_val32, E.d_feat};
	int const gr = grid_of(R, kWG), gt = grid_of(R, kT);
	// u = N^T N x, z = y + P Delta^-1/2 u
	k_encode<true, float, u16><<<gr, kWG, 0, st>>>(R, B, r, ps, E.d_feat, m->d_params, m->h[0], m->x, nullptr);
	net_forward<u16>(st, g, R, B, m, m->h, m->m, nullptr);
	net_adjoint<u16>(st, g, R, B, m, m->h[3], m->adj[0], m->adj[1], m->adj[2], m->adj[3], nullptr);
	k_decode<true, false, float, float, u16><<<gt, kWG, 0, st>>>(R, B, r, ps, g.slice_ptr, g.row_len, g.col, g.val32,
	                                                        g.feat, m->d_params, m->adj[0], m->adj[1], nullptr,
	                                                        nullptr, m->z, nullptr, nullptr, nullptr);
	// loss and g = P Delta^-1/2 dL/dz (into xg)
	k_loss<<<gr, kWG, 0, st>>>(R, B, g.slice_ptr, g.row_len, g.col, g.val32, g.feat, m->z, e, ps, E.d_den + p0,
	                           m->xg, m->lpart);
	k_sum_loss<<<1, kWG, 0, st>>>(gr, m->lpart, m->loss);
	// N g
	k_encode<false, float, u16><<<gr, kWG, 0, st>>>(R, B, m->xg, 0, E.d_feat, m->d_params, m->hg[0], nullptr,
	                                                nullptr);
	net_forward<u16>(st, g, R, B, m, m->hg, m->mg, nullptr);
	// d<N g, N x>/d theta + d<N x, N g>/d theta
	k_bf16_to_f32<<<grid_of(R * kH, kWG), kWG, 0, st>>>(R * kH, m->hg[3], m->up);
	net_weight_grad(st, g, R, B, m, m->h, m->m, m->x, m->up, 0);
	k_bf16_to_f32<<<grid_of(R * kH, kWG), kWG, 0, st>>>(R * kH, m->h[3], m->up);
	net_weight_grad(st, g, R, B, m, m->hg, m->mg, m->xg, m->up, 1);
	BFMX_HIP(hipGetLastError());
	return BFMX_OK;
}

int model_adam(bfmx_model* m, float lr) {
	m->adam_t += 1;
	float const b1 = 0.9f, b2 = 0.999f;
	float const c1 = 1.f / (1.f - std::pow(b1, static_cast<float>(m->adam_t)));
	float const c2 = 1.f / (1.f - std::pow(b2, static_cast<float>(m->adam_t)));
	k_adam<<<grid_of(kNParams, kWG), kWG, 0, m->stream>>>(kNParams, m->d_params, m->d_grad, m->d_m, m->d_v, lr, b1,
	                                                      b2, 1e-8f, c1, c2);
	k_pack<<<grid_of(3 * kLayer, kWG), kWG, 0, m->stream>>>(m->d_params, m->d_W, m->d_Wt);
	BFMX_HIP(hipGetLastError());
	return BFMX_OK;
}

int model_train(bfmx_model* m, bfmx_trainset* ts, int steps, int batch, float lr, uint64_t seed, double* losses) {
	if (ts->entries.empty()) {
		set_error("empty training set");
		return BFMX_E_ARG;
	}
	std::mt19937_64 rng(seed);
	for (int k = 0; k < steps; k++) {
		int const ei = static_cast<int>(rng() % ts->entries.size());
		auto const& E = ts->entries[ei];
		int const B = std::min(batch, E.P);
		int const p0 = static_cast<int>(rng() % static_cast<uint64_t>(E.P - B + 1));
		BFMX_CHECK(model_loss_grad(m, ts, ei, p0, B));
		BFMX_CHECK(model_adam(m, lr));
		if (losses) {
			BFMX_HIP(hipMemcpyAsync(&losses[k], m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
		}
	}
	BFMX_HIP(hipStreamSynchronize(m->stream));
	return BFMX_OK;
}

void trainset_destroy(bfmx_trainset* ts) {
	if (!ts) {
		return;
	}
	for (auto& E : ts->entries) {
		dfree(E.d_val32);
		dfree(E.d_feat);
		dfree(E.d_r);
		dfree(E.d_e);
		dfree(E.d_s);
		dfree(E.d_den);
	}
	delete ts;
}

// Records up to max_pairs (r, e) pairs of a plain Jacobi-PCG run on the
// system's current right-hand side (one every `every` iterations), after a
// first solve to rtol gives x*. The FP32 matrix and the node features are
// snapshotted with the pairs.
int trainset_add(bfmx_trainset* ts, bfmx_system* s, int max_pairs, int every, double rtol) {
	hipStream_t const st = s->stream;
	int64_t const n = s->n_nodes;
	BFMX_CHECK(finalize_conditions(s));
	bfmx_stats_t stats{};
	int const rc = solve(s, nullptr, BFMX_PRECOND_JACOBI, BFMX_PRECISION_FP64, rtol, 0, &stats);
	if (rc != BFMX_OK && rc != BFMX_E_STAGNATION) {
		return rc;
	}
	int64_t const iters = stats.iterations;
	if (every == 0) {   // spread the pairs over the whole run
		every = static_cast<int>(std::max<int64_t>(1, iters / std::max(1, max_pairs)));
	}
	int const P = static_cast<int>(std::min<int64_t>(max_pairs, std::max<int64_t>(1, iters / every)));
	bfmx_trainset::Entry E;
	E.sys = s;
	E.n = n;
	E.P = P;
	double* d_xs = nullptr;
	double* d_part = nullptr;
	int const rc2 = [&]() -> int {
		BFMX_CHECK(dalloc(&d_xs, 2 * n));
		BFMX_HIP(hipMemcpyAsync(d_xs, s->d_x, 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
		BFMX_CHECK(dalloc(&E.d_val32, 4 * s->nnzb_pad));
		BFMX_CHECK(dalloc(&E.d_feat, 8 * n));
		BFMX_CHECK(dalloc(&E.d_r, 2 * n * P));
		BFMX_CHECK(dalloc(&E.d_e, 2 * n * P));
		BFMX_CHECK(dalloc(&E.d_s, P));
		BFMX_CHECK(dalloc(&E.d_den, P));
		BFMX_CHECK(dalloc(&d_part, 2 * static_cast<size_t>(grid_of(n, kWG))));
		BFMX_CHECK(record_pcg(s, P, every, [&](int p) -> int {
			k_record<<<grid_of(2 * n, kWG), kWG, 0, st>>>(2 * n, s->d_r, s->d_x, d_xs, E.d_r + 2 * n * p,
			                                              E.d_e + 2 * n * p);
			BFMX_HIP(hipGetLastError());
			return BFMX_OK;
		}));
		BFMX_CHECK(snapshot_f32(s, E.d_val32));
		BFMX_CHECK(mpnn_features(s, E.d_feat, st));
		std::vector<float> sv(P), dv(P);
		std::vector<double> part(2 * static_cast<size_t>(grid_of(n, kWG)));
		for (int p = 0; p < P; p++) {
			k_pair_norms<<<grid_of(n, kWG), kWG, 0, st>>>(static_cast<int>(n), s->d_slice_ptr, s->d_row_len, s->d_col,
			                                              s->d_val, E.d_r + 2 * n * p, E.d_e + 2 * n * p, d_part);
			BFMX_HIP(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
			BFMX_HIP(hipStreamSynchronize(st));
			double rr = 0, eae = 0;
			for (size_t w = 0; w < part.size() / 2; w++) {
				rr += part[2 * w];
				eae += part[2 * w + 1];
			}
			sv[p] = static_cast<float>(rr > 0 ? std::sqrt(rr / (2.0 * n)) : 1.0);
			dv[p] = static_cast<float>(eae > 0 ? eae : 1.0);
		}
		BFMX_HIP(hipMemcpyAsync(E.d_s, sv.data(), P * sizeof(float), hipMemcpyHostToDevice, st));
		BFMX_HIP(hipMemcpyAsync(E.d_den, dv.data(), P * sizeof(float), hipMemcpyHostToDevice, st));
		BFMX_HIP(hipStreamSynchronize(st));
		return BFMX_OK;
	}();
	dfree(d_xs);
	dfree(d_part);
	if (rc2 != BFMX_OK) {
		dfree(E.d_val32);
		dfree(E.d_feat);
		dfree(E.d_r);
		dfree(E.d_e);
		dfree(E.d_s);
		dfree(E.d_den);
		return rc2;
	}
	ts->entries.push_back(E);
	ts->pairs += P;
	return BFMX_OK;
}

namespace {

__global__ __launch_bounds__(kWG) void k_f32_to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kWG + threadIdx.x;
	if (i < n) {
		b[i] = static_cast<double>(a[i]);
	}
}

// e = e - v + D^-1 r for one pair (node i: D^-1 from its FP32 features).
__global__ __launch_bounds__(kWG) void k_rebase(int64_t n, const float* __restrict__ feat, const float* __restrict__ r,
                                                const float* __restrict__ v, float* __restrict__ e) {
	int64_t const i = static_cast<int64_t>(blockIdx.x) * kWG + threadIdx.x;
	if (i >= n) {
		return;
	}
	const float* f = feat + 8 * i;
	float const r0 = r[2 * i], r1 = r[2 * i + 1];
	e[2 * i] = e[2 * i] - v[2 * i] + (f[0] * r0 + f[1] * r1);
	e[2 * i + 1] = e[2 * i + 1] - v[2 * i + 1] + (f[2] * r0 + f[3] * r1);
}

}  // namespace

int trainset_rebase(bfmx_trainset* ts) {
	for (auto& E : ts->entries) {
		bfmx_system* s = E.sys;
		hipStream_t const st = s->stream;
		int64_t const n = E.n;
		BFMX_CHECK(ensure_workspace_public(s));
		BFMX_CHECK(prepare_amg(s));
		// the V-cycle's kernels run while the scalars say running (done = 0)
		BFMX_HIP(hipMemsetAsync(s->d_scal, 0, sizeof(Scalars), st));
		for (int p = 0; p < E.P; p++) {
			float* rp = E.d_r + 2 * n * p;
			k_f32_to_f64<<<grid_of(2 * n, kWG), kWG, 0, st>>>(2 * n, rp, s->d_r);
			int nw = 0;
			float* v = amg_vcycle(s, s->d_r, nullptr, nullptr, s->d_scal, &nw);
			k_rebase<<<grid_of(n, kWG), kWG, 0, st>>>(n, E.d_feat, rp, v, E.d_e + 2 * n * p);
		}
		BFMX_HIP(hipGetLastError());
		BFMX_HIP(hipStreamSynchronize(st));
	}
	return BFMX_OK;
}

int trainset_get(bfmx_trainset* ts, int entry, float* feat, float* r, float* e, float* s_den) {
	if (entry < 0 || entry >= static_cast<int>(ts->entries.size())) {
		set_error("no training entry %d", entry);
		return BFMX_E_ARG;
	}
	auto const& E = ts->entries[entry];
	hipStream_t const st = E.sys->stream;
	BFMX_HIP(hipStreamSynchronize(st));
	size_t const v = 2 * static_cast<size_t>(E.n) * E.P;
	if (feat) {
		BFMX_HIP(hipMemcpy(feat, E.d_feat, 8 * E.n * sizeof(float), hipMemcpyDeviceToHost));
	}
	if (r) {
		BFMX_HIP(hipMemcpy(r, E.d_r, v * sizeof(float), hipMemcpyDeviceToHost));
	}
	if (e) {
		BFMX_HIP(hipMemcpy(e, E.d_e, v * sizeof(float), hipMemcpyDeviceToHost));
	}
	if (s_den) {
		BFMX_HIP(hipMemcpy(s_den, E.d_s, E.P * sizeof(float), hipMemcpyDeviceToHost));
		BFMX_HIP(hipMemcpy(s_den + E.P, E.d_den, E.P * sizeof(float), hipMemcpyDeviceToHost));
	}
	return BFMX_OK;
}

// ------------------------------------------------------------------ solve

int mpnn_prepare(bfmx_system* s) {
	int64_t const n = s->n_nodes;
	if (!s->model) {
		set_error("no model attached to the system");
		return BFMX_E_ARG;
	}
	if (s->partial_cap < grid_of(n, kT)) {   // the decoder writes one partial per 64 rows
		set_error("solver partial-sum buffer too small for the learned preconditioner");
		return BFMX_E_ARG;
	}
	if (!s->d_mp_feat) {
		BFMX_CHECK(dalloc(&s->d_mp_feat, 8 * n));
		for (auto& p : s->d_mp_h) {
			BFMX_CHECK(dalloc(&p, kH * n));
		}
		BFMX_CHECK(dalloc(&s->d_mp_z, 2 * n));
	}
	return BFMX_OK;
}

int mpnn_setup_numeric(bfmx_system* s) {
	// the model's weights are written on its own stream
	BFMX_HIP(hipStreamSynchronize(s->model->stream));
	return mpnn_features(s, s->d_mp_feat, s->stream);
}

double* mpnn_apply(bfmx_system* s, const double* r, const double* q, double* partial, const Scalars* S,
                   int* n_partials, const float* base) {
	hipStream_t const st = s->stream;
	bfmx_model* m = s->model;
	int64_t const R = s->n_nodes;
	Net const g{s->d_slice_ptr, s->d_row_len, s->d_col, s->d_val32, s->d_mp_feat};
	float** h = s->d_mp_h;   // four FP32 buffers of 64 per node
	int const gr = grid_of(R, kWG), gt = grid_of(R, kT);
	k_encode<true, double, float><<<gr, kWG, 0, st>>>(R, 1, r, 0, g.feat, m->d_params, h[0], nullptr, S);
	float* hf[4] = {h[0], h[1], h[0], h[1]};   // ping-pong: h3 lands in h[1]
	net_forward<float>(st, g, R, 1, m, hf, nullptr, S);
	net_adjoint<float>(st, g, R, 1, m, h[1], h[2], h[3], h[0], h[1], S);
	k_decode<false, true, double, double, float><<<gt, kWG, 0, st>>>(R, 1, r, 0, g.slice_ptr, g.row_len, g.col, g.val32,
	                                                          g.feat, m->d_params, h[2], h[3], s->d_dinv, base,
	                                                          s->d_mp_z, q, partial, S);
	*n_partials = gt;
	return s->d_mp_z;
}

void mpnn_release(bfmx_system* s) {
	dfree(s->d_mp_feat);
	for (auto& p : s->d_mp_h) {
		dfree(p);
	}
	dfree(s->d_mp_z);
}

}  // namespace bfmx
