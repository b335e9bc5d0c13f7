// Tile-pass kernels: several consecutive gates per HBM pass (gfx950).
//
// A "pass" is a run of up to kMaxPassGates consecutive gates of the circuit
// whose wires all lie in a set Q of K "tile" qubits.  A tile is the 2^K x C
// sub-array of a column block X[a * B + c] (column fastest) obtained by
// fixing the other L - K row bits ("rest") and a chunk of C consecutive
// columns.  C * sizeof(element) = 128 bytes, so every tile row is one fully
// used 128-byte segment.  A workgroup loads a tile into LDS, applies all gates
// of the pass there, and writes it back: one HBM read and one write per pass
// instead of per gate.
//
// Forward pass (k_pass_forward): psi -> G psi (and dpsi -> G dpsi + Z psi),
// checkpoints only at pass boundaries.
//
// Reverse pass (k_pass_reverse): reloads the pass-input checkpoint, recomputes
// the per-gate states psi_j, dpsi_j of the pass in LDS, then walks the gates
// backwards:
//   gradient / HVP sums  grad_j += phi (x) psi_j,  hess_j += phi (x) dpsi_j + dphi (x) psi_j
//   adjoint update       phi <- G_j^T phi,  dphi <- G_j^T dphi + Z_j^T phi
// The outer-product sums run on the FP64 matrix cores: for 4 quads at a time a
// wave forms C(16x16) += A(16x4) B(4x16) with A the real components of
// (phi, dphi) and B those of (psi_j, dpsi_j); the 16x16 accumulator holds the
// real and imaginary parts of all products phi_r psi_c, phi_r dpsi_c,
// dphi_r psi_c (dphi_r dpsi_c is unused) in 4 registers per lane and gate.
// The adjoint update runs on the vector ALUs.  Each wave owns a fixed range of
// quads, so the matrix-core reads of a gate and the vector update of the same
// quads need no barrier between them; gates are separated by barriers.
// Partial sums per (workgroup, wave, gate) are reduced in a fixed order by
// k_pass_reduce: no atomics, bitwise reproducible.
#pragma once

#include "zc.hpp"

namespace rqc {

constexpr int kPassThreads = 256;
constexpr int kPassWaves = kPassThreads / 64;
constexpr int kMaxPassGates = 4;

typedef double double4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct PassArgs
{
	int K;             // tile row bits
	int logC;          // log2 of tile columns
	int nrest;         // L - K
	int ng;            // gates in this pass
	int rb[16];        // row bit of tile bit t, t < K
	int restb[24];     // row bit of rest bit u, u < L - K
	int lp[kMaxPassGates], lq[kMaxPassGates];  // tile bits of wire i and wire j of each gate
	int g0;            // first gate of the pass in the device gate table
};

// gate table in LDS: G_j at [32 j], Z_j at [32 j + 16]; filled from the device
// table (G_k at [32 k], Z_k at [32 k + 16]) at kernel start.  Gate entries are
// read from LDS at the point of use (a broadcast read per entry) instead of
// being held in scalar registers, which cannot hold 8 complex FP64 4x4 matrices.
template <typename T>
__device__ __forceinline__ void load_gate_table(const cx<T>* __restrict__ gtab, int g0, int ng, cx<T>* lgt)
{
	for (int i = threadIdx.x; i < 32 * ng; i += blockDim.x) { lgt[i] = gtab[32L * g0 + i]; }
}

// zero bits inserted at tile bit positions lo < hi
__device__ __forceinline__ int tile_quad_row(int q, int lo, int hi)
{
	int x = ((q >> lo) << (lo + 1)) | (q & ((1 << lo) - 1));
	x = ((x >> hi) << (hi + 1)) | (x & ((1 << hi) - 1));
	return x;
}

// element offset (within a tile's LDS state) of amplitude k (= 2 qi + qj) of quad qd of gate (p, q)
struct TileQuad
{
	int e[4];
};

__device__ __forceinline__ TileQuad tile_quad(int qd, int logC, int p, int q)
{
	const int col = qd & ((1 << logC) - 1);
	const int qr = qd >> logC;
	const int base = tile_quad_row(qr, min(p, q), max(p, q));
	TileQuad t;
	t.e[0] = (base << logC) | col;
	t.e[1] = ((base | (1 << q)) << logC) | col;
	t.e[2] = ((base | (1 << p)) << logC) | col;
	t.e[3] = ((base | (1 << p) | (1 << q)) << logC) | col;
	return t;
}

// row-index contribution of the tile bits of tile row t_row
__device__ __forceinline__ long tile_row_bits(const int* rb, int K, int t_row)
{
	long r = 0;
	for (int t = 0; t < K; t++) { r |= (long)((t_row >> t) & 1) << rb[t]; }
	return r;
}

__device__ __forceinline__ long rest_row_bits(const int* restb, int nrest, long rest)
{
	long r = 0;
	for (int u = 0; u < nrest; u++) { r |= ((rest >> u) & 1L) << restb[u]; }
	return r;
}

// elements a thread loads/stores per tile state (E / threads, E <= 2048)
constexpr int kMaxEpt = 8;

// row-index contributions of the tile bits for the elements e = tid + i * threads
// a thread moves between HBM and LDS; fixed for all tiles of a pass
__device__ __forceinline__ void thread_tile_rows(const int* rb, int K, int logC, int E, long* trow)
{
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kPassThreads;
		trow[i] = (e < E) ? tile_row_bits(rb, K, e >> logC) : 0;
	}
}

// apply gate j of the pass to one quad held in LDS (in place): x <- G x,
// with DUAL also dx <- G dx + Z x (dx read before x is overwritten)
template <typename T, bool DUAL>
__device__ __forceinline__ void lds_gate_forward(cx<T>* S, cx<T>* D, const TileQuad& o, const cx<T>* G, const cx<T>* Z)
{
	cx<T> x[4], y[4];
#pragma unroll
	for (int k = 0; k < 4; k++) { x[k] = S[o.e[k]]; }
#pragma unroll
	for (int r = 0; r < 4; r++) {
		cx<T> s = {0, 0};
#pragma unroll
		for (int k = 0; k < 4; k++) { cmac(s, G[4*r + k], x[k]); }
		y[r] = s;
	}
	if (DUAL) {
		cx<T> dx[4];
#pragma unroll
		for (int k = 0; k < 4; k++) { dx[k] = D[o.e[k]]; }
#pragma unroll
		for (int r = 0; r < 4; r++) {
			cx<T> s = {0, 0};
#pragma unroll
			for (int k = 0; k < 4; k++) { cmac(s, G[4*r + k], dx[k]); cmac(s, Z[4*r + k], x[k]); }
			D[o.e[r]] = s;
		}
	}
#pragma unroll
	for (int k = 0; k < 4; k++) { S[o.e[k]] = y[k]; }
}

// out-of-place variant (per-gate states of the reverse recompute)
template <typename T, bool DUAL>
__device__ __forceinline__ void lds_gate_forward_oop(const cx<T>* S0, const cx<T>* D0, cx<T>* S1, cx<T>* D1,
	const TileQuad& o, const cx<T>* G, const cx<T>* Z)
{
	cx<T> x[4];
#pragma unroll
	for (int k = 0; k < 4; k++) { x[k] = S0[o.e[k]]; }
#pragma unroll
	for (int r = 0; r < 4; r++) {
		cx<T> s = {0, 0};
#pragma unroll
		for (int k = 0; k < 4; k++) { cmac(s, G[4*r + k], x[k]); }
		S1[o.e[r]] = s;
	}
	if (DUAL) {
		cx<T> dx[4];
#pragma unroll
		for (int k = 0; k < 4; k++) { dx[k] = D0[o.e[k]]; }
#pragma unroll
		for (int r = 0; r < 4; r++) {
			cx<T> s = {0, 0};
#pragma unroll
			for (int k = 0; k < 4; k++) { cmac(s, G[4*r + k], dx[k]); cmac(s, Z[4*r + k], x[k]); }
			D1[o.e[r]] = s;
		}
	}
}

// ---------------------------------------------------------------------------
// forward pass.  flags: 1 = input is the identity block (columns b0 + c),
// 2 = input dpsi is zero
template <typename T, bool DUAL>
__global__ __launch_bounds__(kPassThreads) void k_pass_forward(
	const cx<T>* __restrict__ in, cx<T>* __restrict__ out, const cx<T>* __restrict__ din, cx<T>* __restrict__ dout,
	const PassArgs<T> P, const cx<T>* __restrict__ gtab, long B, long b0, long ntiles, int flags)
{
	extern __shared__ char lds_raw[];
	const int E = 1 << (P.K + P.logC);
	const int C = 1 << P.logC;
	cx<T>* S = reinterpret_cast<cx<T>*>(lds_raw);
	cx<T>* D = S + E;
	cx<T>* lgt = S + (DUAL ? 2 : 1) * E;
	load_gate_table(gtab, P.g0, P.ng, lgt);
	const long nchunks = B >> P.logC;
	const int tid = threadIdx.x;
	long trow[kMaxEpt];
	thread_tile_rows(P.rb, P.K, P.logC, E, trow);
	for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const long chunk = tile % nchunks, rest = tile / nchunks;
		const long rbase = rest_row_bits(P.restb, P.nrest, rest);
		const long c0 = chunk << P.logC;
#pragma unroll
		for (int i = 0; i < kMaxEpt; i++) {
			const int e = tid + i * kPassThreads;
			if (e >= E) { break; }
			const long row = rbase | trow[i];
			const long col = c0 + (e & (C - 1));
			if (flags & 1) {
				S[e] = cx<T>{(row == b0 + col) ? (T)1 : (T)0, (T)0};
			} else {
				S[e] = in[row * B + col];
			}
			if (DUAL) { D[e] = (flags & 2) ? cx<T>{0, 0} : din[row * B + col]; }
		}
		__syncthreads();
		for (int j = 0; j < P.ng; j++) {
			for (int qd = tid; qd < E / 4; qd += kPassThreads) {
				const TileQuad o = tile_quad(qd, P.logC, P.lp[j], P.lq[j]);
				lds_gate_forward<T, DUAL>(S, D, o, lgt + 32 * j, lgt + 32 * j + 16);
			}
			__syncthreads();
		}
#pragma unroll
		for (int i = 0; i < kMaxEpt; i++) {
			const int e = tid + i * kPassThreads;
			if (e >= E) { break; }
			const long row = rbase | trow[i];
			const long col = c0 + (e & (C - 1));
			out[row * B + col] = S[e];
			if (DUAL) { dout[row * B + col] = D[e]; }
		}
		__syncthreads();
	}
}

// value of the matrix-core operand for lane component m (8 s + 2 r + p) of quad o:
// s selects the primary (0) or tangent (1) state, r the amplitude, p re/im
template <typename T, bool DUAL>
__device__ __forceinline__ double mfma_operand(const cx<T>* X, const cx<T>* DX, const TileQuad& o, int m)
{
	const int s = m >> 3, r = (m >> 1) & 3, p = m & 1;
	if (!DUAL && s) { return 0.0; }
	const cx<T> v = (s ? DX : X)[o.e[r]];
	return (double)(p ? v.y : v.x);
}

// ---------------------------------------------------------------------------
// reverse pass.  LDS: psi_j (and dpsi_j) for j < ng, phi (and dphi).
// flags: 1 = pass input is the identity block, 2 = input dpsi is zero,
//        4 = input dphi (at the pass output) is zero, 8 = do not write phi_in
template <typename T, bool DUAL>
__global__ __launch_bounds__(kPassThreads) void k_pass_reverse(
	const cx<T>* __restrict__ psi_in, const cx<T>* __restrict__ dpsi_in,
	const cx<T>* __restrict__ phi_out, const cx<T>* __restrict__ dphi_out,
	cx<T>* __restrict__ phi_in, cx<T>* __restrict__ dphi_in,
	const PassArgs<T> P, const cx<T>* __restrict__ gtab, long B, long b0, long ntiles, int flags, double* __restrict__ partial)
{
	extern __shared__ char lds_raw[];
	const int E = 1 << (P.K + P.logC);
	const int C = 1 << P.logC;
	const int ng = P.ng;
	cx<T>* base = reinterpret_cast<cx<T>*>(lds_raw);
	// slots: S_0 .. S_{ng-1}, [D_0 .. D_{ng-1}], F, [DF]
	auto Sj = [&](int j) { return base + (long)j * E; };
	auto Dj = [&](int j) { return base + (long)(ng + j) * E; };
	cx<T>* F = base + (long)(DUAL ? 2 * ng : ng) * E;
	cx<T>* DF = F + E;
	cx<T>* lgt = base + (long)(DUAL ? 2 * ng + 2 : ng + 1) * E;
	load_gate_table(gtab, P.g0, ng, lgt);
	const long nchunks = B >> P.logC;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int qpw = (E / 4) / kPassWaves;          // quads owned by a wave
	const int q0 = wave * qpw;

	// two accumulators per gate (even / odd matrix-core steps) halve the dependent chain
	double4_t acc[kMaxPassGates][2];
#pragma unroll
	for (int j = 0; j < kMaxPassGates; j++) { acc[j][0] = double4_t{0, 0, 0, 0}; acc[j][1] = double4_t{0, 0, 0, 0}; }

	long trow[kMaxEpt];
	thread_tile_rows(P.rb, P.K, P.logC, E, trow);
	for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const long chunk = tile % nchunks, rest = tile / nchunks;
		const long rbase = rest_row_bits(P.restb, P.nrest, rest);
		const long c0 = chunk << P.logC;
#pragma unroll
		for (int i = 0; i < kMaxEpt; i++) {
			const int e = tid + i * kPassThreads;
			if (e >= E) { break; }
			const long row = rbase | trow[i];
			const long col = c0 + (e & (C - 1));
			const long g = row * B + col;
			Sj(0)[e] = (flags & 1) ? cx<T>{(row == b0 + col) ? (T)1 : (T)0, (T)0} : psi_in[g];
			F[e] = phi_out[g];
			if (DUAL) {
				Dj(0)[e] = (flags & 2) ? cx<T>{0, 0} : dpsi_in[g];
				DF[e] = (flags & 4) ? cx<T>{0, 0} : dphi_out[g];
			}
		}
		__syncthreads();
		// per-gate states of the pass
		for (int j = 0; j + 1 < ng; j++) {
			for (int qd = tid; qd < E / 4; qd += kPassThreads) {
				const TileQuad o = tile_quad(qd, P.logC, P.lp[j], P.lq[j]);
				lds_gate_forward_oop<T, DUAL>(Sj(j), Dj(j), Sj(j + 1), Dj(j + 1), o, lgt + 32 * j, lgt + 32 * j + 16);
			}
			__syncthreads();
		}
		// backwards through the gates (unrolled over the maximum so that the
		// per-gate matrix-core accumulators stay in registers)
#pragma unroll
		for (int j = kMaxPassGates - 1; j >= 0; j--) {
			if (j >= ng) { continue; }
			const int p = P.lp[j], q = P.lq[j];
			// matrix-core sums over the wave's quads, 4 quads per instruction
			const int ngroups = qpw / 4;   // 4-quad groups owned by this wave
			int it = 0;
			for (; it + 1 < ngroups; it += 2) {
				const TileQuad o0 = tile_quad(q0 + 4 * it + (lane >> 4), P.logC, p, q);
				const TileQuad o1 = tile_quad(q0 + 4 * it + 4 + (lane >> 4), P.logC, p, q);
				const double a0 = mfma_operand<T, DUAL>(F, DF, o0, lane & 15);
				const double b0v = mfma_operand<T, DUAL>(Sj(j), Dj(j), o0, lane & 15);
				const double a1 = mfma_operand<T, DUAL>(F, DF, o1, lane & 15);
				const double b1v = mfma_operand<T, DUAL>(Sj(j), Dj(j), o1, lane & 15);
				acc[j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0v, acc[j][0], 0, 0, 0);
				acc[j][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1v, acc[j][1], 0, 0, 0);
			}
			if (it < ngroups) {
				const TileQuad o0 = tile_quad(q0 + 4 * it + (lane >> 4), P.logC, p, q);
				const double a0 = mfma_operand<T, DUAL>(F, DF, o0, lane & 15);
				const double b0v = mfma_operand<T, DUAL>(Sj(j), Dj(j), o0, lane & 15);
				acc[j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0v, acc[j][0], 0, 0, 0);
			}
			// adjoint update of the same quads (vector ALUs)
			if (j > 0 || !(flags & 8)) {
				for (int qd = q0 + lane; qd < q0 + qpw; qd += 64) {
					const TileQuad o = tile_quad(qd, P.logC, p, q);
					cx<T> f[4], y[4];
#pragma unroll
					for (int k = 0; k < 4; k++) { f[k] = F[o.e[k]]; y[k] = cx<T>{0, 0}; }
					const cx<T>* G = lgt + 32 * j;
#pragma unroll
					for (int c = 0; c < 4; c++) {
#pragma unroll
						for (int r = 0; r < 4; r++) { cmac(y[c], G[4*r + c], f[r]); }
					}
					if (DUAL) {
						const cx<T>* Z = lgt + 32 * j + 16;
						cx<T> df[4], dy[4];
#pragma unroll
						for (int k = 0; k < 4; k++) { df[k] = DF[o.e[k]]; dy[k] = cx<T>{0, 0}; }
#pragma unroll
						for (int c = 0; c < 4; c++) {
#pragma unroll
							for (int r = 0; r < 4; r++) { cmac(dy[c], G[4*r + c], df[r]); cmac(dy[c], Z[4*r + c], f[r]); }
						}
#pragma unroll
						for (int k = 0; k < 4; k++) { DF[o.e[k]] = dy[k]; }
					}
#pragma unroll
					for (int k = 0; k < 4; k++) { F[o.e[k]] = y[k]; }
				}
			}
			__syncthreads();
		}
		if (!(flags & 8)) {
#pragma unroll
			for (int i = 0; i < kMaxEpt; i++) {
				const int e = tid + i * kPassThreads;
				if (e >= E) { break; }
				const long row = rbase | trow[i];
				const long col = c0 + (e & (C - 1));
				phi_in[row * B + col] = F[e];
				if (DUAL) { dphi_in[row * B + col] = DF[e]; }
			}
		}
		__syncthreads();
	}
	// partial[(part * ng + j) * 256 + lane * 4 + v], part = block * waves + wave
	const long part = (long)blockIdx.x * kPassWaves + wave;
#pragma unroll
	for (int j = 0; j < kMaxPassGates; j++) {
		if (j < ng) {
			double* d = partial + ((part * ng + j) * 256) + lane * 4;
			const double4_t a = acc[j][0] + acc[j][1];
			d[0] = a[0]; d[1] = a[1]; d[2] = a[2]; d[3] = a[3];
		}
	}
}

// Sum of the matrix-core partials over all (workgroup, wave) parts in a fixed
// order, decoded into 16 gradient and 16 HVP entries per gate and added to
// acc[1 + 32 * (gate0 + j) + e].  One workgroup of 256 threads per gate.
// The accumulator of lane l, register v holds D[(l >> 4) + 4 v][l & 15].
__global__ __launch_bounds__(256) void k_pass_reduce(const double* __restrict__ partial, long nparts, int ng, int gate0,
	zc* __restrict__ acc)
{
	__shared__ double Dm[16][16];
	const int j = blockIdx.x, i = threadIdx.x;
	double s = 0;
	for (long p = 0; p < nparts; p++) { s += partial[(p * ng + j) * 256 + i]; }
	const int l = i >> 2, v = i & 3;
	Dm[(l >> 4) + 4 * v][l & 15] = s;
	__syncthreads();
	if (i < 32) {
		const int hess = i >> 4, r = (i >> 2) & 3, c = i & 3;
		auto M = [&](int sidx, int rr, int pp) { return 8 * sidx + 2 * rr + pp; };
		double re, im;
		if (!hess) {
			re = Dm[M(0, r, 0)][M(0, c, 0)] - Dm[M(0, r, 1)][M(0, c, 1)];
			im = Dm[M(0, r, 0)][M(0, c, 1)] + Dm[M(0, r, 1)][M(0, c, 0)];
		} else {
			re = Dm[M(0, r, 0)][M(1, c, 0)] - Dm[M(0, r, 1)][M(1, c, 1)] + Dm[M(1, r, 0)][M(0, c, 0)] - Dm[M(1, r, 1)][M(0, c, 1)];
			im = Dm[M(0, r, 0)][M(1, c, 1)] + Dm[M(0, r, 1)][M(1, c, 0)] + Dm[M(1, r, 0)][M(0, c, 1)] + Dm[M(1, r, 1)][M(0, c, 0)];
		}
		zc* a = acc + 1 + 32L * (gate0 + j) + 16 * hess + 4 * r + c;
		a->x += re;
		a->y += im;
	}
}

}  // namespace rqc
