// Tile-pass kernels (gfx950): several gates per HBM pass, gates applied on the
// matrix cores.
//
// A "pass" is a set of gates, taken in an order consistent with the circuit
// (gates on disjoint wires commute), whose wires all lie in a set Q of K
// "tile" qubits.  A tile is the 2^K x C sub-array of a column block
// X[a * B + c] (column fastest) obtained by fixing the other L - K row bits
// ("rest") and a chunk of C consecutive columns; C * sizeof(element) = 128
// bytes, so every tile row is one fully used 128-byte segment.  A workgroup
// loads a tile into LDS, applies all gates of the pass there and writes it
// back: one HBM read and one write per pass instead of per gate.
//
// Gate application on the matrix cores.  A dual quad (psi and dpsi on the four
// amplitudes of a gate) has 16 real components m = 4 v + r, v = 2 s + p
// (s: primary / tangent state, p: re / im, r: amplitude).  The update
//   psi' = G psi,  dpsi' = G dpsi + Z psi
// is a real 16 x 16 matrix M (complex blocks [[G, 0], [Z, G]]) applied to the
// 16 x 16 array of 16 quads, i.e. four chained 16x16x4 steps
//   D = sum_kk M[:, 4kk:4kk+4] X[4kk:4kk+4, :].
// Lane l holds quad (l & 15), amplitude r = l >> 4: its B operand of step kk
// is component v = kk of that amplitude, its A operand M[l % 16][4 kk + l / 16]
// (the host lays the four A values of every lane out in a table).  The gate
// thus costs four registers per lane, and each state element is read and
// written once per gate.  Without the tangent state (plain sweeps) the second
// half of the components is a second quad (quad + 16) and M = [[G, 0], [0, G]].
//
// Reverse pass (k_tile_reverse): reloads the pass-output checkpoint and walks
// the pass's gates backwards, recovering the state before each gate with the
// inverse update (uncompute: psi = G^-1 psi', dpsi = G^-1 dpsi' - G^-1 Z G^-1 psi'),
// so that LDS holds only psi, dpsi, phi, dphi whatever the number of gates:
//   gradient / HVP sums  grad_j += phi (x) psi_j,  hess_j += phi (x) dpsi_j + dphi (x) psi_j
//   adjoint update       phi <- G_j^T phi,  dphi <- G_j^T dphi + Z_j^T phi
// The engine uses it only for gates that are unitary to 1e-6 (then G^-1 has
// norm 1 and the recovery adds rounding only).  The outer-product sums run on
// the matrix cores as 16x16x4 steps over 4 quads, in FP64 (FP64 states) or in
// FP32 per 16-quad group added into FP64 accumulators (FP32 states).  Partial
// sums per (workgroup, wave, gate) are reduced in a fixed order by
// k_tile_reduce: no atomics, bitwise reproducible.
#pragma once

#include "zc.hpp"

namespace rqc {

constexpr int kTileThreads = 256;
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kMaxTileGates = 8;
// elements a thread moves per tile state (E / threads, E <= 2048)
constexpr int kMaxEpt = 8;

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

struct TileArgs
{
	int K;             // tile row bits
	int logC;          // log2 of tile columns
	int ng;            // gates in this pass
	int x0;            // position of the pass's first gate in the operand table
	unsigned tmask;    // row bits of the tile (tile bit t = t-th lowest set bit)
	unsigned rmask;    // the other row bits (rest)
	int lp[kMaxTileGates], lq[kMaxTileGates];  // tile bits of wire i and wire j of each gate
};

// operand table: A values of gate x (execution order), kind (0 forward, 1 inverse,
// 2 adjoint), step kk, lane l at amat[((x * 3 + kind) * 4 + kk) * 64 + l]
__device__ __forceinline__ long amat_index(int x, int kind, int kk, int lane)
{
	return (((long)x * 3 + kind) * 4 + kk) * 64 + lane;
}

// zero bits inserted at tile bit positions lo < hi
__device__ __forceinline__ int tile_quad_row(int q, int lo, int hi)
{
	int x = ((q >> lo) << (lo + 1)) | (q & ((1 << lo) - 1));
	x = ((x >> hi) << (hi + 1)) | (x & ((1 << hi) - 1));
	return x;
}

// element offset (within a tile's LDS state) of amplitude r = 2 qi + qj of quad
// qd of the gate on tile bits (p, q): the quad's row with bit p = qi, bit q = qj
__device__ __forceinline__ int quad_elem(int qd, int logC, int p, int q, int r)
{
	const int col = qd & ((1 << logC) - 1);
	const int row = tile_quad_row(qd >> logC, min(p, q), max(p, q)) | ((r >> 1) << p) | ((r & 1) << q);
	return (row << logC) | col;
}

// scatter the low bits of x to the set bits of mask, lowest first
__device__ __forceinline__ long deposit_bits(long x, unsigned mask)
{
	long r = 0;
	for (unsigned m = mask; m != 0; m &= m - 1) {
		if (x & 1) { r |= (long)(m & (~m + 1)); }
		x >>= 1;
	}
	return r;
}

// row-index contributions of the tile bits for the elements e = tid + i * threads
// a thread moves between HBM and LDS; fixed for all tiles of a pass
__device__ __forceinline__ void thread_tile_rows(unsigned tmask, int logC, int E, long* trow)
{
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		trow[i] = (e < E) ? deposit_bits(e >> logC, tmask) : 0;
	}
}

// one 16x16x4 matrix-core step, FP64 or FP32
__device__ __forceinline__ double4_t mfma_step(double a, double b, double4_t c)
{
	return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float4_t mfma_step(float a, float b, float4_t c)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef double4_t type; };
template <> struct Vec4<float> { typedef float4_t type; };

// row of the 16x16 result held by lane l in register v:
// FP64 16x16x4: (l >> 4) + 4 v; FP32 16x16x4: 4 (l >> 4) + v (column: l & 15)
template <typename T>
__device__ __forceinline__ int mfma_row(int lane, int v)
{
	return (sizeof(T) == 8) ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
}

// apply a gate to one group of 16 quads in LDS (in place).  Quad qx of X and
// quad qy of Y hold the lane's operands (dual: Y = tangent state, qy = qx;
// plain: Y = X, qy = qx + 16); the lane's B operands are amplitude
// r = lane >> 4 of them.  a[kk] = M[lane % 16][4 kk + lane / 16].
template <typename T>
__device__ __forceinline__ void mfma_gate_apply(cx<T>* X, cx<T>* Y, int qx, int qy, int logC, int p, int q, const T* a)
{
	const int lane = threadIdx.x & 63;
	const int r = lane >> 4;
	const int ex = quad_elem(qx, logC, p, q, r);
	const int ey = quad_elem(qy, logC, p, q, r);
	const cx<T> x = X[ex];
	const cx<T> y = Y[ey];
	typename Vec4<T>::type acc = {0, 0, 0, 0};
	acc = mfma_step(a[0], x.x, acc);
	acc = mfma_step(a[1], x.y, acc);
	acc = mfma_step(a[2], y.x, acc);
	acc = mfma_step(a[3], y.y, acc);
	if constexpr (sizeof(T) == 8) {
		// register v = component v (x.re, x.im, y.re, y.im) of amplitude r
		X[ex] = cx<T>{acc[0], acc[1]};
		Y[ey] = cx<T>{acc[2], acc[3]};
	} else {
		// register v = component r (x.re, x.im, y.re, y.im) of amplitude v
		T* base = reinterpret_cast<T*>(r < 2 ? X : Y);
		const int qd = (r < 2) ? qx : qy;
#pragma unroll
		for (int v = 0; v < 4; v++) { base[2 * quad_elem(qd, logC, p, q, v) + (r & 1)] = acc[v]; }
	}
}

// element (in units of T) of component m = 8 s + 2 r + p of quad qd for the
// outer-product steps (dual: s selects the primary / tangent array; plain: s
// selects quad qd / qd + 16 of the same array)
template <bool DUAL>
__device__ __forceinline__ int outer_elem(int qd, int logC, int pb, int qb, int m)
{
	const int s = m >> 3, r = (m >> 1) & 3, p = m & 1;
	return 2 * quad_elem(DUAL ? qd : qd + 16 * s, logC, pb, qb, r) + p;
}

// load / store one tile state between HBM and LDS
template <typename T>
__device__ __forceinline__ void tile_load(cx<T>* S, const cx<T>* __restrict__ src, const long* trow, long rbase, long c0,
	int E, int C, long B)
{
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		if (e >= E) { break; }
		S[e] = src[(rbase | trow[i]) * B + c0 + (e & (C - 1))];
	}
}

template <typename T>
__device__ __forceinline__ void tile_fill(cx<T>* S, int E, cx<T> v)
{
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		if (e >= E) { break; }
		S[e] = v;
	}
}

template <typename T>
__device__ __forceinline__ void tile_identity(cx<T>* S, const long* trow, long rbase, long c0, long b0, int E, int C)
{
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		if (e >= E) { break; }
		const long row = rbase | trow[i];
		S[e] = cx<T>{(row == b0 + c0 + (e & (C - 1))) ? (T)1 : (T)0, (T)0};
	}
}

template <typename T>
__device__ __forceinline__ void tile_store(cx<T>* __restrict__ dst, const cx<T>* S, const long* trow, long rbase, long c0,
	int E, int C, long B)
{
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		if (e >= E) { break; }
		dst[(rbase | trow[i]) * B + c0 + (e & (C - 1))] = S[e];
	}
}

// ---------------------------------------------------------------------------
// forward pass: psi -> (gates) psi, with DUAL also dpsi.  The input is
// column block b0 of the identity (flags & 1) or `in`; dpsi is zero (flags & 2)
// or `din`.
template <typename T, bool DUAL>
__global__ __launch_bounds__(kTileThreads) void k_tile_forward(
	const cx<T>* __restrict__ in, cx<T>* __restrict__ out, const cx<T>* __restrict__ din, cx<T>* __restrict__ dout,
	const TileArgs P, const T* __restrict__ amat, long B, long b0, long ntiles, int flags)
{
	extern __shared__ __align__(16) char lds_raw[];
	const int E = 1 << (P.K + P.logC);
	const int C = 1 << P.logC;
	cx<T>* S = reinterpret_cast<cx<T>*>(lds_raw);
	cx<T>* D = S + E;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	constexpr int gq = DUAL ? 16 : 32;            // quads per matrix-core group
	const int qpw = (E / 4) / kTileWaves;         // quads owned by a wave
	const int q0 = wave * qpw;

	T a[kMaxTileGates][4];
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) {
#pragma unroll
		for (int kk = 0; kk < 4; kk++) { a[j][kk] = (j < P.ng) ? amat[amat_index(P.x0 + j, 0, kk, lane)] : (T)0; }
	}
	const long nchunks = B >> P.logC;
	long trow[kMaxEpt];
	thread_tile_rows(P.tmask, P.logC, E, trow);
	for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const long chunk = tile % nchunks, rest = tile / nchunks;
		const long rbase = deposit_bits(rest, P.rmask);
		const long c0 = chunk << P.logC;
		if (flags & 1) { tile_identity(S, trow, rbase, c0, b0, E, C); }
		else { tile_load(S, in, trow, rbase, c0, E, C, B); }
		if (DUAL) {
			if (flags & 2) { tile_fill(D, E, cx<T>{0, 0}); }
			else { tile_load(D, din, trow, rbase, c0, E, C, B); }
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < kMaxTileGates; j++) {
			if (j < P.ng) {
				const int p = P.lp[j], q = P.lq[j];
				for (int g = q0; g < q0 + qpw; g += gq) {
					const int qx = g + (lane & 15);
					mfma_gate_apply<T>(S, DUAL ? D : S, qx, DUAL ? qx : qx + 16, P.logC, p, q, a[j]);
				}
				__syncthreads();
			}
		}
		tile_store(out, S, trow, rbase, c0, E, C, B);
		if (DUAL) { tile_store(dout, D, trow, rbase, c0, E, C, B); }
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------
// reverse pass.  Inputs: psi (and dpsi) at the pass output (checkpoint), phi
// (and dphi) at the pass output; outputs phi (and dphi) at the pass input.
// flags: 4 = dphi at the pass output is zero, 8 = do not write phi at the pass input.
// partial[((part * ng + j) * 256) + row * 16 + col], part = block * waves + wave:
// the 16 x 16 outer-product sums of gate j (rows: phi-side component
// 8 s + 2 r + p, columns: psi-side component).
template <typename T, bool DUAL>
__global__ __launch_bounds__(kTileThreads) void k_tile_reverse(
	const cx<T>* __restrict__ psi_out, const cx<T>* __restrict__ dpsi_out,
	const cx<T>* __restrict__ phi_out, const cx<T>* __restrict__ dphi_out,
	cx<T>* __restrict__ phi_in, cx<T>* __restrict__ dphi_in,
	const TileArgs P, const T* __restrict__ amat, long B, long ntiles, int flags, double* __restrict__ partial)
{
	extern __shared__ __align__(16) char lds_raw[];
	const int E = 1 << (P.K + P.logC);
	const int C = 1 << P.logC;
	cx<T>* S = reinterpret_cast<cx<T>*>(lds_raw);
	cx<T>* D = S + E;
	cx<T>* F = S + (DUAL ? 2 : 1) * E;
	cx<T>* DF = F + E;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	constexpr int gq = DUAL ? 16 : 32;
	const int qpw = (E / 4) / kTileWaves;
	const int q0 = wave * qpw;

	double4_t acc[kMaxTileGates];
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) { acc[j] = double4_t{0, 0, 0, 0}; }

	const long nchunks = B >> P.logC;
	long trow[kMaxEpt];
	thread_tile_rows(P.tmask, P.logC, E, trow);
	for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const long chunk = tile % nchunks, rest = tile / nchunks;
		const long rbase = deposit_bits(rest, P.rmask);
		const long c0 = chunk << P.logC;
		tile_load(S, psi_out, trow, rbase, c0, E, C, B);
		tile_load(F, phi_out, trow, rbase, c0, E, C, B);
		if (DUAL) {
			tile_load(D, dpsi_out, trow, rbase, c0, E, C, B);
			if (flags & 4) { tile_fill(DF, E, cx<T>{0, 0}); }
			else { tile_load(DF, dphi_out, trow, rbase, c0, E, C, B); }
		}
		__syncthreads();
#pragma unroll
		for (int j = kMaxTileGates - 1; j >= 0; j--) {
			if (j < P.ng) {
				const int p = P.lp[j], q = P.lq[j];
				T au[4], aa[4];
#pragma unroll
				for (int kk = 0; kk < 4; kk++) {
					au[kk] = amat[amat_index(P.x0 + j, 1, kk, lane)];
					aa[kk] = amat[amat_index(P.x0 + j, 2, kk, lane)];
				}
				const bool adjoint = (j > 0) || !(flags & 8);
				const int m = lane & 15;
				const T* Fa = reinterpret_cast<const T*>((DUAL && (m >> 3)) ? DF : F);
				const T* Sa = reinterpret_cast<const T*>((DUAL && (m >> 3)) ? D : S);
				for (int g = q0; g < q0 + qpw; g += gq) {
					const int qx = g + (lane & 15);
					const int qy = DUAL ? qx : qx + 16;
					// state before gate j
					mfma_gate_apply<T>(S, DUAL ? D : S, qx, qy, P.logC, p, q, au);
					// outer products over the group's quads, 4 quads per step
					typename Vec4<T>::type grp = {0, 0, 0, 0};
#pragma unroll
					for (int t = 0; t < 4; t++) {
						const int el = outer_elem<DUAL>(g + 4 * t + (lane >> 4), P.logC, p, q, m);
						grp = mfma_step(Fa[el], Sa[el], grp);
					}
#pragma unroll
					for (int v = 0; v < 4; v++) { acc[j][v] += (double)grp[v]; }
					// adjoint at the gate input
					if (adjoint) { mfma_gate_apply<T>(F, DUAL ? DF : F, qx, qy, P.logC, p, q, aa); }
				}
				__syncthreads();
			}
		}
		if (!(flags & 8)) {
			tile_store(phi_in, F, trow, rbase, c0, E, C, B);
			if (DUAL) { tile_store(dphi_in, DF, trow, rbase, c0, E, C, B); }
		}
		__syncthreads();
	}
	const long part = (long)blockIdx.x * kTileWaves + wave;
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) {
		if (j < P.ng) {
			double* d = partial + (part * P.ng + j) * 256;
#pragma unroll
			for (int v = 0; v < 4; v++) { d[mfma_row<T>(lane, v) * 16 + (lane & 15)] = acc[j][v]; }
		}
	}
}

// gate indices of a pass for the reduction
struct TileGates
{
	int k[kMaxTileGates];
};

// Sum of the outer-product partials over all (workgroup, wave) parts in a fixed
// order, decoded into 16 gradient and (DUAL) 16 HVP entries per gate and added
// to acc[1 + 32 * gate + e].  One workgroup of 256 threads per gate of the pass.
//   dual:  grad = phi (x) psi,  hess = phi (x) dpsi + dphi (x) psi
//   plain: grad = phi (x) psi summed over the two quads of each operand pair
__global__ __launch_bounds__(256) void k_tile_reduce(const double* __restrict__ partial, long nparts, int ng,
	const TileGates gk, int dual, zc* __restrict__ acc)
{
	__shared__ double Dm[16][16];
	const int j = blockIdx.x, i = threadIdx.x;
	double s = 0;
	for (long p = 0; p < nparts; p++) { s += partial[(p * ng + j) * 256 + i]; }
	Dm[i >> 4][i & 15] = s;
	__syncthreads();
	if (i < 32) {
		const int hess = i >> 4, r = (i >> 2) & 3, c = i & 3;
		auto M = [&](int sidx, int rr, int pp) { return 8 * sidx + 2 * rr + pp; };
		// sum_{s in S} phi_r(s) * psi_c(s') over the listed block pairs
		auto blk = [&](int sa, int sb, double& re, double& im) {
			re += Dm[M(sa, r, 0)][M(sb, c, 0)] - Dm[M(sa, r, 1)][M(sb, c, 1)];
			im += Dm[M(sa, r, 0)][M(sb, c, 1)] + Dm[M(sa, r, 1)][M(sb, c, 0)];
		};
		double re = 0, im = 0;
		if (dual) {
			if (!hess) { blk(0, 0, re, im); }
			else { blk(0, 1, re, im); blk(1, 0, re, im); }
		} else {
			if (hess) { return; }
			blk(0, 0, re, im);
			blk(1, 1, re, im);
		}
		zc* a = acc + 1 + 32L * gk.k[j] + 16 * hess + 4 * r + c;
		a->x += re;
		a->y += im;
	}
}

}  // namespace rqc
