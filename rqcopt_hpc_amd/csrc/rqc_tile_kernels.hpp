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
// matrix-core groups a wave processes together (independent chains interleave)
constexpr int kGroupStep = 2;
// elements a thread moves per tile state (E / threads, E <= 2048)
constexpr int kMaxEpt = 8;
// LDS states of a tile are E elements + 64 bytes apart, so that equal positions of
// two states fall on different banks
constexpr int kLdsPadBytes = 64;

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// LDS layout of a tile: element (row, col) at (row << logC) | (col ^ f(row)),
// with f the XOR of per-tile-bit column masks h[t] over the set bits of row
// (bit i of f = parity(row & swz[i])).  The host picks h per pass (a colouring
// of the pass's gate graph) so that the four amplitude rows of every gate's
// quads fall on different banks.
struct TileArgs
{
	int K;             // tile row bits
	int logC;          // log2 of tile columns
	int ng;            // gates in this pass
	int x0;            // position of the pass's first gate in the operand table
	unsigned rmask;    // row bits outside the tile (rest), deposited lowest first
	unsigned swz[3];   // column swizzle masks over tile rows
	unsigned barrier;  // bit j: the waves own different quads for gates j - 1 and j
	int rb[16];        // row bit of tile bit t, t < K
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

__device__ __forceinline__ int swz_f(int row, const TileArgs& P)
{
	return (__popc(row & P.swz[0]) & 1) | ((__popc(row & P.swz[1]) & 1) << 1) | ((__popc(row & P.swz[2]) & 1) << 2);
}

__device__ __forceinline__ int lds_elem(int row, int col, const TileArgs& P)
{
	return (row << P.logC) | (col ^ swz_f(row, P));
}

// Quad qd of the gate on tile bits (p, q): compressed row qd >> logC (the row
// without bits p, q), column qd mod C.  LDS position of its amplitude
// r = 2 qi + qj (bit p = qi, bit q = qj).  The position is XOR-linear in qd
// for disjoint bit parts, so pos(gbase | j, r) = pos(gbase, 0) ^ pos(j, r): a
// wave-uniform group term and a per-lane term fixed for the gate.
__device__ __forceinline__ int quad_pos(int qd, int r, int p, int q, const TileArgs& P)
{
	const int row = tile_quad_row(qd >> P.logC, min(p, q), max(p, q)) | ((r >> 1) << p) | ((r & 1) << q);
	return lds_elem(row, qd & ((1 << P.logC) - 1), P);
}

// row-index contributions of the tile bits and LDS positions of the elements
// e = tid + i * threads a thread moves between HBM and LDS; fixed for a pass
__device__ __forceinline__ void thread_tile_elems(const TileArgs& P, int E, int* trow, int* lidx)
{
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		int r = 0;
		const int tr = e >> P.logC;
		for (int t = 0; t < P.K; t++) { r |= ((tr >> t) & 1) << P.rb[t]; }
		trow[i] = (e < E) ? r : 0;
		lidx[i] = (e < E) ? lds_elem(tr, e & ((1 << P.logC) - 1), P) : 0;
	}
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

// per-lane LDS terms of a gate: gate application (lane: quad lane & 15,
// amplitude lane >> 4; FP32 stores: amplitudes 0..3) and outer-product steps
// (step t: quad 4 t + lane >> 4, amplitude of component lane & 15)
struct GateLanes
{
	int p, q;
	int ea;          // gate application, amplitude lane >> 4
	int ew[4];       // amplitudes 0..3 of quad lane & 15
	int eo[4];       // outer-product step t
};

__device__ __forceinline__ GateLanes gate_lanes(int p, int q, const TileArgs& P)
{
	const int lane = threadIdx.x & 63;
	GateLanes g;
	g.p = p;
	g.q = q;
#pragma unroll
	for (int v = 0; v < 4; v++) { g.ew[v] = quad_pos(lane & 15, v, p, q, P); }
	g.ea = quad_pos(lane & 15, lane >> 4, p, q, P);
	const int ra = ((lane & 15) >> 1) & 3;
#pragma unroll
	for (int t = 0; t < 4; t++) { g.eo[t] = quad_pos(4 * t + (lane >> 4), ra, p, q, P); }
	return g;
}

// apply a gate to kGroupStep groups of 16 quads in LDS (in place).  Group n
// starts at quad gb[n] (a multiple of 16; group term eg[n] = pos(gb[n], 0)); its
// second operand quads are the same quads of Y (dual: tangent state) or the 16
// quads after them in X (plain, group terms eg2[n]).  a[kk] = M[lane % 16][4 kk + lane / 16].
// All loads are issued before the matrix-core chains and all stores after, so
// the groups' chains overlap.
template <typename T>
__device__ __forceinline__ void mfma_gate_apply(cx<T>* X, cx<T>* Y, const int* eg, const int* eg2, const GateLanes& gl,
	const T* a)
{
	const int lane = threadIdx.x & 63;
	const int r = lane >> 4;
	int ex[kGroupStep], ey[kGroupStep];
	cx<T> x[kGroupStep], y[kGroupStep];
	typename Vec4<T>::type acc[kGroupStep];
#pragma unroll
	for (int n = 0; n < kGroupStep; n++) {
		ex[n] = eg[n] ^ gl.ea;
		ey[n] = eg2[n] ^ gl.ea;
		x[n] = X[ex[n]];
		y[n] = Y[ey[n]];
		acc[n] = typename Vec4<T>::type{0, 0, 0, 0};
	}
#pragma unroll
	for (int n = 0; n < kGroupStep; n++) { acc[n] = mfma_step(a[0], x[n].x, acc[n]); }
#pragma unroll
	for (int n = 0; n < kGroupStep; n++) { acc[n] = mfma_step(a[1], x[n].y, acc[n]); }
#pragma unroll
	for (int n = 0; n < kGroupStep; n++) { acc[n] = mfma_step(a[2], y[n].x, acc[n]); }
#pragma unroll
	for (int n = 0; n < kGroupStep; n++) { acc[n] = mfma_step(a[3], y[n].y, acc[n]); }
#pragma unroll
	for (int n = 0; n < kGroupStep; n++) {
		if constexpr (sizeof(T) == 8) {
			// register v = component v (x.re, x.im, y.re, y.im) of amplitude r
			X[ex[n]] = cx<T>{acc[n][0], acc[n][1]};
			Y[ey[n]] = cx<T>{acc[n][2], acc[n][3]};
		} else {
			// register v = component r (x.re, x.im, y.re, y.im) of amplitude v
			T* base = reinterpret_cast<T*>(r < 2 ? X : Y);
			const int e0 = (r < 2) ? eg[n] : eg2[n];
#pragma unroll
			for (int v = 0; v < 4; v++) { base[2 * (e0 ^ gl.ew[v]) + (r & 1)] = acc[n][v]; }
		}
	}
}

// load / store one tile state between HBM and LDS
template <typename T>
__device__ __forceinline__ void tile_load(cx<T>* S, const cx<T>* __restrict__ src, const int* trow, const int* lidx,
	long rbase, long c0, int E, int logC, long B)
{
	const int C = 1 << logC;
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		if (e >= E) { break; }
		S[lidx[i]] = src[(rbase | (long)trow[i]) * B + c0 + (e & (C - 1))];
	}
}

// (a uniform fill needs no swizzle)
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
__device__ __forceinline__ void tile_identity(cx<T>* S, const int* trow, const int* lidx, long rbase, long c0, long b0,
	int E, int logC)
{
	const int C = 1 << logC;
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		if (e >= E) { break; }
		const long row = rbase | (long)trow[i];
		S[lidx[i]] = cx<T>{(row == b0 + c0 + (e & (C - 1))) ? (T)1 : (T)0, (T)0};
	}
}

template <typename T>
__device__ __forceinline__ void tile_store(cx<T>* __restrict__ dst, const cx<T>* S, const int* trow, const int* lidx,
	long rbase, long c0, int E, int logC, long B)
{
	const int C = 1 << logC;
#pragma unroll
	for (int i = 0; i < kMaxEpt; i++) {
		const int e = threadIdx.x + i * kTileThreads;
		if (e >= E) { break; }
		dst[(rbase | (long)trow[i]) * B + c0 + (e & (C - 1))] = S[lidx[i]];
	}
}

// Wave w owns the quads [w qpw, (w + 1) qpw) of every gate, i.e. those whose two
// highest tile bits outside the gate equal w; consecutive gates with the same
// two bits need no barrier between them (TileArgs::barrier).

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
	const int pad = kLdsPadBytes / (int)sizeof(cx<T>);
	cx<T>* S = reinterpret_cast<cx<T>*>(lds_raw);
	cx<T>* D = S + E + pad;
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	constexpr int gq = DUAL ? 16 : 32;            // quads per matrix-core group
	const int qpw = (E / 4) / kTileWaves;         // quads owned by a wave
	const int q0 = wave * qpw;

	const long nchunks = B >> P.logC;
	int trow[kMaxEpt], lidx[kMaxEpt];
	thread_tile_elems(P, E, trow, lidx);
	for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const long chunk = tile % nchunks, rest = tile / nchunks;
		const long rbase = deposit_bits(rest, P.rmask);
		const long c0 = chunk << P.logC;
		if (flags & 1) { tile_identity(S, trow, lidx, rbase, c0, b0, E, P.logC); }
		else { tile_load(S, in, trow, lidx, rbase, c0, E, P.logC, B); }
		if (DUAL) {
			if (flags & 2) { tile_fill(D, E, cx<T>{0, 0}); }
			else { tile_load(D, din, trow, lidx, rbase, c0, E, P.logC, B); }
		}
		__syncthreads();
		for (int j = 0; j < P.ng; j++) {
			if (j > 0 && ((P.barrier >> j) & 1)) { __syncthreads(); }
			const int p = P.lp[j], q = P.lq[j];
			const GateLanes gl = gate_lanes(p, q, P);
			T a[4];
#pragma unroll
			for (int kk = 0; kk < 4; kk++) { a[kk] = amat[amat_index(P.x0 + j, 0, kk, lane)]; }
			for (int g = q0; g < q0 + qpw; g += kGroupStep * gq) {
				int eg[kGroupStep], eg2[kGroupStep];
#pragma unroll
				for (int n = 0; n < kGroupStep; n++) {
					eg[n] = quad_pos(g + n * gq, 0, p, q, P);
					eg2[n] = DUAL ? eg[n] : quad_pos(g + n * gq + 16, 0, p, q, P);
				}
				mfma_gate_apply<T>(S, DUAL ? D : S, eg, eg2, gl, a);
			}
		}
		__syncthreads();
		tile_store(out, S, trow, lidx, rbase, c0, E, P.logC, B);
		if (DUAL) { tile_store(dout, D, trow, lidx, rbase, c0, E, P.logC, B); }
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
__global__ __launch_bounds__(kTileThreads, 2) void k_tile_reverse(
	const cx<T>* __restrict__ psi_out, const cx<T>* __restrict__ dpsi_out,
	const cx<T>* __restrict__ phi_out, const cx<T>* __restrict__ dphi_out,
	cx<T>* __restrict__ phi_in, cx<T>* __restrict__ dphi_in,
	const TileArgs P, const T* __restrict__ amat, long B, long ntiles, int flags, double* __restrict__ partial)
{
	extern __shared__ __align__(16) char lds_raw[];
	const int E = 1 << (P.K + P.logC);
	const int pad = kLdsPadBytes / (int)sizeof(cx<T>);
	cx<T>* S = reinterpret_cast<cx<T>*>(lds_raw);
	cx<T>* D = S + E + pad;
	cx<T>* F = S + (DUAL ? 2 : 1) * (E + pad);
	cx<T>* DF = F + E + pad;
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	constexpr int gq = DUAL ? 16 : 32;
	const int qpw = (E / 4) / kTileWaves;
	const int q0 = wave * qpw;
	const int m = lane & 15;
	// outer-product operands: component m = 8 s + 2 r + p of the quad (dual: s
	// selects the primary / tangent array; plain: the quad or the one 16 after it)
	const T* Fa = reinterpret_cast<const T*>((DUAL && (m >> 3)) ? DF : F);
	const T* Sa = reinterpret_cast<const T*>((DUAL && (m >> 3)) ? D : S);
	const int osel = DUAL ? 0 : (m >> 3);

	double4_t acc[kMaxTileGates];
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) { acc[j] = double4_t{0, 0, 0, 0}; }

	const long nchunks = B >> P.logC;
	int trow[kMaxEpt], lidx[kMaxEpt];
	thread_tile_elems(P, E, trow, lidx);
	for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const long chunk = tile % nchunks, rest = tile / nchunks;
		const long rbase = deposit_bits(rest, P.rmask);
		const long c0 = chunk << P.logC;
		tile_load(S, psi_out, trow, lidx, rbase, c0, E, P.logC, B);
		tile_load(F, phi_out, trow, lidx, rbase, c0, E, P.logC, B);
		if (DUAL) {
			tile_load(D, dpsi_out, trow, lidx, rbase, c0, E, P.logC, B);
			if (flags & 4) { tile_fill(DF, E, cx<T>{0, 0}); }
			else { tile_load(DF, dphi_out, trow, lidx, rbase, c0, E, P.logC, B); }
		}
		__syncthreads();
		for (int j = P.ng - 1; j >= 0; j--) {
			{
				if (j + 1 < P.ng && ((P.barrier >> (j + 1)) & 1)) { __syncthreads(); }
				const int p = P.lp[j], q = P.lq[j];
				const GateLanes gl = gate_lanes(p, q, P);
				T au[4], aa[4];
#pragma unroll
				for (int kk = 0; kk < 4; kk++) {
					au[kk] = amat[amat_index(P.x0 + j, 1, kk, lane)];
					aa[kk] = amat[amat_index(P.x0 + j, 2, kk, lane)];
				}
				const bool adjoint = (j > 0) || !(flags & 8);
				double4_t gsum = {0, 0, 0, 0};
				for (int g = q0; g < q0 + qpw; g += kGroupStep * gq) {
					int eg[kGroupStep], eg2[kGroupStep];
#pragma unroll
					for (int n = 0; n < kGroupStep; n++) {
						eg[n] = quad_pos(g + n * gq, 0, p, q, P);
						eg2[n] = DUAL ? eg[n] : quad_pos(g + n * gq + 16, 0, p, q, P);
					}
					// states before gate j
					mfma_gate_apply<T>(S, DUAL ? D : S, eg, eg2, gl, au);
					// outer products over the groups' quads, 4 quads per step
					typename Vec4<T>::type grp[kGroupStep];
					T av[kGroupStep][4], bv[kGroupStep][4];
#pragma unroll
					for (int n = 0; n < kGroupStep; n++) {
						grp[n] = typename Vec4<T>::type{0, 0, 0, 0};
						const int eb = osel ? eg2[n] : eg[n];
#pragma unroll
						for (int t = 0; t < 4; t++) {
							const int el = 2 * (eb ^ gl.eo[t]) + (m & 1);
							av[n][t] = Fa[el];
							bv[n][t] = Sa[el];
						}
					}
#pragma unroll
					for (int t = 0; t < 4; t++) {
#pragma unroll
						for (int n = 0; n < kGroupStep; n++) { grp[n] = mfma_step(av[n][t], bv[n][t], grp[n]); }
					}
#pragma unroll
					for (int n = 0; n < kGroupStep; n++) {
#pragma unroll
						for (int v = 0; v < 4; v++) { gsum[v] += (double)grp[n][v]; }
					}
					// adjoints at the gate input
					if (adjoint) { mfma_gate_apply<T>(F, DUAL ? DF : F, eg, eg2, gl, aa); }
				}
				// accumulator of gate j (constant register indices in every case)
				static_assert(kMaxTileGates == 8, "one case per gate of a pass");
				switch (j) {
				case 0: acc[0] += gsum; break;
				case 1: acc[1] += gsum; break;
				case 2: acc[2] += gsum; break;
				case 3: acc[3] += gsum; break;
				case 4: acc[4] += gsum; break;
				case 5: acc[5] += gsum; break;
				case 6: acc[6] += gsum; break;
				default: acc[7] += gsum; break;
				}
			}
		}
		__syncthreads();
		if (!(flags & 8)) {
			tile_store(phi_in, F, trow, lidx, rbase, c0, E, P.logC, B);
			if (DUAL) { tile_store(dphi_in, DF, trow, lidx, rbase, c0, E, P.logC, B); }
		}
		__syncthreads();
	}
	const long part = (long)blockIdx.x * kTileWaves + (threadIdx.x >> 6);
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
