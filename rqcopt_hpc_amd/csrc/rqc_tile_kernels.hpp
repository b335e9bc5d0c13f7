// Tile-pass kernels (gfx950): several gates per HBM pass, gates applied on the
// matrix cores.
//
// A "pass" is a set of gates, taken in an order consistent with the circuit
// (gates on disjoint wires commute), whose wires all lie in a set Q of K
// "tile" qubits.  A tile is the 2^K x C sub-array of a column block
// X[a * B + c] (column fastest) obtained by fixing the other L - K row bits
// ("rest") and a chunk of C consecutive columns (at the planned tiles, rows of
// 256 bytes).  A workgroup loads a tile into LDS, applies all gates of the pass
// there and writes it back: one HBM read and one write per pass instead of per
// gate.
//
// Gate application on the matrix cores.  A dual quad (psi and dpsi on the four
// amplitudes of a gate) has 16 real components; the update
//   psi' = G psi,  dpsi' = G dpsi + Z psi
// is a real 16 x 16 matrix with complex blocks [[G, 0], [Z, G]].  Lane l holds
// amplitude l >> 4 of quad l & 15 of a 16-quad group.
// - FP64: 4 x 4 block products (v_mfma_f64_4x4x4_4b_f64, see gate_mfma44): each
//   complex 4 x 4 gate is 4 real block products, the dual matrix 12, with no
//   zero block; plain sweeps (no tangent state) apply G to two groups of quads.
// - FP32: the real 16 x 16 form, four chained 16x16x4 steps
//     D = sum_kk M[:, 4kk:4kk+4] X[4kk:4kk+4, :],
//   lane l's A operand M[l % 16][4 kk + l / 16] read from the pass's gate
//   matrices in LDS through a per-lane map fixed for the launch; plain sweeps
//   take M = [[G, 0], [0, G]] on two quads.
// Each state element is read and written once per gate.
//
// Reverse pass (k_tile_reverse): reloads the pass-output checkpoint and walks
// the pass's gates backwards, recovering the state before each gate with the
// inverse update (uncompute: psi = G^-1 psi', dpsi = G^-1 dpsi' - G^-1 Z G^-1 psi'),
// so that LDS holds only psi, dpsi, phi, dphi whatever the number of gates:
//   gradient / HVP sums  grad_j += phi (x) psi_j,  hess_j += phi (x) dpsi_j + dphi (x) psi_j
//   adjoint update       phi <- G_j^T phi,  dphi <- G_j^T dphi + Z_j^T phi
// (phi, dphi: the adjoint at the gate's output).  The engine uses it only for
// gates that are unitary to 1e-6 (then G^-1 has norm 1 and the recovery adds
// rounding only).  Both updates run transposed (state as the A operand), so
// that their results are the operands of the outer products, which thus pair
// the adjoint at the gate's input with the recovered state; the reduction maps
// these sums back to the output adjoint with G^-T and (-G^-1 Z G^-1)^T.  The
// outer-product sums run on the matrix cores: FP64 as 4 x 4 block products into
// four FP64 accumulators per gate, FP32 as 16x16x4 steps chained in FP32 over a
// gate's quads of the tile and added into FP64 accumulators.  Partial sums per
// (workgroup, wave, gate) are reduced in a fixed order by k_tile_reduce_chunks
// and k_tile_reduce: no atomics, bitwise reproducible.
#pragma once

#include <cstdint>
#include <type_traits>

#include "zc.hpp"

namespace rqc {

constexpr int kTileThreads = 256;
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kMaxTileGates = 8;
// LDS states of a tile are E elements + 64 bytes apart, so that equal positions of
// two states fall on different banks
constexpr int kLdsPadBytes = 64;
// The reverse pass stores a quad's transposed results into two states with
// one instruction (8-byte stores: 16 lanes, 4-byte stores: 32 lanes, over a
// 128-byte bank window); the 64-byte pad puts the second state's amplitudes on
// the other half of the window.
template <typename T>
constexpr int lds_pad_rev() { return 64; }

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// LDS layout of a tile: element (row, col) at (row << logC) | (col ^ f(row)),
// with f the XOR of per-tile-bit column masks h[t] over the set bits of row
// (bit i of f = parity(row & swz[i])).  The host picks h per pass (a colouring
// of the pass's gate graph) so that the four amplitude rows of every gate's
// quads fall on different banks.
// Wave-uniform LDS position terms of a gate (see quad_pos): pos(0, r) of the
// amplitudes, pos(4 t, 0) and pos(k, 0) (t, k < 4) of the quads within a
// 16-quad group, pos(w qpw, 0) of the waves' quad ranges and pos(16 * 2^b, 0)
// of the groups within a range (pos(16 i, 0) is the XOR over the set bits b of i).  Computed by the host per gate in execution order
// (a read-only device table read with scalar loads), so that the gate loop
// forms every position with a few selects and XORs.
struct GateTerms
{
	int u[4], w[4], q4[4], wv[kTileWaves], gbit[3];
};

struct TileArgs
{
	int K;             // tile row bits
	int logC;          // log2 of tile columns
	int ng;            // gates in this pass
	int x0;            // position of the pass's first gate in the matrix table
	int lognc;         // log2 of the column chunks of the block (B >> logC)
	unsigned rmask;    // row bits outside the tile (rest), deposited lowest first
	unsigned swz[5];   // column swizzle masks over tile rows (bit i of the column term)
	unsigned barrier;  // bit j: the waves own different quads for gates j - 1 and j
	int wide;          // FP32 dual reverse pass: four matrix-core groups per step
	int rb[16];        // row bit of tile bit t, t < K
};

// Gate matrices of the sweep in execution order: gate x has kMats complex 4x4
// matrices at mats[(x * kMats + i) * 16], i = G, Z, G^-1, -G^-1 Z G^-1, G^T, Z^T.
// A kernel copies its pass's matrices into LDS once; each lane then forms its
// matrix-core A operands from them (no global loads inside the pipelined tile
// loop, whose asynchronous tile loads complete in issue order with any other
// vector-memory access).
constexpr int kMats = 6;
enum { kMatG = 0, kMatZ = 1, kMatGinv = 2, kMatW = 3, kMatGT = 4, kMatZT = 5 };

// LDS layout of a pass's gate matrices: each complex 4x4 matrix is stored
// transposed (entry (r, c) at 4 c + r), and the second matrix Y of a block pair
// [[X, 0], [Y, X]] sits kMatPairOff entries after X.  A 16-lane phase of an
// operand read takes one column of X and of Y (re and im): 8 consecutive doubles
// of X and 8 of Y, which the offset puts on the other half of a 128-byte bank
// window.  Per gate: forward X, [Y]; reverse X_rec, [Y_rec], X_adj, [Y_adj].
constexpr int kMatPairOff = 20;
template <bool DUAL> constexpr int fwd_mat_stride() { return DUAL ? 2 * kMatPairOff : 16; }
template <bool DUAL> constexpr int rev_mat_stride() { return DUAL ? 4 * kMatPairOff : 2 * kMatPairOff; }
// after the reverse pass's gate matrices: a running trace sum (two doubles) per wave
constexpr int kRevTraceBytes = kTileWaves * 16;
// after the forward pass's gate matrices: one word per wave (pass 0's zero-tile vote)
constexpr int kFwdVoteBytes = kTileWaves * 4;

// A operand of step kk of the real 16 x 16 matrix with complex blocks
// [[X, 0], [Y, X]] (Y = 0 without the tangent state), components m = 4 v + r,
// v = 2 s + p: lane l takes M[l % 16][4 kk + l / 16], i.e. element aoff[kk] (in
// T units from X, in the layout above) times afac[kk] (the sign of -Im above
// the diagonal, 0 in the zero block).  Fixed per lane for the kernel.
// AMP_MAJOR orders the rows m = 4 r + v instead: the FP32 result rows of lane
// l, 4 (l >> 4) + v, are then the four components of one amplitude, which
// the forward pass stores as two 8-byte elements (as in FP64, whose result
// rows (l >> 4) + 4 v are that amplitude's components with the default order).
template <typename T, bool DUAL>
struct AOperandMap
{
	int aoff[4];
	T afac[4];
};

template <typename T, bool DUAL, bool AMP_MAJOR = false>
__device__ __forceinline__ AOperandMap<T, DUAL> a_operand_map()
{
	const int lane = threadIdx.x & 63;
	const int i = lane & 15, ri = lane >> 4;
	const int vo = AMP_MAJOR ? (i & 3) : (i >> 2), ro = AMP_MAJOR ? (i >> 2) : (i & 3);
	const int so = vo >> 1, po = vo & 1;
	AOperandMap<T, DUAL> m;
#pragma unroll
	for (int kk = 0; kk < 4; kk++) {
		const int si = kk >> 1, pi = kk & 1;
		const int blk = (DUAL && so != si) ? 2 * kMatPairOff : 0;   // Y of [[X, 0], [Y, X]]
		m.aoff[kk] = blk + 2 * (4 * ri + ro) + (po == pi ? 0 : 1);
		const int keep = (so == si) | (DUAL & (so == 1));
		m.afac[kk] = (T)((1 - 2 * (po < pi)) * keep);
	}
	return m;
}

// copy the pass's gate matrices (those listed in `which`) into LDS, transposed,
// matrix w of gate j at lm + j * stride + w * kMatPairOff
template <typename T>
__device__ __forceinline__ void load_pass_mats(const cx<T>* __restrict__ mats, int x0, int ng, const int* which, int nw,
	int stride, cx<T>* lm)
{
	for (int e = threadIdx.x; e < ng * nw * 16; e += kTileThreads) {
		const int j = e / (nw * 16), w = (e / 16) % nw, t = e & 15;
		lm[j * stride + w * kMatPairOff + 4 * (t & 3) + (t >> 2)] = mats[((long)(x0 + j) * kMats + which[w]) * 16 + t];
	}
}

// zero bits inserted at tile bit positions lo < hi
__host__ __device__ __forceinline__ int tile_quad_row(int q, int lo, int hi)
{
	int x = ((q >> lo) << (lo + 1)) | (q & ((1 << lo) - 1));
	x = ((x >> hi) << (hi + 1)) | (x & ((1 << hi) - 1));
	return x;
}

__host__ __device__ __forceinline__ int swz_f(int row, const TileArgs& P)
{
	return (__builtin_popcount(row & P.swz[0]) & 1) | ((__builtin_popcount(row & P.swz[1]) & 1) << 1) |
	       ((__builtin_popcount(row & P.swz[2]) & 1) << 2) | ((__builtin_popcount(row & P.swz[3]) & 1) << 3) |
	       ((__builtin_popcount(row & P.swz[4]) & 1) << 4);
}

__host__ __device__ __forceinline__ int lds_elem(int row, int col, const TileArgs& P)
{
	return (row << P.logC) | (col ^ swz_f(row, P));
}

// Quad qd of the gate on tile bits (p, q): compressed row qd >> logC (the row
// without bits p, q), column qd mod C.  LDS position of its amplitude
// r = 2 qi + qj (bit p = qi, bit q = qj).  The position is XOR-linear in qd
// for disjoint bit parts, so pos(gbase | j, r) = pos(gbase, 0) ^ pos(j, r): a
// wave-uniform group term and a per-lane term fixed for the gate.
__host__ __device__ __forceinline__ int quad_pos(int qd, int r, int p, int q, const TileArgs& P)
{
	const int row = tile_quad_row(qd >> P.logC, min(p, q), max(p, q)) | ((r >> 1) << p) | ((r & 1) << q);
	return lds_elem(row, qd & ((1 << P.logC) - 1), P);
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
// amplitude lane >> 4, for its loads and stores) and the transposed
// application's results (register v: amplitude lane & 3 of quad
// (lane >> 4) + 4 v in FP64, 4 (lane >> 4) + v in FP32)
struct GateLanes
{
	int ea;          // gate application, amplitude lane >> 4
	int et[4];       // transposed results, register v
};

// (by linearity: pos(j, r) = pos(j, 0) ^ pos(0, r), pos(4 t + k, 0) =
// pos(4 t, 0) ^ pos(k, 0))
// v[i] for a per-lane i < 4 of four wave-uniform values, as selects (the values
// are passed by value so that they stay in scalar registers)
__device__ __forceinline__ int sel4(int v0, int v1, int v2, int v3, int i)
{
	const int lo = (i & 1) ? v1 : v0;
	const int hi = (i & 1) ? v3 : v2;
	return (i & 2) ? hi : lo;
}

__device__ __forceinline__ int sel4(const int (&v)[4], int i) { return sel4(v[0], v[1], v[2], v[3], i); }

// position term of the groups starting at quad 16 i of a wave's range
__device__ __forceinline__ int group_term(const GateTerms& t, int i)
{
	return ((i & 1) ? t.gbit[0] : 0) ^ ((i & 2) ? t.gbit[1] : 0) ^ ((i & 4) ? t.gbit[2] : 0);
}

// v[i] for a per-lane i < 4 of four wave-uniform terms that are XOR-affine in
// i (v[3] = v[0] ^ v[1] ^ v[2], checked when the terms are made): masks of the
// two index bits select the differences, no per-value selects
__device__ __forceinline__ int sel4_affine(const int (&v)[4], int i)
{
	const int m1 = -(i & 1), m2 = -((i >> 1) & 1);
	return v[0] ^ ((v[1] ^ v[0]) & m1) ^ ((v[2] ^ v[0]) & m2);
}

// transposed results: et[v] = bt ^ tv(v), the lane's term bt and the register's
// wave-uniform term tv(v) (quad q of a group: position q4[q & 3] ^ w[(q >> 2) & 3])
template <typename T>
__device__ __forceinline__ int transposed_lane_term(const GateTerms& t)
{
	const int lane = threadIdx.x & 63;
	return (sizeof(T) == 8 ? sel4_affine(t.q4, lane >> 4) : sel4_affine(t.w, lane >> 4)) ^ sel4_affine(t.u, lane & 3);
}

template <typename T>
__device__ __forceinline__ int transposed_reg_term(const GateTerms& t, int v)
{
	return (sizeof(T) == 8) ? t.w[v] : t.q4[v];
}

template <typename T = double>
__device__ __forceinline__ GateLanes gate_lanes(const GateTerms& t)
{
	const int lane = threadIdx.x & 63;
	GateLanes g;
	const int base = sel4_affine(t.q4, lane & 3) ^ sel4_affine(t.w, (lane >> 2) & 3);
	g.ea = base ^ sel4_affine(t.u, lane >> 4);
	const int bt = transposed_lane_term<T>(t);
#pragma unroll
	for (int v = 0; v < 4; v++) { g.et[v] = bt ^ transposed_reg_term<T>(t, v); }
	return g;
}

// The three parts of a gate application on GS groups, so that independent
// applications (the state and the adjoint of the reverse pass) can interleave.
template <typename T, int GS>
struct GateOperands
{
	cx<T> x[GS], y[GS];
	typename Vec4<T>::type acc[GS];
};

template <typename T, int GS>
__device__ __forceinline__ void gate_load(const cx<T>* X, const cx<T>* Y, const int* eg, const int* eg2,
	const GateLanes& gl, GateOperands<T, GS>& o)
{
#pragma unroll
	for (int n = 0; n < GS; n++) {
		o.x[n] = X[eg[n] ^ gl.ea];
		o.y[n] = Y[eg2[n] ^ gl.ea];
	}
}

template <typename T, int GS>
__device__ __forceinline__ void gate_mfma(const T* a, GateOperands<T, GS>& o)
{
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(a[0], o.x[n].x, typename Vec4<T>::type{0, 0, 0, 0}); }
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(a[1], o.x[n].y, o.acc[n]); }
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(a[2], o.y[n].x, o.acc[n]); }
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(a[3], o.y[n].y, o.acc[n]); }
}

template <typename T, int GS>
__device__ __forceinline__ void gate_store(cx<T>* X, cx<T>* Y, const int* eg, const int* eg2, const GateLanes& gl,
	const GateOperands<T, GS>& o)
{
#pragma unroll
	for (int n = 0; n < GS; n++) {
		// register v = component v (x.re, x.im, y.re, y.im) of amplitude lane >> 4
		// (FP64 rows (l >> 4) + 4 v; FP32 rows 4 (l >> 4) + v with the amplitude-major operand map)
		X[eg[n] ^ gl.ea] = cx<T>{o.acc[n][0], o.acc[n][1]};
		Y[eg2[n] ^ gl.ea] = cx<T>{o.acc[n][2], o.acc[n][3]};
	}
}

// Transposed application D^T = S^T M^T: the same operand registers with the
// state as the A operand and the gate as the B operand.  The result holds, in
// lane l, component l & 15 (4 s + r: s = x.re, x.im, y.re, y.im, r the
// amplitude) of quad (l >> 4) + 4 v (FP64) or 4 (l >> 4) + v (FP32) in register
// v, which is the operand layout of an outer-product step over those quads.
template <typename T, int GS>
__device__ __forceinline__ void gate_mfma_t(const T* a, GateOperands<T, GS>& o)
{
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(o.x[n].x, a[0], typename Vec4<T>::type{0, 0, 0, 0}); }
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(o.x[n].y, a[1], o.acc[n]); }
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(o.y[n].x, a[2], o.acc[n]); }
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = mfma_step(o.y[n].y, a[3], o.acc[n]); }
}

template <typename T, int GS>
__device__ __forceinline__ void gate_store_t(cx<T>* X, cx<T>* Y, const int* eg, const int* eg2, const GateLanes& gl,
	const GateOperands<T, GS>& o)
{
	const int sc = (threadIdx.x & 15) >> 2;   // x.re, x.im, y.re, y.im
#pragma unroll
	for (int n = 0; n < GS; n++) {
		T* base = reinterpret_cast<T*>(sc < 2 ? X : Y);
		const int e0 = (sc < 2) ? eg[n] : eg2[n];
#pragma unroll
		for (int v = 0; v < 4; v++) { base[2 * (e0 ^ gl.et[v]) + (sc & 1)] = o.acc[n][v]; }
	}
}

// ---------------------------------------------------------------------------
// FP64 gate applications on 4 x 4 blocks.  The 4x4x4 FP64 matrix-core
// instruction computes four independent 4 x 4 x 4 products, block b from lanes
// 16 k + 4 b + i (A row i, column k), 16 k + 4 b + j (B row k, column j) into
// lane 16 i + 4 b + j, at the 16x16x4 instruction's rate per product.  A
// complex 4 x 4 gate on the states' real and imaginary parts is then 4 such
// products, and the dual matrix [[X, 0], [Y, X]] 12, where the real 16 x 16 form
// costs the time of 16 (its zero block included); the outer products likewise
// need 12 instead of 16.  Signs come from the instruction's operand negation.
__device__ __forceinline__ double mfma44(double a, double b, double c)
{
	return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// -a b + c (the last argument of the FP64 instruction negates its operands: 1 = A)
__device__ __forceinline__ double mfma44_na(double a, double b, double c)
{
	return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 1);
}

// Entry (l & 3, l >> 4) of a block pair [[X, 0], [Y, X]] held by lane l (the
// pair's matrices in the transposed LDS layout, entry (r, c) at 4 c + r): the A
// operand of the direct application and the B operand of the transposed one.  Real and
// imaginary parts, and for the dual sweeps the sums xa = Re X + Im X, xb = Im X - Re X (ya,
// yb of Y) of the three-product form below.
struct Frag44
{
	double xr, xi, yr, yi;
	double xa, xb, ya, yb;
};

__device__ __forceinline__ Frag44 frag44(const cx<double>* m)
{
	const int lane = threadIdx.x & 63;
	const int fi = 4 * (lane >> 4) + (lane & 3);
	const cx<double> x = m[fi], y = m[fi + kMatPairOff];
	return Frag44{x.x, x.y, y.x, y.y, x.x + x.y, x.y - x.x, y.x + y.y, y.y - y.x};
}

// The dual sweeps' complex block products take three real products each (Gauss's form):
// for M = Mr + i Mi and s = sr + i si,
//   t = Mr (sr + si),  Re(M s) = t - (Mr + Mi) si,  Im(M s) = t + (Mi - Mr) sr,
// and the same with the operands' order swapped (the transposed applications).  A dual gate
// pair, x' = X x and y' = Y x + X y, is 9 block products instead of 12; the sums sr + si cost
// one vector add per state and group.  (The outer-product sums below keep four products: their
// long sums then round as the direct complex products do.)

// Direct application to GS groups of 16 quads in the gate layout (lane l:
// amplitude l >> 4 of quad l & 15; x = primary, y = tangent state): the results
// land in the same lanes, acc[n] = (x.re, x.im, y.re, y.im).
template <int GS>
__device__ __forceinline__ void gate_mfma44(const Frag44& f, GateOperands<double, GS>& o)
{
	double t[GS][2], d[GS][4];
#pragma unroll
	for (int n = 0; n < GS; n++) {
		const double xs = o.x[n].x + o.x[n].y;
		t[n][0] = mfma44(f.xr, xs, 0.0);
		t[n][1] = mfma44(f.yr, xs, 0.0);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		const double ys = o.y[n].x + o.y[n].y;
		t[n][1] = mfma44(f.xr, ys, t[n][1]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][0] = mfma44_na(f.xa, o.x[n].y, t[n][0]);
		d[n][1] = mfma44(f.xb, o.x[n].x, t[n][0]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][2] = mfma44_na(f.ya, o.x[n].y, t[n][1]);
		d[n][3] = mfma44(f.yb, o.x[n].x, t[n][1]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][2] = mfma44_na(f.xa, o.y[n].y, d[n][2]);
		d[n][3] = mfma44(f.xb, o.y[n].x, d[n][3]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = double4_t{d[n][0], d[n][1], d[n][2], d[n][3]}; }
}

// Transposed application (the state as the A operand, the gate as B): the
// results of lane l are amplitude l & 3 of quad 4 ((l >> 2) & 3) + (l >> 4) of
// the group, which is the operand layout of the outer products below.
template <int GS>
__device__ __forceinline__ void gate_mfma44_t(const Frag44& f, GateOperands<double, GS>& o)
{
	double t[GS][2], d[GS][4];
#pragma unroll
	for (int n = 0; n < GS; n++) {
		const double xs = o.x[n].x + o.x[n].y;
		t[n][0] = mfma44(xs, f.xr, 0.0);
		t[n][1] = mfma44(xs, f.yr, 0.0);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		const double ys = o.y[n].x + o.y[n].y;
		t[n][1] = mfma44(ys, f.xr, t[n][1]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][0] = mfma44_na(o.x[n].y, f.xa, t[n][0]);
		d[n][1] = mfma44(o.x[n].x, f.xb, t[n][0]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][2] = mfma44_na(o.x[n].y, f.ya, t[n][1]);
		d[n][3] = mfma44(o.x[n].x, f.yb, t[n][1]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][2] = mfma44_na(o.y[n].y, f.xa, d[n][2]);
		d[n][3] = mfma44(o.y[n].x, f.xb, d[n][3]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = double4_t{d[n][0], d[n][1], d[n][2], d[n][3]}; }
}

// Outer-product sums of GS groups' transposed results (a: adjoint, s: state)
// onto acc = (A1.re, A1.im, A2.re, A2.im), A1 = sum a (x) s and
// A2 = sum a (x) ds + da (x) s (complex, no conjugation): lane 16 r + 4 b + c
// holds entry (r, c) summed over the quads of block b.
template <int GS>
__device__ __forceinline__ void outer44(const GateOperands<double, GS>& s, const GateOperands<double, GS>& a,
	double4_t& acc)
{
	double c0 = acc[0], c1 = acc[1], c2 = acc[2], c3 = acc[3];
#pragma unroll
	for (int n = 0; n < GS; n++) {
		c0 = mfma44(a.acc[n][0], s.acc[n][0], c0);
		c1 = mfma44(a.acc[n][0], s.acc[n][1], c1);
		c2 = mfma44(a.acc[n][0], s.acc[n][2], c2);
		c3 = mfma44(a.acc[n][0], s.acc[n][3], c3);
		c0 = mfma44_na(a.acc[n][1], s.acc[n][1], c0);
		c1 = mfma44(a.acc[n][1], s.acc[n][0], c1);
		c2 = mfma44_na(a.acc[n][1], s.acc[n][3], c2);
		c3 = mfma44(a.acc[n][1], s.acc[n][2], c3);
		c2 = mfma44(a.acc[n][2], s.acc[n][0], c2);
		c3 = mfma44(a.acc[n][2], s.acc[n][1], c3);
		c2 = mfma44_na(a.acc[n][3], s.acc[n][1], c2);
		c3 = mfma44(a.acc[n][3], s.acc[n][0], c3);
	}
	acc = double4_t{c0, c1, c2, c3};
}

// Plain sweeps (no tangent state): x and y are two groups of quads, both
// multiplied by the block pair's X, in the three-product form (6 block products
// instead of the real 16 x 16 form's 16); Frag44 from frag44x, whose y part is unused.
__device__ __forceinline__ Frag44 frag44x(const cx<double>* m)
{
	const int lane = threadIdx.x & 63;
	const cx<double> x = m[4 * (lane >> 4) + (lane & 3)];
	return Frag44{x.x, x.y, 0.0, 0.0, x.x + x.y, x.y - x.x, 0.0, 0.0};
}

template <int GS>
__device__ __forceinline__ void gate_mfma44_pair(const Frag44& f, GateOperands<double, GS>& o)
{
	double t[GS][2], d[GS][4];
#pragma unroll
	for (int n = 0; n < GS; n++) {
		t[n][0] = mfma44(f.xr, o.x[n].x + o.x[n].y, 0.0);
		t[n][1] = mfma44(f.xr, o.y[n].x + o.y[n].y, 0.0);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][0] = mfma44_na(f.xa, o.x[n].y, t[n][0]);
		d[n][1] = mfma44(f.xb, o.x[n].x, t[n][0]);
		d[n][2] = mfma44_na(f.xa, o.y[n].y, t[n][1]);
		d[n][3] = mfma44(f.xb, o.y[n].x, t[n][1]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = double4_t{d[n][0], d[n][1], d[n][2], d[n][3]}; }
}

template <int GS>
__device__ __forceinline__ void gate_mfma44_pair_t(const Frag44& f, GateOperands<double, GS>& o)
{
	double t[GS][2], d[GS][4];
#pragma unroll
	for (int n = 0; n < GS; n++) {
		t[n][0] = mfma44(o.x[n].x + o.x[n].y, f.xr, 0.0);
		t[n][1] = mfma44(o.y[n].x + o.y[n].y, f.xr, 0.0);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) {
		d[n][0] = mfma44_na(o.x[n].y, f.xa, t[n][0]);
		d[n][1] = mfma44(o.x[n].x, f.xb, t[n][0]);
		d[n][2] = mfma44_na(o.y[n].y, f.xa, t[n][1]);
		d[n][3] = mfma44(o.y[n].x, f.xb, t[n][1]);
	}
#pragma unroll
	for (int n = 0; n < GS; n++) { o.acc[n] = double4_t{d[n][0], d[n][1], d[n][2], d[n][3]}; }
}

// outer-product sums of the plain sweeps: acc = (A1.re, A1.im, -, -), A1 summed
// over both groups of quads
template <int GS>
__device__ __forceinline__ void outer44_pair(const GateOperands<double, GS>& s, const GateOperands<double, GS>& a,
	double4_t& acc)
{
	double c0 = acc[0], c1 = acc[1];
#pragma unroll
	for (int n = 0; n < GS; n++) {
		c0 = mfma44(a.acc[n][0], s.acc[n][0], c0);
		c1 = mfma44(a.acc[n][0], s.acc[n][1], c1);
		c0 = mfma44_na(a.acc[n][1], s.acc[n][1], c0);
		c1 = mfma44(a.acc[n][1], s.acc[n][0], c1);
		c0 = mfma44(a.acc[n][2], s.acc[n][2], c0);
		c1 = mfma44(a.acc[n][2], s.acc[n][3], c1);
		c0 = mfma44_na(a.acc[n][3], s.acc[n][3], c0);
		c1 = mfma44(a.acc[n][3], s.acc[n][2], c1);
	}
	acc = double4_t{c0, c1, acc[2], acc[3]};
}

// LDS position term of lane l's transposed result (amplitude l & 3 of group
// quad 4 ((l >> 2) & 3) + (l >> 4))
__device__ __forceinline__ int transposed44_lane_term(const GateTerms& t)
{
	const int lane = threadIdx.x & 63;
	return sel4_affine(t.q4, lane >> 4) ^ sel4_affine(t.w, (lane >> 2) & 3) ^ sel4_affine(t.u, lane & 3);
}

// LDS accesses by byte offset into the kernel's dynamic LDS.  The tile kernels
// declare no static LDS, so their dynamic LDS starts at offset 0 (the host
// checks this when it plans the passes) and an offset is the LDS address: the
// state strides fold into the instructions' offset fields and no base is added.
// (a 16-byte vector type, so that each access is one 16-byte instruction)
typedef double lds_d2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) lds_d2 lds_d2_as3;

__device__ __forceinline__ cx<double> lds_load_cxd(unsigned off)
{
	const lds_d2 v = *reinterpret_cast<const lds_d2_as3*>(off);
	return cx<double>{v.x, v.y};
}

__device__ __forceinline__ void lds_store_cxd(unsigned off, cx<double> v)
{
	*reinterpret_cast<lds_d2_as3*>(off) = lds_d2{v.x, v.y};
}

// FP64 on 4 x 4 blocks (DUAL: the pair [[G, 0], [Z, G]] on (psi, dpsi); plain: G
// on two groups of quads)
template <int GS, bool DUAL>
__device__ __forceinline__ void mfma_gate_apply44(cx<double>* X, cx<double>* Y, const int* eg, const int* eg2,
	const GateLanes& gl, const Frag44& f)
{
	GateOperands<double, GS> o;
	gate_load<double, GS>(X, Y, eg, eg2, gl, o);
	__builtin_amdgcn_s_setprio(1);
	if constexpr (DUAL) { gate_mfma44<GS>(f, o); }
	else { gate_mfma44_pair<GS>(f, o); }
	__builtin_amdgcn_s_setprio(0);
	gate_store<double, GS>(X, Y, eg, eg2, gl, o);
}

// all loads are issued before the matrix-core chains and all stores after, so
// the groups' chains overlap
template <typename T, int GS>
__device__ __forceinline__ void mfma_gate_apply(cx<T>* X, cx<T>* Y, const int* eg, const int* eg2, const GateLanes& gl,
	const T* a)
{
	GateOperands<T, GS> o;
	gate_load<T, GS>(X, Y, eg, eg2, gl, o);
	// matrix-core work at raised wave priority, so that the other waves' LDS
	// traffic issues in its shadow
	__builtin_amdgcn_s_setprio(1);
	gate_mfma<T, GS>(a, o);
	__builtin_amdgcn_s_setprio(0);
	gate_store<T, GS>(X, Y, eg, eg2, gl, o);
}

// ---------------------------------------------------------------------------
// HBM <-> LDS.  A tile state is E * sizeof(cx<T>) / 16 slots of 16 bytes (one
// FP64 or two FP32 elements; the FP32 swizzle keeps the two elements of a slot
// in adjacent columns).  Thread tid handles slots tid + k * 256.  Loads go
// straight from HBM into LDS (global_load_lds_dwordx4: lane-contiguous slots
// from the wave's base), so the next tile streams in while the current one is
// computed; the lane picks the global element whose LDS position is its slot.
constexpr int kMaxSlots = 4;   // slots per thread and state (E * sizeof(cx<T>) <= 16 KiB, checked by the host)

template <typename T>
struct TileIO
{
	static constexpr int epl = 16 / (int)sizeof(cx<T>);   // elements per slot
	long goff[kMaxSlots];   // global element offset (from rbase * B + c0) of the slot's first element
	int trow[kMaxSlots];    // row-index contribution of the tile bits
	int col[kMaxSlots];     // column of the slot's first element
	int nk;                 // slots per thread
};

template <typename T>
__device__ __forceinline__ void tile_io_init(const TileArgs& P, int E, long B, TileIO<T>& io)
{
	constexpr int epl = TileIO<T>::epl;
	io.nk = E / epl / kTileThreads;
#pragma unroll
	for (int k = 0; k < kMaxSlots; k++) {
		const int pe = (threadIdx.x + k * kTileThreads) * epl;   // LDS position of the slot
		const int row = pe >> P.logC;
		const int col = (pe & ((1 << P.logC) - 1)) ^ swz_f(row, P);
		int tr = 0;
		for (int t = 0; t < P.K; t++) { tr |= ((row >> t) & 1) << P.rb[t]; }
		io.trow[k] = tr;
		io.col[k] = col;
		io.goff[k] = (long)tr * B + col;
	}
}

// start the asynchronous load of a tile state (completion: s_waitcnt vmcnt)
template <typename T>
__device__ __forceinline__ void tile_issue(cx<T>* S, const cx<T>* __restrict__ src, long base, const TileIO<T>& io)
{
#if defined(__HIP_DEVICE_COMPILE__)
	// (the builtin exists in the device compilation only)
	constexpr int epl = TileIO<T>::epl;
	const int w0 = threadIdx.x & ~63;
#pragma unroll
	for (int k = 0; k < kMaxSlots; k++) {
		if (k < io.nk) {
			__builtin_amdgcn_global_load_lds(src + base + io.goff[k],
				reinterpret_cast<__attribute__((address_space(3))) void*>(
					reinterpret_cast<uintptr_t>(S + (k * kTileThreads + w0) * epl)), 16, 0, 0);
		}
	}
#endif
}

template <typename T>
__device__ __forceinline__ void tile_fill(cx<T>* S, const TileIO<T>& io, cx<T> v)
{
	constexpr int epl = TileIO<T>::epl;
#pragma unroll
	for (int k = 0; k < kMaxSlots; k++) {
		if (k < io.nk) {
#pragma unroll
			for (int i = 0; i < epl; i++) { S[(threadIdx.x + k * kTileThreads) * epl + i] = v; }
		}
	}
}

// column block b0 of the identity: element (row, col) = [row == b0 + c0 + col]
// (returns whether the thread's slots hold a nonzero element)
template <typename T>
__device__ __forceinline__ int tile_identity(cx<T>* S, const TileIO<T>& io, long rbase, long c0, long b0)
{
	constexpr int epl = TileIO<T>::epl;
	int nz = 0;
#pragma unroll
	for (int k = 0; k < kMaxSlots; k++) {
		if (k < io.nk) {
			const long row = rbase | (long)io.trow[k];
#pragma unroll
			for (int i = 0; i < epl; i++) {
				const bool one = row == b0 + c0 + io.col[k] + i;
				nz |= one;
				S[(threadIdx.x + k * kTileThreads) * epl + i] = cx<T>{one ? (T)1 : (T)0, (T)0};
			}
		}
	}
	return nz;
}

template <typename T>
__device__ __forceinline__ void tile_store(cx<T>* __restrict__ dst, const cx<T>* S, long base, const TileIO<T>& io)
{
	constexpr int epl = TileIO<T>::epl;
	typedef double slot_t __attribute__((ext_vector_type(2)));
#pragma unroll
	for (int k = 0; k < kMaxSlots; k++) {
		if (k < io.nk) {
			const slot_t v = *reinterpret_cast<const slot_t*>(S + (threadIdx.x + k * kTileThreads) * epl);
			*reinterpret_cast<slot_t*>(dst + base + io.goff[k]) = v;
		}
	}
}

// Wave w owns the quads [w qpw, (w + 1) qpw) of every gate, i.e. those whose two
// highest tile bits outside the gate equal w; consecutive gates with the same
// two bits need no barrier between them (TileArgs::barrier).
//
// Both kernels are software-pipelined over the tiles of a workgroup: LDS holds
// two buffers of the tile's states; while the gates of tile t run in one, tile
// t + grid streams into the other.

// ---------------------------------------------------------------------------
// forward pass: psi -> (gates) psi, with DUAL also dpsi.  The input is
// column block b0 of the identity (flags & 1) or `in`; dpsi is zero (flags & 2)
// or `din`.
// Forward-pass buffering: one buffer (a workgroup stores a tile and loads the
// next into the same slots) and several workgroups per CU, whose gates cover
// each other's tile loads (measured slightly faster than two double-buffered
// workgroups fewer): four for the 16 KiB states of the planned dual tiles (LDS
// holds no more), which leaves 128 registers for the pipelined steps, else 5
// (FP64) or 6 (FP32).
template <typename T, bool DUAL> constexpr int fwd_buffers() { return 1; }
template <typename T, bool DUAL, int LOGE = 0> constexpr int fwd_waves_per_simd()
{
	return (DUAL && LOGE > 0 && (1 << LOGE) * (int)sizeof(cx<T>) >= 16384) ? 4 : (sizeof(T) == 4 ? 6 : 5);
}

// LOGE > 0: the tile holds 2^LOGE elements (checked by the host), so the LDS state stride is a
// compile-time constant and every state's LDS base is an instruction offset: one XOR of the
// lane's and the group's byte terms addresses the element in every state.  LOGE = 0: any tile.
template <typename T, bool DUAL, int LOGE = 0>
__global__ __launch_bounds__(kTileThreads, (fwd_waves_per_simd<T, DUAL, LOGE>())) void k_tile_forward(
	const cx<T>* __restrict__ in, cx<T>* __restrict__ out, const cx<T>* __restrict__ din, cx<T>* __restrict__ dout,
	const TileArgs P, const cx<T>* __restrict__ mats, const GateTerms* __restrict__ gterms, long B, long b0, long ntiles,
	int flags, const int* skip)
{
	if (skipped(skip)) { return; }
	extern __shared__ __align__(16) char lds_raw[];
	constexpr int NS = DUAL ? 2 : 1;              // states per buffer
	constexpr int GS = DUAL ? 2 : 1;              // matrix-core groups per step
	constexpr int gq = DUAL ? 16 : 32;            // quads per matrix-core group
	const int E = (LOGE > 0) ? (1 << LOGE) : (1 << (P.K + P.logC));
	const int sbytes = E * (int)sizeof(cx<T>) + kLdsPadBytes;
	auto state = [&](int b, int i) { return reinterpret_cast<cx<T>*>(lds_raw + (b * NS + i) * sbytes); };
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int qpw = (E / 4) / kTileWaves;         // quads owned by a wave
	const long cmask = (1L << P.lognc) - 1;
	TileIO<T> io;
	tile_io_init(P, E, B, io);
	// gate matrices G, Z of the pass after the buffers
	cx<T>* lm = reinterpret_cast<cx<T>*>(lds_raw + fwd_buffers<T, DUAL>() * NS * sbytes);
	const int which[2] = {kMatG, kMatZ};
	load_pass_mats(mats, P.x0, P.ng, which, DUAL ? 2 : 1, fwd_mat_stride<DUAL>(), lm);
	// FP32: amplitude-major rows, so that a result lane holds one amplitude (gate_store)
	const AOperandMap<T, DUAL> amap = a_operand_map<T, DUAL, sizeof(T) == 4>();
	// pass 0 (flags & 1): whether the thread's slots of the prefetched identity tile hold a
	// nonzero; most identity tiles are zero, and a zero tile's output is zero (its gates skipped)
	int nz = 1;
	static_assert(kTileWaves == 4, "one vote word per wave");
	volatile int* tnz = reinterpret_cast<volatile int*>(reinterpret_cast<char*>(lm) +
		kMaxTileGates * fwd_mat_stride<DUAL>() * (int)sizeof(cx<T>));
	auto prefetch = [&](long tile, int b) {
		const long rbase = deposit_bits(tile >> P.lognc, P.rmask);
		const long c0 = (tile & cmask) << P.logC;
		if (flags & 1) { nz = tile_identity(state(b, 0), io, rbase, c0, b0); }
		else { tile_issue(state(b, 0), in, rbase * B + c0, io); }
		if (DUAL) {
			if (flags & 2) { tile_fill(state(b, 1), io, cx<T>{0, 0}); }
			else { tile_issue(state(b, 1), din, rbase * B + c0, io); }
		}
	};
	// store tile t - 1 from buffer b ^ 1 (after the barrier that ends its gates); each
	// thread stores and then prefetches the same slots, so its own LDS reads precede
	// its asynchronous LDS writes and no further barrier is needed
	auto store = [&](long tile, int b) {
		const long base = deposit_bits(tile >> P.lognc, P.rmask) * B + ((tile & cmask) << P.logC);
		tile_store(out, state(b, 0), base, io);
		if (DUAL) { tile_store(dout, state(b, 1), base, io); }
	};
	// the pass's gates on the tile in buffer b
	// Dual sweeps with a planned tile: the gates' position terms are fixed over the
	// launch's tiles, so they are formed once: the lane's byte term (eaj) and the
	// wave's byte terms (sgj: the wave's range, then the group bits).
	constexpr int ES = (int)sizeof(cx<T>);
	constexpr bool FIXED = DUAL && LOGE > 0;
	int eaj[kMaxTileGates];
	int sgj[kMaxTileGates][4];
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) {
		eaj[j] = 0;
		sgj[j][0] = sgj[j][1] = sgj[j][2] = sgj[j][3] = 0;
		if (FIXED && j < P.ng) {
			const GateTerms gt = gterms[P.x0 + j];
			eaj[j] = gate_lanes(gt).ea * ES;
			sgj[j][0] = sel4(gt.wv, wave) * ES;
			sgj[j][1] = gt.gbit[0] * ES;
			sgj[j][2] = gt.gbit[1] * ES;
			sgj[j][3] = gt.gbit[2] * ES;
		}
	}
	// the pass's gates on the tile in buffer b
	auto run_tile = [&](int b) {
		cx<T>* S = state(b, 0);
		cx<T>* D = state(b, DUAL ? 1 : 0);
		auto gate = [&](auto jc) {
			constexpr int j = decltype(jc)::value;
			const T* mj = reinterpret_cast<const T*>(lm + j * fwd_mat_stride<DUAL>());
			// FP64: 4 x 4 block products (F44: the planned dual tiles); FP32: the real 16 x 16 form
			constexpr bool F44 = FIXED && sizeof(T) == 8;
			T a[4];
			Frag44 f44;
			if constexpr (sizeof(T) == 8) {
				f44 = DUAL ? frag44(reinterpret_cast<const cx<double>*>(mj)) : frag44x(reinterpret_cast<const cx<double>*>(mj));
			} else {
#pragma unroll
				for (int kk = 0; kk < 4; kk++) { a[kk] = mj[amap.aoff[kk]] * amap.afac[kk]; }
			}
			if constexpr (FIXED) {
				// psi and dpsi of a group at one byte offset (the lane's XOR the group's), the
				// states' bases as instruction offsets (b = 0: one buffer).  Software-pipelined
				// over the gate's NST steps: step s + 1's operands are loaded before step s's
				// matrix-core chain (the steps' quads are disjoint), so that the chain covers
				// the loads' latency.
				constexpr int NST = ((1 << LOGE) / 4 / kTileWaves) / (GS * gq);
				static_assert(NST >= 1 && NST <= 4, "one to four steps per gate and wave");
				GateOperands<T, GS> o[NST];
				int la[NST][GS];
				// (an empty asm statement makes the lane term opaque here, so that the
				// compiler forms the groups' offsets per tile instead of holding every
				// gate's offsets in registers across the tile loop)
				int ea = eaj[j];
				asm volatile("" : "+v"(ea));
				auto load_step = [&](auto sc) {
					constexpr int SI = decltype(sc)::value;
#pragma unroll
					for (int n = 0; n < GS; n++) {
						const int gi = (SI * GS * gq + n * gq) >> 4;   // group index in the wave's range
						const int gb = sgj[j][0] ^ ((gi & 1) ? sgj[j][1] : 0) ^ ((gi & 2) ? sgj[j][2] : 0) ^
						               ((gi & 4) ? sgj[j][3] : 0);
						la[SI][n] = ea ^ gb;
						o[SI].x[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la[SI][n]);
						o[SI].y[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la[SI][n] + sbytes);
					}
				};
				auto step = [&](auto sc) {
					constexpr int SI = decltype(sc)::value;
					if constexpr (SI + 1 < NST) { load_step(std::integral_constant<int, SI + 1>{}); }
					// (a scheduling fence: those loads issue before the chain)
					__builtin_amdgcn_sched_barrier(0);
					__builtin_amdgcn_s_setprio(1);
					if constexpr (F44) { gate_mfma44<GS>(f44, o[SI]); }
					else { gate_mfma<T, GS>(a, o[SI]); }
					__builtin_amdgcn_s_setprio(0);
#pragma unroll
					for (int n = 0; n < GS; n++) {
						*reinterpret_cast<cx<T>*>(lds_raw + la[SI][n]) = cx<T>{o[SI].acc[n][0], o[SI].acc[n][1]};
						*reinterpret_cast<cx<T>*>(lds_raw + la[SI][n] + sbytes) = cx<T>{o[SI].acc[n][2], o[SI].acc[n][3]};
					}
				};
				load_step(std::integral_constant<int, 0>{});
				step(std::integral_constant<int, 0>{});
				if constexpr (NST > 1) { step(std::integral_constant<int, 1>{}); }
				if constexpr (NST > 2) { step(std::integral_constant<int, 2>{}); }
				if constexpr (NST > 3) { step(std::integral_constant<int, 3>{}); }
			} else {
				const GateTerms gt = gterms[P.x0 + j];
				const GateLanes gl = gate_lanes(gt);
				const int ewv = sel4(gt.wv, wave);
				for (int g = 0; g < qpw; g += GS * gq) {
					int eg[GS], eg2[GS];
#pragma unroll
					for (int n = 0; n < GS; n++) {
						eg[n] = ewv ^ group_term(gt, (g + n * gq) >> 4);
						eg2[n] = DUAL ? eg[n] : ewv ^ group_term(gt, ((g + n * gq) >> 4) + 1);
					}
					if constexpr (sizeof(T) == 8) {
						mfma_gate_apply44<GS, DUAL>(reinterpret_cast<cx<double>*>(S), reinterpret_cast<cx<double>*>(D), eg, eg2,
							gl, f44);
					} else {
						mfma_gate_apply<T, GS>(S, D, eg, eg2, gl, a);
					}
				}
			}
		};
		// the pass's gates as straight-line code under wave-uniform guards
		auto gate_at = [&](auto jc) {
			constexpr int j = decltype(jc)::value;
			if (j < P.ng) {
				if (j > 0 && ((P.barrier >> j) & 1)) { __syncthreads(); }
				gate(jc);
			}
		};
		static_assert(kMaxTileGates == 8, "one guarded body per gate of a pass");
		gate_at(std::integral_constant<int, 0>{});
		gate_at(std::integral_constant<int, 1>{});
		gate_at(std::integral_constant<int, 2>{});
		gate_at(std::integral_constant<int, 3>{});
		gate_at(std::integral_constant<int, 4>{});
		gate_at(std::integral_constant<int, 5>{});
		gate_at(std::integral_constant<int, 6>{});
		gate_at(std::integral_constant<int, 7>{});
	};
	long tile = blockIdx.x;
	if (tile < ntiles) { prefetch(tile, 0); }
	if constexpr (fwd_buffers<T, DUAL>() == 2) {
		int b = 0;
		for (; tile < ntiles; tile += gridDim.x, b ^= 1) {
			__builtin_amdgcn_s_waitcnt(0);
			__syncthreads();
			if (tile != (long)blockIdx.x) { store(tile - gridDim.x, b ^ 1); }
			if (tile + gridDim.x < ntiles) { prefetch(tile + gridDim.x, b ^ 1); }
			run_tile(b);
		}
		if ((long)blockIdx.x < ntiles) {
			__syncthreads();
			store(tile - gridDim.x, b ^ 1);
		}
	} else {
		for (; tile < ntiles; tile += gridDim.x) {
			__builtin_amdgcn_s_waitcnt(0);
			if (flags & 1) {
				// each wave's vote in LDS after the gate matrices (read by all before the
				// barrier that ends the tile, written again only after it)
				const int wnz = __any(nz);   // (a vote of all the wave's lanes)
				if ((threadIdx.x & 63) == 0) { tnz[wave] = wnz ? 1 : 0; }
				__syncthreads();
				if (tnz[0] | tnz[1] | tnz[2] | tnz[3]) { run_tile(0); }
			} else {
				__syncthreads();
				run_tile(0);
			}
			__syncthreads();
			store(tile, 0);
			if (tile + gridDim.x < ntiles) { prefetch(tile + gridDim.x, 0); }
		}
	}
}

// ---------------------------------------------------------------------------
// reverse pass.  Inputs: psi (and dpsi) at the pass output (checkpoint), phi
// (and dphi) at the pass output; outputs phi (and dphi) at the pass input.
// flags: 4 = dphi at the pass output is zero, 8 = do not write phi at the pass input,
// 16 (SEED, planned tiles) = psi and dpsi are given at the pass input and the pass's
// gates are applied to them first (G, Z after the trace sums in LDS), 32 (with 8) = the
// sweep started from identity columns b0 + j: tiles without an identity element are skipped.
// SEED (dual sweeps, the circuit's last pass): phi_out is the target U's column block (FP64 in
// both precisions; FP32 reads it from HBM in the seed step, FP64 loads it as the adjoint tile);
// the seed -conj(U) is formed in LDS and the trace sum(-conj(U) psi) goes to
// trace[2 (block * waves + wave)] (a separate instantiation, so that the other passes run
// exactly the kernel without it: the extra loop code measured 8 % slower on every pass).
// partial[((part * ng + j) * 256) + row * 16 + col], part = block * waves + wave:
// the 16 x 16 outer-product sums of gate j (rows: component 4 (2 s + p) + r of
// the adjoint at the gate input, columns: that of the state before the gate).
// Reverse-pass buffering.  Dual sweeps hold one buffer of the four tile states
// (a workgroup stores a tile's adjoints and loads the next tile into the same
// slots), and the other workgroups' gates cover the loads.  With the default
// 16 KiB tiles (FP64 64 x 16, FP32 64 x 32) two workgroups fit a CU, and each
// wave runs two independent 32-quad steps per gate: FP64 measured 9 % faster
// than double-buffered 512-element tiles at the same two workgroups (FP64
// needs 183 registers, so a third wave would spill).  FP32 with 1024-element
// tiles fits 3 workgroups (168 registers).  The plain sweeps (two states)
// double-buffer, the next tile streaming in while the gates run.
template <typename T, bool DUAL> constexpr int rev_buffers() { return DUAL ? 1 : 2; }
template <typename T, bool DUAL> constexpr int rev_waves_per_simd() { return 2; }

template <typename T, bool DUAL, int LOGE = 0, bool SEED = false>
__global__ __launch_bounds__(kTileThreads, (rev_waves_per_simd<T, DUAL>())) void k_tile_reverse(
	const cx<T>* __restrict__ psi_out, const cx<T>* __restrict__ dpsi_out,
	const cx<T>* __restrict__ phi_out, const cx<T>* __restrict__ dphi_out,
	cx<T>* __restrict__ phi_in, cx<T>* __restrict__ dphi_in,
	const TileArgs P, const cx<T>* __restrict__ mats, const GateTerms* __restrict__ gterms, long B, long b0, long ntiles,
	int flags, double* __restrict__ partial, double* __restrict__ trace, const int* skip)
{
	if (skipped(skip)) { return; }
	extern __shared__ __align__(16) char lds_raw[];
	constexpr int NS = DUAL ? 4 : 2;              // states per buffer: psi, [dpsi], phi, [dphi]
	constexpr int GS = DUAL ? 2 : 1;
	constexpr int gq = DUAL ? 16 : 32;
	const int E = (LOGE > 0) ? (1 << LOGE) : (1 << (P.K + P.logC));
	const int sbytes = E * (int)sizeof(cx<T>) + lds_pad_rev<T>();
	auto state = [&](int b, int i) { return reinterpret_cast<cx<T>*>(lds_raw + (b * NS + i) * sbytes); };
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int qpw = (E / 4) / kTileWaves;

	double4_t acc[kMaxTileGates];
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) { acc[j] = double4_t{0, 0, 0, 0}; }
	// FP32 states: the last processed gate's tile sums (FP32 matrix-core chain over the
	// wave's quads of the tile), added into its FP64 accumulator by the next gate
	typename Vec4<T>::type pend = {0, 0, 0, 0};
	auto flush32 = [&](double4_t& a) {
#pragma unroll
		for (int v = 0; v < 4; v++) { a[v] += (double)pend[v]; }
		pend = typename Vec4<T>::type{0, 0, 0, 0};
	};

	const long cmask = (1L << P.lognc) - 1;
	TileIO<T> io;
	tile_io_init(P, E, B, io);
	// gate matrices G^-1, -G^-1 Z G^-1, G^T, Z^T of the pass after the two buffers
	constexpr int NM = DUAL ? 4 : 2;
	cx<T>* lm = reinterpret_cast<cx<T>*>(lds_raw + rev_buffers<T, DUAL>() * NS * sbytes);
	const int which[4] = {kMatGinv, DUAL ? kMatW : kMatGT, kMatGT, kMatZT};
	const int which_plain[2] = {kMatGinv, kMatGT};
	load_pass_mats(mats, P.x0, P.ng, DUAL ? which : which_plain, NM, rev_mat_stride<DUAL>(), lm);
	const AOperandMap<T, DUAL> amap = a_operand_map<T, DUAL>();
	// transposed results: lane l holds component (l & 15) >> 2 = x.re, x.im, y.re, y.im
	const int sc_lane = (threadIdx.x & 15) >> 2;
	const int ybo = (sc_lane < 2) ? 0 : sbytes;
	auto prefetch = [&](long tile, int b) {
		const long base = deposit_bits(tile >> P.lognc, P.rmask) * B + ((tile & cmask) << P.logC);
		tile_issue(state(b, 0), psi_out, base, io);
		// (the FP32 seed pass reads the FP64 target's tile itself, in its seed step)
		if constexpr (!(SEED && sizeof(T) == 4)) { tile_issue(state(b, NS / 2), phi_out, base, io); }
		if (DUAL) {
			tile_issue(state(b, 1), dpsi_out, base, io);
			if (flags & 4) { tile_fill(state(b, 3), io, cx<T>{0, 0}); }
			else { tile_issue(state(b, 3), dphi_out, base, io); }
		}
	};
	// phi (and dphi) of tile t - 1 from buffer b ^ 1, stored after the barrier that ends
	// its gates, by the threads that then prefetch the same slots
	auto store = [&](long tile, int b) {
		if (flags & 8) { return; }
		const long base = deposit_bits(tile >> P.lognc, P.rmask) * B + ((tile & cmask) << P.logC);
		tile_store(phi_in, state(b, NS / 2), base, io);
		if (DUAL) { tile_store(dphi_in, state(b, 3), base, io); }
	};
	// Dual sweeps with a planned tile: the gates' position terms are fixed over the
	// launch's tiles, so they are formed once: the lane's load term (eabj) and
	// transposed-store term (btbj; FP32: with the re / im byte of its component), and
	// the wave's terms (sgj: the wave's range, the group bits, then FP32's result
	// registers' terms).
	constexpr int ES = (int)sizeof(cx<T>);
	constexpr bool FIXED = DUAL && LOGE > 0;
	// FP64 sweeps run on 4 x 4 block products (gate_mfma44_t, outer44; plain sweeps the
	// _pair forms): acc[J] holds (A1.re, A1.im[, A2.re, A2.im]) of the lane's block entry.
	// F44: the dual sweeps with a planned tile, btbj the lane's term of its transposed
	// results (one amplitude of each state)
	constexpr bool F44 = DUAL && LOGE > 0 && sizeof(T) == 8;
	int eabj[kMaxTileGates], btbj[kMaxTileGates];
	int sgj[kMaxTileGates][8];
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) {
		eabj[j] = btbj[j] = 0;
#pragma unroll
		for (int i = 0; i < 8; i++) { sgj[j][i] = 0; }
		if (FIXED && j < P.ng) {
			const GateTerms gt = gterms[P.x0 + j];
			eabj[j] = gate_lanes<T>(gt).ea * ES;
			btbj[j] = F44 ? transposed44_lane_term(gt) * ES
			              : (transposed_lane_term<T>(gt) * ES) | ((sc_lane & 1) * (int)sizeof(T));
			sgj[j][0] = sel4(gt.wv, wave) * ES;
			sgj[j][1] = gt.gbit[0] * ES;
			sgj[j][2] = gt.gbit[1] * ES;
			sgj[j][3] = gt.gbit[2] * ES;
#pragma unroll
			for (int v = 0; v < 4; v++) { sgj[j][4 + v] = transposed_reg_term<T>(gt, v) * ES; }
		}
	}
	// the pass's gates on the tile in buffer b
	auto run_tile = [&](int b) {
		cx<T>* S = state(b, 0);
		cx<T>* D = state(b, DUAL ? 1 : 0);
		cx<T>* F = state(b, NS / 2);
		cx<T>* DF = state(b, DUAL ? 3 : NS / 2);
		// body of gate J (compile-time, so that its outer products accumulate in
		// the registers acc[J] directly)
		auto gate_body = [&](auto jc) {
			constexpr int J = decltype(jc)::value;
			if constexpr (sizeof(T) == 4 && J + 1 < kMaxTileGates) {
				// FP32: the previous gate (J + 1) left its tile sums in pend
				if (J + 1 < P.ng) { flush32(acc[J + 1]); }
			}
			// recovery: [[G^-1, 0], [-G^-1 Z G^-1, G^-1]]; adjoint: [[G^T, 0], [Z^T, G^T]]
			const T* mj = reinterpret_cast<const T*>(lm + J * rev_mat_stride<DUAL>());
			T au[4], aa[4];
			Frag44 fu, fa;
			if constexpr (sizeof(T) == 8) {
				// recovery G^-1 [and W], adjoint G^T [and Z^T]: pairs after 2 kMatPairOff, plain X
				// matrices after kMatPairOff
				const cx<double>* m = reinterpret_cast<const cx<double>*>(mj);
				fu = DUAL ? frag44(m) : frag44x(m);
				fa = DUAL ? frag44(m + 2 * kMatPairOff) : frag44x(m + kMatPairOff);
			} else {
#pragma unroll
				for (int kk = 0; kk < 4; kk++) {
					au[kk] = mj[amap.aoff[kk]] * amap.afac[kk];
					aa[kk] = mj[amap.aoff[kk] + 2 * kMatPairOff * (DUAL ? 2 : 1)] * amap.afac[kk];
				}
			}
			const bool adjoint = (J > 0) || !(flags & 8);
			// the outer products of a step's transposed results: phi_in (x) psi,
			// mapped by the reduction to the gate's phi_out (x) psi
			auto outer = [&](auto gsc, const auto& os, const auto& oa) {
				constexpr int GSX = decltype(gsc)::value;
				if constexpr (sizeof(T) == 8) {
					// FP64: one chain over the groups' steps accumulating onto acc[J]
#pragma unroll
					for (int t = 0; t < 4; t++) {
#pragma unroll
						for (int n = 0; n < GSX; n++) { acc[J] = mfma_step(oa.acc[n][t], os.acc[n][t], acc[J]); }
					}
				} else {
					// FP32: one chain over the gate's quads of this tile into pend,
					// added into the FP64 accumulator by the next gate (no wait here)
#pragma unroll
					for (int t = 0; t < 4; t++) {
#pragma unroll
						for (int n = 0; n < GSX; n++) { pend = mfma_step(oa.acc[n][t], os.acc[n][t], pend); }
					}
				}
			};
			// GSX matrix-core groups of gq quads per step
			auto steps = [&](auto gsc) {
				constexpr int GSX = decltype(gsc)::value;
				if constexpr (DUAL && LOGE > 0) {
					// The four states (psi, dpsi, phi, dphi) of a group at one byte offset, their
					// bases as instruction offsets (b = 0: one buffer).  Software-pipelined over
					// the gate's NST steps: step s + 1's operands are loaded after step s's
					// stores (a wave's LDS accesses complete in issue order, and the steps'
					// quads are disjoint) and before step s's outer products, whose
					// matrix-core time covers the loads' latency.
					constexpr int NST = ((1 << LOGE) / 4 / kTileWaves) / (GSX * gq);
					static_assert(NST >= 1, "a wave owns at least one step of quads");
					GateOperands<T, GSX> os[NST], oa[NST];
					int egb[NST][GSX];
					// (an empty asm statement makes the lane terms opaque here, so that the
					// compiler forms the groups' offsets per tile instead of holding every
					// gate's offsets in registers across the tile loop)
					int eab, btb, sg[8];
					eab = eabj[J];
					btb = btbj[J];
					asm volatile("" : "+v"(eab), "+v"(btb));
					// (16-byte element terms: whole-element LDS accesses with the state
					// strides in the instructions' offset fields)
					if constexpr (F44) { __builtin_assume(((eab | btb) & 15) == 0); }
#pragma unroll
					for (int i = 0; i < 8; i++) { sg[i] = sgj[J][i]; }
					auto load_step = [&](auto sc) {
						constexpr int S = decltype(sc)::value;
#pragma unroll
						for (int n = 0; n < GSX; n++) {
							const int gi = (S * GSX * gq + n * gq) >> 4;   // group index in the wave's range
							egb[S][n] = sg[0] ^ ((gi & 1) ? sg[1] : 0) ^ ((gi & 2) ? sg[2] : 0) ^ ((gi & 4) ? sg[3] : 0);
							const int la = eab ^ egb[S][n];
							if constexpr (F44) {
								os[S].x[n] = lds_load_cxd((unsigned)la);
								os[S].y[n] = lds_load_cxd((unsigned)la + sbytes);
								oa[S].x[n] = lds_load_cxd((unsigned)la + 2 * sbytes);
								oa[S].y[n] = lds_load_cxd((unsigned)la + 3 * sbytes);
							} else {
								os[S].x[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la);
								os[S].y[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la + sbytes);
								oa[S].x[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la + 2 * sbytes);
								oa[S].y[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la + 3 * sbytes);
							}
						}
					};
					auto step = [&](auto sc) {
						constexpr int S = decltype(sc)::value;
						// matrix-core work at raised wave priority, so that the other wave's
						// LDS traffic issues in its shadow
						__builtin_amdgcn_s_setprio(1);
						if constexpr (F44) {
							// the transposed results of lane l: one amplitude of each state
							gate_mfma44_t<GSX>(fu, os[S]);
							gate_mfma44_t<GSX>(fa, oa[S]);
							// (the states recovered by the pass's first gate, J = 0, are its
							// outer products' operands only: no gate reads them from LDS)
							unsigned st[GSX];
#pragma unroll
							for (int n = 0; n < GSX; n++) {
								st[n] = (unsigned)(btb ^ egb[S][n]);
								if constexpr (J > 0) {
									lds_store_cxd(st[n], cx<double>{os[S].acc[n][0], os[S].acc[n][1]});
									lds_store_cxd(st[n] + sbytes, cx<double>{os[S].acc[n][2], os[S].acc[n][3]});
								}
							}
							if (adjoint) {
#pragma unroll
								for (int n = 0; n < GSX; n++) {
									lds_store_cxd(st[n] + 2 * sbytes, cx<double>{oa[S].acc[n][0], oa[S].acc[n][1]});
									lds_store_cxd(st[n] + 3 * sbytes, cx<double>{oa[S].acc[n][2], oa[S].acc[n][3]});
								}
							}
							if constexpr (S + 1 < NST) { load_step(std::integral_constant<int, S + 1>{}); }
							__builtin_amdgcn_sched_barrier(0);
							outer44<GSX>(os[S], oa[S], acc[J]);
							__builtin_amdgcn_s_setprio(0);
							return;
						}
						gate_mfma_t<T, GSX>(au, os[S]);
						gate_mfma_t<T, GSX>(aa, oa[S]);
						// component l & 15 of the transposed results: state psi / dpsi (phi /
						// dphi) by the lane's ybo, re / im in etb; one address for both pairs
						int sa[GSX][4];
#pragma unroll
						for (int n = 0; n < GSX; n++) {
#pragma unroll
							for (int v = 0; v < 4; v++) {
								sa[n][v] = (btb ^ (sg[4 + v] ^ egb[S][n])) + ybo;
								*reinterpret_cast<T*>(lds_raw + sa[n][v]) = os[S].acc[n][v];
							}
						}
						if (adjoint) {
#pragma unroll
							for (int n = 0; n < GSX; n++) {
#pragma unroll
								for (int v = 0; v < 4; v++) {
									*reinterpret_cast<T*>(lds_raw + sa[n][v] + 2 * sbytes) = oa[S].acc[n][v];
								}
							}
						}
						if constexpr (S + 1 < NST) { load_step(std::integral_constant<int, S + 1>{}); }
						// (a scheduling fence: the stores and loads issue before the outer
						// products, which would otherwise hold them back behind the busy
						// matrix core)
						__builtin_amdgcn_sched_barrier(0);
						outer(gsc, os[S], oa[S]);
						__builtin_amdgcn_s_setprio(0);
					};
					load_step(std::integral_constant<int, 0>{});
					step(std::integral_constant<int, 0>{});
					if constexpr (NST > 1) { step(std::integral_constant<int, 1>{}); }
					if constexpr (NST > 2) { step(std::integral_constant<int, 2>{}); }
					if constexpr (NST > 3) { step(std::integral_constant<int, 3>{}); }
					static_assert(NST <= 4, "at most four steps per gate and wave");
				} else {
					const GateTerms gt = gterms[P.x0 + J];
					const GateLanes gl = gate_lanes<T>(gt);
					const int es44 = (sizeof(T) == 8) ? transposed44_lane_term(gt) : 0;
					const int ewv = sel4(gt.wv, wave);
					for (int g = 0; g < qpw; g += GSX * gq) {
						int eg[GSX], eg2[GSX];
#pragma unroll
						for (int n = 0; n < GSX; n++) {
							eg[n] = ewv ^ group_term(gt, (g + n * gq) >> 4);
							eg2[n] = DUAL ? eg[n] : ewv ^ group_term(gt, ((g + n * gq) >> 4) + 1);
						}
						// The state's recovery (uncompute) and the adjoint's update are
						// independent: load both, run both chains (transposed, so that the
						// results are outer-product operands), store both; the outer
						// products pair the adjoint at the gate input with the recovered
						// state straight from the registers.
						GateOperands<T, GSX> os, oa;
						gate_load<T, GSX>(S, D, eg, eg2, gl, os);
						gate_load<T, GSX>(F, DF, eg, eg2, gl, oa);
						__builtin_amdgcn_s_setprio(1);
						if constexpr (sizeof(T) == 8) {
							// 4 x 4 block products; lane l's results: one amplitude of each state
							if constexpr (DUAL) {
								gate_mfma44_t<GSX>(fu, os);
								gate_mfma44_t<GSX>(fa, oa);
							} else {
								gate_mfma44_pair_t<GSX>(fu, os);
								gate_mfma44_pair_t<GSX>(fa, oa);
							}
#pragma unroll
							for (int n = 0; n < GSX; n++) {
								if constexpr (J > 0) {
									S[eg[n] ^ es44] = cx<T>{os.acc[n][0], os.acc[n][1]};
									D[eg2[n] ^ es44] = cx<T>{os.acc[n][2], os.acc[n][3]};
								}
								if (adjoint) {
									F[eg[n] ^ es44] = cx<T>{oa.acc[n][0], oa.acc[n][1]};
									DF[eg2[n] ^ es44] = cx<T>{oa.acc[n][2], oa.acc[n][3]};
								}
							}
							if constexpr (DUAL) { outer44<GSX>(os, oa, acc[J]); }
							else { outer44_pair<GSX>(os, oa, acc[J]); }
						} else {
							gate_mfma_t<T, GSX>(au, os);
							gate_mfma_t<T, GSX>(aa, oa);
							gate_store_t<T, GSX>(S, D, eg, eg2, gl, os);
							if (adjoint) { gate_store_t<T, GSX>(F, DF, eg, eg2, gl, oa); }
							outer(gsc, os, oa);
						}
						__builtin_amdgcn_s_setprio(0);
					}
				}
			};
			if constexpr (DUAL && sizeof(T) == 4) {
				// FP32 dual sweeps: four groups (64 quads) per step when the wave's quads allow
				constexpr bool wide_fits = (LOGE == 0) || (((1 << LOGE) / 4 / kTileWaves) % (4 * gq)) == 0;
				if constexpr (wide_fits) {
					if (P.wide && (qpw % (4 * gq)) == 0) { steps(std::integral_constant<int, 4>{}); }
					else { steps(std::integral_constant<int, GS>{}); }
				} else {
					steps(std::integral_constant<int, GS>{});
				}
			} else {
				steps(std::integral_constant<int, GS>{});
			}
		};
		// The pass's gates, last first, as straight-line code under wave-uniform
		// guards: each accumulator acc[J] stays in its registers (a loop over a switch
		// on the gate index makes the compiler move every accumulator on each gate).
		auto gate_at = [&](auto jc) {
			constexpr int J = decltype(jc)::value;
			if (J < P.ng) {
				if (J + 1 < P.ng && ((P.barrier >> (J + 1)) & 1)) { __syncthreads(); }
				gate_body(jc);
			}
		};
		static_assert(kMaxTileGates == 8, "one guarded body per gate of a pass");
		gate_at(std::integral_constant<int, 7>{});
		gate_at(std::integral_constant<int, 6>{});
		gate_at(std::integral_constant<int, 5>{});
		gate_at(std::integral_constant<int, 4>{});
		gate_at(std::integral_constant<int, 3>{});
		gate_at(std::integral_constant<int, 2>{});
		gate_at(std::integral_constant<int, 1>{});
		gate_at(std::integral_constant<int, 0>{});
		if constexpr (sizeof(T) == 4) { flush32(acc[0]); }
	};
	// flags & 32 (the sweep's first pass, from identity columns b0 + j, no adjoint written):
	// psi and dpsi at the pass output are zero on every tile without an identity element (the
	// pass's gates act inside a tile), so such tiles add nothing to the sums and are skipped,
	// loads included.  Tile t holds element (b0 + c0 + i, b0 + c0 + i) for some column i < C
	// iff the column chunk's row bits outside the tile equal the tile's (its low log2 C bits
	// range over every value).
	auto tile_nonzero = [&](long t) {
		if (!(flags & 32)) { return true; }
		const long rbase = deposit_bits(t >> P.lognc, P.rmask);
		const long hi = ~((1L << P.logC) - 1);
		return ((b0 + ((t & cmask) << P.logC)) & (long)P.rmask & hi) == (rbase & hi);
	};
	auto next_tile = [&](long t) {
		while (t < ntiles && !tile_nonzero(t)) { t += gridDim.x; }
		return t;
	};
	long tile = next_tile(blockIdx.x);
	if (tile < ntiles) { prefetch(tile, 0); }
	if constexpr (rev_buffers<T, DUAL>() == 2) {
		int b = 0;
		long prev = -1;
		while (tile < ntiles) {
			__builtin_amdgcn_s_waitcnt(0);
			__syncthreads();
			if (prev >= 0) { store(prev, b ^ 1); }
			const long next = next_tile(tile + gridDim.x);
			if (next < ntiles) { prefetch(next, b ^ 1); }
			run_tile(b);
			prev = tile;
			tile = next;
			b ^= 1;
		}
		if (prev >= 0) {
			__syncthreads();
			store(prev, b ^ 1);
		}
	} else {
		// SEED: phi_out holds the target U's tile; the adjoint seed is -conj(U), formed in
		// place, and the target's trace sum(-conj(U) psi) is taken over the tile on the way
		// (per thread and wave, then added to the wave's running sum)
		static_assert(!SEED || DUAL, "the seed pass is a dual sweep");
		// (the running trace sum of a wave sits in LDS after the gate matrices, so that no
		// register is held across the gates for it)
		double* tracc = reinterpret_cast<double*>(lds_raw + rev_buffers<T, DUAL>() * NS * sbytes +
			kMaxTileGates * rev_mat_stride<DUAL>() * (int)sizeof(cx<T>)) + 2 * (threadIdx.x >> 6);
		if constexpr (SEED) {
			if (lane == 0) { tracc[0] = 0; tracc[1] = 0; }
		}
		// SEED with flags & 16 (planned tiles): psi_out and dpsi_out hold the states at
		// the pass input, and the pass's gates are applied to them here first, so that the
		// circuit's output states never go through HBM (its last forward pass, fused)
		cx<T>* lmf = reinterpret_cast<cx<T>*>(reinterpret_cast<char*>(tracc - 2 * (threadIdx.x >> 6)) + kRevTraceBytes);
		if constexpr (SEED && FIXED) {
			if (flags & 16) {
				const int whichf[2] = {kMatG, kMatZ};
				load_pass_mats(mats, P.x0, P.ng, whichf, 2, fwd_mat_stride<true>(), lmf);
			}
		}
		auto forward_gates = [&]() {
			if constexpr (SEED && FIXED) {
				// the forward pass's steps: FP64 the 4 x 4 blocks, FP32 the real 16 x 16 form with
				// amplitude-major rows (the forward kernel's gate code, so that the states are the
				// ones a forward launch would write)
				constexpr int GSF = 2, gqf = 16;
				constexpr int NSTF = ((1 << LOGE) / 4 / kTileWaves) / (GSF * gqf);
				static_assert(NSTF >= 1 && NSTF <= 4, "one to four steps per gate and wave");
				auto fgate = [&](auto jc) {
					constexpr int j = decltype(jc)::value;
					if (j < P.ng) {
						if (j > 0 && ((P.barrier >> j) & 1)) { __syncthreads(); }
						Frag44 f;
						T af[4];
						if constexpr (F44) {
							f = frag44(reinterpret_cast<const cx<double>*>(lmf + j * fwd_mat_stride<true>()));
						} else {
							const AOperandMap<T, true> amf = a_operand_map<T, true, true>();
							const T* mjf = reinterpret_cast<const T*>(lmf + j * fwd_mat_stride<true>());
#pragma unroll
							for (int kk = 0; kk < 4; kk++) { af[kk] = mjf[amf.aoff[kk]] * amf.afac[kk]; }
						}
						int ea = eabj[j];
						asm volatile("" : "+v"(ea));
						if constexpr (F44) { __builtin_assume((ea & 15) == 0); }
						GateOperands<T, GSF> o[NSTF];
						unsigned la[NSTF][GSF];
						auto load_step = [&](auto sc) {
							constexpr int SI = decltype(sc)::value;
#pragma unroll
							for (int n = 0; n < GSF; n++) {
								const int gi = (SI * GSF * gqf + n * gqf) >> 4;
								const int gb = sgj[j][0] ^ ((gi & 1) ? sgj[j][1] : 0) ^ ((gi & 2) ? sgj[j][2] : 0) ^
								               ((gi & 4) ? sgj[j][3] : 0);
								la[SI][n] = (unsigned)(ea ^ gb);
								if constexpr (F44) {
									o[SI].x[n] = lds_load_cxd(la[SI][n]);
									o[SI].y[n] = lds_load_cxd(la[SI][n] + sbytes);
								} else {
									o[SI].x[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la[SI][n]);
									o[SI].y[n] = *reinterpret_cast<const cx<T>*>(lds_raw + la[SI][n] + sbytes);
								}
							}
						};
						auto step = [&](auto sc) {
							constexpr int SI = decltype(sc)::value;
							if constexpr (SI + 1 < NSTF) { load_step(std::integral_constant<int, SI + 1>{}); }
							__builtin_amdgcn_sched_barrier(0);
							__builtin_amdgcn_s_setprio(1);
							if constexpr (F44) { gate_mfma44<GSF>(f, o[SI]); }
							else { gate_mfma<T, GSF>(af, o[SI]); }
							__builtin_amdgcn_s_setprio(0);
#pragma unroll
							for (int n = 0; n < GSF; n++) {
								if constexpr (F44) {
									lds_store_cxd(la[SI][n], cx<double>{o[SI].acc[n][0], o[SI].acc[n][1]});
									lds_store_cxd(la[SI][n] + sbytes, cx<double>{o[SI].acc[n][2], o[SI].acc[n][3]});
								} else {
									*reinterpret_cast<cx<T>*>(lds_raw + la[SI][n]) = cx<T>{o[SI].acc[n][0], o[SI].acc[n][1]};
									*reinterpret_cast<cx<T>*>(lds_raw + la[SI][n] + sbytes) =
										cx<T>{o[SI].acc[n][2], o[SI].acc[n][3]};
								}
							}
						};
						load_step(std::integral_constant<int, 0>{});
						step(std::integral_constant<int, 0>{});
						if constexpr (NSTF > 1) { step(std::integral_constant<int, 1>{}); }
						if constexpr (NSTF > 2) { step(std::integral_constant<int, 2>{}); }
						if constexpr (NSTF > 3) { step(std::integral_constant<int, 3>{}); }
					}
				};
				fgate(std::integral_constant<int, 0>{});
				fgate(std::integral_constant<int, 1>{});
				fgate(std::integral_constant<int, 2>{});
				fgate(std::integral_constant<int, 3>{});
				fgate(std::integral_constant<int, 4>{});
				fgate(std::integral_constant<int, 5>{});
				fgate(std::integral_constant<int, 6>{});
				fgate(std::integral_constant<int, 7>{});
			}
		};
		while (tile < ntiles) {
			__builtin_amdgcn_s_waitcnt(0);
			__syncthreads();
			if constexpr (SEED && FIXED) {
				if (flags & 16) {
					forward_gates();
					__syncthreads();
				}
			}
			// FP32 seed pass: the target's FP64 elements of the thread's slots, from HBM (held
			// across the forward gates they would take registers the gates need)
			constexpr int EPL = TileIO<T>::epl;
			cx<double> useed[(SEED && sizeof(T) == 4) ? kMaxSlots * EPL : 1];
			if constexpr (SEED && sizeof(T) == 4) {
				const cx<double>* Ud = reinterpret_cast<const cx<double>*>(phi_out);
				const long base = deposit_bits(tile >> P.lognc, P.rmask) * B + ((tile & cmask) << P.logC);
#pragma unroll
				for (int k = 0; k < kMaxSlots; k++) {
					if (k < io.nk) {
#pragma unroll
						for (int i = 0; i < EPL; i++) { useed[k * EPL + i] = Ud[base + io.goff[k] + i]; }
					}
				}
			}
			if constexpr (SEED) {
				double tr_re = 0, tr_im = 0;
#pragma unroll
				for (int k = 0; k < kMaxSlots; k++) {
					if (k < io.nk) {
#pragma unroll
						for (int i = 0; i < EPL; i++) {
							const int e = (threadIdx.x + k * kTileThreads) * EPL + i;   // the slot's elements
							cx<T>* ph = state(0, NS / 2) + e;
							cx<double> u;
							if constexpr (sizeof(T) == 4) { u = useed[k * EPL + i]; }
							else { u = cx<double>{(double)ph->x, (double)ph->y}; }
							const cx<double> ps = {(double)state(0, 0)[e].x, (double)state(0, 0)[e].y};
							const cx<double> sd = {-u.x, u.y};
							*ph = cx<T>{(T)sd.x, (T)sd.y};
							tr_re += sd.x * ps.x - sd.y * ps.y;
							tr_im += sd.x * ps.y + sd.y * ps.x;
						}
					}
				}
#pragma unroll
				for (int o = 32; o > 0; o >>= 1) {
					tr_re += __shfl_xor(tr_re, o);
					tr_im += __shfl_xor(tr_im, o);
				}
				if (lane == 0) {
					tracc[0] += tr_re;
					tracc[1] += tr_im;
				}
				__syncthreads();
			}
			run_tile(0);
			__syncthreads();
			// each thread stores the slots that it then refills
			store(tile, 0);
			const long next = next_tile(tile + gridDim.x);
			if (next < ntiles) { prefetch(next, 0); }
			tile = next;
		}
		if constexpr (SEED) {
			if (lane == 0) {
				const long w = (long)blockIdx.x * kTileWaves + (threadIdx.x >> 6);
				trace[2 * w] = tracc[0];
				trace[2 * w + 1] = tracc[1];
			}
		}
	}
	const long part = (long)blockIdx.x * kTileWaves + (threadIdx.x >> 6);
#pragma unroll
	for (int j = 0; j < kMaxTileGates; j++) {
		if (j < P.ng) {
			double* d = partial + (part * P.ng + j) * 256;
			if constexpr (sizeof(T) == 8) {
				// the four blocks' sums (lane bits 2, 3) into the 16 x 16 layout: A1 and A2 as
				// the (re, re), (re, im) blocks of phi (x) psi and phi (x) dpsi, zeros elsewhere
				double4_t v = acc[j];
#pragma unroll
				for (int q = 0; q < 4; q++) {
					v[q] += __shfl_xor(v[q], 4);
					v[q] += __shfl_xor(v[q], 8);
				}
				if (((lane >> 2) & 3) == 0) {
#pragma unroll
					for (int q = 0; q < 4; q++) { d[(lane >> 4) * 16 + 4 * q + (lane & 3)] = v[q]; }
				}
#pragma unroll
				for (int z = 1; z < 4; z++) { d[64 * z + lane] = 0.0; }
			} else {
#pragma unroll
				for (int v = 0; v < 4; v++) { d[mfma_row<T>(lane, v) * 16 + (lane & 15)] = acc[j][v]; }
			}
		}
	}
}

// gate indices of a pass for the reduction
struct TileGates
{
	int k[kMaxTileGates];
};

// First stage of the partial-sum reduction: chunk c of nchunks sums the parts
// [c * ceil(nparts / nchunks), ...) in order, for every gate j (grid ng x nchunks);
// out has the partials' layout with nchunks parts.
constexpr int kReduceChunks = 16;

__global__ __launch_bounds__(256) void k_tile_reduce_chunks(const double* __restrict__ partial, long nparts, int ng,
	int nchunks, double* __restrict__ out, const int* skip)
{
	if (skipped(skip)) { return; }
	const int j = blockIdx.x, c = blockIdx.y, i = threadIdx.x;
	const long per = (nparts + nchunks - 1) / nchunks;
	const long p0 = (long)c * per, p1 = min(nparts, p0 + per);
	double s = 0;
	for (long p = p0; p < p1; p++) { s += partial[(p * ng + j) * 256 + i]; }
	out[((long)c * ng + j) * 256 + i] = s;
}

// Sum of the outer-product partials over all parts in a fixed order, decoded
// into 16 gradient and (DUAL) 16 HVP entries per gate and added to
// acc[1 + 32 * gate + e].  One workgroup of 256 threads per gate of the pass.
// The partials pair the adjoint at the gate input (phi_in = G^T phi,
// dphi_in = G^T dphi + Z^T phi) with the state before the gate; component
// 4 (2 s + p) + r is array s (primary / tangent; plain sweeps: the quad of the
// pair), re / im p, amplitude r.  With A1 = sum phi_in (x) psi,
// A2 = sum phi_in (x) dpsi, A3 = sum dphi_in (x) psi and Q = G^-T, the gate's
// sums over the adjoint at its output are
//   grad = Q A1,   hess = Q (A2 + A3) + W^T A1   (W = -G^-1 Z G^-1);
// rm holds Q and W^T of the pass's gates (FP64, 32 entries per gate).
__global__ __launch_bounds__(256) void k_tile_reduce(const double* __restrict__ partial, long nparts, int ng,
	const TileGates gk, const zc* __restrict__ rm, int dual, zc* __restrict__ acc, const int* skip)
{
	if (skipped(skip)) { return; }
	__shared__ double Dm[16][16];
	__shared__ zc A[2][16];
	const int j = blockIdx.x, i = threadIdx.x;
	double s = 0;
	for (long p = 0; p < nparts; p++) { s += partial[(p * ng + j) * 256 + i]; }
	Dm[i >> 4][i & 15] = s;
	__syncthreads();
	const int which = (i >> 4) & 1, r = (i >> 2) & 3, c = i & 3;
	if (i < 32) {
		auto M = [&](int sidx, int rr, int pp) { return 4 * (2 * sidx + pp) + rr; };
		// sum phi_r(s) * psi_c(s') over the listed block pairs
		auto blk = [&](int sa, int sb, double& re, double& im) {
			re += Dm[M(sa, r, 0)][M(sb, c, 0)] - Dm[M(sa, r, 1)][M(sb, c, 1)];
			im += Dm[M(sa, r, 0)][M(sb, c, 1)] + Dm[M(sa, r, 1)][M(sb, c, 0)];
		};
		double re = 0, im = 0;
		if (dual) {
			if (!which) { blk(0, 0, re, im); }
			else { blk(0, 1, re, im); blk(1, 0, re, im); }
		} else if (!which) {
			blk(0, 0, re, im);
			blk(1, 1, re, im);
		}
		A[which][4 * r + c] = zc{re, im};
	}
	__syncthreads();
	if (i < 32 && (dual || !which)) {
		const zc* Q = rm + 32L * j;
		const zc* WT = Q + 16;
		zc v = {0, 0};
		for (int k = 0; k < 4; k++) { cmac(v, Q[4 * r + k], A[which][4 * k + c]); }
		if (which) {
			for (int k = 0; k < 4; k++) { cmac(v, WT[4 * r + k], A[0][4 * k + c]); }
		}
		zc* a = acc + 1 + 32L * gk.k[j] + 16 * which + 4 * r + c;
		a->x += v.x;
		a->y += v.y;
	}
}

}  // namespace rqc
