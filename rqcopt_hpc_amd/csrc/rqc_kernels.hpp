// HIP kernels of the gate-sweep engine (gfx950).
//
// Data layout of a column block: X[a * B + c] — amplitude (row) index a over
// the 2^L circuit qubits, column c over B consecutive basis columns, column
// fastest.  A gate on wires (i, j) touches rows a with stride 2^(L-1-i) and
// 2^(L-1-j), so every load and store of a wavefront covers 64 consecutive
// columns: fully coalesced for any wire pair, including the wrap pair
// (L-1, 0) of a periodic brickwall.  The same layout is the reference's
// interleaved statevector_array (data[amp * nstates + k]).
//
// All reductions (trace, gate gradients, Hessian-vector products) accumulate
// in FP64, first per thread, then by wavefront shuffles and an LDS tree in a
// fixed order; per-workgroup partial sums are combined by a second kernel in
// workgroup order.  No atomics: results are bitwise reproducible run to run.
#pragma once

#include "zc.hpp"

namespace rqc {

constexpr int kBlock = 256;    // threads per workgroup of the sweep kernels

// skip flag of a batched launch (a device word read when the kernel starts):
// the whole grid returns at once when it is set
__device__ __forceinline__ bool skipped(const int* skip) { return skip != nullptr && *skip != 0; }
constexpr int kWidth = 32;     // partial-sum width of the reverse kernel (16 grad + 16 hess)

template <typename T>
struct Gate4
{
	cx<T> g[16];
};

// bit positions of the gate wires: wire i -> bit (L-1-i) carries the more
// significant index of the 4x4 gate, wire j -> bit (L-1-j) the less significant
struct Wires
{
	int bi, bj;
};

// index of quad q with zero bits inserted at positions lo < hi
__device__ __forceinline__ long quad_base(long q, int lo, int hi)
{
	long x = ((q >> lo) << (lo + 1)) | (q & ((1L << lo) - 1));
	x = ((x >> hi) << (hi + 1)) | (x & ((1L << hi) - 1));
	return x;
}

// element offsets of the 4 amplitudes (|00>, |01>, |10>, |11> of (wire i, wire j))
// of work item t = (quad, column)
struct Quad
{
	long off[4];
};

__device__ __forceinline__ Quad quad_offsets(long t, long B, int logB, Wires w)
{
	long q, c;
	if (logB >= 0) { q = t >> logB; c = t & (B - 1); }
	else { q = t / B; c = t - q * B; }
	const int lo = min(w.bi, w.bj), hi = max(w.bi, w.bj);
	const long base = quad_base(q, lo, hi);
	const long si = 1L << w.bi, sj = 1L << w.bj;
	Quad r;
	r.off[0] = base * B + c;
	r.off[1] = (base + sj) * B + c;
	r.off[2] = (base + si) * B + c;
	r.off[3] = (base + si + sj) * B + c;
	return r;
}

// y = G x for one quad
template <typename T>
__device__ __forceinline__ void gate_apply(const Gate4<T>& G, const cx<T>* x, cx<T>* y)
{
#pragma unroll
	for (int r = 0; r < 4; r++) {
		cx<T> s = {0, 0};
#pragma unroll
		for (int k = 0; k < 4; k++) { cmac(s, G.g[4*r + k], x[k]); }
		y[r] = s;
	}
}

// y += G^T x for one quad
template <typename T>
__device__ __forceinline__ void gate_apply_t_acc(const Gate4<T>& G, const cx<T>* x, cx<T>* y)
{
#pragma unroll
	for (int c = 0; c < 4; c++) {
#pragma unroll
		for (int r = 0; r < 4; r++) { cmac(y[c], G.g[4*r + c], x[r]); }
	}
}

// ---------------------------------------------------------------------------
// forward sweep of one gate: out = G in; with DUAL also dout = G din + Z in
// (din treated as zero when dzero is set)
template <typename T, bool DUAL>
__global__ __launch_bounds__(kBlock) void k_gate_forward(
	const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
	const cx<T>* __restrict__ din, cx<T>* __restrict__ dout,
	Gate4<T> G, Gate4<T> Z, Wires w, long B, int logB, long nitems, int dzero)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t >= nitems) { return; }
	const Quad o = quad_offsets(t, B, logB, w);
	cx<T> x[4], y[4];
#pragma unroll
	for (int k = 0; k < 4; k++) { x[k] = in[o.off[k]]; }
	gate_apply(G, x, y);
#pragma unroll
	for (int k = 0; k < 4; k++) { out[o.off[k]] = y[k]; }
	if (DUAL) {
		cx<T> dy[4];
		if (dzero) {
			gate_apply(Z, x, dy);
		} else {
			cx<T> dx[4];
#pragma unroll
			for (int k = 0; k < 4; k++) { dx[k] = din[o.off[k]]; }
			gate_apply(G, dx, dy);
#pragma unroll
			for (int r = 0; r < 4; r++) {
#pragma unroll
				for (int k = 0; k < 4; k++) { cmac(dy[r], Z.g[4*r + k], x[k]); }
			}
		}
#pragma unroll
		for (int k = 0; k < 4; k++) { dout[o.off[k]] = dy[k]; }
	}
}

// ---------------------------------------------------------------------------
// block-wide fixed-order sum of NV complex FP64 values per thread; thread 0..NV-1
// of the block end up holding... the results are written to out[0..NV)
template <int NV>
__device__ __forceinline__ void block_sum_store(zc* v, zc* out)
{
	__shared__ double red[kBlock / 64][NV][2];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
	for (int e = 0; e < NV; e++) {
		double re = v[e].x, im = v[e].y;
#pragma unroll
		for (int s = 32; s >= 1; s >>= 1) {
			re += __shfl_xor(re, s, 64);
			im += __shfl_xor(im, s, 64);
		}
		if (lane == 0) { red[wave][e][0] = re; red[wave][e][1] = im; }
	}
	__syncthreads();
	if (threadIdx.x < NV) {
		double re = 0, im = 0;
		for (int k = 0; k < kBlock / 64; k++) { re += red[k][threadIdx.x][0]; im += red[k][threadIdx.x][1]; }
		out[threadIdx.x] = {re, im};
	}
}

// ---------------------------------------------------------------------------
// reverse sweep of one gate (adjoint state phi after the gate, psi before it):
//   grad_rc += phi_r psi_c                                    (16 entries)
//   hess_rc += phi_r dpsi_c + dphi_r psi_c           (DUAL)   (16 entries)
//   phi_out  = G^T phi,  dphi_out = G^T dphi + Z^T phi (DUAL)
// flags: 1 = dpsi is zero, 2 = dphi is zero, 4 = do not write phi_out/dphi_out
template <typename T, bool DUAL>
__global__ __launch_bounds__(kBlock) void k_gate_reverse(
	const cx<T>* __restrict__ psi, const cx<T>* __restrict__ dpsi,
	const cx<T>* __restrict__ phi, const cx<T>* __restrict__ dphi,
	cx<T>* __restrict__ phi_out, cx<T>* __restrict__ dphi_out,
	Gate4<T> G, Gate4<T> Z, Wires w, long B, int logB, long nitems, int flags, zc* __restrict__ partial)
{
	zc ga[16], ha[DUAL ? 16 : 1];
#pragma unroll
	for (int e = 0; e < 16; e++) { ga[e] = {0, 0}; }
	if (DUAL) {
#pragma unroll
		for (int e = 0; e < 16; e++) { ha[e] = {0, 0}; }
	}
	const bool dpsi_zero = flags & 1, dphi_zero = flags & 2, no_write = flags & 4;
	const long stride = (long)gridDim.x * kBlock;
	for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < nitems; t += stride) {
		const Quad o = quad_offsets(t, B, logB, w);
		cx<T> x[4], f[4];
#pragma unroll
		for (int k = 0; k < 4; k++) { x[k] = psi[o.off[k]]; f[k] = phi[o.off[k]]; }
		zc xz[4], fz[4];
#pragma unroll
		for (int k = 0; k < 4; k++) { xz[k] = to_z(x[k]); fz[k] = to_z(f[k]); }
#pragma unroll
		for (int r = 0; r < 4; r++) {
#pragma unroll
			for (int c = 0; c < 4; c++) { cmac(ga[4*r + c], fz[r], xz[c]); }
		}
		if (DUAL) {
			cx<T> dx[4], df[4];
#pragma unroll
			for (int k = 0; k < 4; k++) {
				dx[k] = dpsi_zero ? cx<T>{0, 0} : dpsi[o.off[k]];
				df[k] = dphi_zero ? cx<T>{0, 0} : dphi[o.off[k]];
			}
			if (!dpsi_zero) {
#pragma unroll
				for (int r = 0; r < 4; r++) {
#pragma unroll
					for (int c = 0; c < 4; c++) { cmac(ha[4*r + c], fz[r], to_z(dx[c])); }
				}
			}
			if (!dphi_zero) {
#pragma unroll
				for (int r = 0; r < 4; r++) {
#pragma unroll
					for (int c = 0; c < 4; c++) { cmac(ha[4*r + c], to_z(df[r]), xz[c]); }
				}
			}
			if (!no_write) {
				cx<T> dy[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
				if (!dphi_zero) { gate_apply_t_acc(G, df, dy); }
				gate_apply_t_acc(Z, f, dy);
#pragma unroll
				for (int k = 0; k < 4; k++) { dphi_out[o.off[k]] = dy[k]; }
			}
		}
		if (!no_write) {
			cx<T> y[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
			gate_apply_t_acc(G, f, y);
#pragma unroll
			for (int k = 0; k < 4; k++) { phi_out[o.off[k]] = y[k]; }
		}
	}
	zc* out = partial + (long)blockIdx.x * kWidth;
	if (DUAL) {
		zc all[32];
#pragma unroll
		for (int e = 0; e < 16; e++) { all[e] = ga[e]; all[16 + e] = ha[e]; }
		block_sum_store<32>(all, out);
	} else {
		block_sum_store<16>(ga, out);
		if (threadIdx.x < 16) { out[16 + threadIdx.x] = {0, 0}; }
	}
}

// ---------------------------------------------------------------------------
// Single-qubit gates.  A 2x2 gate V on wire i acts on the pairs of rows that
// differ in bit b = L-1-i (the same qubit-to-bit map as the two-qubit gates).
// Work item t = (pair, column): the pair's row with bit b clear is the pair
// index with a zero bit inserted at b.
template <typename T>
struct Gate2
{
	cx<T> g[4];
};

__device__ __forceinline__ void pair_offsets(long t, long B, int logB, int b, long* o0, long* o1)
{
	long q, c;
	if (logB >= 0) { q = t >> logB; c = t & (B - 1); }
	else { q = t / B; c = t - q * B; }
	const long base = ((q >> b) << (b + 1)) | (q & ((1L << b) - 1));
	*o0 = base * B + c;
	*o1 = (base + (1L << b)) * B + c;
}

// out = V in on wire bit b (in and out may be the same array)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gate1_forward(const cx<T>* in, cx<T>* out, Gate2<T> V, int b, long B,
	int logB, long nitems)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t >= nitems) { return; }
	long o0, o1;
	pair_offsets(t, B, logB, b, &o0, &o1);
	const cx<T> x0 = in[o0], x1 = in[o1];
	cx<T> y0 = {0, 0}, y1 = {0, 0};
	cmac(y0, V.g[0], x0);
	cmac(y0, V.g[1], x1);
	cmac(y1, V.g[2], x0);
	cmac(y1, V.g[3], x1);
	out[o0] = y0;
	out[o1] = y1;
}

// reverse step of a single-qubit gate (phi after the gate, psi before it):
//   dV_rc += phi_r psi_c (FP64 partial sums per workgroup, 4 entries),  phi_out = V^T phi
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gate1_reverse(const cx<T>* __restrict__ psi, const cx<T>* __restrict__ phi,
	cx<T>* __restrict__ phi_out, Gate2<T> V, int b, long B, int logB, long nitems, zc* __restrict__ partial)
{
	zc ga[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
	const long stride = (long)gridDim.x * kBlock;
	for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < nitems; t += stride) {
		long o0, o1;
		pair_offsets(t, B, logB, b, &o0, &o1);
		const cx<T> x[2] = {psi[o0], psi[o1]};
		const cx<T> f[2] = {phi[o0], phi[o1]};
#pragma unroll
		for (int r = 0; r < 2; r++) {
#pragma unroll
			for (int c = 0; c < 2; c++) { cmac(ga[2 * r + c], to_z(f[r]), to_z(x[c])); }
		}
		cx<T> y0 = {0, 0}, y1 = {0, 0};
		cmac(y0, V.g[0], f[0]);
		cmac(y0, V.g[2], f[1]);
		cmac(y1, V.g[1], f[0]);
		cmac(y1, V.g[3], f[1]);
		phi_out[o0] = y0;
		phi_out[o1] = y1;
	}
	block_sum_store<4>(ga, partial + (long)blockIdx.x * 4);
}

// ---------------------------------------------------------------------------
// trace of the column block: f += sum_{a,c} -conj(U[a, b0+c]) psi[a, c]; when
// phi is given the adjoint seed -conj(U[:, b0+c]) is also stored there
template <typename T>
__global__ __launch_bounds__(kBlock) void k_target_trace(
	const zc* __restrict__ U, long ldu, long b0, const cx<T>* __restrict__ psi, cx<T>* __restrict__ phi,
	long B, int logB, long nitems, zc* __restrict__ partial, const int* skip)
{
	if (skipped(skip)) { return; }
	zc acc = {0, 0};
	const long stride = (long)gridDim.x * kBlock;
	for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < nitems; t += stride) {
		long a, c;
		if (logB >= 0) { a = t >> logB; c = t & (B - 1); }
		else { a = t / B; c = t - a * B; }
		const zc u = U[a * ldu + b0 + c];
		const zc s = {-u.x, u.y};
		cmac(acc, s, to_z(psi[t]));
		if (phi != nullptr) { phi[t] = cx<T>{(T)s.x, (T)s.y}; }
	}
	block_sum_store<1>(&acc, partial + blockIdx.x);
}

// Block-encoding residual (reference src/target.c:1162-1180) over the columns
// b0 .. b0 + B - 1 of the embedded basis P e_b, where P puts bit i of an
// index of L/2 qubits on bit 2i + 1 (the even bits |0>).  The embedded rows
// c(a) are those with no even bit set: there d = psi - H[a][b], |d|^2 / 2 goes
// to the partial sums and the adjoint seed is conj(d); all other entries, and
// the padding columns b >= 2^(L/2), seed zero.  With J (HVP sweeps) the
// projected tangent P^dagger dpsi of the sweep's direction is stored at
// J[b 2^(L/2) + a].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_blockenc_residual(const zc* __restrict__ H, int half, long b0,
	const cx<T>* __restrict__ psi, cx<T>* __restrict__ phi, const cx<T>* __restrict__ dpsi, zc* __restrict__ J, long B,
	int logB, long nitems, zc* __restrict__ partial, const int* skip)
{
	if (skipped(skip)) { return; }
	const long s = 1L << half;
	const long even = 0x5555555555555555L & ((1L << (2 * half)) - 1);
	double acc = 0;
	const long stride = (long)gridDim.x * kBlock;
	for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < nitems; t += stride) {
		long row, c;
		if (logB >= 0) { row = t >> logB; c = t & (B - 1); }
		else { row = t / B; c = t - row * B; }
		const long b = b0 + c;
		cx<T> seed = {(T)0, (T)0};
		if (b < s && (row & even) == 0) {
			long a = 0;
			for (int i = 0; i < half; i++) { a |= ((row >> (2 * i + 1)) & 1) << i; }
			const zc h = H[a * s + b];
			const zc p = to_z(psi[t]);
			const double dx = p.x - h.x, dy = p.y - h.y;
			acc += 0.5 * (dx * dx + dy * dy);
			seed = cx<T>{(T)dx, (T)(-dy)};
			if (J != nullptr) { J[b * s + a] = to_z(dpsi[t]); }
		}
		if (phi != nullptr) { phi[t] = seed; }
	}
	zc r = {acc, 0};
	block_sum_store<1>(&r, partial + blockIdx.x);
}

// the embedded basis columns P e_b as blocks of N x B amplitudes (column b in
// block b / B), zero past b = 2^(L/2) - 1; sz = N B, nitems = blocks * sz
template <typename T>
__global__ __launch_bounds__(kBlock) void k_blockenc_columns(cx<T>* __restrict__ X, int half, long B, int logB, long sz,
	long nitems)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t >= nitems) { return; }
	const long blk = t / sz, r = t - blk * sz;
	long row, c;
	if (logB >= 0) { row = r >> logB; c = r & (B - 1); }
	else { row = r / B; c = r - row * B; }
	const long b = blk * B + c;
	long e = 0;
	for (int i = 0; i < half; i++) { e |= ((b >> i) & 1) << (2 * i + 1); }
	X[t] = cx<T>{(b < (1L << half) && row == e) ? (T)1 : (T)0, (T)0};
}

// out[x][y] = sum_t conj(J_x[t]) J_y[t] over n entries for x >= y in a fixed
// order, out[y][x] its conjugate and the diagonal real: an exactly Hermitian
// Gram matrix of the m columns J_x = J + x n; one workgroup per pair
__global__ __launch_bounds__(kBlock) void k_gram_hermitian(const zc* __restrict__ J, long n, int m, zc* __restrict__ out)
{
	const int x = blockIdx.x, y = blockIdx.y;
	if (y > x) { return; }
	const zc* jx = J + (long)x * n;
	const zc* jy = J + (long)y * n;
	zc acc = {0, 0};
	for (long t = threadIdx.x; t < n; t += kBlock) { cmac(acc, zc{jx[t].x, -jx[t].y}, jy[t]); }
	__shared__ zc res[1];
	block_sum_store<1>(&acc, res);
	__syncthreads();
	if (threadIdx.x == 0) {
		zc v = res[0];
		if (x == y) { v.y = 0; }
		out[(long)x * m + y] = v;
		out[(long)y * m + x] = zc{v.x, -v.y};
	}
}

// acc[e] += sum_p partial[p * width + e] in a fixed order; one workgroup per e
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const zc* __restrict__ partial, int nparts, int width,
	zc* __restrict__ acc, const int* skip)
{
	if (skipped(skip)) { return; }
	const int e = blockIdx.x;
	zc s = {0, 0};
	for (int p = threadIdx.x; p < nparts; p += kBlock) { s += partial[(long)p * width + e]; }
	zc r[1] = {s};
	__shared__ zc res[1];
	block_sum_store<1>(r, res);
	__syncthreads();
	if (threadIdx.x == 0) { acc[e] += res[0]; }
}

// X[a * B + c] = (a == b0 + c)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_identity_block(cx<T>* __restrict__ X, long B, int logB, long b0, long nitems)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t >= nitems) { return; }
	long a, c;
	if (logB >= 0) { a = t >> logB; c = t & (B - 1); }
	else { a = t / B; c = t - a * B; }
	X[t] = cx<T>{(a == b0 + c) ? (T)1 : (T)0, (T)0};
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(zc* __restrict__ p, long n, const int* skip)
{
	if (skipped(skip)) { return; }
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t < n) { p[t] = {0, 0}; }
}

// Gate placeholder (reference src/gate.c:374-460): the 16 states E_rc psi, r, c
// in the gate basis of (wire i, wire j), interleaved as out[a * 16 + 4 r + c]:
// amplitude a is psi at a's index with its (i, j) bits set to c when those bits
// equal r, else zero.  One thread per output element.
__global__ __launch_bounds__(kBlock) void k_gate_placeholder(const zc* __restrict__ psi, zc* __restrict__ out, Wires w,
	long n16)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t >= n16) { return; }
	const long a = t >> 4;
	const int s = (int)(t & 15), r = s >> 2, c = s & 3;
	const int ra = (int)((((a >> w.bi) & 1) << 1) | ((a >> w.bj) & 1));
	if (ra != r) { out[t] = {0, 0}; return; }
	const long src = (a & ~((1L << w.bi) | (1L << w.bj))) | ((long)(c >> 1) << w.bi) | ((long)(c & 1) << w.bj);
	out[t] = psi[src];
}

// Gate gradients of a statevector array (reference src/gate.c:269-366):
// entry 4 r + c of state k = sum over quads of dpsi_out[amp r] * psi[amp c, state k],
// stored (accumulate = 0) or added (accumulate = 1) at out[k * stride + 4 r + c];
// one workgroup per output entry, summed in a fixed order.
__global__ __launch_bounds__(kBlock) void k_gate_backward_array(const zc* __restrict__ psi, const zc* __restrict__ dpsi,
	zc* __restrict__ out, Wires w, long nq, int nstates, long stride, int accumulate)
{
	const int o = blockIdx.x;
	const int k = o >> 4, r = (o >> 2) & 3, c = o & 3;
	const int lo = min(w.bi, w.bj), hi = max(w.bi, w.bj);
	zc acc = {0, 0};
	for (long q = threadIdx.x; q < nq; q += kBlock) {
		const long base = quad_base(q, lo, hi);
		const long ar = base | ((long)(r >> 1) << w.bi) | ((long)(r & 1) << w.bj);
		const long ac = base | ((long)(c >> 1) << w.bi) | ((long)(c & 1) << w.bj);
		cmac(acc, dpsi[ar], psi[ac * nstates + k]);
	}
	__shared__ zc res[1];
	block_sum_store<1>(&acc, res);
	__syncthreads();
	if (threadIdx.x == 0) {
		zc* d = out + (long)k * stride + (o & 15);
		*d = accumulate ? *d + res[0] : res[0];
	}
}

// partial sums of |x|^2 (real parts of the partials; fixed grid, fixed order)
__global__ __launch_bounds__(kBlock) void k_norm2_partials(const zc* __restrict__ x, long n, zc* __restrict__ partial)
{
	zc acc = {0, 0};
	for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < n; t += (long)gridDim.x * kBlock) {
		acc.x += x[t].x * x[t].x + x[t].y * x[t].y;
	}
	block_sum_store<1>(&acc, partial + blockIdx.x);
}

// x /= sqrt(s[0].x) unless it is zero
__global__ __launch_bounds__(kBlock) void k_scale_inv_norm(zc* __restrict__ x, long n, const zc* __restrict__ s)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	const double nrm = sqrt(s[0].x);
	if (t < n && nrm > 0) {
		const double inv = 1.0 / nrm;
		x[t] = zc{x[t].x * inv, x[t].y * inv};
	}
}

// qubit permutation: output bit q of the index reads input bit src[q]
struct BitMap
{
	int src[32];
};

__global__ __launch_bounds__(kBlock) void k_transpose_amplitudes(const zc* __restrict__ in, zc* __restrict__ out, BitMap m,
	int L, long n)
{
	const long j = (long)blockIdx.x * kBlock + threadIdx.x;
	if (j >= n) { return; }
	long k = 0;
	for (int q = 0; q < L; q++) {
		if ((j >> q) & 1) { k |= 1L << m.src[q]; }
	}
	out[j] = in[k];
}

__global__ __launch_bounds__(kBlock) void k_to_float(const zc* __restrict__ x, cx<float>* __restrict__ y, long n)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t < n) { y[t] = cx<float>{(float)x[t].x, (float)x[t].y}; }
}

__global__ __launch_bounds__(kBlock) void k_add(const zc* __restrict__ a, const zc* __restrict__ b, zc* __restrict__ c, long n)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t < n) { c[t] = a[t] + b[t]; }
}

// y = U x for a dense row-major N x N matrix (one row per thread)
__global__ __launch_bounds__(kBlock) void k_matvec(const zc* __restrict__ U, const zc* __restrict__ x, zc* __restrict__ y, long n)
{
	const long a = (long)blockIdx.x * kBlock + threadIdx.x;
	if (a >= n) { return; }
	zc s = {0, 0};
	for (long b = 0; b < n; b++) { cmac(s, U[a * n + b], x[b]); }
	y[a] = s;
}

// plain streaming copy (16-byte elements), used to calibrate HBM counters
__global__ __launch_bounds__(kBlock) void k_copy16(const zc* __restrict__ src, zc* __restrict__ dst, long n)
{
	const long stride = (long)gridDim.x * kBlock;
	for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < n; t += stride) { dst[t] = src[t]; }
}

// streaming copy tuned for HBM bandwidth: four 16-byte elements per thread in flight
// (block-strided, so that each load instruction is coalesced), non-temporal loads and stores
__global__ __launch_bounds__(kBlock) void k_copy_stream(const zc* __restrict__ src, zc* __restrict__ dst, long n)
{
	constexpr int U = 4;
	const long step = (long)gridDim.x * kBlock * U;
	for (long t0 = (long)blockIdx.x * kBlock * U + threadIdx.x; t0 < n; t0 += step) {
		zc v[U];
#pragma unroll
		for (int k = 0; k < U; k++) {
			const long i = t0 + (long)k * kBlock;
			if (i < n) {
				v[k].x = __builtin_nontemporal_load(&src[i].x);
				v[k].y = __builtin_nontemporal_load(&src[i].y);
			}
		}
#pragma unroll
		for (int k = 0; k < U; k++) {
			const long i = t0 + (long)k * kBlock;
			if (i < n) {
				__builtin_nontemporal_store(v[k].x, &dst[i].x);
				__builtin_nontemporal_store(v[k].y, &dst[i].y);
			}
		}
	}
}

// copy column block X (n x B, column fastest) into columns [b0, b0 + B) of the row-major n x n U
__global__ __launch_bounds__(kBlock) void k_store_columns(const zc* __restrict__ X, zc* __restrict__ U, long n, long B, long b0)
{
	const long t = (long)blockIdx.x * kBlock + threadIdx.x;
	if (t >= n * B) { return; }
	const long a = t / B, c = t - a * B;
	U[a * n + b0 + c] = X[t];
}

}  // namespace rqc
