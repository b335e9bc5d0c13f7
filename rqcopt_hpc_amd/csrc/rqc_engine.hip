// Gate-sweep engine of librqcopt_hip and the GPU entry points of its C ABI
// (gate.h, statevector.h, quantum_circuit.h, brickwall_circuit.h, target.h,
// rqc_device.h).
//
// Target evaluation (reference src/target.c:158-436) is restructured for the
// GPU: instead of one basis column at a time, a block of B columns is swept
// at once as an N x B array (N = 2^L), gate by gate:
//   forward  psi_{k+1} = G_k psi_k                     (+ dpsi_{k+1} = G_k dpsi_k + Z_k psi_k)
//   trace    f += sum -conj(U) . psi_G,  phi = -conj(U e_b)
//   reverse  grad_k += phi_k (x) psi_k, phi_{k-1} = G_k^T phi_k
//            (+ hess_k += phi_k (x) dpsi_k + dphi_k (x) psi_k, dphi_{k-1} = G_k^T dphi_k + Z_k^T phi_k)
// which is the reference's per-column forward / backward / tangent sweeps
// (src/quantum_circuit.c:147-236) evaluated for all columns of the block, with
// the tangent-forward folded into the forward sweep (dual numbers).
// The column blocks cover all N columns; B is the largest power of two whose
// checkpoints fit the device workspace.

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <set>
#include <vector>

#include "rqc_internal.hpp"
#include "rqc_eval.hpp"
#include "rqc_kernels.hpp"
#include "rqc_tile_kernels.hpp"
#include "mat4.hpp"
#include "rqcopt/gate.h"
#include "rqcopt/statevector.h"
#include "rqcopt/quantum_circuit.h"
#include "rqcopt/brickwall_circuit.h"
#include "rqcopt/target.h"
#include "rqcopt/rqc_device.h"
#include "rqcopt/rqc_gate1q.h"
#include "rqcopt/rng.h"
#include <rocblas/rocblas.h>

namespace rqc {

// ============================================================== context
// Workspace of the no-allocation evaluation entry point: a file-scope device
// array, allocated with the code object when it loads on a device and held
// until the process ends.  Host code never passes its name to a kernel (on the
// host it is a stand-in, not a device address): the context looks up its
// device address once, with the context's device current, and keeps it.
static __device__ __align__(256) char scratch[kScratchBytes];

void DeviceBuffer::reserve(size_t n)
{
	if (n <= bytes) { return; }
	if (ptr != nullptr) {
		RQC_HIP(hipDeviceSynchronize());
		RQC_HIP(hipFree(ptr));
		ptr = nullptr;
		bytes = 0;
	}
	const size_t r = (n + 4095) & ~(size_t)4095;
	RQC_HIP(hipMalloc(&ptr, r));
	bytes = r;
}

void DeviceBuffer::release()
{
	if (ptr != nullptr) {
		(void)hipFree(ptr);
		ptr = nullptr;
		bytes = 0;
	}
}

Context& ctx()
{
	static Context* c = nullptr;
	static std::mutex init_mu;
	std::lock_guard<std::mutex> lk(init_mu);
	if (c == nullptr) {
		int n = 0;
		if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
			throw Error(-1, "librqcopt_hip: no HIP GPU visible (this library has no CPU path)");
		}
		auto* cc = new Context();
		RQC_HIP(hipGetDevice(&cc->device));
		hipDeviceProp_t p;
		RQC_HIP(hipGetDeviceProperties(&p, cc->device));
		cc->cu_count = p.multiProcessorCount;
		cc->hbm_bytes = p.totalGlobalMem;
		cc->name = p.gcnArchName;
		cc->lds_per_block = p.sharedMemPerBlock;
		const char* eng = getenv("RQC_SWEEP_ENGINE");
		if (eng != nullptr && std::string(eng) == "pergate") { cc->engine = RQC_ENGINE_PER_GATE; }
		const char* at = getenv("RQC_AUTOTUNE");
		if (at != nullptr && std::string(at) == "0") { cc->autotune = false; }
		const char* ea = getenv("RQC_EVAL_AUTOTUNE");
		if (ea != nullptr && std::string(ea) == "1") { cc->eval_autotune = true; }
		RQC_HIP(hipStreamCreateWithFlags(&cc->stream, hipStreamNonBlocking));
		void* sp = nullptr;
		const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
		if (se != hipSuccess || sp == nullptr) {
			throw Error(-1, std::string("librqcopt_hip: hipGetSymbolAddress of the scratch array failed: ") +
				hipGetErrorString(se));
		}
		cc->scratch = static_cast<char*>(sp);
		c = cc;
	}
	return *c;
}

LaunchTimer::LaunchTimer(Slot& sl, int kind, double bytes, double flops, double issued) : s(sl)
{
	if (!s.profile) { return; }
	if (s.events_used == s.events.size()) {
		Slot::Ev e;
		RQC_HIP(hipEventCreate(&e.a));
		RQC_HIP(hipEventCreate(&e.b));
		s.events.push_back(e);
	}
	ev = &s.events[s.events_used++];
	ev->kind = kind;
	ev->bytes = bytes;
	ev->flops = flops;
	ev->issued = issued;
	RQC_HIP(hipEventRecord(ev->a, s.stream));
}

LaunchTimer::~LaunchTimer()
{
	if (ev != nullptr) { (void)hipEventRecord(ev->b, s.stream); }
}

void profile_flush(Context& c)
{
	if (c.events_used == 0) { return; }
	RQC_HIP(hipStreamSynchronize(c.stream));
	for (size_t k = 0; k < c.events_used; k++) {
		float ms = 0;
		RQC_HIP(hipEventElapsedTime(&ms, c.events[k].a, c.events[k].b));
		c.prof_ms[c.events[k].kind] += ms;
		c.prof_n[c.events[k].kind] += 1;
		c.prof_bytes[c.events[k].kind] += c.events[k].bytes;
		c.prof_flops[c.events[k].kind] += c.events[k].flops;
		c.prof_issued[c.events[k].kind] += c.events[k].issued;
	}
	c.events_used = 0;
}

bool is_device_pointer(const void* p)
{
	if (p == nullptr) { return false; }
	hipPointerAttribute_t a;
	if (hipPointerGetAttributes(&a, p) != hipSuccess) {
		(void)hipGetLastError();
		return false;
	}
	return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void fatal(const char* where, const std::exception& e)
{
	fprintf(stderr, "librqcopt_hip: %s failed: %s\n", where, e.what());
	fflush(stderr);
	abort();
}

// ============================================================== helpers
static inline unsigned nblocks(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

static inline int ilog2_or_neg(long b)
{
	if (b <= 0 || (b & (b - 1)) != 0) { return -1; }
	int l = 0;
	while ((1L << l) < b) { l++; }
	return l;
}

template <typename T>
static Gate4<T> to_gate(const zc* g)
{
	Gate4<T> r;
	for (int e = 0; e < 16; e++) { r.g[e] = cx<T>{(T)g[e].x, (T)g[e].y}; }
	return r;
}

static Wires wires_of(int i, int j, int L) { return Wires{L - 1 - i, L - 1 - j}; }

static void check_wires(int i, int j, int L)
{
	if (L < 2 || i < 0 || j < 0 || i >= L || j >= L || i == j) {
		throw Error(-3, "invalid gate wires (" + std::to_string(i) + ", " + std::to_string(j) + ") for " +
			std::to_string(L) + " qubits");
	}
}

// grid of the grid-stride reduction kernels: fixed for given sizes (bitwise reproducibility)
static inline unsigned reduce_grid(long nitems)
{
	const long want = (nitems + 4L * kBlock - 1) / (4L * kBlock);
	return (unsigned)std::max(1L, std::min(want, 4096L));
}

Circuit make_circuit(const struct mat4x4* gates, const struct mat4x4* dirs, int ngates, const int* wires, int L)
{
	if (ngates < 1) { throw Error(-3, "a circuit needs at least one gate"); }
	Circuit C;
	C.L = L;
	C.ngates = ngates;
	C.gates.resize((size_t)ngates * 16);
	std::memcpy(C.gates.data(), gates, (size_t)ngates * 16 * sizeof(zc));
	if (dirs != nullptr) {
		C.dirs.resize((size_t)ngates * 16);
		std::memcpy(C.dirs.data(), dirs, (size_t)ngates * 16 * sizeof(zc));
	}
	for (int k = 0; k < ngates; k++) {
		check_wires(wires[2*k], wires[2*k + 1], L);
		C.wi.push_back(wires[2*k]);
		C.wj.push_back(wires[2*k + 1]);
	}
	return C;
}

// expand a brickwall (reference src/target.c:443-458)
Circuit make_brickwall(const struct mat4x4* vlist, const struct mat4x4* vdirs, int nlayers, const int* const* perms, int L)
{
	if (L < 2 || L % 2 != 0) { throw Error(-3, "brickwall circuits need an even number of qubits >= 2"); }
	if (nlayers < 1) { throw Error(-3, "a brickwall circuit needs at least one layer"); }
	const int per = L / 2;
	std::vector<struct mat4x4> g((size_t)nlayers * per), d;
	std::vector<int> w((size_t)nlayers * per * 2);
	if (vdirs != nullptr) { d.resize((size_t)nlayers * per); }
	for (int i = 0; i < nlayers; i++) {
		std::vector<int> seen(L, 0);
		for (int q = 0; q < L; q++) {
			const int p = perms[i][q];
			if (p < 0 || p >= L || seen[p]) { throw Error(-3, "perms[" + std::to_string(i) + "] is not a permutation"); }
			seen[p] = 1;
		}
		for (int k = 0; k < per; k++) {
			g[(size_t)i * per + k] = vlist[i];
			if (vdirs != nullptr) { d[(size_t)i * per + k] = vdirs[i]; }
			w[2 * ((size_t)i * per + k)] = perms[i][2*k];
			w[2 * ((size_t)i * per + k) + 1] = perms[i][2*k + 1];
		}
	}
	return make_circuit(g.data(), vdirs != nullptr ? d.data() : nullptr, nlayers * per, w.data(), L);
}

// ============================================================== column-block sweeps
static int states_needed(Mode m, int G)
{
	switch (m) {
	case Mode::Target: return 2;
	case Mode::Grad:   return (G + 1) + 2;
	default:           return (G + 1) + G + 4;
	}
}

// largest power-of-two column block whose nstates state arrays fit the budget
// (0 = 85% of the free device memory plus the `held` bytes the caller would release)
static long choose_block(long N, int nstates, size_t elem, size_t budget, size_t held, long Bmax = 0)
{
	if (budget == 0) {
		size_t fr = 0, tot = 0;
		RQC_HIP(hipMemGetInfo(&fr, &tot));
		budget = (size_t)(0.85 * (double)(fr + held));
	}
	long B = (Bmax > 0) ? std::min(N, Bmax) : N;
	while (B > 1 && (size_t)nstates * N * B * elem > budget) { B /= 2; }
	if ((size_t)nstates * N * B * elem > budget) {
		throw Error(-1, "device memory too small for " + std::to_string(nstates) + " statevectors of " +
			std::to_string(N) + " amplitudes");
	}
	return B;
}

// per-gate engine: one gate per pass over the column block.
// acc[0] = f, acc[1 + k*32 + e]: e < 16 gradient of gate k, e >= 16 its HVP
template <typename T>
static void run_gates(Slot& s, const EvalPlan& pl, const Circuit& C, const zc* dU, zc* d_acc, zc* d_partial, cx<T>* S,
	const int* skip, const cx<T>* X, const zc* W, const BlockEnc* be)
{
	const Mode mode = pl.mode;
	const int L = C.L, G = C.ngates;
	const long N = pl.N, B = pl.B;
	const int logB = ilog2_or_neg(B);
	const size_t sz = (size_t)N * B;
	auto psi = [&](int k) { return S + (size_t)k * sz; };
	auto dpsi = [&](int k) { return S + (size_t)(G + k) * sz; };            // k = 1 .. G
	cx<T>* phi[2] = {S + (size_t)(mode == Mode::Hvp ? 2 * G + 1 : G + 1) * sz,
	                 S + (size_t)(mode == Mode::Hvp ? 2 * G + 2 : G + 2) * sz};
	cx<T>* dphi[2] = {S + (size_t)(2 * G + 3) * sz, S + (size_t)(2 * G + 4) * sz};

	std::vector<Gate4<T>> Gt(G), Zt(G);
	for (int k = 0; k < G; k++) {
		Gt[k] = to_gate<T>(&C.gates[16 * (size_t)k]);
		Zt[k] = (mode == Mode::Hvp) ? to_gate<T>(&C.dirs[16 * (size_t)k]) : Gate4<T>{};
	}
	const long nq = (N / 4) * B;
	const unsigned grev = pl.grev_pg;
	const unsigned gtr = pl.gtr;
	const double ab = (double)sz * sizeof(cx<T>);   // bytes of one state block
	const bool hvp = (mode == Mode::Hvp);

	for (long blk = 0; blk < pl.nblk; blk++) {
		const long b0 = blk * B;
		// the block's input columns and the trace's seed columns: basis columns and U, or samples and U X
		const zc* seed = (be != nullptr) ? nullptr : pl.sampled ? W + (size_t)blk * sz : dU;
		const long ldu = pl.sampled ? B : N, u0 = pl.sampled ? 0 : b0;
		if (pl.sampled) {
			RQC_HIP(hipMemcpyAsync(psi(0), X + (size_t)blk * sz, sz * sizeof(cx<T>), hipMemcpyDeviceToDevice, s.stream));
		} else {
			LaunchTimer lt(s, RQC_KERNEL_OTHER, ab);
			k_identity_block<T><<<nblocks(N * B), kBlock, 0, s.stream>>>(psi(0), B, logB, b0, N * B);
		}
		// forward sweep
		for (int k = 0; k < G; k++) {
			const Wires w = wires_of(C.wi[k], C.wj[k], L);
			cx<T>* in = (mode == Mode::Target) ? psi(k % 2) : psi(k);
			cx<T>* out = (mode == Mode::Target) ? psi((k + 1) % 2) : psi(k + 1);
			LaunchTimer lt(s, RQC_KERNEL_FORWARD, hvp ? (3 + (k > 0)) * ab : 2 * ab);
			if (mode == Mode::Hvp) {
				k_gate_forward<T, true><<<nblocks(nq), kBlock, 0, s.stream>>>(in, out, k > 0 ? dpsi(k) : nullptr, dpsi(k + 1),
					Gt[k], Zt[k], w, B, logB, nq, k == 0 ? 1 : 0);
			} else {
				k_gate_forward<T, false><<<nblocks(nq), kBlock, 0, s.stream>>>(in, out, nullptr, nullptr,
					Gt[k], Gt[k], w, B, logB, nq, 0);
			}
		}
		cx<T>* psiG = (mode == Mode::Target) ? psi(G % 2) : psi(G);
		if (be != nullptr) {
			LaunchTimer lt(s, RQC_KERNEL_TARGET, (mode == Mode::Target ? 1 : 2) * ab);
			k_blockenc_residual<T><<<gtr, kBlock, 0, s.stream>>>(be->H, L / 2, b0, psiG,
				mode == Mode::Target ? nullptr : phi[0], hvp ? dpsi(G) : nullptr, hvp ? be->J : nullptr, B, logB, N * B,
				d_partial, skip);
		} else {
			LaunchTimer lt(s, RQC_KERNEL_TARGET, (double)sz * sizeof(zc) + (mode == Mode::Target ? 1 : 2) * ab);
			k_target_trace<T><<<gtr, kBlock, 0, s.stream>>>(seed, ldu, u0, psiG, mode == Mode::Target ? nullptr : phi[0],
				B, logB, N * B, d_partial, skip);
		}
		{
			LaunchTimer lt(s, RQC_KERNEL_OTHER);
			k_reduce_partials<<<1, kBlock, 0, s.stream>>>(d_partial, (int)gtr, 1, d_acc, skip);
		}
		if (mode == Mode::Target) { continue; }
		// reverse sweep
		int cur = 0;
		for (int k = G - 1; k >= 0; k--) {
			const Wires w = wires_of(C.wi[k], C.wj[k], L);
			int flags = (k == 0) ? 4 : 0;
			{
				LaunchTimer lt(s, RQC_KERNEL_REVERSE, hvp ? (2 + (k > 0) + (k < G - 1) + (k > 0 ? 2 : 0)) * ab
				                                          : (2 + (k > 0)) * ab);
				if (mode == Mode::Hvp) {
					if (k == 0) { flags |= 1; }
					if (k == G - 1) { flags |= 2; }
					k_gate_reverse<T, true><<<grev, kBlock, 0, s.stream>>>(psi(k), k > 0 ? dpsi(k) : nullptr,
						phi[cur], dphi[cur], phi[1 - cur], dphi[1 - cur], Gt[k], Zt[k], w, B, logB, nq, flags, d_partial);
				} else {
					k_gate_reverse<T, false><<<grev, kBlock, 0, s.stream>>>(psi(k), nullptr,
						phi[cur], nullptr, phi[1 - cur], nullptr, Gt[k], Gt[k], w, B, logB, nq, flags, d_partial);
				}
			}
			{
				LaunchTimer lt(s, RQC_KERNEL_OTHER);
				k_reduce_partials<<<kWidth, kBlock, 0, s.stream>>>(d_partial, (int)grev, kWidth, d_acc + 1 + 32L * k, skip);
			}
			cur = 1 - cur;
		}
	}
}

// ---------------------------------------------------------------- tile passes
struct Pass
{
	std::vector<int> gates;  // circuit gate indices, in execution order
	std::vector<int> rb;     // tile row bits, ascending
};

// Pass planning over the circuit's dependency order: gates on disjoint wires
// commute, so a pass may take any gate whose predecessors on its two wires are
// done, as long as its wires lie in the pass's K tile qubits.  Candidate tile
// sets for a pass: every window of K consecutive wires (cyclic) and the wires of
// the ready gates in circuit order; a candidate takes its ready gates in circuit
// order, at most gmax.  A beam search over these choices keeps, after each pass,
// the kPlanBeam states with the most gates done (ties to the earlier state and
// candidate), and the first plan to finish has the fewest passes found; width 1
// is the greedy plan (the candidate that takes the most gates).  At L = 14,
// depth 9, K = 6 and 8 gates the beam finds 9 passes where the greedy plan
// needs 10 (its last passes take 3 and 1 gates).  Tiles are padded with the
// lowest unused row bits.  Any topological order of the gates gives the same
// states at each gate up to rounding, so the sweeps of the passes compute the
// reference's per-gate quantities.
static std::vector<Pass> plan_passes(const Circuit& C, int K, int gmax)
{
	const int L = C.L, G = C.ngates;
	auto bitmask = [&](int k) { return (1u << (L - 1 - C.wi[k])) | (1u << (L - 1 - C.wj[k])); };
	std::vector<int> prev_i(G, -1), prev_j(G, -1), last(L, -1);
	for (int k = 0; k < G; k++) {
		prev_i[k] = last[C.wi[k]];
		prev_j[k] = last[C.wj[k]];
		last[C.wi[k]] = k;
		last[C.wj[k]] = k;
	}
	const int beam_width = (G <= 256) ? 32 : (G <= 2048) ? 4 : 1;
	std::vector<char> taken(G, 0);
	auto ok = [&](const std::vector<char>& done, int k) {
		return (prev_i[k] < 0 || done[prev_i[k]] || taken[prev_i[k]]) &&
		       (prev_j[k] < 0 || done[prev_j[k]] || taken[prev_j[k]]);
	};
	// the ready gates of the tile set `mask`, in circuit order, at most gmax
	auto take = [&](const std::vector<char>& done, unsigned mask, std::vector<int>& out) {
		std::fill(taken.begin(), taken.end(), 0);
		out.clear();
		for (int k = 0; k < G && (int)out.size() < gmax; k++) {
			if (done[k] || (bitmask(k) & ~mask) != 0 || !ok(done, k)) { continue; }
			taken[k] = 1;
			out.push_back(k);
		}
		std::fill(taken.begin(), taken.end(), 0);
	};
	auto candidates = [&](const std::vector<char>& done) {
		std::vector<unsigned> cands;
		if (K >= L) {
			cands.push_back((L >= 32) ? ~0u : ((1u << L) - 1));
			return cands;
		}
		for (int s = 0; s < L; s++) {
			unsigned m = 0;
			for (int t = 0; t < K; t++) { m |= 1u << (L - 1 - (s + t) % L); }
			cands.push_back(m);
		}
		unsigned m = 0;
		for (int k = 0; k < G; k++) {
			if (done[k] || !ok(done, k)) { continue; }
			const unsigned m2 = m | bitmask(k);
			if (__builtin_popcount(m2) <= K) { m = m2; }
		}
		cands.push_back(m);
		return cands;
	};
	struct State
	{
		std::vector<char> done;
		int ndone;
		std::vector<std::vector<int>> passes;
	};
	std::vector<State> beam(1);
	beam[0].done.assign(G, 0);
	beam[0].ndone = 0;
	std::vector<std::vector<int>> plan;
	if (G == 0) { beam.clear(); }
	std::vector<int> lst;
	while (!beam.empty()) {
		std::vector<State> next;
		std::set<std::vector<char>> seen;
		bool finished = false;
		for (const State& st : beam) {
			// expansions in candidate order, the one taking the most gates first
			std::vector<std::pair<int, std::vector<int>>> exps;
			for (unsigned m : candidates(st.done)) {
				take(st.done, m, lst);
				if (!lst.empty()) { exps.emplace_back((int)lst.size(), lst); }
			}
			std::stable_sort(exps.begin(), exps.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
			for (auto& e : exps) {
				State t;
				t.done = st.done;
				for (int k : e.second) { t.done[k] = 1; }
				if (!seen.insert(t.done).second) { continue; }
				t.ndone = st.ndone + e.first;
				t.passes = st.passes;
				t.passes.push_back(std::move(e.second));
				if (t.ndone == G) {
					plan = std::move(t.passes);
					finished = true;
					break;
				}
				next.push_back(std::move(t));
			}
			if (finished) { break; }
		}
		if (finished) { break; }
		if (next.empty()) { throw Error(-4, "pass planning made no progress"); }
		std::stable_sort(next.begin(), next.end(), [](const State& a, const State& b) { return a.ndone > b.ndone; });
		if ((int)next.size() > beam_width) { next.resize(beam_width); }
		beam = std::move(next);
	}
	std::vector<Pass> out;
	for (auto& gl : plan) {
		Pass p;
		p.gates = gl;
		unsigned used = 0;
		for (int k : p.gates) { used |= bitmask(k); }
		for (int b = 0; b < L; b++) {
			if (used & (1u << b)) { p.rb.push_back(b); }
		}
		for (int b = 0; b < L && (int)p.rb.size() < K; b++) {
			if (!(used & (1u << b))) { p.rb.push_back(b); }
		}
		std::sort(p.rb.begin(), p.rb.end());
		out.push_back(p);
	}
	return out;
}

// Kernel arguments of a pass.  Tile-bit order: the two least used bits of the
// pass go to the top tile positions (wave ownership of quads changes, and a
// barrier is needed, only between consecutive gates whose two highest non-gate
// tile bits differ) and the third least used to position 0 (the two rows of a
// 256-byte LDS bank window differ in tile bit 0, so gates away from it read
// their matrix-core groups without bank conflicts).  Column swizzle: a greedy
// colouring of the pass's gate graph (tile bits as vertices) with three colours,
// column masks h in {1, 2, 3} (FP64) or {2, 4, 6} (FP32), gives the four
// amplitude rows of every gate distinct bank slots in the reverse pass's
// transposed stores.
static TileArgs make_tile_args(const Circuit& C, const Pass& p, int K, int logC, int lognc, int x0, int elem_bytes,
	GateTerms* gterms)
{
	TileArgs a;
	std::memset(&a, 0, sizeof(a));
	a.K = K;
	a.logC = logC;
	a.lognc = lognc;
	a.ng = (int)p.gates.size();
	a.x0 = x0;
	{
		const char* w = getenv("RQC_REV_WIDE");
		a.wide = (w == nullptr) ? 1 : atoi(w);
	}
	int lp[kMaxTileGates], lq[kMaxTileGates];
	// usage of the tile's row bits by the pass's gates
	std::vector<int> use(C.L, 0);
	for (int k : p.gates) { use[C.L - 1 - C.wi[k]]++; use[C.L - 1 - C.wj[k]]++; }
	std::vector<int> bits = p.rb;
	std::stable_sort(bits.begin(), bits.end(), [&](int x, int y) { return use[x] < use[y]; });
	std::vector<int> order(K);
	int lo = 0, hi = K - 1;
	for (int i = 0; i < K; i++) {
		// least used -> K-1, second -> K-2, third -> 0, then upwards from 1
		if (i < 2 && K >= 3) { order[hi--] = bits[i]; }
		else { order[lo++] = bits[i]; }
	}
	std::vector<int> tile_of(C.L, -1);
	unsigned tmask = 0;
	for (int t = 0; t < K; t++) { a.rb[t] = order[t]; tile_of[order[t]] = t; tmask |= 1u << order[t]; }
	a.rmask = (unsigned)(((1UL << C.L) - 1) & ~(unsigned long)tmask);
	for (int j = 0; j < a.ng; j++) {
		const int k = p.gates[j];
		lp[j] = tile_of[C.L - 1 - C.wi[k]];
		lq[j] = tile_of[C.L - 1 - C.wj[k]];
	}
	// three-colouring of the gate graph, highest degree first
	std::vector<std::vector<int>> nb(K);
	for (int j = 0; j < a.ng; j++) { nb[lp[j]].push_back(lq[j]); nb[lq[j]].push_back(lp[j]); }
	std::vector<int> vs(K), col(K, 0);
	for (int t = 0; t < K; t++) { vs[t] = t; }
	std::stable_sort(vs.begin(), vs.end(), [&](int x, int y) { return nb[x].size() > nb[y].size(); });
	for (int t : vs) {
		int c = 1;
		for (; c <= 3; c++) {
			bool used = false;
			for (int u : nb[t]) { used |= (col[u] == c); }
			if (!used) { break; }
		}
		col[t] = (c <= 3) ? c : 1;
	}
	const int shift = (elem_bytes == 16) ? 0 : 1;
	// FP32: an 8-byte LDS read serves 32 lanes over 256 bytes (32 elements), i.e. two
	// amplitudes of a 16-quad group, whose rows differ in one tile bit; column bit 4 in every
	// mask (bit 4 of f = parity of the row) puts them in different halves of the window
	const int bit4 = (elem_bytes == 8 && logC >= 5) ? 16 : 0;
	for (int t = 0; t < K; t++) {
		const int h = ((col[t] << shift) | bit4) & ((1 << logC) - 1);
		for (int i = 0; i < 5; i++) {
			if ((h >> i) & 1) { a.swz[i] |= 1u << t; }
		}
	}
	// barriers between gates whose quads are owned differently
	auto top2 = [&](int j) {
		int b1 = -1, b2 = -1;
		for (int t = K - 1; t >= 0 && b2 < 0; t--) {
			if (t == lp[j] || t == lq[j]) { continue; }
			if (b1 < 0) { b1 = t; } else { b2 = t; }
		}
		return std::make_pair(b1, b2);
	};
	for (int j = 1; j < a.ng; j++) {
		if (top2(j) != top2(j - 1)) { a.barrier |= 1u << j; }
	}
	// uniform LDS position terms of each gate (after the swizzle is fixed)
	const int qpw = (1 << (K + logC)) / 4 / kTileWaves;
	for (int j = 0; j < a.ng; j++) {
		GateTerms& g = gterms[x0 + j];
		for (int i = 0; i < 4; i++) {
			g.u[i] = quad_pos(0, i, lp[j], lq[j], a);
			if (elem_bytes == 16) {
				g.w[i] = quad_pos(4 * i, 0, lp[j], lq[j], a);
				g.q4[i] = quad_pos(i, 0, lp[j], lq[j], a);
			} else {
				// FP32: matrix-core row q of a group is quad q with bits 0 and 2
				// exchanged, so that the two quads of a transposed-store phase
				// (rows v and 4 + v) differ in column bit 0 and, with the 64-byte
				// state padding, the phase's 16 elements take all 16 bank slots
				g.w[i] = quad_pos((i & 1) | ((i & 2) << 2), 0, lp[j], lq[j], a);
				g.q4[i] = quad_pos(((i & 1) << 2) | (i & 2), 0, lp[j], lq[j], a);
			}
		}
		for (int w = 0; w < kTileWaves; w++) { g.wv[w] = quad_pos(w * qpw, 0, lp[j], lq[j], a); }
		for (int i = 0; i < 3; i++) { g.gbit[i] = (16 << i < qpw) ? quad_pos(16 << i, 0, lp[j], lq[j], a) : 0; }
		// the kernels select among u, w, q4 by index bits (XOR-affine terms)
		for (const int* v : {g.u, g.w, g.q4}) {
			if (v[3] != (v[0] ^ v[1] ^ v[2])) { throw Error(-4, "tile position terms are not XOR-affine"); }
		}
	}
	return a;
}

template <typename KernelT>
static void allow_lds(KernelT* fn, size_t bytes)
{
	if (bytes > 65536) {
		RQC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
	}
	// the tile kernels address their dynamic LDS by offset from 0, which holds while they
	// declare no static LDS
	hipFuncAttributes fa;
	RQC_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(fn)));
	if (fa.sharedSizeBytes != 0) { throw Error(-4, "tile kernel with static LDS"); }
}

// gates whose inverse is their adjoint up to 1e-6 (the reverse tile pass
// recovers earlier states with G^-1, which then has norm 1)
bool gates_near_unitary(const Circuit& C)
{
	for (int k = 0; k < C.ngates; k++) {
		const zc* g = &C.gates[16 * (size_t)k];
		zc gh[16], p[16];
		m4_adjoint(g, gh);
		m4_mul(gh, g, p);
		for (int e = 0; e < 16; e++) {
			const zc d = p[e] - zc{(e % 5 == 0) ? 1.0 : 0.0, 0.0};
			if (std::fabs(d.x) > 1e-6 || std::fabs(d.y) > 1e-6) { return false; }
		}
	}
	return true;
}

bool sweep_unitary(const Circuit& C, Mode mode, bool fp32)
{
	return (mode == Mode::Target && !fp32) || gates_near_unitary(C);
}

// tile shape for element type T: 2^K rows x 2^logC columns, 128-byte rows
struct TileShape
{
	int K, logC;
	// >= 512 elements: 4 waves x one matrix-core group of 32 quads (plain sweeps);
	// >= 2 columns: the two FP32 elements of a 16-byte slot are adjacent columns
	bool usable() const { return K + logC >= 9 && logC >= 1; }
};

template <typename T>
static TileShape tile_shape(int L, long B, int logE_choice = 0)
{
	const int elem = (int)sizeof(cx<T>);
	// 16 KiB per state, as 64 rows of 256 bytes (FP64 16 columns, FP32 32): the
	// reverse pass holds one buffer of four states, two workgroups per CU.
	// Measured per L=14 eval against 8 KiB tiles of 128-byte rows: FP64 7 %
	// faster (8 KiB double-buffered), FP32 3 %.
	int logE = (elem == 16) ? 10 : 11;
	if (const char* v = getenv(elem == 16 ? "RQC_TILE_LOGE_FP64" : "RQC_TILE_LOGE_FP32")) { logE = atoi(v); }
	if (logE_choice > 0) { logE = logE_choice; }
	logE = std::max(9, std::min(logE, elem == 16 ? 10 : 11));   // the reverse pass's states fit 160 KiB
	// rows of 256 bytes for the 16 KiB tiles, 128 bytes for the 8 KiB ones
	int logCdef = (elem == 16) ? (logE == 10 ? 4 : 3) : (logE == 11 ? 5 : 4);
	if (const char* v = getenv(elem == 16 ? "RQC_TILE_LOGC_FP64" : "RQC_TILE_LOGC_FP32")) {
		logCdef = std::max(1, std::min(atoi(v), 6));
	}
	int logB = 0;
	while ((1L << (logB + 1)) <= B) { logB++; }
	TileShape s;
	s.K = std::min(L, logE - logCdef);
	s.logC = std::min(logB, logE - s.K);
	return s;
}

static int tile_gmax()
{
	int g = kMaxTileGates;
	if (const char* v = getenv("RQC_PASS_GMAX")) { g = std::max(1, std::min(g, atoi(v))); }
	return g;
}

// tile-pass part of a plan: passes, their kernel arguments and the gate
// position terms (all fixed by the circuit structure and the column block)
struct TilePlan
{
	TileShape ts;                   // reverse passes (and forward, unless tsf differs)
	TileShape tsf;                  // forward passes
	std::vector<Pass> passes;
	std::vector<TileArgs> args;
	std::vector<TileArgs> args_fwd;
	std::vector<TileGates> gks;
	std::vector<GateTerms> terms;   // per gate, in execution order
	std::vector<GateTerms> terms_fwd;   // forward tiles' terms (when tsf differs from ts)
	size_t fterms_off = 0;          // byte offset of the forward terms in the table's terms region
	long ntiles_fwd = 0;
	std::vector<int> order;         // circuit gate index of each execution slot
	unsigned gfwd = 0, grev = 0;
	size_t lds_fwd = 0, lds_rev = 0;
	size_t lds_rev_seed = 0;        // the seed pass: also the fused forward pass's G, Z
	size_t partial_doubles = 0;
};

// Forward-pass tile: the forward pass holds half the states of the reverse pass
// and no outer-product accumulators, so a tile of twice the elements keeps
// enough workgroups per CU; its HBM rows are 256 bytes ('c': twice the columns)
// or it spans one more row bit ('k').  The passes stay those planned for the
// reverse tile: their gates lie in the forward tile's row bits either way.
static TileShape forward_shape(const TileShape& ts, int L, long B, int elem)
{
	char mode = 'c';
	if (const char* v = getenv("RQC_FWD_TILE")) { mode = v[0]; }
	TileShape f = ts;
	const bool fits = (size_t)(2L << (ts.K + ts.logC)) * elem <= (size_t)kMaxSlots * 16 * kTileThreads;
	if (!fits) { return f; }
	if (mode == 'c' && (2L << ts.logC) <= B) { f.logC++; }
	else if (mode == 'k' && ts.K < L) { f.K++; }
	return f;
}

// tile-kernel instantiation for a tile of 2^logE elements: the dual sweeps have variants
// with the LDS state stride fixed at compile time for the planned tile heights
template <typename T>
static decltype(&k_tile_forward<T, true, 0>) fwd_kernel(bool dual, int logE)
{
	if (!dual) { return k_tile_forward<T, false, 0>; }
	if constexpr (sizeof(T) == 8) {
		if (logE == 10) { return k_tile_forward<T, true, 10>; }
		if (logE == 9) { return k_tile_forward<T, true, 9>; }
	} else {
		if (logE == 11) { return k_tile_forward<T, true, 11>; }
		if (logE == 10) { return k_tile_forward<T, true, 10>; }
	}
	return k_tile_forward<T, true, 0>;
}

template <typename T>
static decltype(&k_tile_reverse<T, true, 0>) rev_kernel(bool dual, int logE, bool seed = false)
{
	if (!dual) { return k_tile_reverse<T, false, 0>; }
	if (seed) {
		if constexpr (sizeof(T) == 8) {
			if (logE == 10) { return k_tile_reverse<T, true, 10, true>; }
			if (logE == 9) { return k_tile_reverse<T, true, 9, true>; }
		} else {
			if (logE == 11) { return k_tile_reverse<T, true, 11, true>; }
			if (logE == 10) { return k_tile_reverse<T, true, 10, true>; }
		}
		return k_tile_reverse<T, true, 0, true>;
	}
	if constexpr (sizeof(T) == 8) {
		if (logE == 10) { return k_tile_reverse<T, true, 10>; }
		if (logE == 9) { return k_tile_reverse<T, true, 9>; }
	} else {
		if (logE == 11) { return k_tile_reverse<T, true, 11>; }
		if (logE == 10) { return k_tile_reverse<T, true, 10>; }
	}
	return k_tile_reverse<T, true, 0>;
}

template <typename T>
static std::shared_ptr<TilePlan> plan_tiles(const Context& dev, const Circuit& C, EvalPlan& pl, const std::vector<Pass>& passes,
	const TileShape& ts)
{
	const bool dual = (pl.mode == Mode::Hvp);
	auto tp = std::make_shared<TilePlan>();
	tp->ts = ts;
	tp->tsf = forward_shape(ts, C.L, pl.B, (int)sizeof(cx<T>));
	const bool own_fwd = (tp->tsf.K != ts.K || tp->tsf.logC != ts.logC);
	tp->passes = passes;
	tp->terms.resize((size_t)C.ngates);
	if (own_fwd) {
		tp->terms_fwd.resize((size_t)C.ngates);
		tp->fterms_off = ((size_t)C.ngates * sizeof(GateTerms) + 255) & ~(size_t)255;
	}
	int x = 0, maxng = 1;
	for (const Pass& p : passes) {
		if (own_fwd) {
			Pass pf = p;
			if (tp->tsf.K > ts.K) {
				// one more row bit: the lowest one outside the pass's bits
				int b = 0;
				while (std::find(pf.rb.begin(), pf.rb.end(), b) != pf.rb.end()) { b++; }
				pf.rb.push_back(b);
				std::sort(pf.rb.begin(), pf.rb.end());
			}
			tp->args_fwd.push_back(make_tile_args(C, pf, tp->tsf.K, tp->tsf.logC, ilog2_or_neg(pl.B >> tp->tsf.logC), x,
				(int)sizeof(cx<T>), tp->terms_fwd.data()));
		}
		tp->args.push_back(make_tile_args(C, p, ts.K, ts.logC, ilog2_or_neg(pl.B >> ts.logC), x, (int)sizeof(cx<T>),
			tp->terms.data()));
		if (!own_fwd) { tp->args_fwd.push_back(tp->args.back()); }
		TileGates tg;
		std::memset(&tg, 0, sizeof(tg));
		for (size_t j = 0; j < p.gates.size(); j++) {
			tg.k[j] = p.gates[j];
			tp->order.push_back(p.gates[j]);
			x++;
		}
		tp->gks.push_back(tg);
		maxng = std::max(maxng, (int)p.gates.size());
	}
	const int E = 1 << (ts.K + ts.logC);
	const int Ef = 1 << (tp->tsf.K + tp->tsf.logC);
	for (int e : {E, Ef}) {
		if ((size_t)e * sizeof(cx<T>) > (size_t)kMaxSlots * 16 * kTileThreads) {
			throw Error(-4, "tile of " + std::to_string(e) + " elements exceeds the kernels' load slots");
		}
	}
	if (tp->tsf.K > C.L || (1L << tp->tsf.logC) > pl.B) { throw Error(-4, "forward tile exceeds the column block"); }
	const long ntiles = (pl.N >> ts.K) * (pl.B >> ts.logC);
	tp->ntiles_fwd = (pl.N >> tp->tsf.K) * (pl.B >> tp->tsf.logC);
	// two buffers of (psi[, dpsi]) forward and (psi, [dpsi,] phi[, dphi]) reverse, then
	// the pass's gate matrices (forward G[, Z]; reverse G^-1, [W,] G^T[, Z^T])
	const size_t fwd_mats = (size_t)kMaxTileGates * (dual ? fwd_mat_stride<true>() : fwd_mat_stride<false>()) * sizeof(cx<T>);
	const size_t rev_mats = (size_t)kMaxTileGates * (dual ? rev_mat_stride<true>() : rev_mat_stride<false>()) * sizeof(cx<T>);
	tp->lds_fwd = (size_t)(dual ? fwd_buffers<T, true>() : fwd_buffers<T, false>()) * (dual ? 2 : 1) *
		(Ef * sizeof(cx<T>) + kLdsPadBytes) + fwd_mats + kFwdVoteBytes;
	tp->lds_rev = (size_t)(dual ? rev_buffers<T, true>() : rev_buffers<T, false>()) * (dual ? 4 : 2) *
		(E * sizeof(cx<T>) + lds_pad_rev<T>()) + rev_mats + kRevTraceBytes;
	auto* kf = fwd_kernel<T>(dual, tp->tsf.K + tp->tsf.logC);
	auto* kr = rev_kernel<T>(dual, ts.K + ts.logC);
	allow_lds(kf, tp->lds_fwd);
	allow_lds(kr, tp->lds_rev);
	tp->lds_rev_seed = tp->lds_rev + (size_t)kMaxTileGates * fwd_mat_stride<true>() * sizeof(cx<T>);
	allow_lds(rev_kernel<T>(dual, ts.K + ts.logC, true), tp->lds_rev_seed);
	int occ_f = 0, occ_r = 0;
	RQC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, reinterpret_cast<const void*>(kf), kTileThreads, tp->lds_fwd));
	RQC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_r, reinterpret_cast<const void*>(kr), kTileThreads, tp->lds_rev));
	tp->gfwd = (unsigned)std::min<long>(tp->ntiles_fwd, (long)std::max(1, occ_f) * dev.cu_count);
	tp->grev = (unsigned)std::max<long>(1, std::min<long>(ntiles, (long)std::max(1, occ_r) * dev.cu_count));
	tp->partial_doubles = (pl.mode == Mode::Target) ? 0 : (size_t)tp->grev * kTileWaves * maxng * 256;
	return tp;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

EvalPlan plan_eval(const Context& dev, const Circuit& C, Mode mode, int per, bool fp32, bool project, bool unitary,
	size_t budget, size_t held, const TileChoice* choice, long ncols)
{
	const int logE_c = (choice != nullptr) ? choice->logE : 0;
	const int gmax = (choice != nullptr && choice->gmax > 0) ? std::min(choice->gmax, (int)kMaxTileGates) : tile_gmax();
	if (per < 1 || C.ngates % per != 0) { throw Error(-3, "gate groups must divide the gate count"); }
	if (mode == Mode::Hvp && C.dirs.size() != C.gates.size()) { throw Error(-3, "an HVP evaluation needs gate directions"); }
	if (fp32 && mode == Mode::Target) { mode = Mode::Grad; }   // gate-rounding correction of f needs the gradient
	EvalPlan pl;
	pl.mode = mode;
	pl.fp32 = fp32;
	pl.project = project;
	pl.L = C.L;
	pl.G = C.ngates;
	pl.per = per;
	pl.ng = C.ngates / per;
	pl.N = 1L << C.L;
	pl.sampled = (ncols > 0);
	pl.ncols = pl.sampled ? ncols : pl.N;
	long Bmax = 1;   // power-of-two cover of the columns
	while (Bmax < pl.ncols) { Bmax <<= 1; }
	const size_t elem = fp32 ? sizeof(cx<float>) : sizeof(cx<double>);
	// tile-pass engine when a tile fits the circuit (and, for reverse sweeps, the gates are unitary)
	bool tiles = dev.engine == RQC_ENGINE_TILE_PASS && C.L <= 30 && (mode == Mode::Target || unitary);
	TileShape ts{0, 0};
	std::vector<Pass> passes;
	if (tiles) {
		ts = fp32 ? tile_shape<float>(C.L, pl.N, logE_c) : tile_shape<double>(C.L, pl.N, logE_c);
		tiles = ts.usable();
	}
	if (tiles) {
		passes = plan_passes(C, ts.K, gmax);
		const int P = (int)passes.size();
		pl.nst = (mode == Mode::Target) ? 2 : (mode == Mode::Grad) ? P + 2 : 2 * P + 4;
		pl.B = choose_block(pl.N, pl.nst, elem, budget, held, Bmax);
		ts = fp32 ? tile_shape<float>(C.L, pl.B, logE_c) : tile_shape<double>(C.L, pl.B, logE_c);
		tiles = ts.usable();
	}
	if (tiles) {
		pl.tiles = true;
		pl.passes = (int)passes.size();
		pl.tile_K = ts.K;
		pl.tile_logC = ts.logC;
		pl.gmax = gmax;
		pl.tp = fp32 ? plan_tiles<float>(dev, C, pl, passes, ts) : plan_tiles<double>(dev, C, pl, passes, ts);
	} else {
		pl.nst = states_needed(mode, C.ngates);
		pl.B = choose_block(pl.N, pl.nst, elem, budget, held, Bmax);
		pl.passes = C.ngates;
		pl.grev_pg = reduce_grid((pl.N / 4) * pl.B);
	}
	pl.nblk = (pl.ncols + pl.B - 1) / pl.B;
	pl.gtr = reduce_grid(pl.N * pl.B);
	pl.n_acc = 1 + 32 * (size_t)pl.G;
	size_t npart = pl.gtr;
	if (pl.tiles) {
		// outer-product partials, the chunk sums of the first reduction stage, then the trace
		// partials of a first reverse pass that takes the target's trace
		npart = std::max(npart, (pl.tp->partial_doubles + (size_t)kReduceChunks * kMaxTileGates * 256 + 1) / 2 +
			(size_t)pl.tp->grev * kTileWaves);
	}
	else { npart = std::max(npart, (size_t)pl.grev_pg * kWidth); }
	pl.n_part = npart;
	if (pl.tiles) {
		pl.mats_bytes = align256((size_t)pl.G * kMats * 16 * elem);
		pl.terms_bytes = align256((size_t)pl.G * sizeof(GateTerms));
		if (!pl.tp->terms_fwd.empty()) { pl.terms_bytes += pl.tp->fterms_off; }   // the forward terms follow
	}
	pl.tab_bytes = pl.mats_bytes + pl.terms_bytes + 2 * 16 * (size_t)pl.ng * sizeof(zc);
	if (pl.tiles && mode != Mode::Target) {
		pl.rmats_off = align256(pl.tab_bytes);
		pl.tab_bytes = pl.rmats_off + (size_t)pl.G * 32 * sizeof(zc);
	}
	pl.off_acc = 0;
	pl.off_part = align256(pl.n_acc * sizeof(zc));
	pl.off_tab = pl.off_part + align256(pl.n_part * sizeof(zc));
	pl.off_res = pl.off_tab + align256(pl.tab_bytes);
	pl.off_states = pl.off_res + align256(pl.res_bytes());
	pl.bytes = pl.off_states + (size_t)pl.nst * pl.N * pl.B * elem;
	return pl;
}

template <typename T>
static void fill_tables(const EvalPlan& pl, const Circuit& C, char* host)
{
	const bool dual = (pl.mode == Mode::Hvp);
	if (pl.tiles) {
		const TilePlan& tp = *pl.tp;
		cx<T>* hm = reinterpret_cast<cx<T>*>(host);
		for (size_t x = 0; x < tp.order.size(); x++) {
			const int k = tp.order[x];
			// G, Z, G^-1, -G^-1 Z G^-1, G^T, Z^T (kMats matrices, zero where unused)
			const zc* g = &C.gates[16 * (size_t)k];
			zc m6[kMats][16];
			std::memset(m6, 0, sizeof(m6));
			std::memcpy(m6[kMatG], g, sizeof(m6[0]));
			if (dual) { std::memcpy(m6[kMatZ], &C.dirs[16 * (size_t)k], sizeof(m6[0])); }
			if (pl.mode != Mode::Target) {
				if (m4_inverse(g, m6[kMatGinv]) != 0) { throw Error(-4, "singular gate in the tile-pass sweep"); }
				m4_transpose(g, m6[kMatGT]);
				if (dual) {
					zc t1[16];
					m4_mul(m6[kMatGinv], m6[kMatZ], t1);
					m4_mul(t1, m6[kMatGinv], m6[kMatW]);
					for (int e = 0; e < 16; e++) { m6[kMatW][e] = zc{-m6[kMatW][e].x, -m6[kMatW][e].y}; }
					m4_transpose(m6[kMatZ], m6[kMatZT]);
				}
			}
			for (int w = 0; w < kMats; w++) {
				for (int e = 0; e < 16; e++) {
					hm[(x * kMats + w) * 16 + e] = cx<T>{(T)m6[w][e].x, (T)m6[w][e].y};
				}
			}
			if (pl.rmats_off > 0) {
				// the reduction's maps from the adjoint at the gate input to its output: G^-T, W^T
				zc* rm = reinterpret_cast<zc*>(host + pl.rmats_off) + 32 * x;
				m4_transpose(m6[kMatGinv], rm);
				m4_transpose(m6[kMatW], rm + 16);
			}
		}
		std::memcpy(host + pl.mats_bytes, tp.terms.data(), tp.terms.size() * sizeof(GateTerms));
		if (!tp.terms_fwd.empty()) {
			std::memcpy(host + pl.mats_bytes + tp.fterms_off, tp.terms_fwd.data(), tp.terms_fwd.size() * sizeof(GateTerms));
		}
	}
	// gates of each group (its first gate) and directions, for the projection
	zc* vin = reinterpret_cast<zc*>(host + pl.mats_bytes + pl.terms_bytes);
	for (int l = 0; l < pl.ng; l++) {
		std::memcpy(&vin[16 * (size_t)l], &C.gates[16 * (size_t)l * pl.per], 16 * sizeof(zc));
		if (dual) { std::memcpy(&vin[16 * ((size_t)pl.ng + l)], &C.dirs[16 * (size_t)l * pl.per], 16 * sizeof(zc)); }
		else { std::memset(&vin[16 * ((size_t)pl.ng + l)], 0, 16 * sizeof(zc)); }
	}
}

void fill_eval_tables(const EvalPlan& pl, const Circuit& C, char* host)
{
	std::memset(host, 0, pl.tab_bytes);
	if (pl.fp32) { fill_tables<float>(pl, C, host); }
	else { fill_tables<double>(pl, C, host); }
}

std::vector<int> table_slot_groups(const EvalPlan& pl)
{
	if (!pl.tiles) { throw Error(-3, "device-side tables need a tile-pass plan"); }
	std::vector<int> g(pl.tp->order.size());
	for (size_t x = 0; x < g.size(); x++) { g[x] = pl.tp->order[x] / pl.per; }
	return g;
}

// The gate-dependent part of fill_tables on the device, one workgroup per
// descriptor: the matrices of every execution slot (from its group's gate and
// direction), the reduction maps and the groups' gates and directions.  The
// arithmetic is fill_tables' (same mat4 routines); a singular gate sets *bad.
template <typename T>
__device__ void fill_slots_device(const DeviceTableDesc& d)
{
	const bool dual = (d.mode == (int)Mode::Hvp);
	cx<T>* hm = reinterpret_cast<cx<T>*>(d.tab);
	for (int x = threadIdx.x; x < d.G; x += blockDim.x) {
		const int l = d.slot_group[x];
		const zc* g = d.gates + 16 * (size_t)l;
		zc m6[kMats][16];
		for (int w = 0; w < kMats; w++) { m4_zero(m6[w]); }
		for (int e = 0; e < 16; e++) { m6[kMatG][e] = g[e]; }
		if (dual) { for (int e = 0; e < 16; e++) { m6[kMatZ][e] = d.dirs[16 * (size_t)l + e]; } }
		if (d.mode != (int)Mode::Target) {
			if (m4_inverse(g, m6[kMatGinv]) != 0) { *d.bad = 1; }
			m4_transpose(g, m6[kMatGT]);
			if (dual) {
				zc t1[16];
				m4_mul(m6[kMatGinv], m6[kMatZ], t1);
				m4_mul(t1, m6[kMatGinv], m6[kMatW]);
				for (int e = 0; e < 16; e++) { m6[kMatW][e] = zc{-m6[kMatW][e].x, -m6[kMatW][e].y}; }
				m4_transpose(m6[kMatZ], m6[kMatZT]);
			}
		}
		for (int w = 0; w < kMats; w++) {
			for (int e = 0; e < 16; e++) { hm[((size_t)x * kMats + w) * 16 + e] = cx<T>{(T)m6[w][e].x, (T)m6[w][e].y}; }
		}
		if (d.rmats_off > 0) {
			zc* rm = reinterpret_cast<zc*>(d.tab + d.rmats_off) + 32 * (size_t)x;
			m4_transpose(m6[kMatGinv], rm);
			m4_transpose(m6[kMatW], rm + 16);
		}
	}
	zc* vin = reinterpret_cast<zc*>(d.tab + d.mats_bytes + d.terms_bytes);
	for (int i = threadIdx.x; i < 16 * d.ng; i += blockDim.x) {
		vin[i] = d.gates[i];
		vin[16 * (size_t)d.ng + i] = dual ? d.dirs[i] : zc{0, 0};
	}
}

__global__ __launch_bounds__(64) void k_fill_tables(const DeviceTableDesc* __restrict__ descs)
{
	const DeviceTableDesc& d = descs[blockIdx.x];
	if (d.skip != nullptr && *d.skip != 0) { return; }
	if (d.fp32) { fill_slots_device<float>(d); }
	else { fill_slots_device<double>(d); }
}

void enqueue_fill_tables(hipStream_t s, const DeviceTableDesc* d_descs, int n)
{
	if (n > 0) { k_fill_tables<<<n, 64, 0, s>>>(d_descs); }
}

// tile-pass engine: several gates per HBM pass through LDS tiles
// tiles of a pass's geometry that hold an element of the identity columns b0 + j of a column
// block of width B (any b0): per column chunk, one value of the row bits outside the tile above
// the chunk's low log2 C bits, and every value of those among the low bits
static long active_tiles(const TileArgs& a, long B)
{
	const unsigned low = a.rmask & ((1u << a.logC) - 1);
	return (B >> a.logC) << __builtin_popcount(low);
}

template <typename T>
static void run_tiles(Slot& s, const EvalPlan& pl, const zc* dU, zc* d_acc, zc* d_partial, const cx<T>* mats,
	const GateTerms* gterms, cx<T>* S, const int* skip, const cx<T>* X, const zc* W, const BlockEnc* be)
{
	const Mode mode = pl.mode;
	const TilePlan& tp = *pl.tp;
	const long N = pl.N, B = pl.B;
	const bool dual = (mode == Mode::Hvp);
	const int P = pl.passes;
	const size_t sz = (size_t)N * B;
	// checkpoints after pass p - 1, p = 1 .. P: psi at PS(p), dpsi at DPS(p)
	auto PS = [&](int p) { return S + (size_t)(p - 1) * sz; };
	auto DPS = [&](int p) { return S + (size_t)(P + p - 1) * sz; };
	cx<T>* buf[2] = {S, S + sz};
	const size_t nck = dual ? 2 * (size_t)P : (mode == Mode::Grad ? (size_t)P : 0);
	cx<T>* phi[2] = {S + nck * sz, S + (nck + 1) * sz};
	cx<T>* dphi[2] = {S + (nck + 2) * sz, S + (nck + 3) * sz};
	const long ntiles = (N >> tp.ts.K) * (B >> tp.ts.logC);
	const double E = (double)(1L << (tp.ts.K + tp.ts.logC)), Ef = (double)(1L << (tp.tsf.K + tp.tsf.logC));
	// matrix-core flops issued per amplitude and gate: a gate application (FP64 4 x 4 blocks in the
	// three-product form: dual pairs 9 block products per 16 quads, plain 3; FP32 the real 16 x 16
	// form) and the outer-product sums
	const double app_issued = (sizeof(T) == 8) ? (dual ? 72.0 : 24.0) : (dual ? 128.0 : 64.0);
	const double outer_issued = (sizeof(T) == 8) ? (dual ? 96.0 : 32.0) : (dual ? 128.0 : 64.0);
	auto* kf = fwd_kernel<T>(dual, tp.tsf.K + tp.tsf.logC);
	auto* kr = rev_kernel<T>(dual, tp.ts.K + tp.ts.logC);
	auto* krs = rev_kernel<T>(dual, tp.ts.K + tp.ts.logC, true);
	const GateTerms* fterms = reinterpret_cast<const GateTerms*>(reinterpret_cast<const char*>(gterms) + tp.fterms_off);
	const zc* rmats = reinterpret_cast<const zc*>(reinterpret_cast<const char*>(mats) + pl.rmats_off);
	double* mpart = reinterpret_cast<double*>(d_partial);   // grev * waves * ng * 256 doubles
	double* cpart = mpart + tp.partial_doubles;             // chunk sums: kReduceChunks * ng * 256 doubles
	zc* tpart = d_partial;                                  // trace partials (reused, sequential)
	const unsigned gtr = pl.gtr;
	const int logB = ilog2_or_neg(B);
	const double ab = (double)sz * sizeof(cx<T>);   // bytes of one state block
	const int ns = dual ? 2 : 1;                     // state arrays per side
	// HVP sweeps over the whole of a dense target in one block: the first reverse pass reads
	// U's block as its output adjoint (FP64: into LDS; FP32: from HBM into registers, U staying
	// FP64), forms the seed -conj(U) in LDS and takes the trace on the way (no separate target
	// pass over U, psi and the seed)
	const bool fuse_trace = dual && be == nullptr && !pl.sampled && B == N && P >= 1;
	double* ttr = cpart + (size_t)kReduceChunks * kMaxTileGates * 256;   // its per-wave trace partials
	// and, with the planned tiles (one tile for both directions), it also runs the circuit's
	// last forward pass on its tiles first: the circuit's output states never go through HBM
	const int logEr = tp.ts.K + tp.ts.logC;
	const bool planned_tile = (sizeof(T) == 8) ? (logEr == 9 || logEr == 10) : (logEr == 10 || logEr == 11);
	const bool fuse_fwd = fuse_trace && P >= 2 && tp.tsf.K == tp.ts.K && tp.tsf.logC == tp.ts.logC && planned_tile;
	const int Pf = fuse_fwd ? P - 1 : P;   // forward launches

	for (long blk = 0; blk < pl.nblk; blk++) {
		const long b0 = blk * B;
		// the block's input columns and the trace's seed columns: basis columns and U, or samples and U X
		const zc* seed = (be != nullptr) ? nullptr : pl.sampled ? W + (size_t)blk * sz : dU;
		const long ldu = pl.sampled ? B : N, u0 = pl.sampled ? 0 : b0;
		for (int p = 0; p < Pf; p++) {
			// pass 0 starts from the identity block (flags 1) or the sample block, with dpsi = 0 (flags 2)
			const int flags = (p == 0) ? (pl.sampled ? 2 : 3) : 0;
			const cx<T>* in = (p == 0) ? (pl.sampled ? X + (size_t)blk * sz : nullptr)
			                           : (mode == Mode::Target ? buf[(p + 1) % 2] : PS(p));
			cx<T>* out = (mode == Mode::Target) ? buf[p % 2] : PS(p + 1);
			// pass 0 applies its gates to the tiles that hold an identity element only
			const double fwd_amps = (double)(flags & 1 ? active_tiles(tp.args_fwd[p], pl.B) : tp.ntiles_fwd) * Ef;
			const double fwd_flops = (dual ? 96.0 : 32.0) * (double)tp.passes[p].gates.size() * fwd_amps;
			LaunchTimer lt(s, RQC_KERNEL_FORWARD, (double)(ns * (p > 0) + ns) * ab, fwd_flops,
				fwd_flops * app_issued / (dual ? 96.0 : 32.0));
			kf<<<tp.gfwd, kTileThreads, tp.lds_fwd, s.stream>>>(in, out, (dual && p > 0) ? DPS(p) : nullptr,
				dual ? DPS(p + 1) : nullptr, tp.args_fwd[p], mats, fterms, B, b0, tp.ntiles_fwd, flags, skip);
		}
		cx<T>* psiG = (mode == Mode::Target) ? buf[(P - 1) % 2] : PS(P);
		if (be != nullptr) {
			LaunchTimer lt(s, RQC_KERNEL_TARGET, (mode == Mode::Target ? 1 : 2) * ab);
			k_blockenc_residual<T><<<gtr, kBlock, 0, s.stream>>>(be->H, pl.L / 2, b0, psiG,
				mode == Mode::Target ? nullptr : phi[0], dual ? DPS(P) : nullptr, dual ? be->J : nullptr, B, logB, N * B,
				tpart, skip);
		} else if (!fuse_trace) {
			LaunchTimer lt(s, RQC_KERNEL_TARGET, (double)sz * sizeof(zc) + (mode == Mode::Target ? 1 : 2) * ab);
			k_target_trace<T><<<gtr, kBlock, 0, s.stream>>>(seed, ldu, u0, psiG, mode == Mode::Target ? nullptr : phi[0],
				B, logB, N * B, tpart, skip);
		}
		if (!fuse_trace) {
			LaunchTimer lt(s, RQC_KERNEL_OTHER);
			k_reduce_partials<<<1, kBlock, 0, s.stream>>>(tpart, (int)gtr, 1, d_acc, skip);
		}
		if (mode == Mode::Target) { continue; }
		int cur = 0;
		for (int p = P - 1; p >= 0; p--) {
			int flags = 0;
			if (p == P - 1) { flags |= 4; }
			if (p == 0) { flags |= 8; }
			// the first pass from identity columns: only tiles with an identity element contribute
			if (p == 0 && !pl.sampled) { flags |= 32; }
			const bool seed_pass = fuse_trace && p == P - 1;
			const bool with_fwd = seed_pass && fuse_fwd;   // psi, dpsi at the pass input (flags 16)
			if (with_fwd) { flags |= 16; }
			const int ng = (int)tp.passes[p].gates.size();
			{
				// (the FP32 seed pass reads U in FP64: twice an FP32 array)
				const int reads = (dual ? 3 + (p < P - 1) : 2) + ((seed_pass && sizeof(T) == 4) ? 1 : 0);
				// algorithmic work of the launch: per amplitude and gate of the tiles it processes,
				// the adjoint update and the outer products (dual 24 complex MACs, plain 8), and a
				// fused forward pass (12); issued besides: the recovery, in the forms above
				const double frac = (flags & 32) ? (double)active_tiles(tp.args[p], pl.B) / (double)ntiles : 1.0;
				const double amps = frac * (double)ntiles * E;
				const double rflops = (dual ? 192.0 : 64.0) * ng * amps;
				const double rissued = (2.0 * app_issued + outer_issued) * ng * amps;
				const double fflops = with_fwd ? 96.0 * ng * (double)ntiles * E : 0.0;
				const double fissued = with_fwd ? app_issued * ng * (double)ntiles * E : 0.0;
				LaunchTimer lt(s, RQC_KERNEL_REVERSE, (double)(reads * frac + ns * (p > 0)) * ab, rflops + fflops,
					rissued + fissued);
				(seed_pass ? krs : kr)<<<tp.grev, kTileThreads, seed_pass ? tp.lds_rev_seed : tp.lds_rev, s.stream>>>(
					with_fwd ? PS(p) : PS(p + 1), dual ? (with_fwd ? DPS(p) : DPS(p + 1)) : nullptr,
					seed_pass ? reinterpret_cast<const cx<T>*>(seed) : phi[cur], dual ? dphi[cur] : nullptr, phi[1 - cur],
					dual ? dphi[1 - cur] : nullptr, tp.args[p], mats, gterms, B, b0, ntiles, flags, mpart, ttr, skip);
			}
			if (seed_pass) {
				LaunchTimer lt(s, RQC_KERNEL_OTHER);
				k_reduce_partials<<<1, kBlock, 0, s.stream>>>(reinterpret_cast<const zc*>(ttr),
					(int)(tp.grev * kTileWaves), 1, d_acc, skip);
			}
			{
				LaunchTimer lt(s, RQC_KERNEL_OTHER);
				const long nparts = (long)tp.grev * kTileWaves;
				const int nch = (int)std::min<long>(kReduceChunks, nparts);
				k_tile_reduce_chunks<<<dim3(ng, nch), 256, 0, s.stream>>>(mpart, nparts, ng, nch, cpart, skip);
				k_tile_reduce<<<ng, 256, 0, s.stream>>>(cpart, nch, ng, tp.gks[p], rmats + 32L * tp.args[p].x0, dual ? 1 : 0,
					d_acc, skip);
			}
			cur = 1 - cur;
		}
	}
}

// per-group sums of per-gate gradients / HVPs and the Riemannian projection
// (reference src/target.c:526-535 and :650-671), one thread per group.
// res: f, dv[16 ng], hz[16 ng] (complex), grad[16 ng], hvp[16 ng] (double)
// fcorr (FP32 sweeps with a gradient): the sweeps applied the gates rounded to
// complex float, fl(V), so the accumulated f is f(fl(V)).  f(V) = f(fl(V)) +
// sum_e dv[e] (V[e] - fl(V)[e]) + O(|V - fl(V)|^2), with dv = df/dV (f is
// holomorphic in the gates for the trace targets: fcorr 1 corrects the complex
// f; the block-encoding f is real with df = Re sum dv dV: fcorr 2 corrects its
// real part).  Every gate of a group is the group's gate vlist[l].
__global__ void k_finalize(const zc* __restrict__ acc, int ngroups, int per, const zc* __restrict__ vlist,
	const zc* __restrict__ vdirs, zc* __restrict__ res, int want_grad, int project, int want_hvp, int fcorr,
	const int* skip)
{
	if (skipped(skip)) { return; }
	const int l = blockIdx.x * blockDim.x + threadIdx.x;
	if (l == 0) {
		zc f = acc[0];
		if (want_grad && fcorr != 0) {
			zc c = {0, 0};
			for (int g = 0; g < ngroups; g++) {
				for (int e = 0; e < 16; e++) {
					const zc v = vlist[16 * g + e];
					const zc dv = {v.x - (double)(float)v.x, v.y - (double)(float)v.y};
					zc s = {0, 0};
					for (int k = 0; k < per; k++) { s += acc[1 + 32L * ((long)g * per + k) + e]; }
					c.x += s.x * dv.x - s.y * dv.y;
					c.y += s.x * dv.y + s.y * dv.x;
				}
			}
			f.x += c.x;
			if (fcorr == 1) { f.y += c.y; }
		}
		res[0] = f;
	}
	if (!want_grad || l >= ngroups) { return; }
	zc* dv_out = res + 1;
	zc* hz_out = dv_out + 16L * ngroups;
	double* grad = reinterpret_cast<double*>(hz_out + 16L * ngroups);
	double* hvp = grad + 16L * ngroups;
	zc dv[16], hz[16];
	for (int e = 0; e < 16; e++) { dv[e] = {0, 0}; hz[e] = {0, 0}; }
	for (int k = 0; k < per; k++) {
		const zc* a = acc + 1 + 32L * ((long)l * per + k);
		for (int e = 0; e < 16; e++) { dv[e] += a[e]; hz[e] += a[16 + e]; }
	}
	for (int e = 0; e < 16; e++) { dv_out[16 * l + e] = dv[e]; hz_out[16 * l + e] = hz[e]; }
	if (project) {
		m4_project_grad_hvp(vlist + 16 * l, want_hvp ? vdirs + 16 * l : nullptr, dv, hz, grad + 16 * l,
			want_hvp ? hvp + 16 * l : nullptr);
	}
}

void enqueue_eval(const Context& dev, Slot& s, const EvalPlan& pl, const Circuit& C, const zc* dU, char* work,
	char* tab, char* res_dev, const char* host_tab, const int* skip, const void* X, const zc* W, const BlockEnc* be)
{
	(void)dev;
	if (skip != nullptr && !pl.tiles) { throw Error(-3, "skip flags need the tile-pass engine"); }
	if (pl.sampled && (X == nullptr || (W == nullptr && be == nullptr))) {
		throw Error(-3, "a sampled evaluation needs its sample and seed blocks");
	}
	if (be != nullptr && (!pl.sampled || be->H == nullptr || pl.ncols != (1L << (pl.L / 2)) || pl.L % 2 != 0)) {
		throw Error(-3, "a block-encoding evaluation needs the 2^(L/2) embedded columns and H");
	}
	zc* acc = reinterpret_cast<zc*>(work + pl.off_acc);
	zc* part = reinterpret_cast<zc*>(work + pl.off_part);
	if (tab == nullptr) { tab = work + pl.off_tab; }
	const zc* vin = reinterpret_cast<const zc*>(tab + pl.mats_bytes + pl.terms_bytes);
	zc* res = reinterpret_cast<zc*>(res_dev != nullptr ? res_dev : work + pl.off_res);
	if (host_tab != nullptr) {
		RQC_HIP(hipMemcpyAsync(tab, host_tab, pl.tab_bytes, hipMemcpyHostToDevice, s.stream));
	}
	s.last_block = pl.B;
	s.last_engine = pl.tiles ? RQC_ENGINE_TILE_PASS : RQC_ENGINE_PER_GATE;
	s.last_K = pl.tiles ? pl.tile_K : 0;
	s.last_logC = pl.tiles ? pl.tile_logC : 0;
	s.last_gmax = pl.tiles ? pl.gmax : 0;
	s.last_passes = pl.passes;
	{
		LaunchTimer lt(s, RQC_KERNEL_OTHER);
		k_fill_zero<<<nblocks((long)pl.n_acc), kBlock, 0, s.stream>>>(acc, (long)pl.n_acc, skip);
	}
	const GateTerms* gterms = reinterpret_cast<const GateTerms*>(tab + pl.mats_bytes);
	if (pl.fp32) {
		cx<float>* S = reinterpret_cast<cx<float>*>(work + pl.off_states);
		const cx<float>* Xs = static_cast<const cx<float>*>(X);
		if (pl.tiles) { run_tiles<float>(s, pl, dU, acc, part, reinterpret_cast<const cx<float>*>(tab), gterms, S, skip, Xs, W, be); }
		else { run_gates<float>(s, pl, C, dU, acc, part, S, skip, Xs, W, be); }
	} else {
		cx<double>* S = reinterpret_cast<cx<double>*>(work + pl.off_states);
		const cx<double>* Xs = static_cast<const cx<double>*>(X);
		if (pl.tiles) { run_tiles<double>(s, pl, dU, acc, part, reinterpret_cast<const cx<double>*>(tab), gterms, S, skip, Xs, W, be); }
		else { run_gates<double>(s, pl, C, dU, acc, part, S, skip, Xs, W, be); }
	}
	LaunchTimer lt(s, RQC_KERNEL_OTHER);
	k_finalize<<<std::max(1, (pl.ng + 63) / 64), 64, 0, s.stream>>>(acc, pl.ng, pl.per, vin, vin + 16 * (size_t)pl.ng, res,
		pl.mode != Mode::Target ? 1 : 0, pl.project ? 1 : 0, pl.mode == Mode::Hvp ? 1 : 0,
		pl.fp32 ? (be != nullptr ? 2 : 1) : 0, skip);
}

void unpack_result(const EvalPlan& pl, const char* res, EvalResult& R)
{
	const size_t n = 16 * (size_t)pl.ng;
	const zc* z = reinterpret_cast<const zc*>(res);
	R.f = z[0];
	R.dv.assign(z + 1, z + 1 + n);
	R.hz.assign(z + 1 + n, z + 1 + 2 * n);
	const double* d = reinterpret_cast<const double*>(z + 1 + 2 * n);
	R.grad.assign(d, d + n);
	R.hvp.assign(d + n, d + 2 * n);
}

// dense U for a target evaluation: device-resident or assembled from the host callback
static const zc* resolve_target(Context& c, linear_func ufunc, void* udata, int L);

// Tile autotuning of single evaluations (rqc_set_eval_autotune): the first evaluation of a
// circuit structure (qubits, wires, groups, kind, precision, projection) times the tile
// candidates on the evaluation itself (one untimed run, then the best of two, HIP events on the
// library's stream) and keeps the fastest for every later evaluation of that structure.  The
// candidates are the tile height (2^10 or 2^9 FP64 elements, 2^11 or 2^10 FP32; the row bits K
// follow from the height and the 256- or 128-byte rows) and the most gates per pass (8 or 6).
static std::string structure_key(const Circuit& C, Mode mode, int per, bool fp32, bool project)
{
	std::string key = std::to_string(C.L) + "/" + std::to_string(per) + "/" + std::to_string((int)mode) + "/" +
		std::to_string(fp32) + "/" + std::to_string(project) + ":";
	for (int k = 0; k < C.ngates; k++) { key += std::to_string(C.wi[k]) + "," + std::to_string(C.wj[k]) + ";"; }
	return key;
}

static std::map<std::string, TileChoice>& tuned_choices()
{
	static std::map<std::string, TileChoice> m;
	return m;
}

static TileChoice tune_eval(Context& c, const Circuit& C, Mode mode, int per, bool project, bool unitary, const zc* dU)
{
	const bool fp32 = c.precision == RQC_PRECISION_MIXED;
	const std::string key = structure_key(C, mode, per, fp32, project);
	auto it = tuned_choices().find(key);
	if (it != tuned_choices().end()) { return it->second; }
	const int base = fp32 ? 11 : 10;
	const TileChoice cands[] = {{base, 8}, {base - 1, 8}, {base, 6}, {base - 1, 6}};
	TileChoice best{0, 0};
	float best_ms = 0;
	hipEvent_t e0, e1;
	RQC_HIP(hipEventCreate(&e0));
	RQC_HIP(hipEventCreate(&e1));
	for (const TileChoice& tc : cands) {
		EvalPlan pl;
		try {
			pl = plan_eval(c, C, mode, per, fp32, project, unitary, c.ws_limit, c.work.bytes, &tc);
		} catch (const Error&) {
			continue;   // a candidate that does not fit is skipped
		}
		if (!pl.tiles) { break; }   // no tile-pass plan: nothing to tune
		std::vector<char> tab(pl.tab_bytes);
		fill_eval_tables(pl, C, tab.data());
		c.work.reserve(pl.bytes);
		float ms = 0;
		for (int r = 0; r < 3; r++) {
			RQC_HIP(hipEventRecord(e0, c.stream));
			enqueue_eval(c, c, pl, C, dU, c.work.as<char>(), nullptr, nullptr, tab.data(), nullptr, nullptr, nullptr);
			RQC_HIP(hipEventRecord(e1, c.stream));
			RQC_HIP(hipEventSynchronize(e1));
			float t = 0;
			RQC_HIP(hipEventElapsedTime(&t, e0, e1));
			if (r == 1 || (r == 2 && t < ms)) { ms = t; }
		}
		if (best.logE == 0 || ms < best_ms) { best = tc; best_ms = ms; }
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	tuned_choices()[key] = best;
	return best;
}

// one evaluation on the main slot (the reference ABI's target functions)
static EvalResult eval_targets(const Circuit& C, linear_func ufunc, void* udata, Mode mode, int per, bool project)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	const zc* dU = resolve_target(c, ufunc, udata, C.L);
	const bool unitary = sweep_unitary(C, mode, c.precision == RQC_PRECISION_MIXED);
	TileChoice tc{0, 0};
	if (c.eval_autotune && c.engine == RQC_ENGINE_TILE_PASS) { tc = tune_eval(c, C, mode, per, project, unitary, dU); }
	const EvalPlan pl = plan_eval(c, C, mode, per, c.precision == RQC_PRECISION_MIXED, project, unitary, c.ws_limit,
		c.work.bytes, tc.logE > 0 ? &tc : nullptr);
	std::vector<char> tab(pl.tab_bytes);
	fill_eval_tables(pl, C, tab.data());
	c.work.reserve(pl.bytes);
	char* work = c.work.as<char>();
	enqueue_eval(c, c, pl, C, dU, work, nullptr, nullptr, tab.data(), nullptr, nullptr, nullptr);
	std::vector<char> res(pl.res_bytes());
	const size_t nres = (mode == Mode::Target) ? sizeof(zc) : res.size();
	RQC_HIP(hipMemcpyAsync(res.data(), work + pl.off_res, nres, hipMemcpyDeviceToHost, c.stream));
	RQC_HIP(hipStreamSynchronize(c.stream));
	RQC_HIP(hipGetLastError());
	EvalResult R;
	unpack_result(pl, res.data(), R);
	return R;
}

// ============================================================== device targets
static constexpr uint64_t kTargetMagic = 0x5251435441524754ULL;  // "RQCTARGT"

}  // namespace rqc

struct rqc_target
{
	uint64_t magic;
	int nqubits;
	zc* d_U;
};

namespace rqc {

static const rqc_target* as_target(void* udata, int L)
{
	const rqc_target* t = static_cast<const rqc_target*>(udata);
	if (t == nullptr || t->magic != kTargetMagic) {
		throw Error(-3, "rqc_device_ufunc needs an rqc_target* as udata");
	}
	if (L >= 0 && t->nqubits != L) {
		throw Error(-3, "rqc_target has " + std::to_string(t->nqubits) + " qubits, circuit has " + std::to_string(L));
	}
	return t;
}

const zc* device_target_matrix(void* udata, int L) { return as_target(udata, L)->d_U; }

static const zc* resolve_target(Context& c, linear_func ufunc, void* udata, int L)
{
	if (ufunc == rqc_device_ufunc) { return as_target(udata, L)->d_U; }
	if (ufunc == nullptr) { throw Error(-3, "ufunc is NULL"); }
	// host callback: evaluate it on every basis column, as the reference does
	// once per column of each evaluation (src/target.c:37-48)
	const long N = 1L << L;
	std::vector<zc> U((size_t)N * N), e((size_t)N), col((size_t)N);
	struct statevector sin = {reinterpret_cast<numeric*>(e.data()), L};
	struct statevector sout = {reinterpret_cast<numeric*>(col.data()), L};
	for (long b = 0; b < N; b++) {
		std::fill(e.begin(), e.end(), zc{0, 0});
		e[b] = {1, 0};
		const int r = ufunc(&sin, udata, &sout);
		if (r < 0) { throw Error(-2, "call of 'ufunc' failed, return value: " + std::to_string(r)); }
		for (long a = 0; a < N; a++) { U[(size_t)a * N + b] = col[a]; }
	}
	c.ubuf.reserve((size_t)N * N * sizeof(zc));
	RQC_HIP(hipMemcpyAsync(c.ubuf.ptr, U.data(), (size_t)N * N * sizeof(zc), hipMemcpyHostToDevice, c.stream));
	RQC_HIP(hipStreamSynchronize(c.stream));
	return c.ubuf.as<zc>();
}

// ============================================================== statevector staging
// Device views of the statevectors of one call: host buffers are copied into
// the library's vector workspace, device buffers are used in place.
struct Staging
{
	struct Item { const void* host; void* user; size_t bytes; bool in, out, dev; size_t off; };
	std::vector<Item> items;
	Context& c;
	explicit Staging(Context& cx) : c(cx) {}

	int add(const void* p, size_t bytes, bool in, bool out)
	{
		Item it{p, const_cast<void*>(p), bytes, in, out, is_device_pointer(p), 0};
		items.push_back(it);
		return (int)items.size() - 1;
	}
	int scratch(size_t bytes) { items.push_back(Item{nullptr, nullptr, bytes, false, false, false, 0}); return (int)items.size() - 1; }

	void upload()
	{
		size_t tot = 0;
		for (auto& it : items) {
			if (!it.dev) { it.off = tot; tot += (it.bytes + 255) & ~(size_t)255; }
		}
		c.vecs.reserve(std::max<size_t>(tot, 256));
		for (auto& it : items) {
			if (!it.dev && it.in) {
				RQC_HIP(hipMemcpyAsync(static_cast<char*>(c.vecs.ptr) + it.off, it.host, it.bytes, hipMemcpyHostToDevice, c.stream));
			}
		}
	}
	template <typename T> T* dev(int k)
	{
		auto& it = items[k];
		return it.dev ? static_cast<T*>(it.user) : reinterpret_cast<T*>(static_cast<char*>(c.vecs.ptr) + it.off);
	}
	void download()
	{
		for (auto& it : items) {
			if (!it.dev && it.out) {
				RQC_HIP(hipMemcpyAsync(it.user, static_cast<char*>(c.vecs.ptr) + it.off, it.bytes, hipMemcpyDeviceToHost, c.stream));
			}
		}
		RQC_HIP(hipStreamSynchronize(c.stream));
		RQC_HIP(hipGetLastError());
	}
};

// copy between two device buffers on the library stream
static void dcopy(Context& c, void* dst, const void* src, size_t bytes)
{
	RQC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream));
}

// copy a device buffer to a user buffer (host or device)
static void copy_out(Context& c, void* user, const void* src, size_t bytes)
{
	RQC_HIP(hipMemcpyAsync(user, src, bytes, is_device_pointer(user) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c.stream));
}

static void copy_in(Context& c, void* dst, const void* user, size_t bytes)
{
	RQC_HIP(hipMemcpyAsync(dst, user, bytes, is_device_pointer(user) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c.stream));
}

// one FP64 gate on a block of B columns
static void launch_forward(Context& c, const zc* G, int i, int j, int L, long B, const zc* in, zc* out)
{
	const long nq = (1L << L) / 4 * B;
	LaunchTimer lt(c, RQC_KERNEL_FORWARD);
	k_gate_forward<double, false><<<nblocks(nq), kBlock, 0, c.stream>>>(in, out, nullptr, nullptr, to_gate<double>(G),
		to_gate<double>(G), wires_of(i, j, L), B, ilog2_or_neg(B), nq, 0);
}

static void launch_forward_dual(Context& c, const zc* G, const zc* Z, int i, int j, int L, const zc* in, zc* out,
	const zc* din, zc* dout)
{
	const long nq = (1L << L) / 4;
	LaunchTimer lt(c, RQC_KERNEL_FORWARD);
	k_gate_forward<double, true><<<nblocks(nq), kBlock, 0, c.stream>>>(in, out, din, dout, to_gate<double>(G),
		to_gate<double>(Z), wires_of(i, j, L), 1, 0, nq, din == nullptr ? 1 : 0);
}

// reverse step of one FP64 gate on a single statevector; gradient (and HVP) summed into acc[0..32)
static void launch_reverse(Context& c, const zc* G, const zc* Z, int i, int j, int L, const zc* psi, const zc* dpsi,
	const zc* phi, const zc* dphi, zc* phi_out, zc* dphi_out, bool dual, zc* acc, zc* part)
{
	const long nq = (1L << L) / 4;
	const unsigned g = reduce_grid(nq);
	int flags = 0;
	if (dpsi == nullptr) { flags |= 1; }
	if (dphi == nullptr) { flags |= 2; }
	if (phi_out == nullptr) { flags |= 4; }
	{
		LaunchTimer lt(c, RQC_KERNEL_REVERSE);
		if (dual) {
			k_gate_reverse<double, true><<<g, kBlock, 0, c.stream>>>(psi, dpsi, phi, dphi, phi_out, dphi_out,
				to_gate<double>(G), to_gate<double>(Z), wires_of(i, j, L), 1, 0, nq, flags, part);
		} else {
			k_gate_reverse<double, false><<<g, kBlock, 0, c.stream>>>(psi, nullptr, phi, nullptr, phi_out, nullptr,
				to_gate<double>(G), to_gate<double>(G), wires_of(i, j, L), 1, 0, nq, flags | 2, part);
		}
	}
	LaunchTimer lt(c, RQC_KERNEL_OTHER);
	k_reduce_partials<<<kWidth, kBlock, 0, c.stream>>>(part, (int)g, kWidth, acc, nullptr);
}

static zc* small_acc(Context& c, size_t n_acc)
{
	c.small.reserve((n_acc + 4096 * (size_t)kWidth) * sizeof(zc));
	zc* acc = c.small.as<zc>();
	LaunchTimer lt(c, RQC_KERNEL_OTHER);
	k_fill_zero<<<nblocks((long)n_acc), kBlock, 0, c.stream>>>(acc, (long)n_acc, nullptr);
	return acc;
}

}  // namespace rqc

using namespace rqc;

// ============================================================== C ABI: statevector.h
extern "C" int allocate_statevector(int nqubits, struct statevector* psi)
{
	psi->nqubits = nqubits;
	const size_t size = ((size_t)1 << nqubits) * sizeof(numeric);
	psi->data = static_cast<numeric*>(aligned_malloc(size));
	if (psi->data == nullptr) {
		fprintf(stderr, "allocating statevector memory (%zu bytes) failed\n", size);
		return -1;
	}
	return 0;
}

extern "C" void free_statevector(struct statevector* psi)
{
	if (psi->data != nullptr) {
		aligned_free(psi->data);
		psi->data = nullptr;
	}
	psi->nqubits = 0;
}

extern "C" int allocate_statevector_array(int nqubits, int nstates, struct statevector_array* psi_array)
{
	psi_array->nqubits = nqubits;
	psi_array->nstates = nstates;
	const size_t size = ((size_t)1 << nqubits) * nstates * sizeof(numeric);
	psi_array->data = static_cast<numeric*>(aligned_malloc(size));
	if (psi_array->data == nullptr) {
		fprintf(stderr, "allocating statevector array memory (%zu bytes) failed\n", size);
		return -1;
	}
	return 0;
}

extern "C" void free_statevector_array(struct statevector_array* psi_array)
{
	if (psi_array->data != nullptr) {
		aligned_free(psi_array->data);
		psi_array->data = nullptr;
	}
	psi_array->nqubits = 0;
	psi_array->nstates = 0;
}

extern "C" void add_statevectors(const struct statevector* RQC_RESTRICT psi1, const struct statevector* RQC_RESTRICT psi2,
	struct statevector* RQC_RESTRICT psi_sum)
{
	try {
		if (psi1->nqubits != psi2->nqubits || psi1->nqubits != psi_sum->nqubits) { throw Error(-3, "qubit counts differ"); }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const long n = 1L << psi1->nqubits;
		Staging s(c);
		const int a = s.add(psi1->data, n * sizeof(zc), true, false);
		const int b = s.add(psi2->data, n * sizeof(zc), true, false);
		const int o = s.add(psi_sum->data, n * sizeof(zc), false, true);
		s.upload();
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_add<<<nblocks(n), kBlock, 0, c.stream>>>(s.dev<zc>(a), s.dev<zc>(b), s.dev<zc>(o), n);
		}
		s.download();
	} catch (const std::exception& e) { fatal("add_statevectors", e); }
}

extern "C" void normalize_statevector(struct statevector* RQC_RESTRICT psi)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const long n = 1L << psi->nqubits;
		const unsigned g = reduce_grid(n);
		c.small.reserve((1 + (size_t)g) * sizeof(zc));
		zc* acc = c.small.as<zc>();
		Staging s(c);
		const int a = s.add(psi->data, n * sizeof(zc), true, true);
		s.upload();
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_fill_zero<<<1, kBlock, 0, c.stream>>>(acc, 1, nullptr);
			k_norm2_partials<<<g, kBlock, 0, c.stream>>>(s.dev<zc>(a), n, acc + 1);
			k_reduce_partials<<<1, kBlock, 0, c.stream>>>(acc + 1, (int)g, 1, acc, nullptr);
			k_scale_inv_norm<<<nblocks(n), kBlock, 0, c.stream>>>(s.dev<zc>(a), n, acc);
		}
		s.download();
	} catch (const std::exception& e) { fatal("normalize_statevector", e); }
}

extern "C" void transpose_statevector(const struct statevector* RQC_RESTRICT psi, const int* perm, struct statevector* RQC_RESTRICT psi_trans)
{
	try {
		const int L = psi->nqubits;
		if (psi_trans->nqubits != L) { throw Error(-3, "qubit counts differ"); }
		if (L < 1 || L > 32) { throw Error(-3, "transpose_statevector supports 1 .. 32 qubits"); }
		BitMap m;
		std::vector<int> seen(L, 0);
		for (int i = 0; i < L; i++) {
			if (perm[i] < 0 || perm[i] >= L || seen[perm[i]]) { throw Error(-3, "perm is not a permutation"); }
			seen[perm[i]] = 1;
		}
		for (int q = 0; q < L; q++) { m.src[q] = L - 1 - perm[L - 1 - q]; }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const long n = 1L << L;
		Staging s(c);
		const int a = s.add(psi->data, n * sizeof(zc), true, false);
		const int o = s.add(psi_trans->data, n * sizeof(zc), false, true);
		s.upload();
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_transpose_amplitudes<<<nblocks(n), kBlock, 0, c.stream>>>(s.dev<zc>(a), s.dev<zc>(o), m, L, n);
		}
		s.download();
	} catch (const std::exception& e) { fatal("transpose_statevector", e); }
}

static void draw_haar_state(long n, struct rng_state* rng, zc* psi);

extern "C" void haar_random_statevector(struct statevector* psi, struct rng_state* rng)
{
	try {
		const long n = 1L << psi->nqubits;
		std::vector<zc> h((size_t)n);
		draw_haar_state(n, rng, h.data());
		if (is_device_pointer(psi->data)) {
			Context& c = ctx();
			std::lock_guard<std::recursive_mutex> lk(c.mu);
			RQC_HIP(hipMemcpyAsync(psi->data, h.data(), n * sizeof(zc), hipMemcpyHostToDevice, c.stream));
			RQC_HIP(hipStreamSynchronize(c.stream));
		} else {
			std::memcpy(psi->data, h.data(), n * sizeof(zc));
		}
	} catch (const std::exception& e) { fatal("haar_random_statevector", e); }
}

// ============================================================== C ABI: gate.h
extern "C" void apply_gate(const struct mat4x4* gate, const int i, const int j, const struct statevector* RQC_RESTRICT psi,
	struct statevector* RQC_RESTRICT psi_out)
{
	try {
		if (psi->nqubits != psi_out->nqubits) { throw Error(-3, "qubit counts differ"); }
		const int L = psi->nqubits;
		check_wires(i, j, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		Staging s(c);
		const int a = s.add(psi->data, bytes, true, false);
		const int o = s.add(psi_out->data, bytes, false, true);
		s.upload();
		launch_forward(c, reinterpret_cast<const zc*>(gate->data), i, j, L, 1, s.dev<zc>(a), s.dev<zc>(o));
		s.download();
	} catch (const std::exception& e) { fatal("apply_gate", e); }
}

extern "C" void apply_gate_backward(const struct mat4x4* gate, const int i, const int j, const struct statevector* RQC_RESTRICT psi,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, struct mat4x4* dgate)
{
	try {
		if (psi->nqubits != dpsi_out->nqubits || psi->nqubits != dpsi->nqubits) { throw Error(-3, "qubit counts differ"); }
		const int L = psi->nqubits;
		check_wires(i, j, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		zc* acc = small_acc(c, 32);
		Staging s(c);
		const int a = s.add(psi->data, bytes, true, false);
		const int d = s.add(dpsi_out->data, bytes, true, false);
		const int o = s.add(dpsi->data, bytes, false, true);
		s.upload();
		const zc* G = reinterpret_cast<const zc*>(gate->data);
		launch_reverse(c, G, G, i, j, L, s.dev<zc>(a), nullptr, s.dev<zc>(d), nullptr, s.dev<zc>(o), nullptr, false, acc, acc + 32);
		copy_out(c, dgate->data, acc, 16 * sizeof(zc));
		s.download();
	} catch (const std::exception& e) { fatal("apply_gate_backward", e); }
}

extern "C" void apply_gate_to_array(const struct mat4x4* gate, const int i, const int j, const struct statevector_array* RQC_RESTRICT psi,
	struct statevector_array* RQC_RESTRICT psi_out)
{
	try {
		if (psi->nqubits != psi_out->nqubits || psi->nstates != psi_out->nstates) { throw Error(-3, "array shapes differ"); }
		const int L = psi->nqubits;
		check_wires(i, j, L);
		if (psi->nstates < 1) { return; }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * psi->nstates * sizeof(zc);
		Staging s(c);
		const int a = s.add(psi->data, bytes, true, false);
		const int o = s.add(psi_out->data, bytes, false, true);
		s.upload();
		launch_forward(c, reinterpret_cast<const zc*>(gate->data), i, j, L, psi->nstates, s.dev<zc>(a), s.dev<zc>(o));
		s.download();
	} catch (const std::exception& e) { fatal("apply_gate_to_array", e); }
}

extern "C" void apply_gate_backward_array(const struct mat4x4* gate, const int i, const int j,
	const struct statevector_array* RQC_RESTRICT psi, const struct statevector* RQC_RESTRICT dpsi_out, struct mat4x4* dgates)
{
	(void)gate;   // the gradients depend on the gate's wires only
	try {
		if (psi->nqubits != dpsi_out->nqubits) { throw Error(-3, "qubit counts differ"); }
		const int L = psi->nqubits;
		check_wires(i, j, L);
		if (psi->nstates < 1) { return; }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t n = (size_t)1 << L;
		c.small.reserve(16 * (size_t)psi->nstates * sizeof(zc));
		Staging s(c);
		const int a = s.add(psi->data, n * psi->nstates * sizeof(zc), true, false);
		const int d = s.add(dpsi_out->data, n * sizeof(zc), true, false);
		s.upload();
		{
			LaunchTimer lt(c, RQC_KERNEL_REVERSE);
			k_gate_backward_array<<<16 * psi->nstates, kBlock, 0, c.stream>>>(s.dev<zc>(a), s.dev<zc>(d), c.small.as<zc>(),
				wires_of(i, j, L), (long)(n / 4), psi->nstates, 16, 0);
		}
		copy_out(c, dgates, c.small.ptr, 16 * (size_t)psi->nstates * sizeof(zc));
		s.download();
	} catch (const std::exception& e) { fatal("apply_gate_backward_array", e); }
}

extern "C" void apply_gate_placeholder(const int i, const int j, const struct statevector* psi, struct statevector_array* psi_out)
{
	try {
		if (psi->nqubits != psi_out->nqubits || psi_out->nstates != 16) { throw Error(-3, "psi_out must hold 16 states of psi's size"); }
		const int L = psi->nqubits;
		check_wires(i, j, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t n = (size_t)1 << L;
		Staging s(c);
		const int a = s.add(psi->data, n * sizeof(zc), true, false);
		const int o = s.add(psi_out->data, n * 16 * sizeof(zc), false, true);
		s.upload();
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_gate_placeholder<<<nblocks((long)n * 16), kBlock, 0, c.stream>>>(s.dev<zc>(a), s.dev<zc>(o), wires_of(i, j, L), (long)n * 16);
		}
		s.download();
	} catch (const std::exception& e) { fatal("apply_gate_placeholder", e); }
}

// Placeholder states of every gate of layer l (from the cached state before it),
// each carried through all later gates of the brickwall and summed (reference
// src/brickwall_circuit.c:264-310): psi_out[a * 16 + 4 r + c] is the derivative
// of the circuit output at amplitude a by entry (r, c) of the layer's gate.
extern "C" int apply_brickwall_unitary_gate_placeholder(const struct mat4x4 vlist[], int nlayers, const int* perms[], int l,
	const struct quantum_circuit_cache* cache, struct statevector_array* RQC_RESTRICT psi_out)
{
	try {
		const int L = psi_out->nqubits;
		if (cache->nqubits != L || psi_out->nstates != 16) { throw Error(-3, "psi_out must hold 16 states of the cache's size"); }
		if (l < 0 || l >= nlayers) { throw Error(-3, "layer index out of range"); }
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, L);
		const int per = L / 2, ng = C.ngates;
		if (cache->ngates != ng) { throw Error(-3, "cache holds " + std::to_string(cache->ngates) + " states, circuit has " + std::to_string(ng) + " gates"); }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t n = (size_t)1 << L;
		const size_t bytes = n * sizeof(zc);
		Staging s(c);
		const int st = s.scratch(bytes);
		const int t0 = s.scratch(bytes * 16), t1 = s.scratch(bytes * 16);
		const int o = s.add(psi_out->data, bytes * 16, false, true);
		s.upload();
		zc* out = s.dev<zc>(o);
		zc* T[2] = {s.dev<zc>(t0), s.dev<zc>(t1)};
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_fill_zero<<<nblocks((long)n * 16), kBlock, 0, c.stream>>>(out, (long)n * 16, nullptr);
		}
		for (int k = 0; k < per; k++) {
			const int g = l * per + k;
			copy_in(c, s.dev<zc>(st), cache->psi_list[g].data, bytes);
			{
				LaunchTimer lt(c, RQC_KERNEL_OTHER);
				k_gate_placeholder<<<nblocks((long)n * 16), kBlock, 0, c.stream>>>(s.dev<zc>(st), T[0],
					wires_of(C.wi[g], C.wj[g], L), (long)n * 16);
			}
			int p = 0;
			for (int h = g + 1; h < ng; h++) {
				launch_forward(c, &C.gates[16 * (size_t)h], C.wi[h], C.wj[h], L, 16, T[p], T[1 - p]);
				p = 1 - p;
			}
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_add<<<nblocks((long)n * 16), kBlock, 0, c.stream>>>(out, T[p], out, (long)n * 16);
		}
		s.download();
		return 0;
	} catch (const std::exception& e) { return report_error("apply_brickwall_unitary_gate_placeholder", e); }
}

// ============================================================== C ABI: quantum_circuit.h
int rqc::report_error(const char* where, const std::exception& e)
{
	fprintf(stderr, "librqcopt_hip: %s failed: %s\n", where, e.what());
	if (auto* r = dynamic_cast<const Error*>(&e)) { return r->code; }
	return -1;
}

extern "C" int apply_quantum_circuit(const struct mat4x4 gates[], const int ngates, const int wires[],
	const struct statevector* RQC_RESTRICT psi, struct statevector* RQC_RESTRICT psi_out)
{
	try {
		if (psi->nqubits != psi_out->nqubits) { throw Error(-3, "qubit counts differ"); }
		const int L = psi->nqubits;
		Circuit C = make_circuit(gates, nullptr, ngates, wires, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		Staging s(c);
		const int a = s.add(psi->data, bytes, true, false);
		const int t0 = s.scratch(bytes), t1 = s.scratch(bytes);
		const int o = s.add(psi_out->data, bytes, false, true);
		s.upload();
		const zc* in = s.dev<zc>(a);
		zc* buf[2] = {s.dev<zc>(t0), s.dev<zc>(t1)};
		for (int k = 0; k < C.ngates; k++) {
			zc* out = (k == C.ngates - 1) ? s.dev<zc>(o) : buf[k % 2];
			launch_forward(c, &C.gates[16 * (size_t)k], C.wi[k], C.wj[k], L, 1, in, out);
			in = out;
		}
		s.download();
		return 0;
	} catch (const std::exception& e) { return report_error("apply_quantum_circuit", e); }
}

extern "C" int allocate_quantum_circuit_cache(const int nqubits, const int ngates, struct quantum_circuit_cache* cache)
{
	if (ngates < 1) {
		fprintf(stderr, "a quantum circuit cache needs at least one gate\n");
		return -3;
	}
	cache->nqubits = nqubits;
	cache->ngates = ngates;
	cache->psi_list = static_cast<struct statevector*>(aligned_malloc(ngates * sizeof(struct statevector)));
	if (cache->psi_list == nullptr) {
		fprintf(stderr, "allocating memory for statevector array failed\n");
		return -1;
	}
	for (int i = 0; i < ngates; i++) { cache->psi_list[i].data = nullptr; cache->psi_list[i].nqubits = 0; }
	for (int i = 0; i < ngates; i++) {
		const int ret = allocate_statevector(nqubits, &cache->psi_list[i]);
		if (ret < 0) { return ret; }
	}
	return 0;
}

extern "C" void free_quantum_circuit_cache(struct quantum_circuit_cache* cache)
{
	if (cache->psi_list != nullptr) {
		for (int i = 0; i < cache->ngates; i++) { free_statevector(&cache->psi_list[i]); }
		aligned_free(cache->psi_list);
	}
	cache->psi_list = nullptr;
	cache->nqubits = 0;
	cache->ngates = 0;
}

extern "C" int quantum_circuit_forward(const struct mat4x4 gates[], const int ngates, const int wires[],
	const struct statevector* RQC_RESTRICT psi, struct quantum_circuit_cache* cache, struct statevector* RQC_RESTRICT psi_out)
{
	try {
		if (psi->nqubits != psi_out->nqubits || cache->nqubits != psi->nqubits || cache->ngates != ngates) {
			throw Error(-3, "shapes of psi, psi_out and cache differ");
		}
		const int L = psi->nqubits;
		Circuit C = make_circuit(gates, nullptr, ngates, wires, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		Staging s(c);
		const int a = s.add(psi->data, bytes, true, false);
		const int st = s.scratch(bytes * (ngates + 1));
		s.upload();
		zc* S = s.dev<zc>(st);
		const size_t n = (size_t)1 << L;
		dcopy(c, S, s.dev<zc>(a), bytes);
		for (int k = 0; k < ngates; k++) {
			launch_forward(c, &C.gates[16 * (size_t)k], C.wi[k], C.wj[k], L, 1, S + k * n, S + (k + 1) * n);
		}
		for (int k = 0; k < ngates; k++) { copy_out(c, cache->psi_list[k].data, S + k * n, bytes); }
		copy_out(c, psi_out->data, S + ngates * n, bytes);
		s.download();
		return 0;
	} catch (const std::exception& e) { return report_error("quantum_circuit_forward", e); }
}

extern "C" int quantum_circuit_backward(const struct mat4x4 gates[], const int ngates, const int wires[], const struct quantum_circuit_cache* cache,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, struct mat4x4 dgates[])
{
	try {
		if (dpsi_out->nqubits != dpsi->nqubits || cache->nqubits != dpsi->nqubits || cache->ngates != ngates) {
			throw Error(-3, "shapes of dpsi_out, dpsi and cache differ");
		}
		const int L = dpsi->nqubits;
		Circuit C = make_circuit(gates, nullptr, ngates, wires, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		const size_t n = (size_t)1 << L;
		zc* acc = small_acc(c, 32 * (size_t)ngates);
		zc* part = acc + 32 * (size_t)ngates;
		Staging s(c);
		const int st = s.scratch(bytes * ngates);
		const int d = s.add(dpsi_out->data, bytes, true, false);
		const int t0 = s.scratch(bytes), t1 = s.scratch(bytes);
		const int o = s.add(dpsi->data, bytes, false, true);
		s.upload();
		zc* S = s.dev<zc>(st);
		for (int k = 0; k < ngates; k++) { copy_in(c, S + k * n, cache->psi_list[k].data, bytes); }
		const zc* cur = s.dev<zc>(d);
		zc* buf[2] = {s.dev<zc>(t0), s.dev<zc>(t1)};
		for (int k = ngates - 1; k >= 0; k--) {
			zc* nxt = (k == 0) ? s.dev<zc>(o) : buf[k % 2];
			const zc* G = &C.gates[16 * (size_t)k];
			launch_reverse(c, G, G, C.wi[k], C.wj[k], L, S + k * n, nullptr, cur, nullptr, nxt, nullptr, false, acc + 32 * (size_t)k, part);
			cur = nxt;
		}
		std::vector<zc> hacc(32 * (size_t)ngates);
		RQC_HIP(hipMemcpyAsync(hacc.data(), acc, hacc.size() * sizeof(zc), hipMemcpyDeviceToHost, c.stream));
		s.download();
		for (int k = 0; k < ngates; k++) { std::memcpy(dgates[k].data, &hacc[32 * (size_t)k], 16 * sizeof(zc)); }
		return 0;
	} catch (const std::exception& e) { return report_error("quantum_circuit_backward", e); }
}

extern "C" int quantum_circuit_gates_hessian_vector_product(const struct mat4x4 gates[], const struct mat4x4 gatedirs[], const int ngates, const int wires[],
	const struct statevector* RQC_RESTRICT psi, const struct statevector* RQC_RESTRICT phi,
	struct statevector* RQC_RESTRICT psi_out, struct mat4x4 dgates[], struct mat4x4 hess_gatedirs[])
{
	try {
		if (psi->nqubits != phi->nqubits || psi->nqubits != psi_out->nqubits) { throw Error(-3, "qubit counts differ"); }
		const int L = psi->nqubits;
		Circuit C = make_circuit(gates, gatedirs, ngates, wires, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		const size_t n = (size_t)1 << L;
		zc* acc = small_acc(c, 32 * (size_t)ngates);
		zc* part = acc + 32 * (size_t)ngates;
		Staging s(c);
		const int a = s.add(psi->data, bytes, true, false);
		const int f = s.add(phi->data, bytes, true, false);
		const int st = s.scratch(bytes * (2 * (size_t)ngates + 5));
		const int o = s.add(psi_out->data, bytes, false, true);
		s.upload();
		zc* S = s.dev<zc>(st);
		auto P = [&](int k) { return S + k * n; };               // psi_k, k = 0 .. G
		auto D = [&](int k) { return S + (ngates + k) * n; };    // dpsi_k, k = 1 .. G
		zc* F[2] = {S + (2 * ngates + 1) * n, S + (2 * ngates + 2) * n};
		zc* DF[2] = {S + (2 * ngates + 3) * n, S + (2 * ngates + 4) * n};
		dcopy(c, P(0), s.dev<zc>(a), bytes);
		for (int k = 0; k < ngates; k++) {
			launch_forward_dual(c, &C.gates[16 * (size_t)k], &C.dirs[16 * (size_t)k], C.wi[k], C.wj[k], L, P(k), P(k + 1),
				k > 0 ? D(k) : nullptr, D(k + 1));
		}
		dcopy(c, s.dev<zc>(o), P(ngates), bytes);
		dcopy(c, F[0], s.dev<zc>(f), bytes);
		int cur = 0;
		for (int k = ngates - 1; k >= 0; k--) {
			launch_reverse(c, &C.gates[16 * (size_t)k], &C.dirs[16 * (size_t)k], C.wi[k], C.wj[k], L, P(k),
				k > 0 ? D(k) : nullptr, F[cur], k < ngates - 1 ? DF[cur] : nullptr,
				k > 0 ? F[1 - cur] : nullptr, k > 0 ? DF[1 - cur] : nullptr, true, acc + 32 * (size_t)k, part);
			cur = 1 - cur;
		}
		std::vector<zc> hacc(32 * (size_t)ngates);
		RQC_HIP(hipMemcpyAsync(hacc.data(), acc, hacc.size() * sizeof(zc), hipMemcpyDeviceToHost, c.stream));
		s.download();
		for (int k = 0; k < ngates; k++) {
			std::memcpy(dgates[k].data, &hacc[32 * (size_t)k], 16 * sizeof(zc));
			std::memcpy(hess_gatedirs[k].data, &hacc[32 * (size_t)k + 16], 16 * sizeof(zc));
		}
		return 0;
	} catch (const std::exception& e) { return report_error("quantum_circuit_gates_hessian_vector_product", e); }
}

// ============================================================== C ABI: brickwall_circuit.h
extern "C" int apply_brickwall_unitary(const struct mat4x4 vlist[], const int nlayers, const int* perms[],
	const struct statevector* RQC_RESTRICT psi, struct statevector* RQC_RESTRICT psi_out)
{
	try {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, psi->nqubits);
		std::vector<int> w(2 * (size_t)C.ngates);
		for (int k = 0; k < C.ngates; k++) { w[2*k] = C.wi[k]; w[2*k + 1] = C.wj[k]; }
		return apply_quantum_circuit(reinterpret_cast<const struct mat4x4*>(C.gates.data()), C.ngates, w.data(), psi, psi_out);
	} catch (const std::exception& e) { return report_error("apply_brickwall_unitary", e); }
}

extern "C" int brickwall_unitary_forward(const struct mat4x4 vlist[], int nlayers, const int* perms[],
	const struct statevector* RQC_RESTRICT psi, struct quantum_circuit_cache* cache, struct statevector* RQC_RESTRICT psi_out)
{
	try {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, psi->nqubits);
		std::vector<int> w(2 * (size_t)C.ngates);
		for (int k = 0; k < C.ngates; k++) { w[2*k] = C.wi[k]; w[2*k + 1] = C.wj[k]; }
		return quantum_circuit_forward(reinterpret_cast<const struct mat4x4*>(C.gates.data()), C.ngates, w.data(), psi, cache, psi_out);
	} catch (const std::exception& e) { return report_error("brickwall_unitary_forward", e); }
}

extern "C" int brickwall_unitary_backward(const struct mat4x4 vlist[], int nlayers, const int* perms[], const struct quantum_circuit_cache* cache,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, struct mat4x4 dvlist[])
{
	try {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, dpsi->nqubits);
		std::vector<int> w(2 * (size_t)C.ngates);
		for (int k = 0; k < C.ngates; k++) { w[2*k] = C.wi[k]; w[2*k + 1] = C.wj[k]; }
		std::vector<struct mat4x4> dg(C.ngates);
		const int r = quantum_circuit_backward(reinterpret_cast<const struct mat4x4*>(C.gates.data()), C.ngates, w.data(), cache,
			dpsi_out, dpsi, dg.data());
		if (r < 0) { return r; }
		// per layer, gradients of its pairs summed in reverse pair order (reference src/brickwall_circuit.c:100-120)
		const int per = dpsi->nqubits / 2;
		for (int i = 0; i < nlayers; i++) {
			zc s[16];
			for (int e = 0; e < 16; e++) { s[e] = {0, 0}; }
			for (int k = per - 1; k >= 0; k--) {
				const zc* g = reinterpret_cast<const zc*>(dg[(size_t)i * per + k].data);
				for (int e = 0; e < 16; e++) { s[e] += g[e]; }
			}
			std::memcpy(dvlist[i].data, s, sizeof(s));
		}
		return 0;
	} catch (const std::exception& e) { return report_error("brickwall_unitary_backward", e); }
}

// Gradient sweep storing the adjoint after every gate; then, per gate position,
// the 16 placeholder states E_rc psi_k are carried through the later gates on
// the GPU (16-column blocks) and contracted with the stored adjoints there
// (reference src/brickwall_circuit.c:134-255); the contractions accumulate
// straight into the Hessian's row block.  Diagonal blocks are symmetrised and
// the lower blocks copied on the host.
extern "C" int brickwall_unitary_backward_hessian(const struct mat4x4 vlist[], int nlayers, const int* perms[],
	const struct quantum_circuit_cache* cache, const struct statevector* RQC_RESTRICT dpsi_out,
	struct statevector* RQC_RESTRICT dpsi, struct mat4x4 dvlist[], numeric* hess)
{
	try {
		const int L = dpsi->nqubits;
		if (dpsi_out->nqubits != L || cache->nqubits != L) { throw Error(-3, "qubit counts differ"); }
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, L);
		const int per = L / 2, ng = C.ngates, m = 16 * nlayers;
		if (cache->ngates != ng) { throw Error(-3, "cache holds " + std::to_string(cache->ngates) + " states, circuit has " + std::to_string(ng) + " gates"); }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t n = (size_t)1 << L;
		const size_t bytes = n * sizeof(zc);
		// small: gradient accumulators (32 per gate), reverse partials, Hessian m x m
		c.small.reserve((32 * (size_t)ng + 4096 * (size_t)kWidth + (size_t)m * m) * sizeof(zc));
		zc* acc = c.small.as<zc>();
		zc* part = acc + 32 * (size_t)ng;
		zc* dh = part + 4096 * (size_t)kWidth;
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_fill_zero<<<nblocks(32L * ng), kBlock, 0, c.stream>>>(acc, 32L * ng, nullptr);
			k_fill_zero<<<nblocks((long)m * m), kBlock, 0, c.stream>>>(dh, (long)m * m, nullptr);
		}
		Staging s(c);
		const int st = s.scratch(bytes * ng);          // forward states (the cache)
		const int ad = s.scratch(bytes * ng);          // adjoint after each gate
		const int t0 = s.scratch(bytes * 16), t1 = s.scratch(bytes * 16);
		const int d = s.add(dpsi_out->data, bytes, true, false);
		const int o = s.add(dpsi->data, bytes, false, true);
		s.upload();
		zc* S = s.dev<zc>(st);
		zc* A = s.dev<zc>(ad);
		for (int k = 0; k < ng; k++) { copy_in(c, S + k * n, cache->psi_list[k].data, bytes); }
		dcopy(c, A + (size_t)(ng - 1) * n, s.dev<zc>(d), bytes);
		for (int k = ng - 1; k >= 0; k--) {
			zc* out = (k > 0) ? A + (size_t)(k - 1) * n : s.dev<zc>(o);
			const zc* G = &C.gates[16 * (size_t)k];
			launch_reverse(c, G, G, C.wi[k], C.wj[k], L, S + k * n, nullptr, A + (size_t)k * n, nullptr, out, nullptr, false,
				acc + 32 * (size_t)k, part);
		}
		zc* T[2] = {s.dev<zc>(t0), s.dev<zc>(t1)};
		for (int g = 0; g < ng; g++) {
			const int i = g / per;
			{
				LaunchTimer lt(c, RQC_KERNEL_OTHER);
				k_gate_placeholder<<<nblocks((long)n * 16), kBlock, 0, c.stream>>>(S + (size_t)g * n, T[0], wires_of(C.wi[g], C.wj[g], L),
					(long)n * 16);
			}
			int p = 0;
			for (int l = g + 1; l < ng; l++) {
				const int j = l / per;
				{
					LaunchTimer lt(c, RQC_KERNEL_REVERSE);
					k_gate_backward_array<<<256, kBlock, 0, c.stream>>>(T[p], A + (size_t)l * n,
						dh + ((size_t)i * 16 * nlayers + j) * 16, wires_of(C.wi[l], C.wj[l], L), (long)(n / 4), 16, m, 1);
				}
				if (l < ng - 1) {
					launch_forward(c, &C.gates[16 * (size_t)l], C.wi[l], C.wj[l], L, 16, T[p], T[1 - p]);
					p = 1 - p;
				}
			}
		}
		std::vector<zc> hacc(32 * (size_t)ng), hh((size_t)m * m);
		RQC_HIP(hipMemcpyAsync(hacc.data(), acc, hacc.size() * sizeof(zc), hipMemcpyDeviceToHost, c.stream));
		RQC_HIP(hipMemcpyAsync(hh.data(), dh, hh.size() * sizeof(zc), hipMemcpyDeviceToHost, c.stream));
		s.download();
		// layer gradients summed in reverse pair order, as brickwall_unitary_backward does
		for (int i = 0; i < nlayers; i++) {
			zc sum[16];
			for (int e = 0; e < 16; e++) { sum[e] = {0, 0}; }
			for (int k = per - 1; k >= 0; k--) {
				for (int e = 0; e < 16; e++) { sum[e] += hacc[32 * ((size_t)i * per + k) + e]; }
			}
			std::memcpy(dvlist[i].data, sum, sizeof(sum));
		}
		for (int i = 0; i < nlayers; i++) {
			for (int x = 0; x < 16; x++) {
				for (int y = 0; y <= x; y++) {
					const size_t a = (size_t)(i * 16 + x) * m + i * 16 + y, b = (size_t)(i * 16 + y) * m + i * 16 + x;
					const zc sum = hh[a] + hh[b];
					hh[a] = sum;
					hh[b] = sum;
				}
			}
			for (int j = 0; j < i; j++) {
				for (int x = 0; x < 16; x++) {
					for (int y = 0; y < 16; y++) { hh[(size_t)(i * 16 + x) * m + j * 16 + y] = hh[(size_t)(j * 16 + y) * m + i * 16 + x]; }
				}
			}
		}
		std::memcpy(hess, hh.data(), hh.size() * sizeof(zc));
		return 0;
	} catch (const std::exception& e) { return report_error("brickwall_unitary_backward_hessian", e); }
}

// ============================================================== C ABI: target.h
static int target_call(const char* where, const std::function<void()>& body)
{
	try {
		body();
		return 0;
	} catch (const Error& e) {
		fprintf(stderr, "librqcopt_hip: %s failed: %s\n", where, e.what());
		return e.code;
	} catch (const std::exception& e) {
		fprintf(stderr, "librqcopt_hip: %s failed: %s\n", where, e.what());
		return -1;
	}
}

static void put_f(numeric* fval, const EvalResult& R) { std::memcpy(fval, &R.f, sizeof(zc)); }

extern "C" int circuit_unitary_target(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const int ngates, const int wires[], const int nqubits, numeric* fval)
{
	return target_call("circuit_unitary_target", [&] {
		Circuit C = make_circuit(gates, nullptr, ngates, wires, nqubits);
		put_f(fval, eval_targets(C, ufunc, udata, Mode::Target, 1, false));
	});
}

extern "C" int circuit_unitary_target_and_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const int ngates, const int wires[], const int nqubits, numeric* fval, struct mat4x4 dgates[])
{
	return target_call("circuit_unitary_target_and_gradient", [&] {
		Circuit C = make_circuit(gates, nullptr, ngates, wires, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Grad, 1, false);
		put_f(fval, R);
		std::memcpy(dgates, R.dv.data(), R.dv.size() * sizeof(zc));
	});
}

extern "C" int circuit_unitary_target_and_projected_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const int ngates, const int wires[], const int nqubits, numeric* fval, double* grad_vec)
{
	return target_call("circuit_unitary_target_and_projected_gradient", [&] {
		Circuit C = make_circuit(gates, nullptr, ngates, wires, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Grad, 1, true);
		put_f(fval, R);
		std::memcpy(grad_vec, R.grad.data(), R.grad.size() * sizeof(double));
	});
}

extern "C" int circuit_unitary_target_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const struct mat4x4 gatedirs[], const int ngates, const int wires[], const int nqubits,
	numeric* fval, struct mat4x4 dgates[], struct mat4x4 hess_gatedirs[])
{
	return target_call("circuit_unitary_target_hessian_vector_product", [&] {
		Circuit C = make_circuit(gates, gatedirs, ngates, wires, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Hvp, 1, false);
		put_f(fval, R);
		std::memcpy(dgates, R.dv.data(), R.dv.size() * sizeof(zc));
		std::memcpy(hess_gatedirs, R.hz.data(), R.hz.size() * sizeof(zc));
	});
}

extern "C" int circuit_unitary_target_projected_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const struct mat4x4 gatedirs[], const int ngates, const int wires[], const int nqubits,
	numeric* fval, double* RQC_RESTRICT grad_vec, double* RQC_RESTRICT hvp_vec)
{
	return target_call("circuit_unitary_target_projected_hessian_vector_product", [&] {
		Circuit C = make_circuit(gates, gatedirs, ngates, wires, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Hvp, 1, true);
		put_f(fval, R);
		std::memcpy(grad_vec, R.grad.data(), R.grad.size() * sizeof(double));
		std::memcpy(hvp_vec, R.hvp.data(), R.hvp.size() * sizeof(double));
	});
}

extern "C" int brickwall_unitary_target(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[], numeric* fval)
{
	return target_call("brickwall_unitary_target", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		put_f(fval, eval_targets(C, ufunc, udata, Mode::Target, nqubits / 2, false));
	});
}

extern "C" int brickwall_unitary_target_and_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[], numeric* fval, struct mat4x4 dvlist[])
{
	return target_call("brickwall_unitary_target_and_gradient", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Grad, nqubits / 2, false);
		put_f(fval, R);
		std::memcpy(dvlist, R.dv.data(), R.dv.size() * sizeof(zc));
	});
}

extern "C" int brickwall_unitary_target_and_projected_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[], numeric* fval, double* grad_vec)
{
	return target_call("brickwall_unitary_target_and_projected_gradient", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Grad, nqubits / 2, true);
		put_f(fval, R);
		std::memcpy(grad_vec, R.grad.data(), R.grad.size() * sizeof(double));
	});
}

extern "C" int brickwall_unitary_target_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const struct mat4x4 vdirs[], const int nlayers, const int* perms[], const int nqubits,
	numeric* fval, struct mat4x4 dvlist[], struct mat4x4 hess_vdirs[])
{
	return target_call("brickwall_unitary_target_hessian_vector_product", [&] {
		Circuit C = make_brickwall(vlist, vdirs, nlayers, perms, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Hvp, nqubits / 2, false);
		put_f(fval, R);
		std::memcpy(dvlist, R.dv.data(), R.dv.size() * sizeof(zc));
		std::memcpy(hess_vdirs, R.hz.data(), R.hz.size() * sizeof(zc));
	});
}

extern "C" int brickwall_unitary_target_projected_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const struct mat4x4 vdirs[], const int nlayers, const int* perms[], const int nqubits,
	numeric* fval, double* RQC_RESTRICT grad_vec, double* RQC_RESTRICT hvp_vec)
{
	return target_call("brickwall_unitary_target_projected_hessian_vector_product", [&] {
		Circuit C = make_brickwall(vlist, vdirs, nlayers, perms, nqubits);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Hvp, nqubits / 2, true);
		put_f(fval, R);
		std::memcpy(grad_vec, R.grad.data(), R.grad.size() * sizeof(double));
		std::memcpy(hvp_vec, R.hvp.data(), R.hvp.size() * sizeof(double));
	});
}

// ============================================================== no-allocation evaluation
// The evaluation above with the context's static scratch array as its whole
// device workspace: the column block is the widest whose plan (state arrays,
// partial sums, tables, results) fits kScratchBytes, and no device memory is
// allocated or freed.  U must be a device target (a host callback would need a
// device copy of U).  Calls hold the library lock and synchronise before they
// return, so they never overlap on the device; the scratch keeps the last
// call's contents, and every buffer that must start at zero is cleared by the
// evaluation's own fill kernels.
static EvalPlan plan_in_scratch(const Context& c, const Circuit& C, Mode mode, int per, bool fp32, bool project,
	bool unitary)
{
	size_t budget = kScratchBytes;
	for (int it = 0; it < 64 && budget > 0; it++) {
		EvalPlan pl;
		try {
			pl = plan_eval(c, C, mode, per, fp32, project, unitary, budget, 0);
		} catch (const Error&) {
			break;   // not even one column's state arrays fit
		}
		if (pl.bytes <= kScratchBytes) { return pl; }
		const size_t over = pl.bytes - kScratchBytes;
		budget = (over < budget) ? budget - over : 0;
	}
	throw Error(-3, "a " + std::to_string(C.L) + "-qubit circuit of " + std::to_string(C.ngates) +
		" gates does not fit the no-allocation scratch of " + std::to_string(kScratchBytes >> 20) + " MiB");
}

extern "C" int rqc_brickwall_unitary_target_projected_hvp_noalloc(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const struct mat4x4 vdirs[], const int nlayers, const int* perms[], const int nqubits,
	numeric* fval, double* RQC_RESTRICT grad_vec, double* RQC_RESTRICT hvp_vec)
{
	return target_call("rqc_brickwall_unitary_target_projected_hvp_noalloc", [&] {
		if (ufunc != rqc_device_ufunc) {
			throw Error(-3, "the no-allocation evaluation needs rqc_device_ufunc with an rqc_target* as udata");
		}
		Circuit C = make_brickwall(vlist, vdirs, nlayers, perms, nqubits);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const zc* dU = as_target(udata, C.L)->d_U;
		const bool fp32 = c.precision == RQC_PRECISION_MIXED;
		const EvalPlan pl = plan_in_scratch(c, C, Mode::Hvp, nqubits / 2, fp32, true, sweep_unitary(C, Mode::Hvp, fp32));
		std::vector<char> tab(pl.tab_bytes);
		fill_eval_tables(pl, C, tab.data());
		enqueue_eval(c, c, pl, C, dU, c.scratch, nullptr, nullptr, tab.data(), nullptr, nullptr, nullptr);
		std::vector<char> res(pl.res_bytes());
		RQC_HIP(hipMemcpyAsync(res.data(), c.scratch + pl.off_res, res.size(), hipMemcpyDeviceToHost, c.stream));
		RQC_HIP(hipStreamSynchronize(c.stream));
		RQC_HIP(hipGetLastError());
		EvalResult R;
		unpack_result(pl, res.data(), R);
		put_f(fval, R);
		std::memcpy(grad_vec, R.grad.data(), R.grad.size() * sizeof(double));
		std::memcpy(hvp_vec, R.hvp.data(), R.hvp.size() * sizeof(double));
	});
}

extern "C" size_t rqc_noalloc_scratch_bytes(void) { return kScratchBytes; }

// ============================================================== single-qubit gates
static Gate2<double> to_gate2(const numeric* g)
{
	Gate2<double> G;
	std::memcpy(G.g, g, sizeof(G.g));
	return G;
}

static void check_wire(int i, int L)
{
	if (L < 1 || i < 0 || i >= L) {
		throw Error(-3, "invalid gate wire " + std::to_string(i) + " for " + std::to_string(L) + " qubits");
	}
}

extern "C" int rqc_apply_single_qubit_gate(const numeric gate[4], int i, const struct statevector* RQC_RESTRICT psi,
	struct statevector* RQC_RESTRICT psi_out)
{
	return target_call("rqc_apply_single_qubit_gate", [&] {
		if (psi->nqubits != psi_out->nqubits) { throw Error(-3, "qubit counts differ"); }
		const int L = psi->nqubits;
		check_wire(i, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		Staging st(c);
		const int a = st.add(psi->data, bytes, true, false);
		const int o = st.add(psi_out->data, bytes, false, true);
		st.upload();
		const long n = (1L << L) / 2;
		{
			LaunchTimer lt(c, RQC_KERNEL_FORWARD, 2.0 * (double)bytes);
			k_gate1_forward<double><<<nblocks(n), kBlock, 0, c.stream>>>(st.dev<zc>(a), st.dev<zc>(o), to_gate2(gate),
				L - 1 - i, 1, 0, n);
		}
		st.download();
	});
}

extern "C" int rqc_apply_single_qubit_gate_to_array(const numeric gate[4], int i,
	const struct statevector_array* RQC_RESTRICT psi, struct statevector_array* RQC_RESTRICT psi_out)
{
	return target_call("rqc_apply_single_qubit_gate_to_array", [&] {
		if (psi->nqubits != psi_out->nqubits || psi->nstates != psi_out->nstates) { throw Error(-3, "array shapes differ"); }
		const int L = psi->nqubits;
		check_wire(i, L);
		if (psi->nstates < 1) { return; }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const long B = psi->nstates;
		const size_t bytes = ((size_t)1 << L) * B * sizeof(zc);
		Staging st(c);
		const int a = st.add(psi->data, bytes, true, false);
		const int o = st.add(psi_out->data, bytes, false, true);
		st.upload();
		const long n = (1L << L) / 2 * B;
		{
			LaunchTimer lt(c, RQC_KERNEL_FORWARD, 2.0 * (double)bytes);
			k_gate1_forward<double><<<nblocks(n), kBlock, 0, c.stream>>>(st.dev<zc>(a), st.dev<zc>(o), to_gate2(gate),
				L - 1 - i, B, ilog2_or_neg(B), n);
		}
		st.download();
	});
}

extern "C" int rqc_apply_single_qubit_gate_backward(const numeric gate[4], int i, const struct statevector* RQC_RESTRICT psi,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, numeric dgate[4])
{
	return target_call("rqc_apply_single_qubit_gate_backward", [&] {
		if (psi->nqubits != dpsi_out->nqubits || psi->nqubits != dpsi->nqubits) { throw Error(-3, "qubit counts differ"); }
		const int L = psi->nqubits;
		check_wire(i, L);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const size_t bytes = ((size_t)1 << L) * sizeof(zc);
		const long n = (1L << L) / 2;
		const unsigned g = reduce_grid(n);
		zc* acc = small_acc(c, 4);
		zc* part = acc + 32;   // small_acc holds 4096 * kWidth further entries, more than 4 per workgroup
		Staging st(c);
		const int a = st.add(psi->data, bytes, true, false);
		const int d = st.add(dpsi_out->data, bytes, true, false);
		const int o = st.add(dpsi->data, bytes, false, true);
		st.upload();
		{
			LaunchTimer lt(c, RQC_KERNEL_REVERSE, 3.0 * (double)bytes);
			k_gate1_reverse<double><<<g, kBlock, 0, c.stream>>>(st.dev<zc>(a), st.dev<zc>(d), st.dev<zc>(o), to_gate2(gate),
				L - 1 - i, 1, 0, n, part);
		}
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_reduce_partials<<<4, kBlock, 0, c.stream>>>(part, (int)g, 4, acc, nullptr);
		}
		copy_out(c, dgate, acc, 4 * sizeof(zc));
		st.download();
	});
}

// A circuit of one- and two-qubit gates as a circuit of two-qubit gates: a one-qubit gate V
// on wire i becomes V (x) I on wires (i, (i + 1) mod L) (first wire the more significant
// index: (V (x) I)[2a + c][2b + d] = V[a][b] delta_cd), its direction Z (x) I.
static Circuit mixed_circuit(const numeric* gates, const numeric* dirs, const int* nq, const int* wires, int ngates, int L,
	std::vector<size_t>& off)
{
	if (ngates < 1) { throw Error(-3, "a circuit needs at least one gate"); }
	if (L < 2) { throw Error(-3, "a circuit needs at least two qubits"); }
	std::vector<struct mat4x4> g4((size_t)ngates), d4((size_t)ngates);
	std::vector<int> w4(2 * (size_t)ngates);
	off.assign((size_t)ngates + 1, 0);
	for (int k = 0; k < ngates; k++) {
		if (nq[k] != 1 && nq[k] != 2) { throw Error(-3, "gate " + std::to_string(k) + " acts on neither 1 nor 2 qubits"); }
		const size_t o = off[k];
		off[k + 1] = o + (nq[k] == 1 ? 4 : 16);
		if (nq[k] == 2) {
			std::memcpy(g4[k].data, gates + o, 16 * sizeof(numeric));
			if (dirs != nullptr) { std::memcpy(d4[k].data, dirs + o, 16 * sizeof(numeric)); }
			w4[2 * k] = wires[2 * k];
			w4[2 * k + 1] = wires[2 * k + 1];
			continue;
		}
		check_wire(wires[2 * k], L);
		for (auto& e : g4[k].data) { e = numeric(0.0, 0.0); }
		for (auto& e : d4[k].data) { e = numeric(0.0, 0.0); }
		for (int a = 0; a < 2; a++) {
			for (int b = 0; b < 2; b++) {
				for (int cc = 0; cc < 2; cc++) {
					g4[k].data[4 * (2 * a + cc) + 2 * b + cc] = gates[o + 2 * a + b];
					if (dirs != nullptr) { d4[k].data[4 * (2 * a + cc) + 2 * b + cc] = dirs[o + 2 * a + b]; }
				}
			}
		}
		w4[2 * k] = wires[2 * k];
		w4[2 * k + 1] = (wires[2 * k] + 1) % L;
	}
	return make_circuit(g4.data(), dirs != nullptr ? d4.data() : nullptr, ngates, w4.data(), L);
}

// per-gate sums of the two-qubit circuit back to the mixed packing: a one-qubit gate's entry
// (a, b) is the partial trace sum_c X[2a + c][2b + c] over the partner wire
static void unpack_mixed(const std::vector<zc>& x4, const int* nq, int ngates, const std::vector<size_t>& off, numeric* out)
{
	for (int k = 0; k < ngates; k++) {
		const zc* x = &x4[16 * (size_t)k];
		if (nq[k] == 2) {
			std::memcpy(out + off[k], x, 16 * sizeof(zc));
			continue;
		}
		for (int a = 0; a < 2; a++) {
			for (int b = 0; b < 2; b++) {
				const zc s = x[4 * (2 * a) + 2 * b] + x[4 * (2 * a + 1) + 2 * b + 1];
				out[off[k] + 2 * a + b] = numeric(s.x, s.y);
			}
		}
	}
}

extern "C" int rqc_mixed_circuit_unitary_target(linear_func ufunc, void* udata, const numeric* gates, const int nq[],
	const int wires[], int ngates, int nqubits, numeric* fval)
{
	return target_call("rqc_mixed_circuit_unitary_target", [&] {
		std::vector<size_t> off;
		Circuit C = mixed_circuit(gates, nullptr, nq, wires, ngates, nqubits, off);
		put_f(fval, eval_targets(C, ufunc, udata, Mode::Target, 1, false));
	});
}

extern "C" int rqc_mixed_circuit_unitary_target_and_gradient(linear_func ufunc, void* udata, const numeric* gates,
	const int nq[], const int wires[], int ngates, int nqubits, numeric* fval, numeric* dgates)
{
	return target_call("rqc_mixed_circuit_unitary_target_and_gradient", [&] {
		std::vector<size_t> off;
		Circuit C = mixed_circuit(gates, nullptr, nq, wires, ngates, nqubits, off);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Grad, 1, false);
		put_f(fval, R);
		unpack_mixed(R.dv, nq, ngates, off, dgates);
	});
}

extern "C" int rqc_mixed_circuit_unitary_target_hessian_vector_product(linear_func ufunc, void* udata, const numeric* gates,
	const numeric* gatedirs, const int nq[], const int wires[], int ngates, int nqubits, numeric* fval, numeric* dgates,
	numeric* hess_gatedirs)
{
	return target_call("rqc_mixed_circuit_unitary_target_hessian_vector_product", [&] {
		std::vector<size_t> off;
		Circuit C = mixed_circuit(gates, gatedirs, nq, wires, ngates, nqubits, off);
		EvalResult R = eval_targets(C, ufunc, udata, Mode::Hvp, 1, false);
		put_f(fval, R);
		unpack_mixed(R.dv, nq, ngates, off, dgates);
		unpack_mixed(R.hz, nq, ngates, off, hess_gatedirs);
	});
}

// ============================================================== sampled (stochastic-trace) targets
// The trace over the 2^L basis columns is replaced by nsamples Haar-random
// states (reference src/target.c:85-149, :807-928): f, gradients and Hessians
// are the sums over the samples scaled by 2^L / nsamples.  The states are drawn
// on the host from the caller's generator in the reference's order (sample by
// sample, entry by entry; crandn then normalisation), so the generator's
// stream and the samples match the reference's.  The trace seeds U psi come
// from one rocBLAS GEMM per column block for a device-resident U, or from the
// caller's callback, one call per sample as in the reference.

// psi_j = crandn, then psi /= ||psi|| (reference src/statevector.c:47-66, :102-127)
static void draw_haar_state(long n, struct rng_state* rng, zc* psi)
{
	double s = 0;
	for (long j = 0; j < n; j++) {
		const numeric z = crandn(rng);
		psi[j] = zc{z.real(), z.imag()};
		s += z.real() * z.real() + z.imag() * z.imag();
	}
	const double nrm = std::sqrt(s);
	if (nrm > 0) {
		const double inv = 1.0 / nrm;
		for (long j = 0; j < n; j++) { psi[j] = zc{psi[j].x * inv, psi[j].y * inv}; }
	}
}

static rocblas_handle blas_handle(Context& c)
{
	static rocblas_handle h = nullptr;
	if (h == nullptr) {
		if (rocblas_create_handle(&h) != rocblas_status_success) { throw Error(-1, "rocblas_create_handle failed"); }
	}
	if (rocblas_set_stream(h, c.stream) != rocblas_status_success) { throw Error(-1, "rocblas_set_stream failed"); }
	return h;
}

// device sample blocks X (sweep precision) and seeds W = U X (FP64), in the
// plan's block layout: block k holds samples k B .. k B + B - 1 as an N x B
// array, zero-padded past nsamples; both live in c.samples
struct SampleBlocks
{
	const void* X = nullptr;
	const zc* W = nullptr;
};

static SampleBlocks prepare_samples(Context& c, const EvalPlan& pl, linear_func ufunc, void* udata, struct rng_state* rng)
{
	const long N = pl.N, B = pl.B, nb = pl.nblk;
	const size_t sz = (size_t)N * B;
	const size_t xbytes = nb * sz * sizeof(zc);
	const size_t x32 = pl.fp32 ? nb * sz * sizeof(cx<float>) : 0;
	c.samples.reserve(2 * xbytes + x32);
	zc* X64 = c.samples.as<zc>();
	zc* W = X64 + nb * sz;
	std::vector<zc> hx(nb * sz, zc{0, 0}), col((size_t)N);
	std::vector<zc> hw;
	const bool device_u = (ufunc == rqc_device_ufunc);
	if (!device_u) {
		if (ufunc == nullptr) { throw Error(-3, "ufunc is NULL"); }
		hw.assign(nb * sz, zc{0, 0});
	}
	std::vector<zc> out((size_t)N);
	for (long b = 0; b < pl.ncols; b++) {
		draw_haar_state(N, rng, col.data());
		const size_t base = (size_t)(b / B) * sz + (size_t)(b % B);
		for (long a = 0; a < N; a++) { hx[base + (size_t)a * B] = col[a]; }
		if (!device_u) {
			struct statevector sin = {reinterpret_cast<numeric*>(col.data()), pl.L};
			struct statevector sout = {reinterpret_cast<numeric*>(out.data()), pl.L};
			const int r = ufunc(&sin, udata, &sout);
			if (r < 0) { throw Error(-2, "call of 'ufunc' failed, return value: " + std::to_string(r)); }
			for (long a = 0; a < N; a++) { hw[base + (size_t)a * B] = out[a]; }
		}
	}
	RQC_HIP(hipMemcpyAsync(X64, hx.data(), xbytes, hipMemcpyHostToDevice, c.stream));
	if (device_u) {
		// W_k = U X_k: row-major N x B = (X_k^T U^T)^T, i.e. the column-major product of X_k (B x N) and U (N x N)
		const zc* U = device_target_matrix(udata, pl.L);
		rocblas_handle h = blas_handle(c);
		const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
		for (long k = 0; k < nb; k++) {
			const rocblas_status st = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)B,
				(rocblas_int)N, (rocblas_int)N, &one, reinterpret_cast<const rocblas_double_complex*>(X64 + k * sz),
				(rocblas_int)B, reinterpret_cast<const rocblas_double_complex*>(U), (rocblas_int)N, &zero,
				reinterpret_cast<rocblas_double_complex*>(W + k * sz), (rocblas_int)B);
			if (st != rocblas_status_success) { throw Error(-1, std::string("rocblas_zgemm failed: ") + rocblas_status_to_string(st)); }
		}
	} else {
		RQC_HIP(hipMemcpyAsync(W, hw.data(), xbytes, hipMemcpyHostToDevice, c.stream));
	}
	SampleBlocks sb;
	sb.W = W;
	if (pl.fp32) {
		cx<float>* X32 = reinterpret_cast<cx<float>*>(W + nb * sz);
		k_to_float<<<nblocks((long)(nb * sz)), kBlock, 0, c.stream>>>(X64, X32, (long)(nb * sz));
		sb.X = X32;
	} else {
		sb.X = X64;
	}
	// the host buffers are pageable and go out of scope: finish the copies here
	RQC_HIP(hipStreamSynchronize(c.stream));
	return sb;
}

static void scale_result(EvalResult& R, double a)
{
	R.f = a * R.f;
	for (auto& z : R.dv) { z = a * z; }
	for (auto& z : R.hz) { z = a * z; }
	for (auto& x : R.grad) { x *= a; }
	for (auto& x : R.hvp) { x *= a; }
}

// ============================================================== block-encoding targets
// f = ||P^dagger W P - H||_F^2 / 2 (reference src/target.c:1100-1720).  The
// sweeps run over the 2^(L/2) embedded basis columns P e_b as a sampled plan,
// with the residual kernel in place of the trace.  H is assembled from the
// caller's hfunc on the unit vectors (2^(L/2) calls, as the reference makes per
// evaluation) or read in place from a device-resident target of L/2 qubits.
struct BlockEncInputs
{
	const void* X = nullptr;  // embedded columns in the plan's blocks, sweep precision
	const zc* H = nullptr;    // 2^(L/2) x 2^(L/2), row-major
	zc* J = nullptr;          // njac projected Jacobian columns of 2^L entries each
};

static BlockEncInputs prepare_blockenc(Context& c, const EvalPlan& pl, linear_func hfunc, void* hdata, int njac)
{
	if (hfunc == nullptr) { throw Error(-3, "hfunc is NULL"); }
	const int half = pl.L / 2;
	const long s = 1L << half, N = pl.N, B = pl.B;
	const size_t sz = (size_t)N * B, nx = (size_t)pl.nblk * sz;
	const size_t xbytes = align256(nx * (pl.fp32 ? sizeof(cx<float>) : sizeof(zc)));
	const size_t hbytes = align256((size_t)s * s * sizeof(zc));
	c.samples.reserve(xbytes + hbytes + (size_t)njac * N * sizeof(zc));
	char* base = c.samples.as<char>();
	const int logB = ilog2_or_neg(B);
	{
		LaunchTimer lt(c, RQC_KERNEL_OTHER);
		if (pl.fp32) {
			k_blockenc_columns<float><<<nblocks((long)nx), kBlock, 0, c.stream>>>(reinterpret_cast<cx<float>*>(base), half, B,
				logB, (long)sz, (long)nx);
		} else {
			k_blockenc_columns<double><<<nblocks((long)nx), kBlock, 0, c.stream>>>(reinterpret_cast<cx<double>*>(base), half,
				B, logB, (long)sz, (long)nx);
		}
	}
	BlockEncInputs bi;
	bi.X = base;
	if (hfunc == rqc_device_ufunc) {
		bi.H = device_target_matrix(hdata, half);
	} else {
		std::vector<zc> h((size_t)s * s), e((size_t)s), out((size_t)s);
		for (long b = 0; b < s; b++) {
			std::fill(e.begin(), e.end(), zc{0, 0});
			e[b] = zc{1, 0};
			struct statevector sin = {reinterpret_cast<numeric*>(e.data()), half};
			struct statevector sout = {reinterpret_cast<numeric*>(out.data()), half};
			const int r = hfunc(&sin, hdata, &sout);
			if (r < 0) { throw Error(-2, "call of 'hfunc' failed, return value: " + std::to_string(r)); }
			for (long a = 0; a < s; a++) { h[(size_t)a * s + b] = out[a]; }
		}
		zc* Hd = reinterpret_cast<zc*>(base + xbytes);
		RQC_HIP(hipMemcpyAsync(Hd, h.data(), h.size() * sizeof(zc), hipMemcpyHostToDevice, c.stream));
		// the host image is pageable and goes out of scope
		RQC_HIP(hipStreamSynchronize(c.stream));
		bi.H = Hd;
	}
	bi.J = (njac > 0) ? reinterpret_cast<zc*>(base + xbytes + hbytes) : nullptr;
	return bi;
}

static EvalResult eval_blockenc(const Circuit& C, linear_func hfunc, void* hdata, Mode mode, bool project)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	const bool unitary = sweep_unitary(C, mode, c.precision == RQC_PRECISION_MIXED);
	const EvalPlan pl = plan_eval(c, C, mode, C.L / 2, c.precision == RQC_PRECISION_MIXED, project, unitary, c.ws_limit,
		c.work.bytes, nullptr, 1L << (C.L / 2));
	const BlockEncInputs bi = prepare_blockenc(c, pl, hfunc, hdata, 0);
	std::vector<char> tab(pl.tab_bytes);
	fill_eval_tables(pl, C, tab.data());
	c.work.reserve(pl.bytes);
	char* work = c.work.as<char>();
	const BlockEnc be{bi.H, nullptr};
	enqueue_eval(c, c, pl, C, nullptr, work, nullptr, nullptr, tab.data(), nullptr, bi.X, nullptr, &be);
	std::vector<char> res(pl.res_bytes());
	RQC_HIP(hipMemcpyAsync(res.data(), work + pl.off_res, res.size(), hipMemcpyDeviceToHost, c.stream));
	RQC_HIP(hipStreamSynchronize(c.stream));
	RQC_HIP(hipGetLastError());
	EvalResult R;
	unpack_result(pl, res.data(), R);
	return R;
}

// Complex Hessian of the brickwall target (reference src/target.c:687-797).  Column
// (j, y) is the HVP with the unit direction E_y on every gate of layer j, so the
// 16 nlayers columns are HVP evaluations of one plan: their tables go up in one
// copy, the evaluations are enqueued back to back into one workspace with a
// result region each, and the results come back in one copy.  The matrix is
// symmetric up to rounding and is symmetrised; f and dvlist come with column 0.
static void eval_hessian(const Circuit& C0, linear_func ufunc, void* udata, int per, zc* fval, zc* dvlist, zc* hess,
	long nsamples = 0, struct rng_state* rng = nullptr, zc* hess2 = nullptr)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	Circuit C = C0;
	C.dirs.assign(C.gates.size(), zc{0, 0});
	const int ng = C.ngates / per, m = 16 * ng;
	// with hess2 the objective is the block encoding of the operator (ufunc, udata) on L/2 qubits
	const bool blockenc = (hess2 != nullptr);
	const bool unitary = gates_near_unitary(C);
	const EvalPlan pl = plan_eval(c, C, Mode::Hvp, per, c.precision == RQC_PRECISION_MIXED, false, unitary, c.ws_limit,
		c.work.bytes, nullptr, blockenc ? (1L << (C.L / 2)) : nsamples);
	const zc* dU = nullptr;
	SampleBlocks sb;
	BlockEncInputs bi;
	if (blockenc) {
		bi = prepare_blockenc(c, pl, ufunc, udata, m);
		sb.X = bi.X;
	} else if (pl.sampled) {
		if (rng == nullptr) { throw Error(-3, "rng is NULL"); }
		sb = prepare_samples(c, pl, ufunc, udata, rng);
	} else {
		dU = resolve_target(c, ufunc, udata, C0.L);
	}
	const double scale = (pl.sampled && !blockenc) ? (double)pl.N / (double)nsamples : 1.0;
	const size_t ts = align256(pl.tab_bytes), rs = align256(pl.res_bytes()), base = align256(pl.bytes);
	std::vector<char> htab((size_t)m * ts);
	std::vector<Circuit> cols((size_t)m, C);
	for (int col = 0; col < m; col++) {
		Circuit& Cc = cols[col];
		const int jl = col / 16, y = col % 16;
		for (int q = 0; q < per; q++) { Cc.dirs[16 * ((size_t)jl * per + q) + y] = zc{1, 0}; }
		fill_eval_tables(pl, Cc, htab.data() + (size_t)col * ts);
	}
	const size_t gbytes = blockenc ? (size_t)m * m * sizeof(zc) : 0;
	c.work.reserve(base + (size_t)m * (ts + rs) + gbytes);
	char* work = c.work.as<char>();
	char* tabs = work + base;
	char* ress = tabs + (size_t)m * ts;
	RQC_HIP(hipMemcpyAsync(tabs, htab.data(), htab.size(), hipMemcpyHostToDevice, c.stream));
	for (int col = 0; col < m; col++) {
		const BlockEnc be{bi.H, blockenc ? bi.J + (size_t)col * pl.N : nullptr};
		enqueue_eval(c, c, pl, cols[col], dU, work, tabs + (size_t)col * ts, ress + (size_t)col * rs, nullptr, nullptr,
			sb.X, sb.W, blockenc ? &be : nullptr);
	}
	std::vector<char> hres((size_t)m * rs);
	RQC_HIP(hipMemcpyAsync(hres.data(), ress, hres.size(), hipMemcpyDeviceToHost, c.stream));
	if (blockenc) {
		// hess2: Gram matrix of the projected Jacobian columns P^dagger dW P e_b of the m unit directions
		zc* gram = reinterpret_cast<zc*>(ress + (size_t)m * rs);
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_gram_hermitian<<<dim3(m, m), kBlock, 0, c.stream>>>(bi.J, pl.N, m, gram);
		}
		RQC_HIP(hipMemcpyAsync(hess2, gram, gbytes, hipMemcpyDeviceToHost, c.stream));
	}
	RQC_HIP(hipStreamSynchronize(c.stream));
	RQC_HIP(hipGetLastError());
	std::vector<zc> H((size_t)m * m);
	for (int col = 0; col < m; col++) {
		EvalResult R;
		unpack_result(pl, hres.data() + (size_t)col * rs, R);
		scale_result(R, scale);
		if (col == 0) {
			*fval = R.f;
			std::memcpy(dvlist, R.dv.data(), R.dv.size() * sizeof(zc));
		}
		for (int r = 0; r < m; r++) { H[(size_t)r * m + col] = R.hz[r]; }
	}
	for (int a = 0; a < m; a++) {
		for (int b = 0; b <= a; b++) {
			const zc v = 0.5 * (H[(size_t)a * m + b] + H[(size_t)b * m + a]);
			hess[(size_t)a * m + b] = v;
			hess[(size_t)b * m + a] = v;
		}
	}
}

extern "C" int brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	numeric* fval, struct mat4x4 dvlist[], numeric* hess)
{
	return target_call("brickwall_unitary_target_gradient_hessian", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		zc f;
		eval_hessian(C, ufunc, udata, nqubits / 2, &f, reinterpret_cast<zc*>(dvlist), reinterpret_cast<zc*>(hess));
		std::memcpy(fval, &f, sizeof(zc));
	});
}

static EvalResult eval_sampled(const Circuit& C, linear_func ufunc, void* udata, Mode mode, int per, bool project, long nsamples,
	struct rng_state* rng)
{
	if (nsamples < 1) { throw Error(-3, "nsamples must be positive"); }
	if (rng == nullptr) { throw Error(-3, "rng is NULL"); }
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	const bool unitary = sweep_unitary(C, mode, c.precision == RQC_PRECISION_MIXED);
	const EvalPlan pl = plan_eval(c, C, mode, per, c.precision == RQC_PRECISION_MIXED, project, unitary, c.ws_limit,
		c.work.bytes, nullptr, nsamples);
	const SampleBlocks sb = prepare_samples(c, pl, ufunc, udata, rng);
	std::vector<char> tab(pl.tab_bytes);
	fill_eval_tables(pl, C, tab.data());
	c.work.reserve(pl.bytes);
	char* work = c.work.as<char>();
	enqueue_eval(c, c, pl, C, nullptr, work, nullptr, nullptr, tab.data(), nullptr, sb.X, sb.W);
	std::vector<char> res(pl.res_bytes());
	RQC_HIP(hipMemcpyAsync(res.data(), work + pl.off_res, res.size(), hipMemcpyDeviceToHost, c.stream));
	RQC_HIP(hipStreamSynchronize(c.stream));
	RQC_HIP(hipGetLastError());
	EvalResult R;
	unpack_result(pl, res.data(), R);
	scale_result(R, (double)pl.N / (double)nsamples);
	return R;
}

extern "C" int circuit_unitary_target_sampling(linear_func ufunc, void* udata, const struct mat4x4 gates[], const int ngates,
	const int wires[], const int nqubits, const long nsamples, struct rng_state* rng, numeric* fval)
{
	return target_call("circuit_unitary_target_sampling", [&] {
		Circuit C = make_circuit(gates, nullptr, ngates, wires, nqubits);
		put_f(fval, eval_sampled(C, ufunc, udata, Mode::Target, 1, false, nsamples, rng));
	});
}

extern "C" int brickwall_unitary_target_sampling(linear_func ufunc, void* udata, const struct mat4x4 vlist[], const int nlayers,
	const int nqubits, const int* perms[], const long nsamples, struct rng_state* rng, numeric* fval)
{
	return target_call("brickwall_unitary_target_sampling", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		put_f(fval, eval_sampled(C, ufunc, udata, Mode::Target, nqubits / 2, false, nsamples, rng));
	});
}

extern "C" int brickwall_unitary_target_gradient_hessian_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[], const long nsamples,
	struct rng_state* rng, numeric* fval, struct mat4x4 dvlist[], numeric* hess)
{
	return target_call("brickwall_unitary_target_gradient_hessian_sampling", [&] {
		if (nsamples < 1) { throw Error(-3, "nsamples must be positive"); }
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		zc f;
		eval_hessian(C, ufunc, udata, nqubits / 2, &f, reinterpret_cast<zc*>(dvlist), reinterpret_cast<zc*>(hess), nsamples, rng);
		std::memcpy(fval, &f, sizeof(zc));
	});
}

extern "C" int brickwall_blockenc_target(linear_func hfunc, void* hdata, const struct mat4x4 vlist[], const int nlayers,
	const int nqubits, const int* perms[], double* fval)
{
	return target_call("brickwall_blockenc_target", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		*fval = eval_blockenc(C, hfunc, hdata, Mode::Target, false).f.x;
	});
}

extern "C" int brickwall_blockenc_target_and_gradient(linear_func hfunc, void* hdata, const struct mat4x4 vlist[],
	const int nlayers, const int nqubits, const int* perms[], double* fval, struct mat4x4 dvlist[])
{
	return target_call("brickwall_blockenc_target_and_gradient", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		const EvalResult R = eval_blockenc(C, hfunc, hdata, Mode::Grad, false);
		*fval = R.f.x;
		std::memcpy(dvlist, R.dv.data(), R.dv.size() * sizeof(zc));
	});
}

extern "C" int brickwall_blockenc_target_and_gradient_vector(linear_func hfunc, void* hdata, const struct mat4x4 vlist[],
	const int nlayers, const int nqubits, const int* perms[], double* fval, double* grad_vec)
{
	return target_call("brickwall_blockenc_target_and_gradient_vector", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		const EvalResult R = eval_blockenc(C, hfunc, hdata, Mode::Grad, true);
		*fval = R.f.x;
		std::memcpy(grad_vec, R.grad.data(), R.grad.size() * sizeof(double));
	});
}

extern "C" int brickwall_blockenc_target_gradient_hessian(linear_func hfunc, void* hdata, const struct mat4x4 vlist[],
	const int nlayers, const int nqubits, const int* perms[], double* fval, struct mat4x4 dvlist[], numeric* hess1,
	numeric* hess2)
{
	return target_call("brickwall_blockenc_target_gradient_hessian", [&] {
		Circuit C = make_brickwall(vlist, nullptr, nlayers, perms, nqubits);
		zc f;
		eval_hessian(C, hfunc, hdata, nqubits / 2, &f, reinterpret_cast<zc*>(dvlist), reinterpret_cast<zc*>(hess1), 0,
			nullptr, reinterpret_cast<zc*>(hess2));
		*fval = f.x;
	});
}

// ============================================================== C ABI: rqc_device.h
extern "C" int rqc_device_ufunc(const struct statevector* RQC_RESTRICT psi, void* fdata, struct statevector* RQC_RESTRICT psi_out)
{
	return target_call("rqc_device_ufunc", [&] {
		if (psi->nqubits != psi_out->nqubits) { throw Error(-3, "qubit counts differ"); }
		const rqc_target* t = as_target(fdata, psi->nqubits);
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const long n = 1L << psi->nqubits;
		Staging s(c);
		const int a = s.add(psi->data, n * sizeof(zc), true, false);
		const int o = s.add(psi_out->data, n * sizeof(zc), false, true);
		s.upload();
		{
			LaunchTimer lt(c, RQC_KERNEL_OTHER);
			k_matvec<<<nblocks(n), kBlock, 0, c.stream>>>(t->d_U, s.dev<zc>(a), s.dev<zc>(o), n);
		}
		s.download();
	});
}

extern "C" void rqc_target_destroy(struct rqc_target* target);

static rqc_target* new_target(int L)
{
	if (L < 1 || L > 15) { throw Error(-3, "dense targets support 1 .. 15 qubits"); }
	Context& c = ctx();
	(void)c;
	auto* t = new rqc_target();
	t->magic = kTargetMagic;
	t->nqubits = L;
	const size_t N = (size_t)1 << L;
	hipError_t e = hipMalloc(&t->d_U, N * N * sizeof(zc));
	if (e != hipSuccess) {
		delete t;
		throw Error(-1, std::string("allocating the device target failed: ") + hipGetErrorString(e));
	}
	return t;
}

extern "C" int rqc_target_create_dense(int nqubits, const numeric* U, struct rqc_target** target)
{
	return target_call("rqc_target_create_dense", [&] {
		rqc_target* t = new_target(nqubits);
		Context& c = ctx();
		const size_t N = (size_t)1 << nqubits;
		copy_in(c, t->d_U, U, N * N * sizeof(zc));
		RQC_HIP(hipStreamSynchronize(c.stream));
		*target = t;
	});
}

extern "C" int rqc_target_create_from_ufunc(linear_func ufunc, void* udata, int nqubits, struct rqc_target** target)
{
	return target_call("rqc_target_create_from_ufunc", [&] {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const zc* dU = resolve_target(c, ufunc, udata, nqubits);
		rqc_target* t = new_target(nqubits);
		const size_t N = (size_t)1 << nqubits;
		dcopy(c, t->d_U, dU, N * N * sizeof(zc));
		RQC_HIP(hipStreamSynchronize(c.stream));
		*target = t;
	});
}

// U = C, formed column block by column block with the per-gate forward kernel
static rqc_target* target_from_circuit(const Circuit& C)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	rqc_target* t = new_target(C.L);
	try {
		const long N = 1L << C.L;
		const long B = choose_block(N, 2, sizeof(zc), c.ws_limit, c.work.bytes);
		c.work.reserve(2 * (size_t)N * B * sizeof(zc));
		zc* S = c.work.as<zc>();
		for (long b0 = 0; b0 < N; b0 += B) {
			zc* buf[2] = {S, S + (size_t)N * B};
			k_identity_block<double><<<nblocks(N * B), kBlock, 0, c.stream>>>(buf[0], B, ilog2_or_neg(B), b0, N * B);
			for (int k = 0; k < C.ngates; k++) {
				launch_forward(c, &C.gates[16 * (size_t)k], C.wi[k], C.wj[k], C.L, B, buf[k % 2], buf[(k + 1) % 2]);
			}
			k_store_columns<<<nblocks(N * B), kBlock, 0, c.stream>>>(buf[C.ngates % 2], t->d_U, N, B, b0);
		}
		RQC_HIP(hipStreamSynchronize(c.stream));
		RQC_HIP(hipGetLastError());
	} catch (...) {
		rqc_target_destroy(t);
		throw;
	}
	return t;
}

extern "C" int rqc_target_create_brickwall(int nqubits, const struct mat4x4 vlist[], int nlayers, const int* perms[], struct rqc_target** target)
{
	return target_call("rqc_target_create_brickwall", [&] {
		*target = target_from_circuit(make_brickwall(vlist, nullptr, nlayers, perms, nqubits));
	});
}

extern "C" int rqc_target_create_circuit(int nqubits, const struct mat4x4 gates[], int ngates, const int wires[], struct rqc_target** target)
{
	return target_call("rqc_target_create_circuit", [&] {
		*target = target_from_circuit(make_circuit(gates, nullptr, ngates, wires, nqubits));
	});
}

extern "C" void rqc_target_destroy(struct rqc_target* target)
{
	if (target == nullptr || target->magic != kTargetMagic) { return; }
	(void)hipFree(target->d_U);
	target->magic = 0;
	delete target;
}

extern "C" int rqc_target_nqubits(const struct rqc_target* target) { return target != nullptr ? target->nqubits : -1; }

extern "C" int rqc_target_download(const struct rqc_target* target, numeric* U)
{
	return target_call("rqc_target_download", [&] {
		as_target(const_cast<rqc_target*>(target), -1);
		Context& c = ctx();
		const size_t N = (size_t)1 << target->nqubits;
		copy_out(c, U, target->d_U, N * N * sizeof(zc));
		RQC_HIP(hipStreamSynchronize(c.stream));
	});
}

extern "C" int rqc_set_precision(int mode)
{
	if (mode != RQC_PRECISION_FP64 && mode != RQC_PRECISION_MIXED) { return -3; }
	return target_call("rqc_set_precision", [&] { ctx().precision = mode; });
}

extern "C" int rqc_set_sweep_engine(int engine)
{
	if (engine != RQC_ENGINE_TILE_PASS && engine != RQC_ENGINE_PER_GATE) { return -3; }
	return target_call("rqc_set_sweep_engine", [&] { ctx().engine = engine; });
}

extern "C" int rqc_set_eval_autotune(int on)
{
	return target_call("rqc_set_eval_autotune", [&] {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		c.eval_autotune = (on != 0);
		if (on == 2) { tuned_choices().clear(); }
	});
}

extern "C" int rqc_last_tiles(int* tile_rows_log2, int* tile_cols_log2, int* gates_per_pass)
{
	return target_call("rqc_last_tiles", [&] {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		if (tile_rows_log2 != nullptr) { *tile_rows_log2 = c.last_K; }
		if (tile_cols_log2 != nullptr) { *tile_cols_log2 = c.last_logC; }
		if (gates_per_pass != nullptr) { *gates_per_pass = c.last_gmax; }
	});
}

extern "C" int rqc_set_autotune(int on)
{
	return target_call("rqc_set_autotune", [&] { ctx().autotune = (on != 0); });
}

extern "C" int rqc_get_precision(void)
{
	try { return ctx().precision; } catch (...) { return -1; }
}

extern "C" int rqc_set_workspace_limit(size_t bytes)
{
	return target_call("rqc_set_workspace_limit", [&] { ctx().ws_limit = bytes; });
}

extern "C" void rqc_release_workspace(void)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		(void)hipStreamSynchronize(c.stream);
		c.work.release();
		c.small.release();
		c.vecs.release();
		c.ubuf.release();
	} catch (...) {}
}

extern "C" long rqc_last_column_block(void)
{
	try { return ctx().last_block; } catch (...) { return -1; }
}

extern "C" int rqc_last_sweep(int* engine, int* passes)
{
	return target_call("rqc_last_sweep", [&] {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		if (engine != nullptr) { *engine = c.last_engine; }
		if (passes != nullptr) { *passes = c.last_passes; }
	});
}

extern "C" void rqc_profile_enable(int on)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		profile_flush(c);
		c.profile = (on != 0);
	} catch (const std::exception& e) { fatal("rqc_profile_enable", e); }
}

extern "C" void rqc_profile_reset(void)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		profile_flush(c);
		for (int k = 0; k < RQC_KERNEL_KINDS; k++) {
			c.prof_ms[k] = 0;
			c.prof_n[k] = 0;
			c.prof_bytes[k] = 0;
			c.prof_flops[k] = 0;
			c.prof_issued[k] = 0;
		}
	} catch (const std::exception& e) { fatal("rqc_profile_reset", e); }
}

extern "C" int rqc_profile_query(int kind, double* total_ms, long* launches, double* alg_bytes)
{
	if (kind < 0 || kind >= RQC_KERNEL_KINDS) { return -3; }
	return target_call("rqc_profile_query", [&] {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		profile_flush(c);
		if (total_ms != nullptr) { *total_ms = c.prof_ms[kind]; }
		if (launches != nullptr) { *launches = c.prof_n[kind]; }
		if (alg_bytes != nullptr) { *alg_bytes = c.prof_bytes[kind]; }
	});
}

extern "C" int rqc_profile_query_flops(int kind, double* alg_flops, double* issued_flops)
{
	if (kind < 0 || kind >= RQC_KERNEL_KINDS) { return -3; }
	return target_call("rqc_profile_query_flops", [&] {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		profile_flush(c);
		if (alg_flops != nullptr) { *alg_flops = c.prof_flops[kind]; }
		if (issued_flops != nullptr) { *issued_flops = c.prof_issued[kind]; }
	});
}

extern "C" int rqc_measure_copy_bandwidth(size_t bytes, int reps, double* gbps)
{
	return target_call("rqc_measure_copy_bandwidth", [&] {
		if (reps < 1 || bytes < 16) { throw Error(-3, "need reps >= 1 and bytes >= 16"); }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const long n = (long)(bytes / 16);
		c.work.reserve(2 * (size_t)n * 16);
		zc* a = c.work.as<zc>();
		zc* b = a + n;
		const unsigned grid = (unsigned)std::min<long>((n + kBlock - 1) / kBlock, 8L * c.cu_count * 8);
		k_fill_zero<<<nblocks(n), kBlock, 0, c.stream>>>(a, n, nullptr);
		hipEvent_t e0, e1;
		RQC_HIP(hipEventCreate(&e0));
		RQC_HIP(hipEventCreate(&e1));
		k_copy16<<<grid, kBlock, 0, c.stream>>>(a, b, n);  // warm-up
		RQC_HIP(hipEventRecord(e0, c.stream));
		for (int r = 0; r < reps; r++) { k_copy16<<<grid, kBlock, 0, c.stream>>>(a, b, n); }
		RQC_HIP(hipEventRecord(e1, c.stream));
		RQC_HIP(hipEventSynchronize(e1));
		float ms = 0;
		RQC_HIP(hipEventElapsedTime(&ms, e0, e1));
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		*gbps = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;
	});
}

extern "C" int rqc_measure_stream_bandwidth(size_t bytes, int reps, double* gbps)
{
	return target_call("rqc_measure_stream_bandwidth", [&] {
		if (reps < 1 || bytes < 16) { throw Error(-3, "need reps >= 1 and bytes >= 16"); }
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const long n = (long)(bytes / 16);
		c.work.reserve(2 * (size_t)n * 16);
		zc* a = c.work.as<zc>();
		zc* b = a + n;
		// 256 blocks per CU of four 4 KiB chunks each (measured best of 1-256 per CU)
		const unsigned grid = (unsigned)std::max<long>(1, std::min<long>((n + 4 * kBlock - 1) / (4 * kBlock), 256L * c.cu_count));
		k_fill_zero<<<nblocks(n), kBlock, 0, c.stream>>>(a, n, nullptr);
		hipEvent_t e0, e1;
		RQC_HIP(hipEventCreate(&e0));
		RQC_HIP(hipEventCreate(&e1));
		k_copy_stream<<<grid, kBlock, 0, c.stream>>>(a, b, n);  // warm-up
		RQC_HIP(hipEventRecord(e0, c.stream));
		for (int r = 0; r < reps; r++) { k_copy_stream<<<grid, kBlock, 0, c.stream>>>(a, b, n); }
		RQC_HIP(hipEventRecord(e1, c.stream));
		RQC_HIP(hipEventSynchronize(e1));
		float ms = 0;
		RQC_HIP(hipEventElapsedTime(&ms, e0, e1));
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		*gbps = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;
	});
}

extern "C" int rqc_synchronize(void)
{
	return target_call("rqc_synchronize", [&] { RQC_HIP(hipStreamSynchronize(ctx().stream)); });
}

extern "C" int rqc_device_info(char* buf, int len)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
		(void)hipGetLastError();
		if (len > 0) { snprintf(buf, len, "no HIP device"); }
		return -1;
	}
	return target_call("rqc_device_info", [&] {
		Context& c = ctx();
		snprintf(buf, len, "%s %d %zu lds/wg=%zu", c.name.c_str(), c.cu_count, c.hbm_bytes, c.lds_per_block);
	});
}
