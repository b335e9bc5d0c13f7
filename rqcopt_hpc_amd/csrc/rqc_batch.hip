// Batched brickwall optimisation (include/rqcopt/rqc_batch.h).
//
// Device memory of a batch is one arena:
//   flags | tables | results | workspaces
// flags[k * P + p] = 1 makes every kernel of problem p in the graph of
// evaluation kind k return at once; tables and results of all problems are
// contiguous so that a step uploads and downloads each in one copy.  Each
// problem has one workspace, sized for the largest of its three plans (a
// problem is evaluated in at most one kind per step).
//
// The graph of kind k forks into nbranch parallel branches; problem p's chain
// (zero accumulators, forward passes, trace, reverse passes and reductions,
// finalisation) runs on branch p.branch, problems being spread over the
// branches largest first onto the least loaded branch.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rqc_eval.hpp"
#include "mat4.hpp"
#include "rqcopt/matrix.h"
#include "rqcopt/rqc_batch.h"
#include "rqcopt/trust_region.h"

namespace rqc {

static constexpr int kKinds = 3;

static Mode kind_mode(int k) { return k == RQC_EVAL_TARGET ? Mode::Target : (k == RQC_EVAL_GRADIENT ? Mode::Grad : Mode::Hvp); }

struct BatchProblem
{
	int L = 0, nlayers = 0, per = 0;
	Circuit C;                  // brickwall expansion; gate values are set per step
	const zc* dU = nullptr;
	EvalPlan plan[kKinds];
	size_t budget = 0;          // state-array budget the plans were made with
	TileChoice tc;              // tile height and gates per pass (autotuned or default)
	size_t work_off = 0, work_bytes = 0;
	size_t tab_off = 0, tab_bytes = 0;
	size_t res_off = 0, res_bytes = 0;
	int branch = 0;
	double cost = 0;
};

struct Batch
{
	Context* dev = nullptr;
	bool fp32 = false;
	int exec = RQC_BATCH_GRAPH;
	std::vector<BatchProblem> probs;
	int nbranch = 1;
	std::vector<Slot> branches;
	Slot io;                     // uploads and downloads
	Slot kstream[kKinds];        // streams the kind graphs are launched on
	char* arena = nullptr;
	size_t arena_bytes = 0;
	size_t off_tabs = 0, off_res = 0, off_work = 0, up_bytes = 0, down_bytes = 0;
	char* h_up = nullptr;        // pinned mirror of flags + tables
	char* h_down = nullptr;      // pinned mirror of the results
	hipGraphExec_t graph[kKinds] = {nullptr, nullptr, nullptr};
	hipEvent_t ev_up = nullptr, ev_fork = nullptr;
	hipEvent_t ev_kind[kKinds] = {nullptr, nullptr, nullptr};
	std::vector<hipEvent_t> ev_join;
	rqc_batch_stats stats{};
	std::mutex mu;
	int rtr = RQC_BATCH_RTR_DEVICE;   // trust region of rqc_batch_optimize_brickwall_hvp

	// device-resident trust region: one table region per problem and kind (the position
	// terms uploaded once, the gate-dependent part refilled on the device), the slot-group
	// maps, the fill descriptors and the graphs captured over these tables
	char* dmem = nullptr;
	size_t dmem_bytes = 0;
	std::vector<size_t> dtab_off[kKinds];
	std::vector<size_t> dgroup_off[kKinds];
	size_t off_desc = 0, off_misc = 0;
	hipGraphExec_t dgraph[kKinds] = {nullptr, nullptr, nullptr};
	int* h_count = nullptr;   // pinned: active problems after a chunk of tCG steps

	int* flags_dev(int k, int p) const { return reinterpret_cast<int*>(arena) + (size_t)k * probs.size() + p; }
	char* work(const BatchProblem& b) const { return arena + off_work + b.work_off; }
	char* tab(const BatchProblem& b) const { return arena + off_tabs + b.tab_off; }
	char* res(const BatchProblem& b) const { return arena + off_res + b.res_off; }
};

static void plan_problem(const Batch& bt, BatchProblem& b)
{
	for (int k = 0; k < kKinds; k++) {
		b.plan[k] = plan_eval(*bt.dev, b.C, kind_mode(k), b.per, bt.fp32, true, true, b.budget, 0, &b.tc);
		if (!b.plan[k].tiles) {
			throw Error(-3, "batched problems need the tile-pass engine (even nqubits >= 6, tile-pass engine selected)");
		}
	}
	b.work_bytes = 0;
	b.tab_bytes = 0;
	b.res_bytes = 0;
	for (int k = 0; k < kKinds; k++) {
		b.work_bytes = std::max(b.work_bytes, b.plan[k].off_states + (size_t)b.plan[k].state_bytes());
		b.tab_bytes = std::max(b.tab_bytes, b.plan[k].tab_bytes);
		b.res_bytes = std::max(b.res_bytes, b.plan[k].res_bytes());
	}
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

static void batch_release(Batch& bt)
{
	for (auto& g : bt.graph) { if (g != nullptr) { (void)hipGraphExecDestroy(g); g = nullptr; } }
	for (auto& g : bt.dgraph) { if (g != nullptr) { (void)hipGraphExecDestroy(g); g = nullptr; } }
	if (bt.dmem != nullptr) { (void)hipFree(bt.dmem); bt.dmem = nullptr; }
	if (bt.h_count != nullptr) { (void)hipHostFree(bt.h_count); bt.h_count = nullptr; }
	if (bt.arena != nullptr) { (void)hipFree(bt.arena); }
	if (bt.h_up != nullptr) { (void)hipHostFree(bt.h_up); }
	if (bt.h_down != nullptr) { (void)hipHostFree(bt.h_down); }
	for (auto& s : bt.branches) { if (s.stream != nullptr) { (void)hipStreamDestroy(s.stream); } }
	if (bt.io.stream != nullptr) { (void)hipStreamDestroy(bt.io.stream); }
	for (auto& s : bt.kstream) { if (s.stream != nullptr) { (void)hipStreamDestroy(s.stream); } }
	if (bt.ev_up != nullptr) { (void)hipEventDestroy(bt.ev_up); }
	if (bt.ev_fork != nullptr) { (void)hipEventDestroy(bt.ev_fork); }
	for (auto& e : bt.ev_kind) { if (e != nullptr) { (void)hipEventDestroy(e); } }
	for (auto& e : bt.ev_join) { (void)hipEventDestroy(e); }
}

static void batch_init(Batch& bt, int n, const rqc_batch_problem* pr, int precision, int nstreams, size_t limit,
	const TileChoice& tc)
{
	if (n < 1) { throw Error(-3, "a batch needs at least one problem"); }
	if (precision != RQC_PRECISION_FP64 && precision != RQC_PRECISION_MIXED) { throw Error(-3, "unknown precision"); }
	bt.dev = &ctx();
	bt.fp32 = (precision == RQC_PRECISION_MIXED);
	bt.probs.resize(n);
	for (int p = 0; p < n; p++) {
		BatchProblem& b = bt.probs[p];
		b.L = pr[p].nqubits;
		b.nlayers = pr[p].nlayers;
		if (b.L < 6 || b.L % 2 != 0) { throw Error(-3, "batched problems need an even number of qubits >= 6"); }
		b.per = b.L / 2;
		b.dU = device_target_matrix(const_cast<rqc_target*>(pr[p].target), b.L);
		std::vector<struct mat4x4> z((size_t)b.nlayers);
		for (auto& m : z) { for (auto& e : m.data) { e = numeric(0.0, 0.0); } }
		for (int i = 0; i < b.nlayers; i++) { z[i].data[0] = z[i].data[5] = z[i].data[10] = z[i].data[15] = numeric(1.0, 0.0); }
		b.C = make_brickwall(z.data(), z.data(), b.nlayers, pr[p].perms, b.L);
		b.budget = (size_t)1 << 62;
		b.cost = std::ldexp((double)b.C.ngates, 2 * b.L);
	}
	for (auto& b : bt.probs) { b.tc = tc; }
	// fit the workspaces: halve the state budget of the largest one until all fit
	size_t avail = limit;
	if (avail == 0) {
		size_t fr = 0, tot = 0;
		RQC_HIP(hipMemGetInfo(&fr, &tot));
		avail = (size_t)(0.85 * (double)fr);
	}
	for (auto& b : bt.probs) { plan_problem(bt, b); }
	for (;;) {
		size_t tot = 0;
		BatchProblem* big = nullptr;
		for (auto& b : bt.probs) {
			tot += al256(b.work_bytes) + al256(b.tab_bytes) + al256(b.res_bytes);
			if (big == nullptr || b.work_bytes > big->work_bytes) { big = &b; }
		}
		if (tot + 3 * sizeof(int) * n + 4096 <= avail) { break; }
		const size_t st = (size_t)big->plan[RQC_EVAL_HVP].state_bytes();
		if (big->plan[RQC_EVAL_HVP].B <= 1 || st < 2) { throw Error(-1, "device memory too small for the batch"); }
		big->budget = st / 2;
		plan_problem(bt, *big);
	}
	// layout of the arena
	size_t tabs = 0, ress = 0, works = 0;
	for (auto& b : bt.probs) {
		b.tab_off = tabs;
		tabs += al256(b.tab_bytes);
		b.res_off = ress;
		ress += al256(b.res_bytes);
		b.work_off = works;
		works += al256(b.work_bytes);
	}
	bt.off_tabs = al256(3 * sizeof(int) * (size_t)n);
	bt.off_res = bt.off_tabs + tabs;
	bt.off_work = bt.off_res + ress;
	bt.up_bytes = bt.off_res;
	bt.down_bytes = ress;
	bt.arena_bytes = bt.off_work + works;
	RQC_HIP(hipMalloc(&bt.arena, bt.arena_bytes));
	RQC_HIP(hipHostMalloc(reinterpret_cast<void**>(&bt.h_up), bt.up_bytes, hipHostMallocDefault));
	RQC_HIP(hipHostMalloc(reinterpret_cast<void**>(&bt.h_down), std::max<size_t>(bt.down_bytes, 256), hipHostMallocDefault));
	std::memset(bt.h_up, 0, bt.up_bytes);
	bt.stats.workspace_bytes = bt.arena_bytes;
	// branches: largest problems first onto the least loaded branch
	bt.nbranch = (nstreams > 0) ? std::min(nstreams, n) : std::min(n, 8);
	std::vector<int> order(n);
	for (int p = 0; p < n; p++) { order[p] = p; }
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bt.probs[a].cost > bt.probs[b].cost; });
	std::vector<double> load(bt.nbranch, 0.0);
	for (int p : order) {
		const int br = (int)(std::min_element(load.begin(), load.end()) - load.begin());
		bt.probs[p].branch = br;
		load[br] += bt.probs[p].cost;
	}
	bt.branches.resize(bt.nbranch);
	for (auto& s : bt.branches) { RQC_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)); }
	RQC_HIP(hipStreamCreateWithFlags(&bt.io.stream, hipStreamNonBlocking));
	for (auto& s : bt.kstream) { RQC_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)); }
	RQC_HIP(hipEventCreateWithFlags(&bt.ev_up, hipEventDisableTiming));
	RQC_HIP(hipEventCreateWithFlags(&bt.ev_fork, hipEventDisableTiming));
	for (auto& e : bt.ev_kind) { RQC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
	bt.ev_join.resize(bt.nbranch);
	for (auto& e : bt.ev_join) { RQC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
}

// capture the graph of evaluation kind k: fork into the branches, every
// problem's chain on its branch with its skip flag, join
static void capture_kind(Batch& bt, int k, bool device_tables = false)
{
	hipStream_t cs = bt.kstream[k].stream;
	RQC_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
	try {
		RQC_HIP(hipEventRecord(bt.ev_fork, cs));
		for (auto& s : bt.branches) { RQC_HIP(hipStreamWaitEvent(s.stream, bt.ev_fork, 0)); }
		for (size_t p = 0; p < bt.probs.size(); p++) {
			BatchProblem& b = bt.probs[p];
			char* tab = device_tables ? bt.dmem + bt.dtab_off[k][p] : bt.tab(b);
			enqueue_eval(*bt.dev, bt.branches[b.branch], b.plan[k], b.C, b.dU, bt.work(b), tab, bt.res(b), nullptr,
				bt.flags_dev(k, (int)p));
		}
		for (int br = 0; br < bt.nbranch; br++) {
			RQC_HIP(hipEventRecord(bt.ev_join[br], bt.branches[br].stream));
			RQC_HIP(hipStreamWaitEvent(cs, bt.ev_join[br], 0));
		}
	} catch (...) {
		hipGraph_t g = nullptr;
		(void)hipStreamEndCapture(cs, &g);
		if (g != nullptr) { (void)hipGraphDestroy(g); }
		throw;
	}
	hipGraph_t g = nullptr;
	RQC_HIP(hipStreamEndCapture(cs, &g));
	const hipError_t e = hipGraphInstantiate(device_tables ? &bt.dgraph[k] : &bt.graph[k], g, nullptr, nullptr, 0);
	(void)hipGraphDestroy(g);
	RQC_HIP(e);
}

// gates of problem b for one step: layer gates (and directions) expanded to its circuit
static void set_gates(BatchProblem& b, const struct mat4x4* vlist, const struct mat4x4* vdirs)
{
	for (int i = 0; i < b.nlayers; i++) {
		for (int q = 0; q < b.per; q++) {
			const size_t g = (size_t)i * b.per + q;
			std::memcpy(&b.C.gates[16 * g], vlist[i].data, 16 * sizeof(zc));
			if (vdirs != nullptr) { std::memcpy(&b.C.dirs[16 * g], vdirs[i].data, 16 * sizeof(zc)); }
			else { std::memset(&b.C.dirs[16 * g], 0, 16 * sizeof(zc)); }
		}
	}
}

// one step: every problem p with kinds[p] >= 0 is evaluated; results unpacked into R[p]
static void batch_step(Batch& bt, const int* kinds, const struct mat4x4* const* vlists, const struct mat4x4* const* vdirs,
	std::vector<EvalResult>& R)
{
	const auto t0 = std::chrono::steady_clock::now();
	const int n = (int)bt.probs.size();
	int* flags = reinterpret_cast<int*>(bt.h_up);
	bool present[kKinds] = {false, false, false};
	long nev = 0;
	for (int p = 0; p < n; p++) {
		const int k = kinds[p];
		if (k < RQC_EVAL_NONE || k > RQC_EVAL_HVP) { throw Error(-3, "unknown evaluation kind"); }
		for (int kk = 0; kk < kKinds; kk++) { flags[(size_t)kk * n + p] = (kk == k) ? 0 : 1; }
		if (k < 0) { continue; }
		BatchProblem& b = bt.probs[p];
		set_gates(b, vlists[p], k == RQC_EVAL_HVP ? vdirs[p] : nullptr);
		if (!sweep_unitary(b.C, kind_mode(k), bt.fp32)) {
			throw Error(-3, "batched gradient, HVP and mixed-precision target evaluations need gates unitary to 1e-6 (problem " +
				std::to_string(p) + ")");
		}
		fill_eval_tables(b.plan[k], b.C, bt.h_up + bt.off_tabs + b.tab_off);
		present[k] = true;
		nev++;
	}
	if (nev == 0) { return; }
	RQC_HIP(hipMemcpyAsync(bt.arena, bt.h_up, bt.up_bytes, hipMemcpyHostToDevice, bt.io.stream));
	RQC_HIP(hipEventRecord(bt.ev_up, bt.io.stream));
	if (bt.exec == RQC_BATCH_GRAPH) {
		for (int k = 0; k < kKinds; k++) {
			if (!present[k]) { continue; }
			if (bt.graph[k] == nullptr) { capture_kind(bt, k); }
			RQC_HIP(hipStreamWaitEvent(bt.kstream[k].stream, bt.ev_up, 0));
			RQC_HIP(hipGraphLaunch(bt.graph[k], bt.kstream[k].stream));
			RQC_HIP(hipEventRecord(bt.ev_kind[k], bt.kstream[k].stream));
			RQC_HIP(hipStreamWaitEvent(bt.io.stream, bt.ev_kind[k], 0));
			bt.stats.graph_launches++;
		}
	} else {
		for (auto& s : bt.branches) { RQC_HIP(hipStreamWaitEvent(s.stream, bt.ev_up, 0)); }
		for (int p = 0; p < n; p++) {
			if (kinds[p] < 0) { continue; }
			BatchProblem& b = bt.probs[p];
			enqueue_eval(*bt.dev, bt.branches[b.branch], b.plan[kinds[p]], b.C, b.dU, bt.work(b), bt.tab(b), bt.res(b),
				nullptr, nullptr);
		}
		for (int br = 0; br < bt.nbranch; br++) {
			RQC_HIP(hipEventRecord(bt.ev_join[br], bt.branches[br].stream));
			RQC_HIP(hipStreamWaitEvent(bt.io.stream, bt.ev_join[br], 0));
		}
	}
	RQC_HIP(hipMemcpyAsync(bt.h_down, bt.arena + bt.off_res, bt.down_bytes, hipMemcpyDeviceToHost, bt.io.stream));
	RQC_HIP(hipStreamSynchronize(bt.io.stream));
	RQC_HIP(hipGetLastError());
	for (int p = 0; p < n; p++) {
		if (kinds[p] < 0) { continue; }
		const BatchProblem& b = bt.probs[p];
		unpack_result(b.plan[kinds[p]], bt.h_down + b.res_off, R[p]);
	}
	bt.stats.steps++;
	bt.stats.evaluations += nev;
	bt.stats.device_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------- lock-step trust region
namespace {

struct Lockstep;

struct SolveCtx
{
	Lockstep* ls;
	int p;
};

struct Request
{
	bool pending = false;
	int kind = RQC_EVAL_NONE;
	const double* x = nullptr;
	std::vector<struct mat4x4> dirs;
	double f = 0;
	double* grad = nullptr;
	double* hvp = nullptr;
};

struct Lockstep
{
	Batch* bt = nullptr;
	std::mutex mu;
	std::condition_variable cv;
	int active = 0;
	int posted = 0;
	bool failed = false;
	std::vector<Request> req;

	// called from problem p's solver thread: wait until the step that includes it has run
	double post(int p, int kind, const double* x, const double* vec, double* grad, double* hvp)
	{
		std::unique_lock<std::mutex> lk(mu);
		Request& r = req[p];
		r.kind = kind;
		r.x = x;
		r.grad = grad;
		r.hvp = hvp;
		if (kind == RQC_EVAL_HVP) {
			const int nl = bt->probs[p].nlayers;
			const auto* v = reinterpret_cast<const struct mat4x4*>(x);
			r.dirs.resize(nl);
			for (int i = 0; i < nl; i++) { real_to_tangent(&vec[16 * i], &v[i], &r.dirs[i]); }
		}
		r.pending = true;
		posted++;
		cv.notify_all();
		cv.wait(lk, [&] { return !r.pending; });
		return r.f;
	}

	void finish()
	{
		std::lock_guard<std::mutex> lk(mu);
		active--;
		cv.notify_all();
	}

	// the calling thread runs the steps until every solver has finished
	void serve()
	{
		const int n = (int)req.size();
		std::vector<int> kinds(n);
		std::vector<const struct mat4x4*> vl(n), vd(n);
		std::vector<EvalResult> R(n);
		std::unique_lock<std::mutex> lk(mu);
		for (;;) {
			cv.wait(lk, [&] { return active == 0 || (posted > 0 && posted == active); });
			if (active == 0) { break; }
			for (int p = 0; p < n; p++) {
				kinds[p] = req[p].pending ? req[p].kind : RQC_EVAL_NONE;
				vl[p] = reinterpret_cast<const struct mat4x4*>(req[p].x);
				vd[p] = req[p].dirs.empty() ? nullptr : req[p].dirs.data();
			}
			lk.unlock();
			bool ok = true;
			try {
				batch_step(*bt, kinds.data(), vl.data(), vd.data(), R);
			} catch (const std::exception& e) {
				fprintf(stderr, "librqcopt_hip: batched step failed: %s\n", e.what());
				ok = false;
			}
			lk.lock();
			if (!ok) { failed = true; }
			for (int p = 0; p < n; p++) {
				Request& r = req[p];
				if (!r.pending) { continue; }
				const double nan = std::nan("");
				r.f = ok ? R[p].f.x : nan;
				const size_t m = 16 * (size_t)bt->probs[p].nlayers;
				if (r.grad != nullptr) { for (size_t i = 0; i < m; i++) { r.grad[i] = ok ? R[p].grad[i] : nan; } }
				if (r.hvp != nullptr) { for (size_t i = 0; i < m; i++) { r.hvp[i] = ok ? R[p].hvp[i] : nan; } }
				r.pending = false;
			}
			posted = 0;
			cv.notify_all();
		}
	}
};

double ls_f(const double* x, void* d)
{
	auto* s = static_cast<SolveCtx*>(d);
	return s->ls->post(s->p, RQC_EVAL_TARGET, x, nullptr, nullptr, nullptr);
}

double ls_f_deriv(const double* RQC_RESTRICT x, void* d, double* RQC_RESTRICT grad)
{
	auto* s = static_cast<SolveCtx*>(d);
	return s->ls->post(s->p, RQC_EVAL_GRADIENT, x, nullptr, grad, nullptr);
}

void ls_f_hvp(const double* RQC_RESTRICT x, void* d, double* RQC_RESTRICT vec, double* RQC_RESTRICT hvp)
{
	auto* s = static_cast<SolveCtx*>(d);
	s->ls->post(s->p, RQC_EVAL_HVP, x, vec, nullptr, hvp);
}

void ls_retract(const double* RQC_RESTRICT x, const double* RQC_RESTRICT eta, void* rdata, double* RQC_RESTRICT xs)
{
	const int nlayers = *static_cast<int*>(rdata);
	const auto* v = reinterpret_cast<const struct mat4x4*>(x);
	auto* w = reinterpret_cast<struct mat4x4*>(xs);
	for (int i = 0; i < nlayers; i++) { retract(&v[i], &eta[16 * i], &w[i]); }
}

}  // namespace

// ---------------------------------------------------------------- device-resident trust region
// The lock-step trust region with its driver on the GPU.  Each problem's state
// (gates, gradient, tCG vectors, radius) lives in device memory and is updated
// by one thread of its own workgroup, in the operation order of the host
// driver (rqc_host.cpp: truncated_cg and trust_region, the reference's
// src/trust_region.c:159-217,306-368), with contraction into fused
// multiply-adds switched off so that the scalar recurrences round as the host's
// do.  Between these kernels the captured evaluation graphs run on tables that
// the device refills from the state (enqueue_fill_tables); a problem that has
// no evaluation pending sets its skip flag, so the graphs stay the same.
struct DevRtr
{
	int n = 0, nl = 0;               // tangent dimension 16 nl; layers
	int active = 0, on_boundary = 0, tcg_iter = 0, pad = 0;
	long evals = 0;                  // evaluations armed for this problem
	double radius = 0, rsq = 0, stoptol = 0, fx = 0, denom = 0;
	double* x = nullptr;             // 32 nl: the layer gates (complex, row-major)
	double* xn = nullptr;            // the candidate
	double* grad = nullptr;          // n
	double* z = nullptr;             // n: the tCG iterate (eta)
	double* r = nullptr;             // n
	double* d = nullptr;             // n
	zc* cur_gates = nullptr;         // 16 nl: gates of the next evaluation
	zc* cur_dirs = nullptr;          // 16 nl: its directions
	const char* res = nullptr;       // the problem's result region
	double* f_iter = nullptr;        // niter + 1
	int* flag[kKinds] = {nullptr, nullptr, nullptr};
};

__device__ __forceinline__ const double* res_grad(const DevRtr& s)
{
	return reinterpret_cast<const double*>(reinterpret_cast<const zc*>(s.res) + 1 + 32 * (size_t)s.nl);
}
__device__ __forceinline__ const double* res_hvp(const DevRtr& s) { return res_grad(s) + 16 * (size_t)s.nl; }
__device__ __forceinline__ double res_f(const DevRtr& s) { return reinterpret_cast<const zc*>(s.res)->x; }

__device__ double dot_seq(int n, const double* a, const double* b)
{
#pragma clang fp contract(off)
	double t = 0;
	for (int i = 0; i < n; i++) { t += a[i] * b[i]; }
	return t;
}

// tangent directions Z_i = V_i A(w_i) of the layer gates V = x for a tangent vector w
__device__ void tangent_dirs(const DevRtr& s, const double* x, const double* w)
{
	for (int i = 0; i < s.nl; i++) {
		m4_real_to_tangent(w + 16 * i, reinterpret_cast<const zc*>(x) + 16 * i, s.cur_dirs + 16 * i);
	}
}

__device__ void set_gates(const DevRtr& s, const double* x)
{
	const zc* v = reinterpret_cast<const zc*>(x);
	for (int e = 0; e < 16 * s.nl; e++) { s.cur_gates[e] = v[e]; }
}

__device__ __forceinline__ void arm(DevRtr& s, int kind, bool on)
{
	*s.flag[kind] = on ? 0 : 1;
	if (on) { s.evals++; }
}

// iteration start: gradient evaluation at x
__global__ void k_rtr_begin(DevRtr* st)
{
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	set_gates(s, s.x);
	arm(s, RQC_EVAL_GRADIENT, true);
}

// f(x), the gradient, and the tCG start: z = 0, r = grad, d = -r (first HVP along d)
__global__ void k_rtr_tcg_init(DevRtr* st, int maxiter, double abstol, double reltol)
{
#pragma clang fp contract(off)
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	s.fx = res_f(s);
	const double* g = res_grad(s);
	for (int i = 0; i < s.n; i++) { s.grad[i] = g[i]; s.r[i] = g[i]; }
	s.rsq = dot_seq(s.n, s.r, s.r);
	s.stoptol = fmax(abstol, reltol * sqrt(s.rsq));
	for (int i = 0; i < s.n; i++) { s.z[i] = 0; s.d[i] = -s.r[i]; }
	s.tcg_iter = 0;
	s.on_boundary = 0;
	s.active = maxiter > 0;
	if (s.active) { tangent_dirs(s, s.x, s.d); }
	arm(s, RQC_EVAL_HVP, s.active);
}

// t > 0 with ||b + t d|| = radius (host move_to_boundary / solve_quadratic)
__device__ double to_boundary(int n, const double* b, const double* d, double radius)
{
#pragma clang fp contract(off)
	const double dsq = dot_seq(n, d, d);
	if (dsq == 0) { return 0; }
	const double p = dot_seq(n, b, d) / dsq;
	const double q = (dot_seq(n, b, b) - radius * radius) / dsq;
	if (p == 0) { return sqrt(-q); }
	const double a = -(p + copysign(1.0, p) * sqrt(p * p - q));
	const double c = q / a;
	return (a <= c) ? c : a;
}

// one tCG step with the HVP of d just evaluated
__global__ void k_rtr_tcg_step(DevRtr* st, int maxiter)
{
#pragma clang fp contract(off)
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	if (!s.active) { return; }
	const double* hd = res_hvp(s);
	const int n = s.n;
	const double dhd = dot_seq(n, s.d, hd);
	const double t = to_boundary(n, s.z, s.d, s.radius);
	const double alpha = s.rsq / dhd;
	if (dhd <= 0 || alpha > t) {
		for (int i = 0; i < n; i++) { s.z[i] += t * s.d[i]; }
		s.on_boundary = 1;
		s.active = 0;
	} else {
		for (int i = 0; i < n; i++) { s.r[i] += alpha * hd[i]; }
		for (int i = 0; i < n; i++) { s.z[i] += alpha * s.d[i]; }
		const double rsq_next = dot_seq(n, s.r, s.r);
		if (sqrt(rsq_next) <= s.stoptol) {
			s.active = 0;
		} else {
			const double beta = rsq_next / s.rsq;
			for (int i = 0; i < n; i++) { s.d[i] = beta * s.d[i] - s.r[i]; }
			s.rsq = rsq_next;
		}
	}
	s.tcg_iter++;
	if (s.active && s.tcg_iter >= maxiter) { s.active = 0; }
	if (s.active) { tangent_dirs(s, s.x, s.d); }
	arm(s, RQC_EVAL_HVP, s.active);
}

// problems still in their tCG loop (one thread, fixed order)
__global__ void k_rtr_count_active(const DevRtr* st, int n, int* count)
{
	if (threadIdx.x != 0 || blockIdx.x != 0) { return; }
	int c = 0;
	for (int p = 0; p < n; p++) { c += st[p].active; }
	*count = c;
}

// candidate x_n = retract(x, eta) per layer; HVP along eta for the model decrease
__global__ void k_rtr_candidate(DevRtr* st)
{
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	const zc* v = reinterpret_cast<const zc*>(s.x);
	zc* w = reinterpret_cast<zc*>(s.xn);
	for (int i = 0; i < s.nl; i++) { m4_retract(v + 16 * i, s.z + 16 * i, w + 16 * i); }
	tangent_dirs(s, s.x, s.z);
	arm(s, RQC_EVAL_HVP, true);
}

// model decrease <grad, eta> + <eta, H eta> / 2; then f at the candidate
__global__ void k_rtr_model(DevRtr* st)
{
#pragma clang fp contract(off)
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	s.denom = dot_seq(s.n, s.grad, s.z) + 0.5 * dot_seq(s.n, s.z, res_hvp(s));
	set_gates(s, s.xn);
	arm(s, RQC_EVAL_TARGET, true);
}

// rho, the radius update and acceptance (reference src/trust_region.c:337-361)
__global__ void k_rtr_accept(DevRtr* st, int k, double rho_trust, double maxradius)
{
#pragma clang fp contract(off)
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	const double rho = (res_f(s) - s.fx) / s.denom;
	if (rho < 0.25) {
		s.radius *= 0.25;
	} else if (rho > 0.75 && s.on_boundary) {
		s.radius = fmin(2 * s.radius, maxradius);
	}
	s.f_iter[k] = s.fx;
	if (rho > rho_trust) {
		for (int e = 0; e < 32 * s.nl; e++) { s.x[e] = s.xn[e]; }
	}
}

__global__ void k_rtr_final_begin(DevRtr* st)
{
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	set_gates(s, s.x);
	arm(s, RQC_EVAL_TARGET, true);
}

__global__ void k_rtr_final(DevRtr* st, int niter)
{
	if (threadIdx.x != 0) { return; }
	DevRtr& s = st[blockIdx.x];
	s.f_iter[niter] = res_f(s);
}

// device tables of every problem and kind (position terms from one host fill at the start
// gates), the slot-group maps and the graphs over these tables; kept for the batch's lifetime
static void ensure_device_tables(Batch& bt)
{
	if (bt.dmem != nullptr) { return; }
	const int n = (int)bt.probs.size();
	size_t off = 0;
	for (int k = 0; k < kKinds; k++) {
		bt.dtab_off[k].resize(n);
		bt.dgroup_off[k].resize(n);
		for (int p = 0; p < n; p++) {
			bt.dtab_off[k][p] = off;
			off += al256(bt.probs[p].plan[k].tab_bytes);
			bt.dgroup_off[k][p] = off;
			off += al256(sizeof(int) * (size_t)bt.probs[p].C.ngates);
		}
	}
	bt.off_desc = off;
	off += al256(sizeof(DeviceTableDesc) * kKinds * (size_t)n);
	bt.off_misc = off;
	off += 256;   // active count, singular-gate flag
	bt.dmem_bytes = off;
	RQC_HIP(hipMalloc(&bt.dmem, bt.dmem_bytes));
	RQC_HIP(hipHostMalloc(reinterpret_cast<void**>(&bt.h_count), 256, hipHostMallocDefault));
	std::vector<char> img(bt.off_desc, 0);
	for (int k = 0; k < kKinds; k++) {
		for (int p = 0; p < n; p++) {
			const auto g = table_slot_groups(bt.probs[p].plan[k]);
			std::memcpy(img.data() + bt.dgroup_off[k][p], g.data(), g.size() * sizeof(int));
		}
	}
	RQC_HIP(hipMemcpy(bt.dmem, img.data(), img.size(), hipMemcpyHostToDevice));
	for (int k = 0; k < kKinds; k++) { capture_kind(bt, k, true); }
}

static void batch_optimize_device(Batch& bt, const struct mat4x4* const* vstart, const struct rtr_params* prm, int niter,
	double* const* f_iters, struct mat4x4* const* vopt)
{
	if (!(0 <= prm->rho_trust && prm->rho_trust < 0.25)) { throw Error(-3, "rho_trust must lie in [0, 0.25)"); }
	const int n = (int)bt.probs.size();
	ensure_device_tables(bt);
	hipStream_t s = bt.io.stream;
	// per-call state: gates, candidate, gradient, tCG vectors, current gates and directions, f_iter
	std::vector<size_t> base(n);
	size_t tot = 0;
	for (int p = 0; p < n; p++) {
		const size_t nl = (size_t)bt.probs[p].nlayers;
		base[p] = tot;
		tot += al256(sizeof(double) * (32 * nl * 2 + 16 * nl * 4 + 32 * nl * 2 + (size_t)niter + 1));
	}
	const size_t st_off = tot;
	tot += al256(sizeof(DevRtr) * (size_t)n);
	struct Pool : DeviceBuffer { ~Pool() { release(); } } pool;
	pool.reserve(tot);
	char* pm = pool.as<char>();
	std::vector<DevRtr> hs(n);
	std::vector<char> himg(st_off, 0);
	for (int p = 0; p < n; p++) {
		const BatchProblem& b = bt.probs[p];
		const int nl = b.nlayers;
		double* q = reinterpret_cast<double*>(pm + base[p]);
		DevRtr& r = hs[p];
		r.n = 16 * nl;
		r.nl = nl;
		r.radius = prm->radius_init;
		r.x = q; q += 32 * nl;
		r.xn = q; q += 32 * nl;
		r.grad = q; q += 16 * nl;
		r.z = q; q += 16 * nl;
		r.r = q; q += 16 * nl;
		r.d = q; q += 16 * nl;
		r.cur_gates = reinterpret_cast<zc*>(q); q += 32 * nl;
		r.cur_dirs = reinterpret_cast<zc*>(q); q += 32 * nl;
		r.f_iter = q;
		r.res = bt.res(b);
		for (int k = 0; k < kKinds; k++) { r.flag[k] = bt.flags_dev(k, p); }
		std::memcpy(himg.data() + base[p], vstart[p], sizeof(double) * 32 * nl);
	}
	DevRtr* dst = reinterpret_cast<DevRtr*>(pm + st_off);
	RQC_HIP(hipMemcpyAsync(pm, himg.data(), himg.size(), hipMemcpyHostToDevice, s));
	RQC_HIP(hipMemcpyAsync(dst, hs.data(), sizeof(DevRtr) * n, hipMemcpyHostToDevice, s));
	// host fill of every table at the start gates (position terms; the rest is refilled on the device),
	// the fill descriptors and zeroed flags
	std::vector<DeviceTableDesc> desc((size_t)kKinds * n);
	int* d_count = reinterpret_cast<int*>(bt.dmem + bt.off_misc);
	int* d_bad = d_count + 1;
	std::vector<std::vector<char>> timg((size_t)kKinds * n);
	for (int k = 0; k < kKinds; k++) {
		for (int p = 0; p < n; p++) {
			BatchProblem& b = bt.probs[p];
			const EvalPlan& pl = b.plan[k];
			set_gates(b, vstart[p], nullptr);
			if (k == 0 && !gates_near_unitary(b.C)) {
				throw Error(-3, "the batched trust region needs start gates unitary to 1e-6 (problem " +
					std::to_string(p) + ")");
			}
			auto& img = timg[(size_t)k * n + p];
			img.assign(pl.tab_bytes, 0);
			fill_eval_tables(pl, b.C, img.data());
			RQC_HIP(hipMemcpyAsync(bt.dmem + bt.dtab_off[k][p], img.data(), img.size(), hipMemcpyHostToDevice, s));
			DeviceTableDesc& d = desc[(size_t)k * n + p];
			d.tab = bt.dmem + bt.dtab_off[k][p];
			d.slot_group = reinterpret_cast<const int*>(bt.dmem + bt.dgroup_off[k][p]);
			d.gates = hs[p].cur_gates;
			d.dirs = hs[p].cur_dirs;
			d.skip = bt.flags_dev(k, p);
			d.bad = d_bad;
			d.mats_bytes = pl.mats_bytes;
			d.terms_bytes = pl.terms_bytes;
			d.rmats_off = pl.rmats_off;
			d.G = b.C.ngates;
			d.ng = pl.ng;
			d.mode = (int)pl.mode;
			d.fp32 = pl.fp32 ? 1 : 0;
		}
	}
	const DeviceTableDesc* d_desc = reinterpret_cast<const DeviceTableDesc*>(bt.dmem + bt.off_desc);
	RQC_HIP(hipMemcpyAsync(const_cast<DeviceTableDesc*>(d_desc), desc.data(), sizeof(DeviceTableDesc) * desc.size(),
		hipMemcpyHostToDevice, s));
	RQC_HIP(hipMemsetAsync(bt.arena, 0xff, 3 * sizeof(int) * (size_t)n, s));   // every kind skipped
	RQC_HIP(hipMemsetAsync(d_count, 0, 2 * sizeof(int), s));
	auto graph = [&](int k) {
		RQC_HIP(hipGraphLaunch(bt.dgraph[k], s));
		bt.stats.graph_launches++;
		bt.stats.steps++;
	};
	auto fill = [&](int k) { enqueue_fill_tables(s, d_desc + (size_t)k * n, n); };
	const auto t0 = std::chrono::steady_clock::now();
	const int maxiter = prm->tcg_params.maxiter;
	constexpr int kChunk = 8;   // tCG steps between reads of the active count
	for (int it = 0; it < niter; it++) {
		k_rtr_begin<<<n, 64, 0, s>>>(dst);
		fill(RQC_EVAL_GRADIENT);
		graph(RQC_EVAL_GRADIENT);
		k_rtr_tcg_init<<<n, 64, 0, s>>>(dst, maxiter, prm->tcg_params.abstol, prm->tcg_params.reltol);
		fill(RQC_EVAL_HVP);
		for (int j = 0; j < maxiter;) {
			const int c = std::min(kChunk, maxiter - j);
			for (int q = 0; q < c; q++) {
				graph(RQC_EVAL_HVP);
				k_rtr_tcg_step<<<n, 64, 0, s>>>(dst, maxiter);
				fill(RQC_EVAL_HVP);
			}
			j += c;
			k_rtr_count_active<<<1, 64, 0, s>>>(dst, n, d_count);
			RQC_HIP(hipMemcpyAsync(bt.h_count, d_count, sizeof(int), hipMemcpyDeviceToHost, s));
			RQC_HIP(hipStreamSynchronize(s));
			if (bt.h_count[0] == 0) { break; }
		}
		k_rtr_candidate<<<n, 64, 0, s>>>(dst);
		fill(RQC_EVAL_HVP);
		graph(RQC_EVAL_HVP);
		k_rtr_model<<<n, 64, 0, s>>>(dst);
		fill(RQC_EVAL_TARGET);
		graph(RQC_EVAL_TARGET);
		k_rtr_accept<<<n, 64, 0, s>>>(dst, it, prm->rho_trust, prm->maxradius);
	}
	k_rtr_final_begin<<<n, 64, 0, s>>>(dst);
	fill(RQC_EVAL_TARGET);
	graph(RQC_EVAL_TARGET);
	k_rtr_final<<<n, 64, 0, s>>>(dst, niter);
	RQC_HIP(hipGetLastError());
	std::vector<char> out(st_off);
	RQC_HIP(hipMemcpyAsync(out.data(), pm, st_off, hipMemcpyDeviceToHost, s));
	RQC_HIP(hipMemcpyAsync(hs.data(), dst, sizeof(DevRtr) * n, hipMemcpyDeviceToHost, s));
	RQC_HIP(hipMemcpyAsync(bt.h_count, d_count, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
	RQC_HIP(hipStreamSynchronize(s));
	if (bt.h_count[1] != 0) { throw Error(-4, "singular gate in a batched trust-region step"); }
	for (int p = 0; p < n; p++) {
		const int nl = bt.probs[p].nlayers;
		const double* q = reinterpret_cast<const double*>(out.data() + base[p]);
		std::memcpy(vopt[p], q, sizeof(double) * 32 * nl);
		std::memcpy(f_iters[p], q + (size_t)32 * nl * 2 + 16 * nl * 4 + 32 * nl * 2, sizeof(double) * (niter + 1));
		bt.stats.evaluations += hs[p].evals;
	}
	bt.stats.device_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace rqc

using namespace rqc;

struct rqc_batch
{
	Batch b;
};

// Tile autotuning of a batch: each candidate (tile height and gates per pass,
// the same for every problem) is timed on whole HVP steps of the batch itself,
// one warm-up and the best of two, since the problems' kernels overlap and a
// choice that is faster for a problem alone can be slower beside the others.
// Results per batch structure (problem shapes, precision, branches) are cached.
static TileChoice tune_batch(int n, const rqc_batch_problem* pr, int precision, int nstreams, size_t limit)
{
	static std::mutex mu;
	static std::map<std::string, TileChoice> cache;
	std::string key = std::to_string(precision) + "/" + std::to_string(nstreams) + ":";
	for (int p = 0; p < n; p++) {
		key += std::to_string(pr[p].nqubits) + "x" + std::to_string(pr[p].nlayers) + "[";
		for (int i = 0; i < pr[p].nlayers; i++) {
			for (int q = 0; q < pr[p].nqubits; q++) { key += std::to_string(pr[p].perms[i][q]) + ","; }
		}
		key += "]";
	}
	std::lock_guard<std::mutex> lk(mu);
	auto it = cache.find(key);
	if (it != cache.end()) { return it->second; }
	// the default tile size (2^10 FP64, 2^11 FP32), half of it, and fewer gates per pass
	const int base = (precision == RQC_PRECISION_MIXED) ? 11 : 10;
	const TileChoice cands[] = {{base, 8}, {base - 1, 8}, {base, 6}, {base, 4}};
	TileChoice best{0, 0};
	double best_ms = 0;
	std::vector<int> kinds(n, RQC_EVAL_HVP);
	for (const TileChoice& c : cands) {
		Batch tmp;
		try {
			batch_init(tmp, n, pr, precision, nstreams, limit, c);
			std::vector<const struct mat4x4*> vl(n), vd(n);
			std::vector<std::vector<struct mat4x4>> gates(n);
			for (int p = 0; p < n; p++) {
				gates[p].resize(pr[p].nlayers);
				for (auto& m : gates[p]) {
					for (int e = 0; e < 16; e++) { m.data[e] = numeric((e % 5 == 0) ? 1.0 : 0.0, 0.0); }
				}
				vl[p] = gates[p].data();
				vd[p] = gates[p].data();
			}
			std::vector<EvalResult> R(n);
			batch_step(tmp, kinds.data(), vl.data(), vd.data(), R);
			double ms = 0;
			for (int r = 0; r < 2; r++) {
				const double before = tmp.stats.device_ms;
				batch_step(tmp, kinds.data(), vl.data(), vd.data(), R);
				const double t = tmp.stats.device_ms - before;
				ms = (r == 0) ? t : std::min(ms, t);
			}
			if (best.logE == 0 || ms < best_ms) { best = c; best_ms = ms; }
		} catch (const Error&) {
			// a candidate that does not fit (memory, tile shape) is skipped
		}
		if (tmp.io.stream != nullptr) { (void)hipStreamSynchronize(tmp.io.stream); }
		batch_release(tmp);
	}
	cache[key] = best;
	return best;
}

extern "C" int rqc_batch_create(int nproblems, const struct rqc_batch_problem problems[], int precision, int nstreams,
	size_t workspace_limit, struct rqc_batch** batch)
{
	rqc_batch* h = nullptr;
	try {
		h = new rqc_batch();
		TileChoice tc{0, 0};
		if (ctx().autotune && nproblems > 0) { tc = tune_batch(nproblems, problems, precision, nstreams, workspace_limit); }
		batch_init(h->b, nproblems, problems, precision, nstreams, workspace_limit, tc);
		*batch = h;
		return 0;
	} catch (const std::exception& e) {
		if (h != nullptr) { batch_release(h->b); delete h; }
		*batch = nullptr;
		return report_error("rqc_batch_create", e);
	}
}

extern "C" void rqc_batch_destroy(struct rqc_batch* batch)
{
	if (batch == nullptr) { return; }
	if (batch->b.io.stream != nullptr) { (void)hipStreamSynchronize(batch->b.io.stream); }
	batch_release(batch->b);
	delete batch;
}

extern "C" int rqc_batch_set_exec(struct rqc_batch* batch, int exec)
{
	if (batch == nullptr || (exec != RQC_BATCH_GRAPH && exec != RQC_BATCH_EAGER)) { return -3; }
	std::lock_guard<std::mutex> lk(batch->b.mu);
	batch->b.exec = exec;
	return 0;
}

extern "C" int rqc_batch_evaluate(struct rqc_batch* batch, const int kinds[], const struct mat4x4* const vlists[],
	const struct mat4x4* const vdirs[], numeric fvals[], double* const grads[], double* const hvps[])
{
	try {
		if (batch == nullptr) { throw Error(-3, "batch is NULL"); }
		Batch& bt = batch->b;
		std::lock_guard<std::mutex> lk(bt.mu);
		const int n = (int)bt.probs.size();
		for (int p = 0; p < n; p++) {
			if (kinds[p] == RQC_EVAL_HVP && (vdirs == nullptr || vdirs[p] == nullptr)) {
				throw Error(-3, "HVP evaluation of problem " + std::to_string(p) + " without directions");
			}
		}
		std::vector<EvalResult> R(n);
		batch_step(bt, kinds, vlists, vdirs, R);
		for (int p = 0; p < n; p++) {
			if (kinds[p] < 0) { continue; }
			if (fvals != nullptr) { std::memcpy(&fvals[p], &R[p].f, sizeof(zc)); }
			if (kinds[p] >= RQC_EVAL_GRADIENT && grads != nullptr && grads[p] != nullptr) {
				std::memcpy(grads[p], R[p].grad.data(), R[p].grad.size() * sizeof(double));
			}
			if (kinds[p] == RQC_EVAL_HVP && hvps != nullptr && hvps[p] != nullptr) {
				std::memcpy(hvps[p], R[p].hvp.data(), R[p].hvp.size() * sizeof(double));
			}
		}
		return 0;
	} catch (const std::exception& e) { return report_error("rqc_batch_evaluate", e); }
}

extern "C" int rqc_batch_optimize_brickwall_hvp(struct rqc_batch* batch, const struct mat4x4* const vlist_start[],
	const struct rtr_params* params, int niter, double* const f_iters[], struct mat4x4* const vlist_opt[])
{
	try {
		if (batch == nullptr) { throw Error(-3, "batch is NULL"); }
		if (niter < 0) { throw Error(-3, "niter must be >= 0"); }
		Batch& bt = batch->b;
		std::lock_guard<std::mutex> lk(bt.mu);
		if (bt.rtr == RQC_BATCH_RTR_DEVICE) {
			batch_optimize_device(bt, vlist_start, params, niter, f_iters, vlist_opt);
			return 0;
		}
		const int n = (int)bt.probs.size();
		Lockstep ls;
		ls.bt = &bt;
		ls.req.resize(n);
		ls.active = n;
		std::vector<SolveCtx> sc(n);
		std::vector<int> nl(n);
		std::vector<std::thread> th;
		th.reserve(n);
		for (int p = 0; p < n; p++) {
			sc[p] = SolveCtx{&ls, p};
			nl[p] = bt.probs[p].nlayers;
			th.emplace_back([&, p] {
				struct rtr_params prm = *params;
				prm.g_func = nullptr;
				prm.g_data = nullptr;
				prm.g_iter = nullptr;
				riemannian_trust_region_optimize_hvp(ls_f, ls_f_deriv, ls_f_hvp, &sc[p], ls_retract, &nl[p],
					nl[p] * num_tangent_params, reinterpret_cast<const double*>(vlist_start[p]), nl[p] * 32, &prm, niter,
					f_iters[p], reinterpret_cast<double*>(vlist_opt[p]));
				ls.finish();
			});
		}
		ls.serve();
		for (auto& t : th) { t.join(); }
		if (ls.failed) { throw Error(-1, "a batched evaluation step failed"); }
		return 0;
	} catch (const std::exception& e) { return report_error("rqc_batch_optimize_brickwall_hvp", e); }
}

extern "C" int rqc_batch_set_rtr(struct rqc_batch* batch, int mode)
{
	if (batch == nullptr || (mode != RQC_BATCH_RTR_DEVICE && mode != RQC_BATCH_RTR_HOST)) { return -3; }
	std::lock_guard<std::mutex> lk(batch->b.mu);
	batch->b.rtr = mode;
	return 0;
}

extern "C" int rqc_batch_get_stats(const struct rqc_batch* batch, struct rqc_batch_stats* stats)
{
	if (batch == nullptr || stats == nullptr) { return -3; }
	*stats = batch->b.stats;
	return 0;
}

extern "C" int rqc_batch_problem_plan(const struct rqc_batch* batch, int p, int kind, long* block, int* engine, int* passes)
{
	if (batch == nullptr || p < 0 || p >= (int)batch->b.probs.size() || kind < 0 || kind >= kKinds) { return -3; }
	const EvalPlan& pl = batch->b.probs[p].plan[kind];
	if (block != nullptr) { *block = pl.B; }
	if (engine != nullptr) { *engine = pl.tiles ? RQC_ENGINE_TILE_PASS : RQC_ENGINE_PER_GATE; }
	if (passes != nullptr) { *passes = pl.passes; }
	return 0;
}

extern "C" int rqc_batch_problem_tiles(const struct rqc_batch* batch, int p, int kind, int* tile_rows_log2, int* tile_cols_log2,
	int* gates_per_pass)
{
	if (batch == nullptr || p < 0 || p >= (int)batch->b.probs.size() || kind < 0 || kind >= kKinds) { return -3; }
	const EvalPlan& pl = batch->b.probs[p].plan[kind];
	if (tile_rows_log2 != nullptr) { *tile_rows_log2 = pl.tile_K; }
	if (tile_cols_log2 != nullptr) { *tile_cols_log2 = pl.tile_logC; }
	if (gates_per_pass != nullptr) { *gates_per_pass = pl.gmax; }
	return 0;
}
