// Request server (include/rqcopt/rqc_server.h).
//
// One device arena holds the workspaces of the requests in flight; a host-side
// first-fit allocator with coalescing hands out 256-byte aligned regions.
// nslots worker threads each own a HIP stream and pinned staging buffers.  A
// free worker takes the oldest queued request once its workspace fits into
// the arena (strict arrival order: a large request at the head is not
// overtaken), uploads its gate tables, enqueues the evaluation, copies the
// results back and releases the region.  Plans are made per session, kind and
// engine on first use with the per-request state budget, so the column block
// and the grids of a request never depend on the load.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rqc_eval.hpp"
#include "rqcopt/matrix.h"
#include "rqcopt/rqc_server.h"
#include "rqcopt/trust_region.h"

namespace rqc {

using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b)
{
	return std::chrono::duration<double, std::milli>(b - a).count();
}

// first-fit allocator over [0, size) with coalescing of freed neighbours
struct ArenaAlloc
{
	size_t size = 0, used = 0, peak = 0;
	std::map<size_t, size_t> free_;  // offset -> length

	void init(size_t n) { size = n; free_.clear(); free_[0] = n; used = peak = 0; }

	bool alloc(size_t n, size_t* off)
	{
		n = (n + 255) & ~(size_t)255;
		for (auto it = free_.begin(); it != free_.end(); ++it) {
			if (it->second < n) { continue; }
			*off = it->first;
			const size_t rest = it->second - n;
			const size_t at = it->first + n;
			free_.erase(it);
			if (rest > 0) { free_[at] = rest; }
			used += n;
			peak = std::max(peak, used);
			return true;
		}
		return false;
	}

	void release(size_t off, size_t n)
	{
		n = (n + 255) & ~(size_t)255;
		used -= n;
		auto it = free_.emplace(off, n).first;
		auto nx = std::next(it);
		if (nx != free_.end() && it->first + it->second == nx->first) {
			it->second += nx->second;
			free_.erase(nx);
		}
		if (it != free_.begin()) {
			auto pv = std::prev(it);
			if (pv->first + pv->second == it->first) {
				pv->second += it->second;
				free_.erase(it);
			}
		}
	}
};

struct Server;

struct Session
{
	Server* srv = nullptr;
	Circuit C;                  // structure and the gates of the last request
	int per = 1, ng = 0;
	bool brickwall = false;
	const zc* dU = nullptr;
	bool fp32 = false;
	int pending = 0;            // requests queued or running
	bool has_gates = false;
	std::mutex mu;              // plans
	std::shared_ptr<const EvalPlan> plan[3][2];   // [kind][unitary]
};

struct Req
{
	long id = 0;
	Session* s = nullptr;
	int kind = 0;
	Circuit C;                  // gates (and directions) of this request
	std::shared_ptr<const EvalPlan> plan;
	numeric* fval = nullptr;
	double* grad = nullptr;
	double* hvp = nullptr;
	int status = 1;
	int slot = -1;
	size_t off = 0;
	Clock::time_point t_sub, t_start, t_end;
	bool done = false;
};

struct Worker
{
	int id = 0;
	Slot slot;
	std::thread th;
	char* h_tab = nullptr;
	size_t tab_cap = 0;
	char* h_res = nullptr;
	size_t res_cap = 0;
};

struct Server
{
	Context* dev = nullptr;
	char* arena = nullptr;
	ArenaAlloc alloc;
	size_t max_req = 0;
	std::vector<std::unique_ptr<Worker>> workers;
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<Req*> queue;
	std::unordered_map<long, std::unique_ptr<Req>> reqs;
	long next_id = 1;
	bool stop = false;
	int in_flight = 0;
	rqc_server_stats stats{};
	std::vector<::rqc_session*> sessions;   // released with the server
	Clock::time_point t0 = Clock::now();
	std::mutex plan_mu;
	std::map<std::string, std::shared_ptr<const EvalPlan>> plans;   // by circuit structure
};

static void ensure_pinned(char** p, size_t* cap, size_t n)
{
	if (n <= *cap) { return; }
	if (*p != nullptr) { (void)hipHostFree(*p); *p = nullptr; *cap = 0; }
	const size_t c = std::max<size_t>(n, (size_t)1 << 20);
	RQC_HIP(hipHostMalloc(reinterpret_cast<void**>(p), c, hipHostMallocDefault));
	*cap = c;
}

static void run_request(Server& S, Worker& w, Req& r)
{
	const EvalPlan& pl = *r.plan;
	ensure_pinned(&w.h_tab, &w.tab_cap, pl.tab_bytes);
	ensure_pinned(&w.h_res, &w.res_cap, pl.res_bytes());
	fill_eval_tables(pl, r.C, w.h_tab);
	char* work = S.arena + r.off;
	enqueue_eval(*S.dev, w.slot, pl, r.C, r.s->dU, work, nullptr, nullptr, w.h_tab, nullptr);
	const size_t nres = (r.kind == RQC_EVAL_TARGET) ? sizeof(zc) : pl.res_bytes();
	RQC_HIP(hipMemcpyAsync(w.h_res, work + pl.off_res, nres, hipMemcpyDeviceToHost, w.slot.stream));
	RQC_HIP(hipStreamSynchronize(w.slot.stream));
	EvalResult R;
	unpack_result(pl, w.h_res, R);
	if (r.fval != nullptr) { std::memcpy(r.fval, &R.f, sizeof(zc)); }
	if (r.kind >= RQC_EVAL_GRADIENT && r.grad != nullptr) { std::memcpy(r.grad, R.grad.data(), R.grad.size() * sizeof(double)); }
	if (r.kind == RQC_EVAL_HVP && r.hvp != nullptr) { std::memcpy(r.hvp, R.hvp.data(), R.hvp.size() * sizeof(double)); }
}

static void worker_loop(Server& S, Worker& w)
{
	(void)hipSetDevice(S.dev->device);
	std::unique_lock<std::mutex> lk(S.mu);
	for (;;) {
		size_t off = 0;
		// the head of the queue, once its workspace fits
		S.cv_work.wait(lk, [&] {
			if (S.queue.empty()) { return S.stop; }
			return S.alloc.alloc(S.queue.front()->plan->bytes, &off);
		});
		if (S.queue.empty()) { break; }
		Req* r = S.queue.front();
		S.queue.pop_front();
		r->off = off;
		r->slot = w.id;
		r->t_start = Clock::now();
		S.in_flight++;
		S.stats.peak_in_flight = std::max(S.stats.peak_in_flight, S.in_flight);
		S.stats.peak_arena_bytes = std::max(S.stats.peak_arena_bytes, S.alloc.peak);
		// the next worker may admit the new head
		S.cv_work.notify_one();
		lk.unlock();
		int status = 0;
		try {
			run_request(S, w, *r);
		} catch (const std::exception& e) {
			status = report_error("rqc_server request", e);
			if (status >= 0) { status = -1; }
		}
		lk.lock();
		S.alloc.release(r->off, r->plan->bytes);
		S.in_flight--;
		r->t_end = Clock::now();
		r->status = status;
		r->done = true;
		r->s->pending--;
		if (status == 0) { S.stats.completed++; } else { S.stats.failed++; }
		S.cv_done.notify_all();
		S.cv_work.notify_all();
	}
}

// plans are shared by every session of the same structure (qubits, wires,
// groups, precision) and made once per evaluation kind and engine
static std::shared_ptr<const EvalPlan> session_plan(Session& s, int kind, bool unitary)
{
	{
		std::lock_guard<std::mutex> lk(s.mu);
		if (s.plan[kind][unitary ? 1 : 0]) { return s.plan[kind][unitary ? 1 : 0]; }
	}
	Server& S = *s.srv;
	std::string key = std::to_string(s.C.L) + "/" + std::to_string(s.per) + "/" + std::to_string(s.fp32) + "/" +
		std::to_string(kind) + "/" + std::to_string(unitary) + ":";
	for (int k = 0; k < s.C.ngates; k++) { key += std::to_string(s.C.wi[k]) + "," + std::to_string(s.C.wj[k]) + ";"; }
	std::shared_ptr<const EvalPlan> p;
	{
		std::lock_guard<std::mutex> lk(S.plan_mu);
		auto it = S.plans.find(key);
		if (it != S.plans.end()) {
			p = it->second;
		} else {
			const Mode m = kind == RQC_EVAL_TARGET ? Mode::Target : (kind == RQC_EVAL_GRADIENT ? Mode::Grad : Mode::Hvp);
			Circuit C = s.C;
			C.dirs.assign(C.gates.size(), zc{0, 0});
			auto np = std::make_shared<EvalPlan>(plan_eval(*S.dev, C, m, s.per, s.fp32, true, unitary, S.max_req, 0));
			if (np->bytes > S.alloc.size) {
				throw Error(-1, "request workspace (" + std::to_string(np->bytes) + " bytes) exceeds the server arena");
			}
			S.plans.emplace(key, np);
			p = np;
		}
	}
	std::lock_guard<std::mutex> lk(s.mu);
	s.plan[kind][unitary ? 1 : 0] = p;
	return p;
}

}  // namespace rqc

using namespace rqc;

struct rqc_server
{
	Server s;
};

struct rqc_session
{
	Session s;
};

extern "C" int rqc_server_create(const struct rqc_server_params* params, struct rqc_server** server)
{
	rqc_server* h = nullptr;
	try {
		Context& c = ctx();
		h = new rqc_server();
		Server& S = h->s;
		S.dev = &c;
		size_t arena = params != nullptr ? params->arena_bytes : 0;
		if (arena == 0) {
			size_t fr = 0, tot = 0;
			RQC_HIP(hipMemGetInfo(&fr, &tot));
			arena = (size_t)(0.8 * (double)fr);
		}
		arena &= ~(size_t)255;
		S.max_req = (params != nullptr && params->max_request_bytes != 0) ? params->max_request_bytes : arena / 4;
		const int nslots = (params != nullptr && params->nslots > 0) ? params->nslots : 8;
		RQC_HIP(hipMalloc(&S.arena, arena));
		S.alloc.init(arena);
		S.stats.arena_bytes = arena;
		for (int i = 0; i < nslots; i++) {
			auto w = std::make_unique<Worker>();
			w->id = i;
			RQC_HIP(hipStreamCreateWithFlags(&w->slot.stream, hipStreamNonBlocking));
			S.workers.push_back(std::move(w));
		}
		for (auto& w : S.workers) {
			Worker* wp = w.get();
			wp->th = std::thread([&S, wp] { worker_loop(S, *wp); });
		}
		*server = h;
		return 0;
	} catch (const std::exception& e) {
		if (h != nullptr) {
			{
				std::lock_guard<std::mutex> lk(h->s.mu);
				h->s.stop = true;
			}
			h->s.cv_work.notify_all();
			for (auto& w : h->s.workers) { if (w->th.joinable()) { w->th.join(); } }
			for (auto& w : h->s.workers) { if (w->slot.stream != nullptr) { (void)hipStreamDestroy(w->slot.stream); } }
			if (h->s.arena != nullptr) { (void)hipFree(h->s.arena); }
			delete h;
		}
		*server = nullptr;
		return report_error("rqc_server_create", e);
	}
}

extern "C" void rqc_server_destroy(struct rqc_server* server)
{
	if (server == nullptr) { return; }
	Server& S = server->s;
	{
		std::lock_guard<std::mutex> lk(S.mu);
		S.stop = true;
	}
	S.cv_work.notify_all();
	for (auto& w : S.workers) { if (w->th.joinable()) { w->th.join(); } }
	for (auto& w : S.workers) {
		if (w->slot.stream != nullptr) { (void)hipStreamDestroy(w->slot.stream); }
		if (w->h_tab != nullptr) { (void)hipHostFree(w->h_tab); }
		if (w->h_res != nullptr) { (void)hipHostFree(w->h_res); }
	}
	if (S.arena != nullptr) { (void)hipFree(S.arena); }
	for (rqc_session* s : S.sessions) { delete s; }
	delete server;
}

static int new_session(rqc_server* server, Circuit&& C, int per, bool brickwall, const rqc_target* target, int precision,
	rqc_session** session)
{
	if (server == nullptr) { throw Error(-3, "server is NULL"); }
	if (precision != RQC_PRECISION_FP64 && precision != RQC_PRECISION_MIXED) { throw Error(-3, "unknown precision"); }
	auto* h = new rqc_session();
	Session& s = h->s;
	s.srv = &server->s;
	s.C = std::move(C);
	s.per = per;
	s.ng = s.C.ngates / per;
	s.brickwall = brickwall;
	s.dU = device_target_matrix(const_cast<rqc_target*>(target), s.C.L);
	s.fp32 = (precision == RQC_PRECISION_MIXED);
	{
		std::lock_guard<std::mutex> lk(server->s.mu);
		server->s.sessions.push_back(h);
	}
	*session = h;
	return 0;
}

extern "C" int rqc_session_create_brickwall(struct rqc_server* server, int nqubits, int nlayers, const int* perms[],
	const struct rqc_target* target, int precision, struct rqc_session** session)
{
	try {
		std::vector<struct mat4x4> z((size_t)std::max(nlayers, 1));
		for (auto& m : z) { for (auto& e : m.data) { e = numeric(0.0, 0.0); } }
		Circuit C = make_brickwall(z.data(), nullptr, nlayers, perms, nqubits);
		return new_session(server, std::move(C), nqubits / 2, true, target, precision, session);
	} catch (const std::exception& e) { return report_error("rqc_session_create_brickwall", e); }
}

extern "C" int rqc_session_create_circuit(struct rqc_server* server, int nqubits, int ngates, const int wires[],
	const struct rqc_target* target, int precision, struct rqc_session** session)
{
	try {
		if (ngates < 1) { throw Error(-3, "a circuit needs at least one gate"); }
		std::vector<struct mat4x4> z((size_t)ngates);
		for (auto& m : z) { for (auto& e : m.data) { e = numeric(0.0, 0.0); } }
		Circuit C = make_circuit(z.data(), nullptr, ngates, wires, nqubits);
		return new_session(server, std::move(C), 1, false, target, precision, session);
	} catch (const std::exception& e) { return report_error("rqc_session_create_circuit", e); }
}

extern "C" int rqc_session_ngroups(const struct rqc_session* session) { return session != nullptr ? session->s.ng : -3; }

extern "C" void rqc_session_destroy(struct rqc_session* session)
{
	if (session == nullptr) { return; }
	Server& S = *session->s.srv;
	std::unique_lock<std::mutex> lk(S.mu);
	S.cv_done.wait(lk, [&] { return session->s.pending == 0; });
	S.sessions.erase(std::remove(S.sessions.begin(), S.sessions.end(), session), S.sessions.end());
	lk.unlock();
	delete session;
}

extern "C" int rqc_server_submit(struct rqc_server* server, struct rqc_session* session, int kind, const struct mat4x4 gates[],
	const struct mat4x4 dirs[], numeric* fval, double* grad, double* hvp, long* ticket)
{
	try {
		if (server == nullptr || session == nullptr || ticket == nullptr) { throw Error(-3, "NULL server, session or ticket"); }
		if (kind < RQC_EVAL_TARGET || kind > RQC_EVAL_HVP) { throw Error(-3, "unknown evaluation kind"); }
		if (kind == RQC_EVAL_HVP && dirs == nullptr) { throw Error(-3, "an HVP request needs directions"); }
		Session& s = session->s;
		Server& S = server->s;
		auto r = std::make_unique<Req>();
		r->t_sub = Clock::now();
		r->s = &s;
		r->kind = kind;
		{
			std::lock_guard<std::mutex> lk(s.mu);
			if (gates != nullptr) {
				for (int l = 0; l < s.ng; l++) {
					for (int q = 0; q < s.per; q++) {
						std::memcpy(&s.C.gates[16 * ((size_t)l * s.per + q)], gates[l].data, 16 * sizeof(zc));
					}
				}
				s.has_gates = true;
			}
			if (!s.has_gates) { throw Error(-3, "the session has no gates yet"); }
			r->C = s.C;
		}
		r->C.dirs.assign(r->C.gates.size(), zc{0, 0});
		if (kind == RQC_EVAL_HVP) {
			for (int l = 0; l < s.ng; l++) {
				for (int q = 0; q < s.per; q++) {
					std::memcpy(&r->C.dirs[16 * ((size_t)l * s.per + q)], dirs[l].data, 16 * sizeof(zc));
				}
			}
		}
		const bool unitary = sweep_unitary(r->C, kind == RQC_EVAL_TARGET ? Mode::Target : Mode::Grad, s.fp32);
		r->plan = session_plan(s, kind, unitary);
		r->fval = fval;
		r->grad = grad;
		r->hvp = hvp;
		std::lock_guard<std::mutex> lk(S.mu);
		if (S.stop) { throw Error(-3, "the server is shutting down"); }
		r->id = S.next_id++;
		*ticket = r->id;
		s.pending++;
		S.stats.submitted++;
		S.queue.push_back(r.get());
		S.reqs.emplace(r->id, std::move(r));
		S.cv_work.notify_one();
		return 0;
	} catch (const std::exception& e) { return report_error("rqc_server_submit", e); }
}

static void fill_info(const Server& S, const Req& r, rqc_request_info* info)
{
	if (info == nullptr) { return; }
	info->start_ms = ms_between(S.t0, r.t_start);
	info->end_ms = ms_between(S.t0, r.t_end);
	info->status = r.status;
	info->slot = r.slot;
	info->queued_ms = ms_between(r.t_sub, r.t_start);
	info->run_ms = ms_between(r.t_start, r.t_end);
	info->latency_ms = ms_between(r.t_sub, r.t_end);
	info->workspace_bytes = r.plan ? r.plan->bytes : 0;
	info->block = r.plan ? r.plan->B : 0;
}

extern "C" int rqc_server_wait(struct rqc_server* server, long ticket, struct rqc_request_info* info)
{
	if (server == nullptr) { return -3; }
	Server& S = server->s;
	std::unique_lock<std::mutex> lk(S.mu);
	auto it = S.reqs.find(ticket);
	if (it == S.reqs.end()) { return -3; }
	Req& r = *it->second;
	S.cv_done.wait(lk, [&] { return r.done; });
	fill_info(S, r, info);
	return r.status;
}

extern "C" int rqc_server_poll(struct rqc_server* server, long ticket, struct rqc_request_info* info)
{
	if (server == nullptr) { return -3; }
	Server& S = server->s;
	std::lock_guard<std::mutex> lk(S.mu);
	auto it = S.reqs.find(ticket);
	if (it == S.reqs.end()) { return -3; }
	if (!it->second->done) { return 0; }
	fill_info(S, *it->second, info);
	return 1;
}

extern "C" int rqc_server_release(struct rqc_server* server, long ticket)
{
	if (server == nullptr) { return -3; }
	Server& S = server->s;
	std::lock_guard<std::mutex> lk(S.mu);
	auto it = S.reqs.find(ticket);
	if (it == S.reqs.end() || !it->second->done) { return -3; }
	S.reqs.erase(it);
	return 0;
}

extern "C" int rqc_server_get_stats(const struct rqc_server* server, struct rqc_server_stats* stats)
{
	if (server == nullptr || stats == nullptr) { return -3; }
	Server& S = const_cast<Server&>(server->s);
	std::lock_guard<std::mutex> lk(S.mu);
	*stats = S.stats;
	stats->peak_arena_bytes = S.alloc.peak;
	return 0;
}

// ---------------------------------------------------------------- trust-region solves through the server
// One optimize_brickwall_circuit_hvp (reference src/brickwall_opt.c:127-150; for a
// general-circuit session optimize_quantum_circuit, src/circuit_opt.c:104-124, whose
// groups are single gates) on the calling thread whose target, gradient and HVP
// evaluations are requests of
// the server: many solves on as many host threads share the device, their
// requests admitted and overlapped like any others.  In mixed precision the
// target request computes the gradient too (its f carries the gate-rounding
// correction), and a gradient request at the same gates reuses it.
namespace {

struct SolveData
{
	rqc_server* server;
	rqc_session* session;
	int nl;
	bool mixed;
	rqc_request_info* rec;
	int cap;
	int n = 0;
	bool failed = false;
	std::vector<double> memo_x, memo_g;
	double memo_f = 0;
};

double solve_request(SolveData& d, int kind, const double* x, const struct mat4x4* dirs, double* grad, double* hvp)
{
	numeric f;
	long ticket = 0;
	int r = rqc_server_submit(d.server, d.session, kind, reinterpret_cast<const struct mat4x4*>(x), dirs, &f, grad, hvp,
		&ticket);
	rqc_request_info info{};
	if (r == 0) { r = rqc_server_wait(d.server, ticket, &info); }
	if (r == 0) {
		if (d.rec != nullptr && d.n < d.cap) { d.rec[d.n] = info; }
		d.n++;
		(void)rqc_server_release(d.server, ticket);
	} else {
		d.failed = true;
		return std::nan("");
	}
	return f.real();
}

double sv_f_deriv(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT grad)
{
	auto* d = static_cast<SolveData*>(fdata);
	const size_t m = 32 * (size_t)d->nl, n = 16 * (size_t)d->nl;
	if (d->mixed && d->memo_x.size() == m && std::memcmp(d->memo_x.data(), x, m * sizeof(double)) == 0) {
		std::memcpy(grad, d->memo_g.data(), n * sizeof(double));
		return d->memo_f;
	}
	const double f = solve_request(*d, RQC_EVAL_GRADIENT, x, nullptr, grad, nullptr);
	if (d->mixed) {
		d->memo_x.assign(x, x + m);
		d->memo_g.assign(grad, grad + n);
		d->memo_f = f;
	}
	return f;
}

double sv_f(const double* x, void* fdata)
{
	auto* d = static_cast<SolveData*>(fdata);
	if (d->mixed) {
		std::vector<double> g(16 * (size_t)d->nl);
		return sv_f_deriv(x, fdata, g.data());
	}
	return solve_request(*d, RQC_EVAL_TARGET, x, nullptr, nullptr, nullptr);
}

void sv_f_hvp(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT vec, double* RQC_RESTRICT hvp)
{
	auto* d = static_cast<SolveData*>(fdata);
	const auto* v = reinterpret_cast<const struct mat4x4*>(x);
	std::vector<struct mat4x4> dirs(d->nl);
	for (int i = 0; i < d->nl; i++) { real_to_tangent(&vec[16 * i], &v[i], &dirs[i]); }
	std::vector<double> grad(16 * (size_t)d->nl);
	(void)solve_request(*d, RQC_EVAL_HVP, x, dirs.data(), grad.data(), hvp);
}

void sv_retract(const double* RQC_RESTRICT x, const double* RQC_RESTRICT eta, void* rdata, double* RQC_RESTRICT xs)
{
	const int nl = *static_cast<int*>(rdata);
	const auto* v = reinterpret_cast<const struct mat4x4*>(x);
	auto* w = reinterpret_cast<struct mat4x4*>(xs);
	for (int i = 0; i < nl; i++) { retract(&v[i], &eta[16 * i], &w[i]); }
}

}  // namespace

extern "C" int rqc_server_optimize_brickwall_hvp(struct rqc_server* server, struct rqc_session* session,
	const struct mat4x4 vlist_start[], const struct rtr_params* params, int niter, double* f_iter,
	struct mat4x4 vlist_opt[], struct rqc_request_info* records, int max_records, int* nrecords)
{
	try {
		if (server == nullptr || session == nullptr || params == nullptr) { throw Error(-3, "NULL server, session or params"); }
		if (niter < 0) { throw Error(-3, "niter must be >= 0"); }
		if (!(0 <= params->rho_trust && params->rho_trust < 0.25)) { throw Error(-3, "rho_trust must lie in [0, 0.25)"); }
		SolveData d{server, session, session->s.ng, session->s.fp32, records, records != nullptr ? max_records : 0};
		struct rtr_params prm = *params;
		prm.g_func = nullptr;
		prm.g_data = nullptr;
		prm.g_iter = nullptr;
		int nl = d.nl;
		riemannian_trust_region_optimize_hvp(sv_f, sv_f_deriv, sv_f_hvp, &d, sv_retract, &nl, nl * num_tangent_params,
			reinterpret_cast<const double*>(vlist_start), nl * 32, &prm, niter, f_iter, reinterpret_cast<double*>(vlist_opt));
		if (nrecords != nullptr) { *nrecords = d.n; }
		if (d.failed) { throw Error(-1, "a request of the solve failed"); }
		return 0;
	} catch (const std::exception& e) { return report_error("rqc_server_optimize_brickwall_hvp", e); }
}
