"""GPU tests of the request server (include/rqcopt/rqc_server.h).

The stated tolerance for "the same alone and in the mix" is zero: a request's
plan (column block, grids, reduction order) is fixed by its session and the
per-request workspace cap, so its result is bitwise identical whatever runs
beside it.  FP64 results also match the oracle to 1e-12 relative and the
single-call path of the reference ABI to 1e-13.
"""
import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from rqcopt_hpc_amd import serving as S
from rqcopt_hpc_amd import workloads as W
from helpers import brickwall_perms, haar_unitary, reldiff

pytestmark = pytest.mark.gpu

TOL = 1e-12
GiB = 1 << 30


@pytest.fixture(scope="module", autouse=True)
def gpu():
    R.require_gpu()
    R.set_precision(R.PRECISION_FP64)
    yield


def _target(L, seed):
    rng = np.random.default_rng(seed)
    g, w = S.chain_trotter_circuit(L, S.xxz_term(0.5, rng.uniform(-1, 1, L)), rng.uniform(0.1, 1.0), 2)
    return R.DeviceTarget.circuit(g, w, L)


def _dirs(rng, V):
    vec = rng.standard_normal(16 * len(V))
    return np.stack([R.real_to_tangent(vec[16 * i:16 * i + 16], V[i]) for i in range(len(V))])


def _setup(server, rng):
    """Sessions of uneven size and precision and one request of each kind per session."""
    specs = [(8, 3, True, R.PRECISION_FP64), (10, 5, True, R.PRECISION_MIXED), (7, 3, False, R.PRECISION_FP64),
             (6, 4, True, R.PRECISION_FP64), (12, 3, True, R.PRECISION_FP64), (9, 2, False, R.PRECISION_MIXED)]
    sess, reqs = [], []
    for i, (L, d, bw, prec) in enumerate(specs):
        T = _target(L, 100 + i)
        if bw:
            P = brickwall_perms(L, d)
            s = server.session_brickwall(P, T, prec)
            V = np.stack([haar_unitary(rng) for _ in range(d)])
            layout = dict(P=P)
        else:
            wires = S.open_brickwall_wires(L, d)
            s = server.session_circuit(wires, L, T, prec)
            V = np.stack([haar_unitary(rng) for _ in range(len(wires))])
            layout = dict(wires=wires)
        sess.append(dict(s=s, T=T, L=L, bw=bw, prec=prec, V=V, **layout))
        for kind in (R.EVAL_TARGET, R.EVAL_GRADIENT, R.EVAL_HVP):
            reqs.append((i, kind, V, _dirs(rng, V) if kind == R.EVAL_HVP else None))
    return sess, reqs


def _run_alone(server, sess, reqs):
    out = []
    for i, kind, V, Z in reqs:
        p = server.submit(sess[i]["s"], kind, V, Z)
        f, g, h, info = server.wait(p)
        assert info["status"] == 0
        out.append((f, None if g is None else g.copy(), None if h is None else h.copy()))
    return out


def test_results_same_alone_and_in_the_mix():
    rng = np.random.default_rng(1)
    srv = R.Server(arena_bytes=8 * GiB, max_request_bytes=GiB, nslots=4)
    sess, reqs = _setup(srv, rng)
    alone = _run_alone(srv, sess, reqs)
    # the mix: every request again, interleaved with background requests of all sessions
    order = rng.permutation(len(reqs))
    pend = []
    for k in order:
        i, kind, V, Z = reqs[k]
        for _ in range(2):
            j = int(rng.integers(len(sess)))
            pend.append((None, srv.submit(sess[j]["s"], R.EVAL_HVP, sess[j]["V"], _dirs(rng, sess[j]["V"]))))
        pend.append((k, srv.submit(sess[i]["s"], kind, V, Z)))
    slots = set()
    for k, p in pend:
        f, g, h, info = srv.wait(p)
        assert info["status"] == 0
        slots.add(info["slot"])
        if k is None:
            continue
        fa, ga, ha = alone[k]
        assert f == fa, (k, f, fa)
        assert (g is None and ga is None) or np.array_equal(g, ga)
        assert (h is None and ha is None) or np.array_equal(h, ha)
    st = srv.stats()
    print("server stats", st, "slots used", sorted(slots), flush=True)
    assert st["failed"] == 0 and st["completed"] == st["submitted"] == len(reqs) + len(pend)
    assert st["peak_in_flight"] >= 2 and len(slots) >= 2
    assert st["peak_arena_bytes"] <= st["arena_bytes"]
    # FP64 sessions: oracle and the single-call path
    for (i, kind, V, Z), (f, g, h) in zip(reqs, alone):
        e = sess[i]
        if e["prec"] != R.PRECISION_FP64 or e["L"] > 8:
            continue
        Uo = O.Ufunc.dense(e["T"].download())
        Zo = Z if Z is not None else np.zeros_like(V)
        if e["bw"]:
            f_o, g_o, h_o = O.brickwall_unitary_target_projected_hvp(Uo, V, Zo, e["P"], e["L"])
            f_s, g_s, h_s = R.brickwall_unitary_target_projected_hessian_vector_product(e["T"], V, Zo, e["P"])
        else:
            f_o, g_o, h_o = O.circuit_unitary_target_projected_hvp(Uo, V, Zo, e["wires"], e["L"])
            f_s, g_s, h_s = R.circuit_unitary_target_projected_hessian_vector_product(e["T"], V, Zo, e["wires"], e["L"])
        assert abs(f - f_o) < TOL * max(1, abs(f_o)) and abs(f - f_s) <= 1e-13 * max(1, abs(f_s))
        if g is not None:
            assert reldiff(g, g_o) < TOL and reldiff(g, g_s) <= 1e-13
        if h is not None:
            assert reldiff(h, h_o) < TOL and reldiff(h, h_s) <= 1e-13
    for e in sess:
        e["s"].close()
    srv.close()


def test_admission_by_footprint():
    """An arena that holds one large request at a time: large requests run one after another,
    small ones beside them; nothing exceeds the arena."""
    rng = np.random.default_rng(2)
    T12 = _target(12, 7)
    P12 = brickwall_perms(12, 5)
    V12 = np.stack([haar_unitary(rng) for _ in range(5)])
    probe = R.Server(arena_bytes=16 * GiB, max_request_bytes=GiB, nslots=1)
    s = probe.session_brickwall(P12, T12)
    need = probe.wait(probe.submit(s, R.EVAL_HVP, V12, _dirs(rng, V12)))[3]["workspace_bytes"]
    probe.close()
    arena = int(1.5 * need)
    srv = R.Server(arena_bytes=arena, max_request_bytes=GiB, nslots=4)
    big = srv.session_brickwall(P12, T12)
    T6 = _target(6, 8)
    small = srv.session_brickwall(brickwall_perms(6, 2), T6)
    V6 = np.stack([haar_unitary(rng) for _ in range(2)])
    pend = []
    for k in range(4):
        pend.append(srv.submit(big, R.EVAL_HVP, V12, _dirs(rng, V12)))
        pend.append(srv.submit(small, R.EVAL_GRADIENT, V6))
    infos = [srv.wait(p)[3] for p in pend]
    st = srv.stats()
    print("arena", arena, "need", need, "stats", st, flush=True)
    assert all(i["status"] == 0 for i in infos)
    assert st["peak_arena_bytes"] <= arena
    big_runs = sorted((i["start_ms"], i["end_ms"]) for i in infos[0::2])
    for a, b in zip(big_runs, big_runs[1:]):
        assert b[0] >= a[1] - 0.5   # a large request is admitted only after the previous one has released its memory
    # a request larger than the arena is refused at submission
    tiny = R.Server(arena_bytes=need // 2, max_request_bytes=GiB, nslots=1)
    sb = tiny.session_brickwall(P12, T12)
    with pytest.raises(R.RqcError):
        tiny.submit(sb, R.EVAL_HVP, V12, _dirs(rng, V12))
    tiny.close()
    srv.close()


def test_invalid_requests_are_rejected():
    rng = np.random.default_rng(3)
    srv = R.Server(arena_bytes=2 * GiB, max_request_bytes=GiB // 4, nslots=2)
    T = _target(8, 9)
    s = srv.session_brickwall(brickwall_perms(8, 3), T)
    V = np.stack([haar_unitary(rng) for _ in range(3)])
    with pytest.raises(R.RqcError):
        srv.submit(s, R.EVAL_TARGET)          # no gates yet
    with pytest.raises(R.RqcError):
        srv.submit(s, R.EVAL_HVP, V)          # HVP without directions
    f1 = srv.wait(srv.submit(s, R.EVAL_GRADIENT, V))
    f2 = srv.wait(srv.submit(s, R.EVAL_GRADIENT))   # the session keeps its gates
    assert f1[0] == f2[0] and np.array_equal(f1[1], f2[1])
    srv.close()


def test_poisson_load_small():
    solves, targets = S.serving_workload(11, nsolves=16, qubits=(8, 9, 10), depths=(3, 5))
    srv = R.Server(arena_bytes=16 * GiB, max_request_bytes=2 * GiB, nslots=4)
    sessions, dev = S.open_sessions(srv, solves, targets)
    res = S.run_poisson(srv, solves, sessions, rate=100.0, duration=0.5, seed=4)
    print("poisson", res, flush=True)
    assert res["requests"] > 10 and res["server"]["failed"] == 0
    assert 0 < res["p50_ms"] <= res["p99_ms"]
    for s in sessions:
        s.close()
    srv.close()


def test_serving_workload_large_sessions_alone_and_in_the_mix():
    """Sessions of the serving workload at its largest sizes (L=13 as an open-brickwall circuit and
    L=14 as a periodic brickwall, both depth 9, FP64 and mixed) with smaller ones beside them: each
    request's result is bitwise the same alone and interleaved with background requests."""
    solves, targets = S.serving_workload(1, nsolves=256)
    big = [next(i for i, s in enumerate(solves) if s["L"] == L and s["depth"] == 9) for L in (13, 14)]
    small = [next(i for i, s in enumerate(solves) if s["L"] == L) for L in (8, 10)]
    chosen = [solves[i] for i in big + small]
    srv = R.Server(arena_bytes=200 * GiB, max_request_bytes=48 * GiB, nslots=4)
    rng = np.random.default_rng(5)
    for prec in (R.PRECISION_FP64, R.PRECISION_MIXED):
        sessions, dev = S.open_sessions(srv, chosen, targets, prec)
        reqs = []
        for j in range(2):   # the two large sessions
            V = chosen[j]["gates"]
            for kind in (R.EVAL_TARGET, R.EVAL_GRADIENT, R.EVAL_HVP):
                reqs.append((j, kind, V, _dirs(rng, V) if kind == R.EVAL_HVP else None))
        alone = []
        for j, kind, V, Z in reqs:
            f, g, h, info = srv.wait(srv.submit(sessions[j], kind, V, Z))
            assert info["status"] == 0
            alone.append((f, None if g is None else g.copy(), None if h is None else h.copy()))
        pend = []
        for k in rng.permutation(len(reqs)):
            j, kind, V, Z = reqs[k]
            b = int(rng.integers(len(chosen)))
            pend.append((None, srv.submit(sessions[b], R.EVAL_HVP, chosen[b]["gates"], _dirs(rng, chosen[b]["gates"]))))
            pend.append((k, srv.submit(sessions[j], kind, V, Z)))
        for k, p in pend:
            f, g, h, info = srv.wait(p)
            assert info["status"] == 0
            if k is None:
                continue
            fa, ga, ha = alone[k]
            print(f"prec {prec} request {k} L={chosen[reqs[k][0]]['L']} kind {reqs[k][1]}: slot {info['slot']}, "
                  f"{info['latency_ms']:.1f} ms", flush=True)
            assert f == fa
            assert (g is None and ga is None) or np.array_equal(g, ga)
            assert (h is None and ha is None) or np.array_equal(h, ha)
        st = srv.stats()
        assert st["failed"] == 0 and st["peak_arena_bytes"] <= st["arena_bytes"]
        for s_ in sessions:
            s_.close()
        for t in dev.values():
            t.close()
    srv.close()


def test_concurrent_trust_region_solves_through_the_server():
    """Trust-region solves on several host threads, every evaluation of each solve's driver a server
    request (rqc_server_optimize_brickwall_hvp): each solve's trajectory equals the library's own
    optimiser on the same problem (FP64: the same plans and kernels, so the same rounding, and
    1e-12 relative in f_iter / 1e-10 in the gates allows for a different column block), also
    while the other solves' requests run beside it; a general-circuit session follows
    optimize_quantum_circuit the same way."""
    import threading

    rng = np.random.default_rng(11)
    server = R.Server(arena_bytes=8 * GiB, nslots=4)
    specs = [(8, 3, True), (6, 4, True), (10, 3, True), (7, 2, False)]
    solves = []
    for i, (L, d, bw) in enumerate(specs):
        T = _target(L, 200 + i)
        if bw:
            P = brickwall_perms(L, d)
            s = server.session_brickwall(P, T)
            V = np.stack([haar_unitary(rng) for _ in range(d)])
            ref = R.optimize_brickwall_circuit_hvp(T, V, P, 3)
        else:
            wires = S.open_brickwall_wires(L, d)
            s = server.session_circuit(wires, L, T)
            V = np.stack([haar_unitary(rng) for _ in range(len(wires))])
            ref = R.optimize_quantum_circuit(T, V, wires, L, 3)
        solves.append(dict(s=s, T=T, V=V, ref=ref))
    out = [None] * len(solves)

    def run(i):
        out[i] = server.optimize(solves[i]["s"], solves[i]["V"], 3)

    ths = [threading.Thread(target=run, args=(i,)) for i in range(len(solves))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for i, sv in enumerate(solves):
        assert out[i] is not None, f"solve {i} raised"
        f_iter, vopt, recs = out[i]
        f_ref, v_ref = sv["ref"]
        ef = np.max(np.abs(f_iter - f_ref) / np.maximum(1, np.abs(f_ref)))
        print(f"solve {i}: {len(recs)} requests, f_iter {f_iter}, rel diff {ef:.2e}", flush=True)
        assert ef < 1e-12 and np.max(np.abs(vopt - v_ref)) < 1e-10
        assert len(recs) > 3 and all(r["status"] == 0 for r in recs)
    assert server.stats()["peak_in_flight"] >= 2
    for sv in solves:
        sv["s"].close()
        sv["T"].close()
    server.close()
