"""Serving workload and open-loop load driver for the request server (rqc_server.h).

Workload: nsolves concurrent trust-region solves, each with qubit count L
uniform in {8, ..., 14}, brickwall depth uniform in {3, 5, 7, 9}, and a target
that is either the Ising chain (J=1, h=0.75, g=0.6, t=0.5; one target per L,
shared) or one of a small pool of seeded XXZ chains (Delta=0.5, fields uniform
in [-1, 1], t uniform in [0.1, 1.0]) per L.  Even L use the periodic brickwall
with one gate per layer (brickwall sessions); odd L use an open brickwall with
one gate per pair, (j, j+1) for j = (layer mod 2), +2, ... (circuit sessions).
Targets are exact-bond Trotter circuits (second-order splitting, 4 steps)
formed once on the GPU and kept resident.

Load: Poisson arrivals at `rate` requests/s for `duration` seconds; each
arrival picks a solve uniformly and a kind (target / gradient / HVP) with the
given weights.  A solve keeps its gates on the server between requests (its
session holds them); with probability `move_prob` a request carries a new
iterate (a retraction of the current gates along a random tangent step).
Latency is measured from the scheduled arrival time to the result on the host,
so a driver that falls behind does not hide queueing.
"""
import time

import numpy as np

from . import (EVAL_GRADIENT, EVAL_HVP, EVAL_TARGET, PRECISION_FP64, DeviceTarget, Rng, real_to_tangent, retract)
from .workloads import (ID2, PAULI_X, PAULI_Y, PAULI_Z, brickwall_perms, expm_hermitian)

ISING = dict(J=1.0, h=0.75, g=0.6, t=0.5)
XXZ_DELTA = 0.5


def chain_bonds(L):
    """Bonds of the chain: periodic for even L, open for odd L; each as (j, j+1 mod L)."""
    n = L if L % 2 == 0 else L - 1
    return [(j, (j + 1) % L) for j in range(n)]


def chain_trotter_circuit(L, bond_term, t, steps):
    """Second-order splitting of H = sum over bonds of bond_term(j, q, wj, wq) into even bonds (j even) and odd
    bonds, adjacent half steps merged; wj = 1 / (number of bonds at site j) splits the on-site terms.
    Returns gates [ngates, 4, 4] and wires [ngates, 2]."""
    bonds = chain_bonds(L)
    deg = np.zeros(L)
    for j, q in bonds:
        deg[j] += 1
        deg[q] += 1
    even = [(j, q) for j, q in bonds if j % 2 == 0]
    odd = [(j, q) for j, q in bonds if j % 2 == 1]
    tau = t / steps
    gates, wires = [], []

    def layer(bs, dt):
        for j, q in bs:
            gates.append(expm_hermitian(bond_term(j, q, 1 / deg[j], 1 / deg[q]), dt))
            wires.append((j, q))

    for s in range(steps):
        layer(even, tau / 2 if s == 0 else tau)
        layer(odd, tau)
    layer(even, tau / 2)
    return np.stack(gates), np.array(wires, dtype=np.int32)


def ising_term(J, h, g):
    def term(j, q, wj, wq):
        return (J * np.kron(PAULI_Z, PAULI_Z) + h * (wj * np.kron(PAULI_X, ID2) + wq * np.kron(ID2, PAULI_X))
                + g * (wj * np.kron(PAULI_Z, ID2) + wq * np.kron(ID2, PAULI_Z)))
    return term


def xxz_term(delta, fields):
    def term(j, q, wj, wq):
        return (np.kron(PAULI_X, PAULI_X) + np.kron(PAULI_Y, PAULI_Y) + delta * np.kron(PAULI_Z, PAULI_Z)
                + fields[j] * wj * np.kron(PAULI_Z, ID2) + fields[q] * wq * np.kron(ID2, PAULI_Z))
    return term


def open_brickwall_wires(L, depth):
    """Odd-L layout: layer i holds the pairs (j, j+1), j = i mod 2, i mod 2 + 2, ... < L - 1."""
    w = []
    for i in range(depth):
        for j in range(i % 2, L - 1, 2):
            w.append((j, j + 1))
    return np.array(w, dtype=np.int32)


def serving_workload(seed, nsolves=256, qubits=tuple(range(8, 15)), depths=(3, 5, 7, 9), xxz_pool=3,
                     trotter_steps=4):
    """Solves (dicts with L, depth, layout, target key, gates) and target specs keyed by (kind, L, index)."""
    r = Rng(seed)
    targets = {}
    for L in qubits:
        targets[("ising", L, 0)] = chain_trotter_circuit(L, ising_term(ISING["J"], ISING["h"], ISING["g"]),
                                                         ISING["t"], trotter_steps)
        for k in range(xxz_pool):
            fields = np.array([2 * r.randu() - 1 for _ in range(L)])
            t = 0.1 + 0.9 * r.randu()
            targets[("xxz", L, k)] = chain_trotter_circuit(L, xxz_term(XXZ_DELTA, fields), t, trotter_steps)
    solves = []
    for _ in range(nsolves):
        L = qubits[int(r.randu() * len(qubits))]
        d = depths[int(r.randu() * len(depths))]
        key = ("ising", L, 0) if r.randu() < 0.5 else ("xxz", L, int(r.randu() * xxz_pool))
        if L % 2 == 0:
            s = dict(L=L, depth=d, brickwall=True, perms=brickwall_perms(L, d))
            ng = d
        else:
            s = dict(L=L, depth=d, brickwall=False, wires=open_brickwall_wires(L, d))
            ng = len(s["wires"])
        s["target"] = key
        s["gates"] = np.stack([r.haar_unitary() for _ in range(ng)])
        solves.append(s)
    return solves, targets


def open_sessions(server, solves, targets, precision=PRECISION_FP64):
    """Device targets (each formed once) and one session per solve."""
    dev = {}
    for s in solves:
        if s["target"] not in dev:
            g, w = targets[s["target"]]
            dev[s["target"]] = DeviceTarget.circuit(g, w, s["L"])
    sessions = []
    for s in solves:
        T = dev[s["target"]]
        if s["brickwall"]:
            sessions.append(server.session_brickwall(s["perms"], T, precision))
        else:
            sessions.append(server.session_circuit(s["wires"], s["L"], T, precision))
    return sessions, dev


def run_poisson(server, solves, sessions, rate, duration, seed, weights=(1 / 3, 1 / 3, 1 / 3), move_prob=0.2,
                max_requests=None):
    """Open-loop Poisson load; returns latency and throughput statistics (milliseconds, requests/s)."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_requests is not None and len(arrivals) >= max_requests):
            break
        arrivals.append(t)
    kinds = [EVAL_TARGET, EVAL_GRADIENT, EVAL_HVP]
    cur = [s["gates"].copy() for s in solves]
    plan = []
    for _ in arrivals:
        j = int(rng.integers(len(solves)))
        kind = kinds[int(rng.choice(3, p=np.asarray(weights) / np.sum(weights)))]
        gates = None
        if rng.random() < move_prob:
            eta = 1e-2 * rng.standard_normal(16 * len(cur[j]))
            cur[j] = np.stack([retract(cur[j][i], eta[16 * i:16 * i + 16]) for i in range(len(cur[j]))])
            gates = cur[j]
        dirs = None
        if kind == EVAL_HVP:
            vec = rng.standard_normal(16 * len(cur[j]))
            dirs = np.stack([real_to_tangent(vec[16 * i:16 * i + 16], cur[j][i]) for i in range(len(cur[j]))])
        plan.append((j, kind, gates, dirs))
    # every session starts from its solve's gates
    first = set()
    pend = []
    t0 = time.perf_counter()
    for (ta, (j, kind, gates, dirs)) in zip(arrivals, plan):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        if gates is None and j not in first:
            gates = solves[j]["gates"]
        first.add(j)
        late = (time.perf_counter() - t0) - ta
        pend.append((ta, late, kind, j, server.submit(sessions[j], kind, gates, dirs)))
    lat, done_at, per_kind = [], [], {EVAL_TARGET: [], EVAL_GRADIENT: [], EVAL_HVP: []}
    for ta, late, kind, j, p in pend:
        _, _, _, info = server.wait(p)
        if info["status"] != 0:
            raise RuntimeError(f"request of solve {j} failed with status {info['status']}")
        l_ms = info["latency_ms"] + 1e3 * late
        lat.append(l_ms)
        per_kind[kind].append(l_ms)
        done_at.append(ta + l_ms / 1e3)
        server.release(p)
    lat = np.array(lat)
    span = max(done_at) - arrivals[0] if arrivals else 0.0
    return dict(requests=len(arrivals), offered_rate=rate, duration_s=duration,
                achieved_rate=len(arrivals) / span if span > 0 else 0.0,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()),
                p50_by_kind={k: (float(np.percentile(v, 50)) if v else None) for k, v in
                             (("target", per_kind[EVAL_TARGET]), ("gradient", per_kind[EVAL_GRADIENT]),
                              ("hvp", per_kind[EVAL_HVP]))},
                server=server.stats())


def run_solves(server, solves, sessions, concurrency, niter, seconds, seed=0):
    """Closed-loop load of real trust-region solves: `concurrency` host threads each run one solve at a
    time (Server.optimize: every target / gradient / HVP evaluation of the solve's driver is a request),
    taking the next solve of the workload, until `seconds` have passed.  Returns request latency
    percentiles (submission to result on the host) and the completed request and solve rates; a
    closed loop admits no backlog, so the achieved rate is the rate the device sustains at that
    concurrency."""
    import threading

    from . import rtr_default_params
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(solves))
    lock = threading.Lock()
    nxt = [0]
    lat, nsolve, fdec, errors = [], [0], [0], []
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                if time.perf_counter() - t0 > seconds:
                    return
                j = int(order[nxt[0] % len(order)])
                nxt[0] += 1
            s = solves[j]
            try:
                f_iter, _, recs = server.optimize(sessions[j], s["gates"], niter, rtr_default_params(16 * len(s["gates"])))
            except Exception as e:  # reported below; a failed solve is a failed benchmark
                errors.append(repr(e))
                return
            with lock:
                lat.extend(r["latency_ms"] for r in recs)
                nsolve[0] += 1
                fdec[0] += int(f_iter[-1] <= f_iter[0])

    ths = [threading.Thread(target=worker) for _ in range(concurrency)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    if errors:
        raise RuntimeError(f"{len(errors)} solves failed: {errors[0]}")
    lat = np.array(lat)
    return dict(concurrency=concurrency, niter=niter, seconds=wall, solves=nsolve[0], f_not_increased=fdec[0],
                requests=int(lat.size), achieved_rate=lat.size / wall, solves_per_s=nsolve[0] / wall,
                p50_ms=float(np.percentile(lat, 50)) if lat.size else None,
                p99_ms=float(np.percentile(lat, 99)) if lat.size else None,
                mean_ms=float(lat.mean()) if lat.size else None, server=server.stats())
