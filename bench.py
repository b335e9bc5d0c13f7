"""Benchmark: grad+HVP evaluations per second of a periodic brickwall circuit.

One "step" is one grad+HVP eval: a call of
brickwall_unitary_target_projected_hessian_vector_product (target f, projected
gradient and projected HVP for one direction) through the C ABI, with U
resident in HBM (rqc_device_ufunc + DeviceTarget).  Default workload: L=14
qubits, depth 9, Haar-random gates (seeded PCG), target U = Strang-Trotter
circuit of the Ising chain (J=1, h=0.75, g=0.6, t=0.5, 4 steps) formed on the
GPU.  The FP64 path is the headline value; the mixed-precision path (complex
FP32 sweeps, FP64 reductions) is reported beside it with its error.

Beside the headline line's single-circuit numbers it reports
  "batched": 64 independent brickwall problems (L in {8, 10, 12}, depth in
    {3, 5, 7}, XXZ Trotter targets with seeded fields and times) evaluated
    together (one graph launch per evaluation kind and step): grad+HVP
    evals/s in FP64 and mixed, and a lock-step trust-region run of all 64;
  "serving": the request server under an open-loop Poisson stream of
    target / gradient / HVP requests from 256 solves (L in {8..14}, depth in
    {3, 5, 7, 9}, Ising or XXZ targets): p50 / p99 latency and completed
    requests/s in FP64 and mixed, at the offered rate and at a quarter of it;
    and under real trust-region solves of the same workload (each solve's
    driver issuing its own requests), at several concurrencies.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--L 14] [--depth 9]

Prints one JSON line.  Multi-GPU runs are not part of this build (--gpus > 1 exits).
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grad+HVP evals/s, L=14 brickwall d9, FP64 vs mixed; served p99 latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
REVERSE_KERNEL = {
    "tile_pass": "k_tile_reverse<double,dual> (reverse tile pass: up to 8 gates per HBM pass; state recovery, "
                 "adjoint update and gradient/HVP outer products on the FP64 matrix cores as 4x4x4 block "
                 "products, the gate applications in the three-real-product complex form; the launch of the circuit's last pass also runs that pass's forward gates, and "
                 "the circuit's first pass skips the tiles that hold no identity element)",
    "per_gate": "k_gate_reverse<double,dual> (per-gate reverse sweep with gradient/HVP sums)",
}
# measured HBM traffic of the reverse kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-kernel calibrated)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_k_tile_reverse_L14_d9.json")


def reverse_alg_flops(L, d):
    """Algorithmic FP64 flops of the reverse sweep of one grad+HVP eval: per gate and amplitude, the
    adjoint update (phi: 4 complex MACs, dphi = G^T dphi + Z^T phi: 8) and the gradient / HVP outer
    products (phi x psi, phi x dpsi, dphi x psi: 4 each): 24 complex MACs, 8 flops each, 192 flops.
    Together with the forward sweep's 12 (psi: 4, dpsi: 8) these are the 36 complex MACs per
    amplitude and gate of the eval's canonical count (144 L d N^2 flops)."""
    return 192.0 * (d * (L // 2)) * (1 << L) ** 2


def canonical_bytes(L, d, s):
    """Algorithmic HBM bytes of one layer-fused grad+HVP eval: (13d - 9) N^2 s."""
    return (13 * d - 9) * (1 << L) ** 2 * s


def canonical_flops(L, d):
    """~144 L d N^2: 36 complex MACs per amplitude per gate, L/2 gates per layer."""
    return 144.0 * L * d * (1 << L) ** 2


def _cpu_worker(args):
    # host-only: the worker loads the oracle and never the GPU library
    L, V, Z, P, Vt, Pt, b0, b1 = args
    import oracle as O
    uf = O.Ufunc.brickwall(Vt, Pt)
    t0 = time.perf_counter()
    O.brickwall_hvp_columns(uf, V, Z, P, L, b0, b1)
    return time.perf_counter() - t0


def cpu_baseline(L, V, Z, P, Vt, Pt, cols_per_thread, threads):
    """The oracle (C restatement of the reference's column loop) on `threads` host cores, each on
    its own range of basis columns; scaled from the sample to one full eval over 2^L columns.
    U is applied matrix-free as the same Trotter circuit the GPU target is built from."""
    N = 1 << L
    jobs = [(L, V, Z, P, Vt, Pt, t * cols_per_thread, (t + 1) * cols_per_thread) for t in range(threads)]
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(threads) as pool:
        per = pool.map(_cpu_worker, jobs)
        wall = time.perf_counter() - t0
        # the oracle's gate kernels timed against the reference's own (oracle/_ref), one thread
        try:
            calib = pool.apply(_calib_worker, (L,))
        except Exception as e:  # the reference subset is built only where the reference is present
            calib = {"skipped": str(e)}
    sec_per_col = float(np.mean(per)) / cols_per_thread
    eval_s = sec_per_col * N / threads
    return {"value": 1.0 / eval_s, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"{cols_per_thread} of {N} basis columns per thread on {threads} threads "
                      f"({wall:.1f} s wall), scaled to a full grad+HVP eval",
            "single_thread_eval_s": sec_per_col * N, "gate_kernels_vs_reference": calib}


def _calib_worker(L):
    from oracle.calibrate import calibrate
    return calibrate(L, reps=5)


def run(args):
    import rqcopt_hpc_amd as R
    from rqcopt_hpc_amd.workloads import config_inputs, ising_trotter_brickwall

    info = R.require_gpu()
    engine = args.engine
    R.set_sweep_engine(R.ENGINE_TILE_PASS if engine == "tile_pass" else R.ENGINE_PER_GATE)
    L, d = args.L, args.depth
    N = 1 << L
    G = d * (L // 2)
    V, P, vec, Z = config_inputs(L, d, args.seed)
    Vt, Pt = ising_trotter_brickwall(L, 1.0, 0.75, 0.6, 0.5, 4)
    T = R.DeviceTarget.brickwall(Vt, Pt)
    # tile shape and gates per pass timed on the first (warm-up) evaluation of each precision
    R.set_eval_autotune(not args.no_autotune)

    def one():
        return R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)

    copy_gbs = R.measure_copy_bandwidth(1 << 30, 5)
    stream_gbs = R.measure_stream_bandwidth(1 << 30, 5)
    print(f"[bench] device copy bandwidth {copy_gbs:.0f} GB/s (1 GiB plain streaming copy, read + write), "
          f"{stream_gbs:.0f} GB/s (copy tuned for HBM: the streaming ceiling)", file=sys.stderr, flush=True)
    precs = [(R.PRECISION_FP64, "fp64"), (R.PRECISION_MIXED, "mixed")]
    if args.precision != "both":
        precs = [p for p in precs if p[1] == args.precision]
    results = {}
    for prec, name in precs:
        R.set_precision(prec)
        # at least one untimed evaluation when the tiles are tuned on the first one
        for _ in range(max(args.warmup, 0 if args.no_autotune else 1)):
            out = one()
        R.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = one()
        R.synchronize()
        dt = time.perf_counter() - t0
        # kernel timing pass (HIP events on the library's stream)
        R.profile_reset()
        if not args.no_profile_pass:
            R.profile_enable(True)
            one()
            R.profile_enable(False)
        prof = {k: R.profile_query(k) for k in range(4)}
        flops = {k: R.profile_query_flops(k) for k in range(4)}
        results[name] = {"eval_s": dt / args.steps, "out": out, "prof": prof, "flops": flops,
                         "block": R.last_column_block(),
                         "tiles": R.last_tiles()}
        print(f"[bench] {name}: {dt / args.steps * 1e3:.2f} ms/eval, column block {R.last_column_block()}, "
              f"tiles (log2 rows, log2 cols, gates per pass) {R.last_tiles()}, "
              f"kernel ms {[round(prof[k][0], 2) for k in range(4)]}, launches {[prof[k][1] for k in range(4)]}, "
              f"GB/s {[round(prof[k][2] / max(prof[k][0], 1e-9) / 1e6) for k in range(4)]}",
              file=sys.stderr, flush=True)
    # the gradient eval (projected gradient only) of the same workload, reported on its own
    grad_eval = {}
    if args.precision == "both" and not args.no_profile_pass:
        for prec, name in precs:
            R.set_precision(prec)
            R.brickwall_unitary_target_and_projected_gradient(T, V, P)
            R.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                R.brickwall_unitary_target_and_projected_gradient(T, V, P)
            R.synchronize()
            dtg = (time.perf_counter() - t0) / args.steps
            grad_eval[name] = {"evals_per_s": 1.0 / dtg, "ms_per_eval": dtg * 1e3}
        print(f"[bench] grad eval: {grad_eval}", file=sys.stderr, flush=True)
    R.set_precision(R.PRECISION_FP64)
    if args.precision != "both" or args.no_profile_pass:
        print(json.dumps({k: {"ms_per_eval": v["eval_s"] * 1e3, "column_block": v["block"]} for k, v in results.items()}))
        return

    r64, r32 = results["fp64"], results["mixed"]
    s = 16
    B = r64["block"]
    # dominant kernel: the reverse sweep, compute-bound on the FP64 matrix cores.  achieved =
    # algorithmic flops of all reverse launches over their HIP-event durations; its HBM rate (state
    # arrays read + written once, summed by the library per launch) is reported beside it
    rev_ms, rev_n, rev_bytes = r64["prof"][R.KERNEL_REVERSE]
    engine_used, passes = R.last_sweep()
    # the library's count per launch: 192 flops (issued on the matrix cores: 240, recovery
    # included, gate applications in the three-product form) per amplitude and gate of the tiles
    # a launch processes (the first pass skips the tiles without an identity element, whose
    # terms vanish), plus the forward pass the last launch carries (96; issued 72)
    rev_flops, rev_issued = r64["flops"][R.KERNEL_REVERSE] if engine == "tile_pass" else (None, None)
    achieved_tf = rev_flops / (rev_ms * 1e-3) / 1e12 if rev_flops else None
    issued_tf = rev_issued / (rev_ms * 1e-3) / 1e12 if rev_flops else None
    hbm_gbs = rev_bytes / (rev_ms * 1e-3) / 1e9
    traffic = None
    if engine == "tile_pass" and os.path.exists(TRAFFIC_FILE) and (L, d) == (14, 9):
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh)["k_tile_reverse_fp64"]["hbm_bytes_per_launch"]
    f64, g64, h64 = r64["out"]
    f32, g32, h32 = r32["out"]
    mixed_err = {"f_abs": float(abs(f32 - f64)), "f_abs_over_N": float(abs(f32 - f64) / N),
                 "f_rel": float(abs(f32 - f64) / max(abs(f64), 1e-300)),
                 "grad_rel": float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))),
                 "hvp_rel": float(np.max(np.abs(h32 - h64)) / np.max(np.abs(h64)))}
    value = 1.0 / r64["eval_s"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["eval_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG Haar gates, Ising-Trotter target formed on the GPU)",
        "copy_bandwidth_gbs": copy_gbs,
        "stream_bandwidth_gbs": stream_gbs,
        "config": {"workload": f"brickwall grad+HVP eval, L={L}, depth {d}, periodic, Ising-Trotter U (t=0.5, 4 Strang steps)",
                   "L": L, "depth": d, "gates": G, "column_block": B, "sweep_engine": engine, "device": info,
                   "tiles": {"fp64": r64["tiles"], "mixed": r32["tiles"],
                             "autotuned": not args.no_autotune,
                             "fields": "log2 tile rows (fused qubits K), log2 tile columns, most gates per pass"}},
        "roofline": ({"bound": "mfma", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": traffic,
                      "alg_flops_per_launch": rev_flops / max(rev_n, 1),
                      "alg_flops_full_sweep": reverse_alg_flops(L, d),
                      "issued": {"flops_per_launch": rev_issued / max(rev_n, 1),
                                 "tflops": issued_tf, "frac": issued_tf / FP64_PEAK_TFLOPS}}
                     if achieved_tf else
                     {"bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": hbm_gbs / HBM_PEAK_GBS, "traffic": None}) | {
                     "kernel": REVERSE_KERNEL[engine], "launches": rev_n, "avg_launch_ms": rev_ms / max(rev_n, 1),
                     "alg_bytes_per_launch": rev_bytes / max(rev_n, 1), "hbm_gbs": hbm_gbs,
                     "hbm_frac": hbm_gbs / HBM_PEAK_GBS, "passes_per_sweep": passes},
        "eval_canonical": {"bytes": canonical_bytes(L, d, s), "flops": canonical_flops(L, d),
                           "hbm_frac": canonical_bytes(L, d, s) / r64["eval_s"] / 1e9 / HBM_PEAK_GBS,
                           "fp64_frac": canonical_flops(L, d) / r64["eval_s"] / 1e12 / FP64_PEAK_TFLOPS},
        "mixed": {"value": 1.0 / r32["eval_s"], "ms_per_step": r32["eval_s"] * 1e3, "speedup_vs_fp64": r64["eval_s"] / r32["eval_s"],
                  "error_vs_fp64": mixed_err, "column_block": r32["block"]},
    }
    line["grad_eval"] = grad_eval
    line["mixed_error_seeds"] = mixed_error_sweep(args, R, T, P) if args.mixed_seeds > 0 else None
    R.release_workspace()
    line["batched"] = None if args.no_batch else run_batched(args, R)
    line["serving"] = None if args.no_serve else run_serving(args, R)
    if line["serving"] is not None:
        line["served_p99_ms"] = {k: line["serving"][k]["offered"]["p99_ms"] for k in ("fp64", "mixed")}
    if not args.no_cpu:
        threads = min(args.cpu_threads, len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = cpu_baseline(L, V, Z, P, Vt, Pt, args.cpu_cols, threads)
        line["cpu_baseline"]["host_cpus"] = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
                                             "used": threads}
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


GOLDEN = os.path.join(ROOT, "tests", "golden", "fixtures", "config12_ising_L14_d9.npz")


def _errors(f, g, h, f0, g0, h0):
    return (abs(f - f0), abs(f - f0) / max(1.0, abs(f0)), np.max(np.abs(g - g0)), np.max(np.abs(g - g0)) / np.max(np.abs(g0)),
            np.max(np.abs(h - h0)), np.max(np.abs(h - h0)) / np.max(np.abs(h0)))


def mixed_error_sweep(args, R, T, P):
    """Errors of f, grad and HVP for seeded gate sets against the oracle's FP64 values at L=14, depth 9
    (golden vectors written by tests/golden/make_config12.py), for the FP64 and the mixed path; at
    other sizes, or for seeds without a golden, against the FP64 path.  Relative errors are
    |df| / max(1, |f|) and max-entry error over max entry."""
    from rqcopt_hpc_amd.workloads import config_inputs

    gold = {}
    if (args.L, args.depth) == (14, 9) and os.path.exists(GOLDEN):
        with np.load(GOLDEN, allow_pickle=False) as z:
            gold = {int(s): (z["f"][k], z["grad"][k], z["hvp"][k]) for k, s in enumerate(z["seeds"])}
    rows32, rows64, vs = [], [], []
    for seed in range(1, args.mixed_seeds + 1):
        V, _, _, Z = config_inputs(args.L, args.depth, seed)
        R.set_precision(R.PRECISION_FP64)
        r64 = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
        R.set_precision(R.PRECISION_MIXED)
        r32 = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
        ref = gold.get(seed)
        vs.append("oracle" if ref is not None else "fp64_gpu")
        if ref is None:
            ref = r64
        else:
            rows64.append(_errors(*r64, *ref))
        rows32.append(_errors(*r32, *ref))
    R.set_precision(R.PRECISION_FP64)
    names = ["f_abs", "f_rel", "grad_abs", "grad_rel", "hvp_abs", "hvp_rel"]

    def summary(rows):
        a = np.array(rows)
        return {"max": dict(zip(names, map(float, a.max(0)))), "median": dict(zip(names, map(float, np.median(a, 0))))}

    out = {"seeds": args.mixed_seeds, "reference": {"oracle": vs.count("oracle"), "fp64_gpu": vs.count("fp64_gpu")},
           "mixed": summary(rows32), "fp64": summary(rows64) if rows64 else None}
    print(f"[bench] errors over {args.mixed_seeds} seeds: {out}", file=sys.stderr, flush=True)
    return out


def run_batched(args, R):
    """64 independent problems: all-HVP steps (evals/s) and a lock-step trust-region run."""
    from rqcopt_hpc_amd import workloads as W

    probs = W.batch_config(args.seed, args.batch_problems)
    targets = [R.DeviceTarget.circuit(p["target_gates"], p["target_wires"], p["L"]) for p in probs]
    V = [p["V0"] for p in probs]
    Z = [W.tangent_direction(1000 + i, V[i])[1] for i in range(len(probs))]
    kinds = [R.EVAL_HVP] * len(probs)
    out = {"problems": len(probs), "qubits": sorted({p["L"] for p in probs}),
           "depths": sorted({p["depth"] for p in probs}), "steps": args.steps}
    for prec, name in ((R.PRECISION_FP64, "fp64"), (R.PRECISION_MIXED, "mixed")):
        B = R.Batch([(p["perms"], t) for p, t in zip(probs, targets)], precision=prec)
        for _ in range(max(1, args.warmup)):
            B.evaluate(kinds, V, Z)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            B.evaluate(kinds, V, Z)
        dt = (time.perf_counter() - t0) / args.steps
        rec = {"evals_per_s": len(probs) / dt, "ms_per_step": dt * 1e3,
               "workspace_gib": B.stats()["workspace_bytes"] / 2**30,
               "tiles": {f"L{p['L']}_d{p['depth']}": B.tiles(i) for i, p in enumerate(probs)}}
        if args.batch_niter > 0:
            st0 = B.stats()
            t0 = time.perf_counter()
            f_iters, _ = B.optimize(V, args.batch_niter)
            tt = time.perf_counter() - t0
            st1 = B.stats()
            rec["trust_region"] = {"niter": args.batch_niter, "seconds": tt,
                                   "steps": st1["steps"] - st0["steps"],
                                   "evaluations": st1["evaluations"] - st0["evaluations"],
                                   "graph_launches": st1["graph_launches"] - st0["graph_launches"],
                                   "evals_per_s": (st1["evaluations"] - st0["evaluations"]) / tt,
                                   "f_decreased": int(sum(f[-1] < f[0] for f in f_iters))}
        out[name] = rec
        print(f"[bench] batched {name}: {rec}", file=sys.stderr, flush=True)
        B.close()
    out["mixed_speedup"] = out["mixed"]["evals_per_s"] / out["fp64"]["evals_per_s"]
    for t in targets:
        t.close()
    return out


def run_serving(args, R):
    """Open-loop Poisson load on the request server, FP64 and mixed."""
    from rqcopt_hpc_amd import serving as S

    solves, targets = S.serving_workload(args.seed, nsolves=256)
    out = {"solves": 256, "offered_rate": args.serve_rate, "seconds": args.serve_seconds,
           "kinds": "target / gradient / HVP, equal weights", "qubits": "8..14", "depths": [3, 5, 7, 9]}
    for prec, name in ((R.PRECISION_FP64, "fp64"), (R.PRECISION_MIXED, "mixed")):
        srv = R.Server(nslots=8)
        sessions, dev = S.open_sessions(srv, solves, targets, prec)
        # plans made and every kernel loaded before the timed load
        S.run_poisson(srv, solves, sessions, rate=args.serve_rate, duration=0.25, seed=args.seed + 99)
        rec = {}
        for label, rate in (("offered", args.serve_rate), ("quarter", args.serve_rate / 4)):
            res = S.run_poisson(srv, solves, sessions, rate=rate, duration=args.serve_seconds, seed=args.seed)
            rec[label] = {k: res[k] for k in ("requests", "offered_rate", "achieved_rate", "p50_ms", "p99_ms",
                                              "mean_ms", "p50_by_kind")}
            print(f"[bench] serving {name} {label}: {rec[label]}", file=sys.stderr, flush=True)
        rec["peak_in_flight"] = srv.stats()["peak_in_flight"]
        rec["arena_gib"] = srv.stats()["arena_bytes"] / 2**30
        # real trust-region solves: every request comes from a solve's driver (closed loop, so the
        # achieved rate is what the device sustains at that many concurrent solves)
        if args.solve_seconds > 0:
            rec["solves"] = {}
            for conc in args.solve_concurrency:
                res = S.run_solves(srv, solves, sessions, conc, 1, args.solve_seconds, seed=args.seed)
                res.pop("server", None)
                rec["solves"][str(conc)] = res
                print(f"[bench] serving {name} solves x{conc}: {res}", file=sys.stderr, flush=True)
        for s_ in sessions:
            s_.close()
        srv.close()
        for t in dev.values():
            t.close()
        out[name] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=14)
    ap.add_argument("--depth", type=int, default=9)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (the GPU host allots 16 cores per GPU)")
    ap.add_argument("--cpu-cols", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", choices=["tile_pass", "per_gate"], default="tile_pass",
                    help="sweep engine: several gates per HBM pass through LDS, or one launch per gate")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="profiling runs: time one precision only and print a short summary")
    ap.add_argument("--no-profile-pass", action="store_true", help="skip the HIP-event kernel timing pass")
    ap.add_argument("--no-autotune", action="store_true",
                    help="default tiles instead of timing the tile candidates on the first evaluation")
    ap.add_argument("--mixed-seeds", type=int, default=20, help="seeded gate sets of the mixed-precision error sweep")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-problems section")
    ap.add_argument("--batch-problems", type=int, default=64)
    ap.add_argument("--batch-niter", type=int, default=2, help="lock-step trust-region iterations (0: skip)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving section")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="offered requests/s of the Poisson stream")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window of the Poisson stream")
    ap.add_argument("--solve-seconds", type=float, default=6.0,
                    help="window of the closed-loop trust-region solves through the server (0: skip)")
    ap.add_argument("--solve-concurrency", type=int, nargs="+", default=[4, 16],
                    help="concurrent solves of the closed-loop serving runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU column sharding is a "
              f"design note, not part of this build", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
