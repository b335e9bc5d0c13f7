"""GPU tests of batched evaluation and lock-step optimisation (include/rqcopt/rqc_batch.h).

Every problem of a batch must give the result it gives alone: the batch plans
each problem as the single-problem path does (same column block, grids and
fixed-order reductions), so batch and single results agree to 1e-13 relative
and FP64 results match the oracle to 1e-12 relative.  Graph replays are
checked against eager launches on every replay.
"""
import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from rqcopt_hpc_amd import workloads as W
from helpers import brickwall_perms, haar_unitary, reldiff

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def gpu():
    R.require_gpu()
    R.set_precision(R.PRECISION_FP64)
    yield
    R.set_precision(R.PRECISION_FP64)


def _problems(seed, shapes):
    """XXZ Trotter targets (seeded fields and times) and Haar gates for (L, depth) pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for L, d in shapes:
        fields = rng.uniform(-1, 1, L)
        t = rng.uniform(0.1, 1.0)
        tg, tw = W.xxz_trotter_circuit(L, 0.5, fields, t, 2)
        T = R.DeviceTarget.circuit(tg, tw, L)
        V = np.stack([haar_unitary(rng) for _ in range(d)])
        vec = rng.standard_normal(16 * d)
        Z = np.stack([R.real_to_tangent(vec[16 * i:16 * i + 16], V[i]) for i in range(d)])
        out.append(dict(L=L, d=d, P=brickwall_perms(L, d), T=T, V=V, Z=Z, vec=vec))
    return out


SHAPES = [(6, 3), (8, 5), (10, 3), (8, 2), (6, 4), (10, 5)]
KINDS = [R.EVAL_HVP, R.EVAL_GRADIENT, R.EVAL_HVP, R.EVAL_TARGET, R.EVAL_NONE, R.EVAL_HVP]


def _single(pr, kind):
    if kind == R.EVAL_TARGET:
        return R.brickwall_unitary_target(pr["T"], pr["V"], pr["P"]), None, None
    if kind == R.EVAL_GRADIENT:
        f, g = R.brickwall_unitary_target_and_projected_gradient(pr["T"], pr["V"], pr["P"])
        return f, g, None
    return R.brickwall_unitary_target_projected_hessian_vector_product(pr["T"], pr["V"], pr["Z"], pr["P"])


def test_batch_matches_single_calls_and_oracle():
    probs = _problems(1, SHAPES)
    print("batch shapes", SHAPES, "kinds", KINDS, flush=True)
    B = R.Batch([(p["P"], p["T"]) for p in probs])
    f, g, h = B.evaluate(KINDS, [p["V"] for p in probs], [p["Z"] for p in probs])
    for i, (pr, k) in enumerate(zip(probs, KINDS)):
        if k == R.EVAL_NONE:
            assert g[i] is None and h[i] is None
            continue
        assert B.plan(i, k)[1] == R.ENGINE_TILE_PASS
        fs, gs, hs = _single(pr, k)
        assert abs(f[i] - fs) <= 1e-13 * max(1, abs(fs)), (i, f[i], fs)
        if k >= R.EVAL_GRADIENT:
            assert reldiff(g[i], gs) <= 1e-13
        if k == R.EVAL_HVP:
            assert reldiff(h[i], hs) <= 1e-13
        if pr["L"] <= 8:
            Uo = O.Ufunc.dense(pr["T"].download())
            f_o, g_o, h_o = O.brickwall_unitary_target_projected_hvp(Uo, pr["V"], pr["Z"], pr["P"], pr["L"])
            assert abs(f[i] - f_o) < TOL * max(1, abs(f_o))
            if k >= R.EVAL_GRADIENT:
                assert reldiff(g[i], g_o) < TOL
            if k == R.EVAL_HVP:
                assert reldiff(h[i], h_o) < TOL
    st = B.stats()
    assert st["steps"] == 1 and st["evaluations"] == 5 and st["graph_launches"] == 3
    B.close()


def test_batch_graph_replays_match_eager():
    probs = _problems(2, SHAPES)
    B = R.Batch([(p["P"], p["T"]) for p in probs], nstreams=3)
    V = [p["V"] for p in probs]
    Z = [p["Z"] for p in probs]
    kinds = [R.EVAL_HVP] * len(probs)
    B.set_exec(R.BATCH_EAGER)
    ref = B.evaluate(kinds, V, Z)
    B.set_exec(R.BATCH_GRAPH)
    for rep in range(4):
        # eager work of the program between the replays
        R.brickwall_unitary_target(probs[0]["T"], probs[0]["V"], probs[0]["P"])
        out = B.evaluate(kinds, V, Z)
        print("replay", rep, flush=True)
        assert np.array_equal(out[0], ref[0])
        for i in range(len(probs)):
            assert np.array_equal(out[1][i], ref[1][i]) and np.array_equal(out[2][i], ref[2][i])
    # problems that sit out a step leave the others unchanged
    sub = [R.EVAL_HVP if i % 2 == 0 else R.EVAL_NONE for i in range(len(probs))]
    out = B.evaluate(sub, V, Z)
    for i in range(0, len(probs), 2):
        assert out[0][i] == ref[0][i] and np.array_equal(out[2][i], ref[2][i])
    B.close()


def test_batch_mixed_precision_error_bounds():
    probs = _problems(3, [(8, 3), (10, 5), (12, 3)])
    pr = [(p["P"], p["T"]) for p in probs]
    V = [p["V"] for p in probs]
    Z = [p["Z"] for p in probs]
    kinds = [R.EVAL_HVP] * len(probs)
    B64 = R.Batch(pr)
    B32 = R.Batch(pr, precision=R.PRECISION_MIXED)
    f64, g64, h64 = B64.evaluate(kinds, V, Z)
    f32, g32, h32 = B32.evaluate(kinds, V, Z)
    for i, p in enumerate(probs):
        N = 1 << p["L"]
        # stated bounds: |df| / N <= 1e-6, gradient and HVP within 1e-4 relative (FP32 sweeps, FP64 sums)
        assert abs(f32[i] - f64[i]) / N < 1e-6
        assert reldiff(g32[i], g64[i]) < 1e-4 and reldiff(h32[i], h64[i]) < 1e-4
        assert B32.plan(i, R.EVAL_HVP)[1] == R.ENGINE_TILE_PASS
    B64.close()
    B32.close()


def test_batch_lockstep_optimisation_matches_single_and_oracle():
    probs = _problems(4, [(6, 3), (8, 3), (6, 2), (8, 4)])
    niter = 4
    B = R.Batch([(p["P"], p["T"]) for p in probs])
    f_iters, vopt = B.optimize([p["V"] for p in probs], niter)
    for i, p in enumerate(probs):
        print("problem", i, "L", p["L"], "d", p["d"], flush=True)
        fs, vs = R.optimize_brickwall_circuit_hvp(p["T"], p["V"], p["P"], niter)
        assert np.max(np.abs(f_iters[i] - fs) / np.maximum(1, np.abs(fs))) < 1e-12
        assert np.max(np.abs(vopt[i] - vs)) < 1e-10
        f_o, v_o = O.optimize_brickwall_circuit_hvp(O.Ufunc.dense(p["T"].download()), p["V"], p["P"], p["L"], niter)
        assert np.max(np.abs(f_iters[i] - f_o) / np.maximum(1, np.abs(f_o))) < 1e-9
        assert np.max(np.abs(vopt[i] - v_o)) < 1e-8
        assert f_iters[i][-1] < f_iters[i][0]
    st = B.stats()
    assert st["steps"] > 0 and st["evaluations"] >= st["steps"]
    B.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_device_trust_region_matches_host_driver(prec):
    """The GPU-resident trust region (tCG recurrences, tangent directions, polar retraction and the
    acceptance test in FP64 device kernels between graph launches) follows the host driver's
    trajectory (one reference-style driver thread per problem over the same batch evaluations):
    f_iter within 1e-12 relative and the gates within 1e-10 in FP64; in mixed precision the same
    evaluations feed both drivers, so the same bounds hold.  The device driver needs no host
    round trip per tCG step: its host syncs are one per chunk of 8 tCG steps."""
    probs = _problems(7, [(6, 3), (8, 4), (10, 3), (6, 2), (8, 2)])
    niter = 5
    precision = R.PRECISION_FP64 if prec == "fp64" else R.PRECISION_MIXED
    Bd = R.Batch([(p["P"], p["T"]) for p in probs], precision=precision)
    Bh = R.Batch([(p["P"], p["T"]) for p in probs], precision=precision)
    Bh.set_rtr(R.BATCH_RTR_HOST)
    V = [p["V"] for p in probs]
    st0 = Bd.stats()
    fd, vd = Bd.optimize(V, niter)
    st1 = Bd.stats()
    fh, vh = Bh.optimize(V, niter)
    launches = st1["graph_launches"] - st0["graph_launches"]
    print(f"{prec}: device driver {launches} graph launches over {niter} iterations "
          f"({launches / niter:.1f} per iteration), {st1['evaluations'] - st0['evaluations']} evaluations", flush=True)
    for i, p in enumerate(probs):
        ef = np.max(np.abs(fd[i] - fh[i]) / np.maximum(1, np.abs(fh[i])))
        ev = np.max(np.abs(vd[i] - vh[i]))
        print(f"problem {i} L={p['L']} d={p['d']}: f_iter {fd[i]} | max rel diff {ef:.2e}, gates {ev:.2e}", flush=True)
        assert ef < 1e-12 and ev < 1e-10
        assert fd[i][-1] < fd[i][0]
    Bd.close()
    Bh.close()


def test_batch_rejects_invalid_problems():
    probs = _problems(5, [(6, 2)])
    with pytest.raises(R.RqcError):
        R.Batch([(brickwall_perms(4, 2), R.DeviceTarget.dense(np.eye(16)))])
    B = R.Batch([(probs[0]["P"], probs[0]["T"])])
    with pytest.raises(R.RqcError):
        B.evaluate([R.EVAL_HVP], [probs[0]["V"]], [None])
    # non-unitary gates are refused by gradient / HVP steps (the tile-pass reverse sweep inverts the gates)
    with pytest.raises(R.RqcError):
        B.evaluate([R.EVAL_GRADIENT], [2 * probs[0]["V"]])
    B.close()


def test_batch_autotuned_tiles_keep_parity():
    """Autotuned tile shapes and pass sizes change only the rounding: FP64 results stay within 1e-12
    of the oracle and of the default-tile batch."""
    probs = _problems(6, [(8, 3), (10, 5), (6, 2)])
    kinds = [R.EVAL_HVP] * len(probs)
    V = [p["V"] for p in probs]
    Z = [p["Z"] for p in probs]
    R.set_autotune(False)
    try:
        Bd = R.Batch([(p["P"], p["T"]) for p in probs])
    finally:
        R.set_autotune(True)
    Ba = R.Batch([(p["P"], p["T"]) for p in probs])
    fd, gd, hd = Bd.evaluate(kinds, V, Z)
    fa, ga, ha = Ba.evaluate(kinds, V, Z)
    for i, p in enumerate(probs):
        K, logC, gmax = Ba.tiles(i)
        print("problem", i, "L", p["L"], "tiles", (K, logC, gmax), "default", Bd.tiles(i), flush=True)
        assert K + logC in (9, 10) and 1 <= gmax <= 8
        assert abs(fa[i] - fd[i]) < TOL * max(1, abs(fd[i]))
        assert reldiff(ga[i], gd[i]) < TOL and reldiff(ha[i], hd[i]) < TOL
        if p["L"] <= 8:
            Uo = O.Ufunc.dense(p["T"].download())
            f_o, g_o, h_o = O.brickwall_unitary_target_projected_hvp(Uo, p["V"], p["Z"], p["P"], p["L"])
            assert abs(fa[i] - f_o) < TOL * max(1, abs(f_o)) and reldiff(ga[i], g_o) < TOL and reldiff(ha[i], h_o) < TOL
    Bd.close()
    Ba.close()


def test_batch_workload_64_problems():
    """The batched workload at its size: 64 problems, L in {8, 10, 12}, depth in {3, 5, 7}, XXZ
    Trotter targets, autotuned tiles.  A sample of problems (every shape present, including the
    largest) equals the single-call path within 1e-12 (the autotuned tile may differ from the single
    call's default one, which changes only the rounding); the L=8 problems match the oracle at 1e-12;
    two lock-step trust-region iterations decrease f for every problem."""
    probs = W.batch_config(1, 64)
    targets = [R.DeviceTarget.circuit(p["target_gates"], p["target_wires"], p["L"]) for p in probs]
    V = [p["V0"] for p in probs]
    Z = [W.tangent_direction(1000 + i, V[i])[1] for i in range(len(probs))]
    shapes = sorted({(p["L"], p["depth"]) for p in probs})
    print("shapes", shapes, flush=True)
    B = R.Batch([(p["perms"], t) for p, t in zip(probs, targets)])
    f, g, h = B.evaluate([R.EVAL_HVP] * len(probs), V, Z)
    sample = [next(i for i, p in enumerate(probs) if (p["L"], p["depth"]) == s) for s in shapes]
    for i in sample:
        p = probs[i]
        fs, gs, hs = R.brickwall_unitary_target_projected_hessian_vector_product(targets[i], V[i], Z[i], p["perms"])
        print(f"problem {i} L={p['L']} d={p['depth']} tiles {B.tiles(i)}", flush=True)
        assert abs(f[i] - fs) < TOL * max(1, abs(fs)) and reldiff(g[i], gs) < TOL and reldiff(h[i], hs) < TOL
    for i, p in enumerate(probs):
        if p["L"] != 8:
            continue
        Uo = O.Ufunc.dense(targets[i].download())
        f_o, g_o, h_o = O.brickwall_unitary_target_projected_hvp(Uo, V[i], Z[i], p["perms"], 8)
        assert abs(f[i] - f_o) < TOL * max(1, abs(f_o)) and reldiff(g[i], g_o) < TOL and reldiff(h[i], h_o) < TOL
    f_iters, _ = B.optimize(V, 2)
    assert all(fi[-1] < fi[0] for fi in f_iters)
    B.close()
    for t in targets:
        t.close()
