"""Golden vectors of the L=14, depth-9 Ising-Trotter workload, written by the oracle.

Workload (the benchmark's headline configuration): L=14, periodic brickwall of
depth 9, Haar-random gates from the seeded PCG generator (seed s), HVP
direction vec ~ N(0,1) from seed s+1 mapped to the tangent space, and the target
U = the Strang-split Trotter circuit of the periodic Ising chain (J=1, h=0.75,
g=0.6, t=0.5, 4 steps = 9 exact bond layers), applied to each basis column
matrix-free by the oracle.

For each seed the oracle evaluates the reference's
brickwall_unitary_target_projected_hessian_vector_product (src/target.c:631-678):
the 2^14 basis columns are split into contiguous ranges over worker processes
(orc_brickwall_hvp_columns), the per-range sums of f, dV and H.Z are added in
range order, and the projection (orc_brickwall_project_grad_hvp) gives grad[144]
and hvp[144].  A single-thread evaluation takes about 35 min; the ranges make it
about 5 min on 8 cores.

Stored per seed: the gates V, the direction Z, f (complex), grad, hvp, and once
the Trotter layers (Uv, Up) that define U.  Seed 1 is the FP64 configuration;
seeds 1..20 are the gate sets of the mixed-precision error study.

    python tests/golden/make_config12.py [--seeds 1-20] [--workers 8]

Seeds already present in the output file are kept and skipped, so an
interrupted run resumes.
"""
import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from rqcopt_hpc_amd.workloads import config_inputs, ising_trotter_brickwall  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures", "config12_ising_L14_d9.npz")
L, DEPTH = 14, 9
ISING = dict(J=1.0, h=0.75, g=0.6, t=0.5, steps=4)


def _columns(args):
    Uv, Up, V, Z, P, b0, b1 = args
    f, dv, h = O.brickwall_hvp_columns(O.Ufunc.brickwall(Uv, Up), V, Z, P, L, b0, b1)
    return f, dv, h


def evaluate(Uv, Up, V, Z, P, pool, nchunks):
    n = 1 << L
    edges = np.linspace(0, n, nchunks + 1).astype(int)
    parts = pool.map(_columns, [(Uv, Up, V, Z, P, int(edges[k]), int(edges[k + 1])) for k in range(nchunks)])
    f = 0j
    dv = np.zeros((DEPTH, 4, 4), np.complex128)
    h = np.zeros((DEPTH, 4, 4), np.complex128)
    for pf, pdv, ph in parts:
        f += pf
        dv += pdv
        h += ph
    g, hv = O.brickwall_project_grad_hvp(V, Z, dv, h)
    return f, g, hv


def parse_seeds(s):
    a, _, b = s.partition("-")
    return list(range(int(a), int(b or a) + 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", default="1-20")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--chunks", type=int, default=64)
    args = ap.parse_args()

    Uv, Up = ising_trotter_brickwall(L, ISING["J"], ISING["h"], ISING["g"], ISING["t"], ISING["steps"])
    have = {}
    if os.path.exists(OUT):
        z = np.load(OUT)
        assert np.array_equal(z["Uv"], Uv) and np.array_equal(z["Up"], Up), "Trotter layers changed"
        for k, s in enumerate(z["seeds"]):
            have[int(s)] = {key: z[key][k] for key in ("V", "Z", "f", "grad", "hvp")}
    with mp.get_context("fork").Pool(args.workers) as pool:
        for s in parse_seeds(args.seeds):
            if s in have:
                continue
            V, P, vec, Z = config_inputs(L, DEPTH, seed=s)
            t0 = time.time()
            f, g, hv = evaluate(Uv, Up, V, Z, P, pool, args.chunks)
            have[s] = dict(V=V, Z=Z, f=f, grad=g, hvp=hv)
            print(f"seed {s}: f = {f:.15g}  |grad|inf = {np.max(np.abs(g)):.6g}  |hvp|inf = {np.max(np.abs(hv)):.6g}"
                  f"  ({time.time() - t0:.0f} s)", flush=True)
            seeds = sorted(have)
            np.savez_compressed(OUT + ".tmp.npz", seeds=np.array(seeds), Uv=Uv, Up=Up, L=L, depth=DEPTH,
                                ising=np.array([ISING[k] for k in ("J", "h", "g", "t", "steps")]),
                                **{key: np.stack([have[q][key] for q in seeds]) for key in ("V", "Z", "f", "grad",
                                                                                           "hvp")})
            os.replace(OUT + ".tmp.npz", OUT)


if __name__ == "__main__":
    main()
