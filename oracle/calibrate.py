"""Speed of the oracle's gate kernels against the reference's own (test infrastructure only).

The CPU baseline of the benchmark times the oracle's per-column loop (a C restatement of the
reference's, reference src/target.c:197-335); its cost is the per-column gate applications and
gate gradients on 2^L-amplitude statevectors.  The reference's target functions need <cblas.h>,
which this image lacks, but its gate kernels (src/gate.c:10-188) build from their sources into
oracle/_ref/libref_subset.so.  This script times both implementations of those two kernels,
single-threaded on the same inputs, and prints the ratio reference / oracle (> 1: the oracle is
faster).

    python oracle/calibrate.py [L] [reps]
"""
import ctypes as ct
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402


class SV(ct.Structure):
    _fields_ = [("data", ct.c_void_p), ("nqubits", ct.c_int)]


def _best(fn, reps):
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        best = min(best, (time.perf_counter() - t0) / reps)
    return best


def calibrate(L=14, reps=20):
    ref_path = os.path.join(HERE, "_ref", "libref_subset.so")
    if not os.path.exists(ref_path):
        raise RuntimeError("oracle/_ref/libref_subset.so is not built (make -C oracle ref)")
    R = ct.CDLL(ref_path)
    lib = O.lib()
    rng = np.random.default_rng(5)
    G = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    psi = rng.standard_normal(1 << L) + 1j * rng.standard_normal(1 << L)
    dout = rng.standard_normal(1 << L) + 1j * rng.standard_normal(1 << L)
    out, dpsi = np.empty_like(psi), np.empty_like(psi)
    dG = np.empty((4, 4), np.complex128)
    p = lambda a: a.ctypes.data_as(ct.c_void_p)
    sv_in, sv_out, sv_d, sv_dp = (ct.byref(SV(a.ctypes.data, L)) for a in (psi, out, dout, dpsi))
    # the wire pairs of one brickwall layer and of its shifted neighbour
    pairs = [(2 * k, 2 * k + 1) for k in range(L // 2)] + [(2 * k + 1, (2 * k + 2) % L) for k in range(L // 2)]
    rows = {}
    for name, ref_fn, orc_fn in (
        ("apply_gate",
         lambda: [R.apply_gate(p(G), i, j, sv_in, sv_out) for i, j in pairs],
         lambda: [lib.orc_apply_gate(p(G), i, j, L, p(psi), p(out)) for i, j in pairs]),
        ("apply_gate_backward",
         lambda: [R.apply_gate_backward(p(G), i, j, sv_in, sv_d, sv_dp, p(dG)) for i, j in pairs],
         lambda: [lib.orc_apply_gate_backward(p(G), i, j, L, p(psi), p(dout), p(dpsi), p(dG)) for i, j in pairs]),
    ):
        tr, to = _best(ref_fn, reps), _best(orc_fn, reps)
        rows[name] = {"reference_us_per_gate": tr / len(pairs) * 1e6, "oracle_us_per_gate": to / len(pairs) * 1e6,
                      "reference_over_oracle": tr / to}
    return {"L": L, "threads": 1, "kernels": rows}


if __name__ == "__main__":
    L = int(sys.argv[1]) if len(sys.argv) > 1 else 14
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    res = calibrate(L, reps)
    for k, v in res["kernels"].items():
        print(f"L={L} {k}: reference {v['reference_us_per_gate']:.1f} us/gate, oracle {v['oracle_us_per_gate']:.1f} "
              f"us/gate, reference/oracle = {v['reference_over_oracle']:.2f}")
