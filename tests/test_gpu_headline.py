"""The benchmark's headline configuration against the oracle: L=14, periodic brickwall of depth 9,
Ising-Trotter target (J=1, h=0.75, g=0.6, t=0.5, 4 Strang steps), projected grad+HVP eval
(reference src/target.c:631-678, brickwall_unitary_target_projected_hessian_vector_product).

The golden f, grad[144] and hvp[144] were written by the oracle over all 2^14 basis columns
(tests/golden/make_config12.py); seed 1 is the FP64 configuration, seeds 1..20 the gate sets of the
mixed-precision study.  Bounds:
  FP64:  |df| / max(1, |f|) < 1e-12, max |dgrad| / max |grad| < 1e-12, the same for the HVP;
  mixed: |df| / max(1, |f|) <= 1e-6 (the stated target error), gradient and HVP within 2e-6 of
         their largest entry (complex FP32 sweeps, FP64 reductions; measured at most ~6e-7).
"""
import os

import numpy as np
import pytest

import rqcopt_hpc_amd as R
from helpers import FIXDIR

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(FIXDIR, "config12_ising_L14_d9.npz")


def _golden():
    if not os.path.exists(GOLDEN):
        pytest.skip("L=14 golden vectors not generated (tests/golden/make_config12.py)")
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _errs(out, f0, g0, h0):
    f, g, h = out
    return (abs(f - f0) / max(1.0, abs(f0)), np.max(np.abs(g - g0)) / np.max(np.abs(g0)),
            np.max(np.abs(h - h0)) / np.max(np.abs(h0)))


@pytest.fixture(scope="module")
def headline():
    z = _golden()
    L, depth = int(z["L"]), int(z["depth"])
    T = R.DeviceTarget.brickwall(z["Uv"], z["Up"])
    from rqcopt_hpc_amd.workloads import brickwall_perms
    yield z, T, brickwall_perms(L, depth)
    T.close()
    R.release_workspace()


def test_fp64_matches_oracle_at_L14_d9(headline):
    z, T, P = headline
    worst = np.zeros(3)
    for k, seed in enumerate(z["seeds"]):
        out = R.brickwall_unitary_target_projected_hessian_vector_product(T, z["V"][k], z["Z"][k], P)
        e = _errs(out, z["f"][k], z["grad"][k], z["hvp"][k])
        worst = np.maximum(worst, e)
        print(f"seed {seed}: FP64 rel errors f {e[0]:.2e} grad {e[1]:.2e} hvp {e[2]:.2e}", flush=True)
        assert e[0] < 1e-12 and e[1] < 1e-12 and e[2] < 1e-12
    print("FP64 worst over", len(z["seeds"]), "seeds:", worst, flush=True)


def test_mixed_precision_within_bounds_at_L14_d9(headline):
    z, T, P = headline
    worst = np.zeros(4)
    R.set_precision(R.PRECISION_MIXED)
    try:
        for k, seed in enumerate(z["seeds"]):
            out = R.brickwall_unitary_target_projected_hessian_vector_product(T, z["V"][k], z["Z"][k], P)
            e = _errs(out, z["f"][k], z["grad"][k], z["hvp"][k])
            # a target-only evaluation carries the same gate-rounding correction of f
            f_t = R.brickwall_unitary_target(T, z["V"][k], P)
            e_t = abs(f_t - z["f"][k]) / max(1.0, abs(z["f"][k]))
            worst = np.maximum(worst, e + (e_t,))
            print(f"seed {seed}: mixed rel errors f {e[0]:.2e} grad {e[1]:.2e} hvp {e[2]:.2e} "
                  f"target-only f {e_t:.2e}", flush=True)
        print("mixed worst over", len(z["seeds"]), "seeds:", worst, flush=True)
    finally:
        R.set_precision(R.PRECISION_FP64)
    assert worst[0] <= 1e-6 and worst[1] <= 2e-6 and worst[2] <= 2e-6 and worst[3] <= 1e-6
