"""GPU tests of the tile shapes the tile-pass engine can use (the target functions of reference
src/target.c:15-678).

Every shape gives the oracle's FP64 target, gradient and HVP to 1e-12.  This covers the
forward pass's own tile, which the engine takes when the reverse pass's tile holds under
16 KiB per state: twice the columns ('c'), one more row bit ('k') or none ('0').  Mixed
precision stays within its stated bounds of FP64 on every shape.  Shapes are forced through
the library's environment overrides, which each evaluation reads when it is planned.
"""
import os

import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from helpers import brickwall_perms, crandn, haar_unitary, reldiff

pytestmark = pytest.mark.gpu

TOL = 1e-12
KEYS = ("RQC_TILE_LOGE_FP64", "RQC_TILE_LOGC_FP64", "RQC_TILE_LOGE_FP32", "RQC_TILE_LOGC_FP32", "RQC_FWD_TILE")


@pytest.fixture(autouse=True)
def clean_env():
    R.require_gpu()
    saved = {k: os.environ.get(k) for k in KEYS}
    yield
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    R.set_precision(R.PRECISION_FP64)


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(31)
    L, nl = 10, 5
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    Z = crandn(rng, (nl, 4, 4))
    P = brickwall_perms(L, nl)
    Uo = O.Ufunc.dense(U)
    ref = {"f": O.brickwall_unitary_target(Uo, V, P, L),
           "grad": O.brickwall_unitary_target_and_gradient(Uo, V, P, L),
           "hvp": O.brickwall_unitary_target_hvp(Uo, V, Z, P, L)}
    T = R.DeviceTarget.dense(U)
    yield L, V, Z, P, T, ref
    T.close()


FP64_SHAPES = {
    "default_64x16": {},
    "rev_64x8_fwd_64x16": {"RQC_TILE_LOGE_FP64": "9"},
    "rev_64x8_fwd_128x8": {"RQC_TILE_LOGE_FP64": "9", "RQC_FWD_TILE": "k"},
    "rev_64x8_fwd_same": {"RQC_TILE_LOGE_FP64": "9", "RQC_FWD_TILE": "0"},
    "128x8": {"RQC_TILE_LOGE_FP64": "10", "RQC_TILE_LOGC_FP64": "3"},
}


@pytest.mark.parametrize("name", list(FP64_SHAPES))
def test_fp64_shapes_match_oracle(problem, name):
    L, V, Z, P, T, ref = problem
    os.environ.update(FP64_SHAPES[name])
    print("tile shape", name, flush=True)
    f = R.brickwall_unitary_target(T, V, P)
    assert R.last_sweep()[0] == R.ENGINE_TILE_PASS
    assert abs(f - ref["f"]) < TOL * max(1, abs(ref["f"]))
    f, dv = R.brickwall_unitary_target_and_gradient(T, V, P)
    assert abs(f - ref["grad"][0]) < TOL * max(1, abs(ref["f"])) and reldiff(dv, ref["grad"][1]) < TOL
    f, dv, hz = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
    f_o, dv_o, hz_o = ref["hvp"]
    assert abs(f - f_o) < TOL * max(1, abs(f_o)) and reldiff(dv, dv_o) < TOL and reldiff(hz, hz_o) < TOL


FP32_SHAPES = {
    "default_64x32": {},
    "rev_64x16_fwd_64x32": {"RQC_TILE_LOGE_FP32": "10"},
    "rev_64x16_fwd_128x16": {"RQC_TILE_LOGE_FP32": "10", "RQC_FWD_TILE": "k"},
    "rev_64x16_fwd_same": {"RQC_TILE_LOGE_FP32": "10", "RQC_FWD_TILE": "0"},
}


@pytest.mark.parametrize("name", list(FP32_SHAPES))
def test_mixed_shapes_within_bounds(problem, name):
    L, V, Z, P, T, ref = problem
    f64, dv64, hz64 = ref["hvp"]
    os.environ.update(FP32_SHAPES[name])
    print("tile shape", name, flush=True)
    R.set_precision(R.PRECISION_MIXED)
    f, dv, hz = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
    assert R.last_sweep()[0] == R.ENGINE_TILE_PASS
    # stated bounds: |df| / N <= 1e-6, gradient and HVP within 1e-4 relative
    assert abs(f - f64) / (1 << L) < 1e-6
    assert reldiff(dv, dv64) < 1e-4 and reldiff(hz, hz64) < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_eval_autotune_keeps_parity(problem, prec):
    """Single-evaluation tile autotuning (rqc_set_eval_autotune): the first HVP evaluation of the
    structure times the tile candidates and keeps the fastest; the tuned evaluation gives the
    oracle's FP64 results to 1e-12 (mixed: its stated bounds), and the choice is one of the
    candidates (2^10 / 2^9 FP64 or 2^11 / 2^10 FP32 elements per tile, 8 or 6 gates per pass) and
    stays fixed for later evaluations of the structure."""
    L, V, Z, P, T, ref = problem
    f_o, dv_o, hz_o = ref["hvp"]
    R.set_precision(R.PRECISION_FP64 if prec == "fp64" else R.PRECISION_MIXED)
    R.set_eval_autotune(True, forget=True)
    try:
        f, dv, hz = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
        K, logC, gmax = R.last_tiles()
        f2, dv2, hz2 = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
        assert R.last_tiles() == (K, logC, gmax)
    finally:
        R.set_eval_autotune(False)
    print(f"{prec}: tuned tile 2^{K} x 2^{logC}, at most {gmax} gates per pass", flush=True)
    base = 10 if prec == "fp64" else 11
    assert K + logC in (base, base - 1) and gmax in (8, 6)
    assert f == f2 and np.array_equal(dv, dv2) and np.array_equal(hz, hz2)
    if prec == "fp64":
        assert abs(f - f_o) < TOL * max(1, abs(f_o)) and reldiff(dv, dv_o) < TOL and reldiff(hz, hz_o) < TOL
    else:
        assert abs(f - f_o) / (1 << L) < 1e-6 and reldiff(dv, dv_o) < 1e-4 and reldiff(hz, hz_o) < 1e-4
