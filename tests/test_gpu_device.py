"""GPU tests of the device measurement helpers (include/rqcopt/rqc_device.h; no reference
counterpart): the plain and the tuned streaming copy report a bandwidth between zero and the
MI355X's 8 TB/s HBM peak, and reject invalid arguments with the reference's -3; the launch
profile's flop counts follow the tile passes' per-amplitude work.
"""
import numpy as np
import pytest

import rqcopt_hpc_amd as R

pytestmark = pytest.mark.gpu

HBM_PEAK_GBS = 8000.0


@pytest.fixture(autouse=True)
def gpu():
    R.require_gpu()
    yield


def test_copy_and_stream_bandwidth_within_peak():
    copy = R.measure_copy_bandwidth(1 << 28, 3)
    stream = R.measure_stream_bandwidth(1 << 28, 3)
    print(f"copy {copy:.0f} GB/s, tuned stream {stream:.0f} GB/s", flush=True)
    assert 0 < copy < 1.05 * HBM_PEAK_GBS
    assert 0 < stream < 1.05 * HBM_PEAK_GBS
    # sizes that are not a whole number of the tuned kernel's four-element chunks
    assert R.measure_stream_bandwidth((1 << 20) + 48, 1) > 0


def test_bandwidth_rejects_invalid_arguments():
    for fn in (R.measure_copy_bandwidth, R.measure_stream_bandwidth):
        with pytest.raises(R.RqcError):
            fn(1 << 20, 0)
        with pytest.raises(R.RqcError):
            fn(8, 1)


def test_profile_flops_of_an_hvp_eval():
    """One L=10 depth-5 HVP eval: the reverse launches count 192 flops per amplitude and gate of
    the tiles they process, the forward launches 96 (complex MACs at 8 flops).  Issued on the
    FP64 matrix cores: a dual gate application takes 9 block products per 16 quads (72 flops per
    amplitude), the outer products 12 (96), so a reverse amplitude-gate issues 240 with the
    recovery and a forward one 72.  The first reverse pass skips the tiles without an identity
    element and the last reverse launch carries the last forward pass, so the reverse counts lie
    between bounds fixed by the sweep."""
    from rqcopt_hpc_amd.workloads import config_inputs
    L, nl = 10, 5
    V, P, vec, Z = config_inputs(L, nl, seed=3)
    rng = np.random.default_rng(3)
    N = 1 << L
    U = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))[0]
    T = R.DeviceTarget.dense(U)
    R.set_precision(R.PRECISION_FP64)
    R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)   # plan and warm up
    R.profile_reset()
    R.profile_enable(True)
    try:
        R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
    finally:
        R.profile_enable(False)
    engine, passes = R.last_sweep()
    G = nl * (L // 2)
    amp_gates = float(G) * N * N
    rev_alg, rev_issued = R.profile_query_flops(R.KERNEL_REVERSE)
    fwd_alg, fwd_issued = R.profile_query_flops(R.KERNEL_FORWARD)
    print(f"passes {passes}: reverse {rev_alg:.4g} / {rev_issued:.4g}, forward {fwd_alg:.4g}", flush=True)
    assert abs(fwd_issued - 0.75 * fwd_alg) <= 1e-12 * fwd_alg
    assert 0 < fwd_alg <= 96 * amp_gates
    # reverse: alg = 192 r + 96 f, issued = 240 r + 72 f (r reverse, f fused forward amplitude-gates)
    r = (rev_issued - 0.75 * rev_alg) / 96.0
    f = (rev_alg - 192.0 * r) / 96.0
    # every reverse gate at least on its tiles with identity elements, at most on all of them
    assert amp_gates / 16 < r <= amp_gates * (1 + 1e-12)
    assert -1e-9 * amp_gates <= f <= amp_gates
