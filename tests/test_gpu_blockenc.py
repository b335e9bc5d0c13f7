"""GPU tests of the block-encoding targets f = ||P^dagger W P - H||_F^2 / 2
(target.h brickwall_blockenc_*, reference src/target.c:1100-1720).

The sweeps run over the 2^(L/2) embedded basis columns; the residual kernel
replaces the trace, and the Hessian comes from HVP columns (hess1) and the
Gram matrix of the projected Jacobian columns (hess2).  FP64 results match the
reference's fixtures and the oracle to 1e-12 relative, on both sweep engines
(the per-gate engine for non-unitary gates, the tile-pass engine otherwise),
for H from a host callback and for a device-resident H.
"""
import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from helpers import cplx, crandn, fixture, haar_unitary, perms_of, random_perms, reldiff, scramble_matrix, stack

pytestmark = pytest.mark.gpu

TOL = 1e-12

SCR_H = R.HostUfunc(lambda x: scramble_matrix(int(np.log2(x.size))) @ x)
SCR_O = O.Ufunc.scramble()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    R.require_gpu()
    R.set_precision(R.PRECISION_FP64)
    yield
    R.set_precision(R.PRECISION_FP64)


def test_fixtures():
    """Every stored value of the reference's block-encoding tests (L = 8 and 6, random non-unitary
    gates, except the Hessian-matrix case)."""
    d = fixture("test_brickwall_blockenc_target_cplx")
    V, P = stack(d, "V", 5), perms_of(d, 5)
    for k, nl in enumerate((4, 5)):
        assert abs(R.brickwall_blockenc_target(SCR_H, V[:nl], P[:nl]) - float(d[f"f{k}"])) < TOL
    d = fixture("test_brickwall_blockenc_target_and_gradient_cplx")
    V, P = stack(d, "V", 5), perms_of(d, 5)
    for k, nl in enumerate((4, 5)):
        f, dv = R.brickwall_blockenc_target_and_gradient(SCR_H, V[:nl], P[:nl])
        assert abs(f - float(d[f"f{k}"])) < TOL and np.max(np.abs(dv - cplx(d[f"dVlist{k}"]))) < TOL
    d = fixture("test_brickwall_blockenc_target_and_gradient_vector_cplx")
    f, g = R.brickwall_blockenc_target_and_gradient_vector(SCR_H, stack(d, "V", 3), perms_of(d, 3))
    assert abs(f - float(d["f"])) < TOL and np.max(np.abs(g - d["grad"])) < TOL
    d = fixture("test_brickwall_blockenc_target_gradient_hessian_cplx")
    V, P = stack(d, "V", 5), perms_of(d, 5)
    for k, nl in enumerate((4, 5)):
        f, dv, h1, h2 = R.brickwall_blockenc_target_gradient_hessian(SCR_H, V[:nl], P[:nl])
        assert abs(f - float(d[f"f{k}"])) < TOL and np.max(np.abs(dv - cplx(d[f"dVlist{k}"]))) < TOL
        assert np.max(np.abs(h1 - cplx(d[f"hess1_{k}"]))) < TOL * max(1, np.max(np.abs(h1)))
        assert np.max(np.abs(h2 - cplx(d[f"hess2_{k}"]))) < TOL * max(1, np.max(np.abs(h2)))
        assert np.array_equal(h1, h1.T) and np.array_equal(h2, h2.conj().T)
    d = fixture("test_brickwall_blockenc_target_gradient_vector_hessian_matrix_cplx")
    f, g, H = R.brickwall_blockenc_target_gradient_vector_hessian_matrix(SCR_H, stack(d, "V", 5), perms_of(d, 5))
    assert abs(f - float(d["f"])) < TOL and np.max(np.abs(g - d["grad"])) < TOL
    assert np.max(np.abs(H - d["H"])) < TOL * np.max(np.abs(d["H"])) and np.array_equal(H, H.T)


@pytest.mark.parametrize("L,nl,unitary", [(2, 2, True), (4, 3, False), (6, 3, False), (8, 4, True), (10, 3, True),
                                         (10, 2, False)])
def test_vs_oracle_both_engines_and_device_h(L, nl, unitary):
    rng = np.random.default_rng(90 + L + nl)
    s = 1 << (L // 2)
    V = np.stack([haar_unitary(rng) if unitary else crandn(rng, (4, 4)) for _ in range(nl)])
    P = random_perms(rng, L, nl)
    H = crandn(rng, (s, s)) / np.sqrt(s)
    Ho = O.Ufunc.dense(H)
    f_o, dv_o, h1_o, h2_o = O.brickwall_blockenc_target_gradient_hessian(Ho, V, P, L)
    _, g_o, Hm_o = O.brickwall_blockenc_target_gradient_vector_hessian_matrix(Ho, V, P, L)
    hd = R.DeviceTarget.dense(H)
    hh = R.HostUfunc(lambda x: H @ x)
    print("blockenc L", L, "layers", nl, "unitary", unitary, flush=True)
    for hop in (hd, hh):
        f = R.brickwall_blockenc_target(hop, V, P)
        assert abs(f - f_o) < TOL * max(1, f_o)
        f, dv = R.brickwall_blockenc_target_and_gradient(hop, V, P)
        engine = R.last_sweep()[0]
        if not unitary:
            assert engine == R.ENGINE_PER_GATE
        elif L >= 8:
            assert engine == R.ENGINE_TILE_PASS
        assert abs(f - f_o) < TOL * max(1, f_o) and reldiff(dv, dv_o) < TOL
        f, g = R.brickwall_blockenc_target_and_gradient_vector(hop, V, P)
        assert reldiff(g, g_o) < TOL
        f, dv, h1, h2 = R.brickwall_blockenc_target_gradient_hessian(hop, V, P)
        assert abs(f - f_o) < TOL * max(1, f_o) and reldiff(dv, dv_o) < TOL
        assert reldiff(h1, h1_o) < TOL and reldiff(h2, h2_o) < TOL
        assert np.array_equal(h1, h1.T) and np.array_equal(h2, h2.conj().T)
        f, g, Hm = R.brickwall_blockenc_target_gradient_vector_hessian_matrix(hop, V, P)
        assert reldiff(g, g_o) < TOL and reldiff(Hm, Hm_o) < TOL and np.array_equal(Hm, Hm.T)
    hd.close()


def test_tile_engine_at_larger_size_and_mixed_precision():
    """L = 12 (64 embedded columns) on the tile-pass engine: the gradient against a finite difference
    of f, and mixed precision within its stated bounds of FP64."""
    rng = np.random.default_rng(95)
    L, nl = 12, 4
    s = 1 << (L // 2)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    P = random_perms(rng, L, nl)
    H = crandn(rng, (s, s)) / np.sqrt(s)
    hd = R.DeviceTarget.dense(H)
    f, dv = R.brickwall_blockenc_target_and_gradient(hd, V, P)
    assert R.last_sweep()[0] == R.ENGINE_TILE_PASS
    D = crandn(rng, (nl, 4, 4))
    eps = 1e-6
    num = (R.brickwall_blockenc_target(hd, V + eps * D, P) - R.brickwall_blockenc_target(hd, V - eps * D, P)) / (2 * eps)
    assert abs(num - np.real(np.sum(dv * D))) < 1e-6 * max(1, abs(num))
    f64, g64 = R.brickwall_blockenc_target_and_gradient_vector(hd, V, P)
    R.set_precision(R.PRECISION_MIXED)
    try:
        f32, g32 = R.brickwall_blockenc_target_and_gradient_vector(hd, V, P)
        _, _, h1, h2 = R.brickwall_blockenc_target_gradient_hessian(hd, V, P)
    finally:
        R.set_precision(R.PRECISION_FP64)
    _, _, h1d, h2d = R.brickwall_blockenc_target_gradient_hessian(hd, V, P)
    print("mixed vs fp64: f", abs(f32 - f64) / f64, "grad", reldiff(g32, g64), "hess1", reldiff(h1, h1d),
          "hess2", reldiff(h2, h2d), flush=True)
    assert abs(f32 - f64) < 1e-5 * f64 and reldiff(g32, g64) < 1e-4
    assert reldiff(h1, h1d) < 1e-4 and reldiff(h2, h2d) < 1e-4
    hd.close()


def test_argument_errors():
    rng = np.random.default_rng(96)
    V = np.stack([haar_unitary(rng) for _ in range(2)])
    with pytest.raises(R.RqcError):
        R.brickwall_blockenc_target(SCR_H, V, random_perms(rng, 7, 2))          # odd qubit count
    hd = R.DeviceTarget.dense(np.eye(16))
    with pytest.raises(R.RqcError):
        R.brickwall_blockenc_target(hd, V, random_perms(rng, 6, 2))             # H on 4 qubits, L/2 = 3
    hd.close()
    bad = R.HostUfunc(lambda x: (_ for _ in ()).throw(ValueError("boom")))
    with pytest.raises(R.RqcError):
        R.brickwall_blockenc_target(bad, V, random_perms(rng, 6, 2))


@pytest.mark.parametrize("L,nl,unitary", [(12, 4, True), (6, 3, False)])
def test_mixed_gate_rounding_correction(L, nl, unitary):
    """The FP32 sweeps apply fl(V); the evaluations with a gradient add Re sum dv (V - fl(V)) to
    the real block-encoding f.  Mixed runs on V and fl(V) share every FP32 operation, so their
    difference is that term alone, and the FP64 path gives its true value f(V) - f(fl(V))."""
    rng = np.random.default_rng(97 + L)
    s = 1 << (L // 2)
    V = np.stack([haar_unitary(rng) if unitary else crandn(rng, (4, 4)) for _ in range(nl)])
    V32 = (V.real.astype(np.float32) + 1j * V.imag.astype(np.float32)).astype(np.complex128)
    P = random_perms(rng, L, nl)
    hd = R.DeviceTarget.dense(crandn(rng, (s, s)) / np.sqrt(s))
    f64, _ = R.brickwall_blockenc_target_and_gradient(hd, V, P)
    f64r, _ = R.brickwall_blockenc_target_and_gradient(hd, V32, P)
    R.set_precision(R.PRECISION_MIXED)
    try:
        fc, dvc = R.brickwall_blockenc_target_and_gradient(hd, V, P)
        assert R.last_sweep()[0] == (R.ENGINE_TILE_PASS if unitary else R.ENGINE_PER_GATE)
        fu, _ = R.brickwall_blockenc_target_and_gradient(hd, V32, P)
    finally:
        R.set_precision(R.PRECISION_FP64)
    exact = f64 - f64r
    print(f"blockenc L={L}: gate-rounding term {abs(exact):.3e}, corrected error {abs(fc - f64):.3e}", flush=True)
    # f reaches 1e5 for the non-unitary gates: a difference of two f values carries ~eps |f|
    rnd = 1e-14 * max(1.0, abs(f64))
    assert abs(exact) > 1e3 * rnd
    assert abs((fc - fu) - exact) < 1e-3 * abs(exact) + rnd
    assert abs((fc - fu) - np.real(np.sum(dvc * (V - V32)))) < rnd
    hd.close()
