"""GPU parity of librqcopt_hip against the oracle and the reference fixtures.

Every call goes through the C ABI (ctypes on librqcopt_hip.so).  Tolerances:
FP64 results match the oracle to 1e-12 relative (the required bar) unless a
tighter reference-test tolerance applies; the mixed-precision path is checked
against the FP64 path with the error bounds stated in each test.
"""
import ctypes as ct

import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from helpers import (brickwall_matrix, brickwall_perms, cplx, crandn, fixture, haar_unitary, perms_of, random_perms,
                     random_wires, reldiff, scramble_matrix, stack)

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def gpu():
    R.require_gpu()
    R.set_precision(R.PRECISION_FP64)
    yield
    R.set_precision(R.PRECISION_FP64)
    R.set_workspace_limit(0)


SCR_H = R.HostUfunc(lambda x: scramble_matrix(int(np.log2(x.size))) @ x)
SCR_O = O.Ufunc.scramble()


def _max(a):
    return float(np.max(np.abs(a)))


# ------------------------------------------------------------------ fixtures
def test_fixture_apply_gate_and_array():
    d = fixture("test_apply_gate_cplx")
    V, psi = cplx(d["V"]), cplx(d["psi"])
    for k, (i, j) in zip((1, 2, 3), ((2, 5), (4, 1), (3, 4))):
        assert _max(R.apply_gate(V, i, j, psi) - cplx(d[f"chi{k}"])) < 1e-12
    d = fixture("test_apply_gate_to_array_cplx")
    V, psi = cplx(d["V"]), cplx(d["psi"])
    for k, (i, j) in zip((1, 2, 3), ((2, 5), (4, 1), (3, 4))):
        assert _max(R.apply_gate_to_array(V, i, j, psi) - cplx(d[f"chi{k}"])) < 1e-12


def test_fixture_circuits():
    d = fixture("test_apply_quantum_circuit_cplx")
    G = stack(d, "G", 4)
    assert _max(R.apply_quantum_circuit(G, d["wires"], cplx(d["psi"])) - cplx(d["psi_out"])) < 1e-12
    d = fixture("test_quantum_circuit_backward_cplx")
    G = stack(d, "G", 5)
    out, cache = R.quantum_circuit_forward(G, d["wires"], cplx(d["psi"]))
    assert _max(out - cplx(d["psi_out"])) < 1e-12
    _, cache_o = O.quantum_circuit_forward(G, d["wires"], cplx(d["psi"]))
    assert _max(cache - cache_o) < 1e-13
    dpsi, dg = R.quantum_circuit_backward(G, d["wires"], cache, cplx(d["dpsi_out"]))
    dpsi_o, dg_o = O.quantum_circuit_backward(G, d["wires"], cache_o, cplx(d["dpsi_out"]))
    assert reldiff(dpsi, dpsi_o) < TOL and reldiff(dg, dg_o) < TOL
    d = fixture("test_quantum_circuit_gates_hessian_vector_product_cplx")
    G, Z = stack(d, "G", 6), stack(d, "Z", 6)
    out, dg, hg = R.quantum_circuit_gates_hessian_vector_product(G, Z, d["wires"], cplx(d["psi"]), cplx(d["phi"]))
    out_o, dg_o, hg_o = O.quantum_circuit_gates_hvp(G, Z, d["wires"], cplx(d["psi"]), cplx(d["phi"]))
    assert _max(out - cplx(d["psi_out"])) < 1e-12
    assert reldiff(dg, dg_o) < TOL and reldiff(hg, hg_o) < TOL


def test_fixture_brickwall_circuits():
    d = fixture("test_apply_brickwall_unitary_cplx")
    V, P = stack(d, "V", 4), perms_of(d, 4)
    for nl in (3, 4):
        assert _max(R.apply_brickwall_unitary(V[:nl], P[:nl], cplx(d["psi"])) - cplx(d[f"chi{nl - 3}"])) < 1e-12
    d = fixture("test_brickwall_unitary_backward_cplx")
    V, P = stack(d, "V", 4), perms_of(d, 4)
    for nl in (3, 4):
        out, cache = R.brickwall_unitary_forward(V[:nl], P[:nl], cplx(d["psi"]))
        assert _max(out - cplx(d[f"psi_out{nl - 3}"])) < 1e-12
        dpsi, dv = R.brickwall_unitary_backward(V[:nl], P[:nl], cache, cplx(d["dpsi_out"]))
        _, cache_o = O.brickwall_unitary_forward(V[:nl], P[:nl], cplx(d["psi"]))
        dpsi_o, dv_o = O.brickwall_unitary_backward(V[:nl], P[:nl], cache_o, cplx(d["dpsi_out"]))
        assert reldiff(dpsi, dpsi_o) < TOL and reldiff(dv, dv_o) < TOL


def test_fixture_circuit_targets():
    d = fixture("test_circuit_unitary_target_cplx")
    assert abs(R.circuit_unitary_target(SCR_H, stack(d, "G", 5), d["wires"], 7) - cplx(d["f"])) < 1e-12
    d = fixture("test_circuit_unitary_target_and_gradient_cplx")
    G = stack(d, "G", 5)
    f, dg = R.circuit_unitary_target_and_gradient(SCR_H, G, d["wires"], 8)
    f_o, dg_o = O.circuit_unitary_target_and_gradient(SCR_O, G, d["wires"], 8)
    assert abs(f - cplx(d["f"])) < 1e-12 and reldiff(dg, dg_o) < TOL
    f, g = R.circuit_unitary_target_and_projected_gradient(SCR_H, G, d["wires"], 8)
    _, g_o = O.circuit_unitary_target_and_projected_gradient(SCR_O, G, d["wires"], 8)
    assert reldiff(g, g_o) < TOL
    for name in ("test_circuit_unitary_target_hessian_vector_product_cplx",
                 "test_circuit_unitary_target_projected_hessian_vector_product_cplx"):
        d = fixture(name)
        G, Z = stack(d, "G", 6), stack(d, "Z", 6)
        f, dg, hg = R.circuit_unitary_target_hessian_vector_product(SCR_H, G, Z, d["wires"], 7)
        f_o, dg_o, hg_o = O.circuit_unitary_target_hvp(SCR_O, G, Z, d["wires"], 7)
        assert abs(f - cplx(d["f"])) < 1e-12 and reldiff(dg, dg_o) < TOL and reldiff(hg, hg_o) < TOL
        f, g, h = R.circuit_unitary_target_projected_hessian_vector_product(SCR_H, G, Z, d["wires"], 7)
        _, g_o, h_o = O.circuit_unitary_target_projected_hvp(SCR_O, G, Z, d["wires"], 7)
        assert reldiff(g, g_o) < TOL and reldiff(h, h_o) < TOL


def test_fixture_brickwall_targets():
    d = fixture("test_brickwall_unitary_target_cplx")
    V, P = stack(d, "V", 5), perms_of(d, 5)
    for k, nl in enumerate((4, 5)):
        assert abs(R.brickwall_unitary_target(SCR_H, V[:nl], P[:nl]) - cplx(d[f"f{k}"])) < 1e-12
    d = fixture("test_brickwall_unitary_target_and_gradient_cplx")
    V, P = stack(d, "V", 5), perms_of(d, 5)
    for k, nl in enumerate((4, 5)):
        f, dv = R.brickwall_unitary_target_and_gradient(SCR_H, V[:nl], P[:nl])
        assert abs(f - cplx(d[f"f{k}"])) < 1e-12 and _max(dv - cplx(d[f"dVlist{k}"])) < 1e-12
    d = fixture("test_brickwall_unitary_target_and_projected_gradient_cplx")
    f, g = R.brickwall_unitary_target_and_projected_gradient(SCR_H, stack(d, "V", 3), perms_of(d, 3))
    assert abs(f - cplx(d["f"])) < 1e-12 and _max(g - d["grad"]) < 1e-14


def test_fixture_hessians_pin_hvp():
    d = fixture("test_brickwall_unitary_target_gradient_hessian_cplx")
    V, Z, P = stack(d, "V", 5), stack(d, "Z", 5), perms_of(d, 5)
    for k, nl in enumerate((4, 5)):
        f, dv, hz = R.brickwall_unitary_target_hessian_vector_product(SCR_H, V[:nl], Z[:nl], P[:nl])
        ref = (cplx(d[f"hess{k}"]) @ Z[:nl].reshape(-1)).reshape(nl, 4, 4)
        assert abs(f - cplx(d[f"f{k}"])) < 1e-12
        assert _max(dv - cplx(d[f"dVlist{k}"])) < 1e-12 and _max(hz - ref) < 1e-12
    d = fixture("test_brickwall_unitary_target_gradient_vector_hessian_matrix_cplx")
    V, P = stack(d, "V", 5), perms_of(d, 5)
    vec = np.random.default_rng(1).standard_normal(80)
    Zs = np.stack([R.real_to_tangent(vec[16 * i:16 * i + 16], V[i]) for i in range(5)])
    f, g, h = R.brickwall_unitary_target_projected_hessian_vector_product(SCR_H, V, Zs, P)
    assert abs(f - cplx(d["f"])) < 1e-12 and _max(g - d["grad"]) < 1e-14 and _max(h - d["H"] @ vec) < 1e-12
    d = fixture("test_brickwall_unitary_target_projected_hessian_vector_product_cplx")
    V, Z, P = stack(d, "V", 4), stack(d, "Z", 4), perms_of(d, 4)
    f, g, h = R.brickwall_unitary_target_projected_hessian_vector_product(SCR_H, V, Z, P)
    f_o, g_o, h_o = O.brickwall_unitary_target_projected_hvp(SCR_O, V, Z, P, 6)
    assert abs(f - cplx(d["f"])) < 1e-12 and reldiff(g, g_o) < TOL and reldiff(h, h_o) < TOL


# ------------------------------------------------------------ random / edges
@pytest.mark.parametrize("L", [2, 3, 5, 9])
def test_apply_gate_all_wire_orders(L):
    rng = np.random.default_rng(100 + L)
    G = crandn(rng, (4, 4))
    psi, dout = crandn(rng, 1 << L), crandn(rng, 1 << L)
    for i in range(L):
        for j in range(L):
            if i == j:
                continue
            assert reldiff(R.apply_gate(G, i, j, psi), O.apply_gate(G, i, j, psi)) < 1e-14
            dpsi, dG = R.apply_gate_backward(G, i, j, psi, dout)
            dpsi_o, dG_o = O.apply_gate_backward(G, i, j, psi, dout)
            assert reldiff(dpsi, dpsi_o) < 1e-14 and reldiff(dG, dG_o) < 1e-13


@pytest.mark.parametrize("nstates", [1, 3, 8, 64])
def test_apply_gate_to_array_widths(nstates):
    rng = np.random.default_rng(nstates)
    L = 6
    G = crandn(rng, (4, 4))
    psi = crandn(rng, (1 << L, nstates))
    for (i, j) in ((0, 5), (5, 0), (2, 3), (4, 1)):
        assert reldiff(R.apply_gate_to_array(G, i, j, psi), O.apply_gate_to_array(G, i, j, psi)) < 1e-14


def test_add_statevectors():
    rng = np.random.default_rng(5)
    a, b = crandn(rng, 1 << 10), crandn(rng, 1 << 10)
    assert np.array_equal(R.add_statevectors(a, b), a + b)


def test_invalid_arguments_are_rejected():
    rng = np.random.default_rng(6)
    G = crandn(rng, (3, 4, 4))
    with pytest.raises(R.RqcError):
        R.apply_quantum_circuit(G, np.array([[0, 1], [1, 1], [2, 0]]), crandn(rng, 8))
    with pytest.raises(R.RqcError):
        R.apply_quantum_circuit(G, np.array([[0, 1], [1, 3], [2, 0]]), crandn(rng, 8))
    with pytest.raises(R.RqcError):
        R.apply_brickwall_unitary(G[:1], np.array([[0, 0, 1, 2]]), crandn(rng, 16))
    bad = R.HostUfunc(lambda x: 1 / 0)
    with pytest.raises(R.RqcError) as e:
        R.brickwall_unitary_target(bad, G[:2], brickwall_perms(4, 2))
    assert e.value.code == -2
    T = R.DeviceTarget.dense(np.eye(16))
    with pytest.raises(R.RqcError):
        R.brickwall_unitary_target(T, G[:2], brickwall_perms(6, 2))


def test_random_circuit_targets_vs_oracle():
    rng = np.random.default_rng(7)
    L, ng = 7, 9
    U = crandn(rng, (1 << L, 1 << L))
    T, Uo = R.DeviceTarget.dense(U), O.Ufunc.dense(U)
    gates, dirs = crandn(rng, (ng, 4, 4)), crandn(rng, (ng, 4, 4))
    wires = random_wires(rng, L, ng)
    assert reldiff(R.circuit_unitary_target(T, gates, wires, L), O.circuit_unitary_target(Uo, gates, wires, L)) < TOL
    f, dg, hg = R.circuit_unitary_target_hessian_vector_product(T, gates, dirs, wires, L)
    f_o, dg_o, hg_o = O.circuit_unitary_target_hvp(Uo, gates, dirs, wires, L)
    assert reldiff(f, f_o) < TOL and reldiff(dg, dg_o) < TOL and reldiff(hg, hg_o) < TOL


@pytest.mark.parametrize("L,nl", [(4, 2), (6, 3), (8, 5), (10, 3)])
def test_random_brickwall_targets_vs_oracle(L, nl):
    rng = np.random.default_rng(L * 10 + nl)
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    T, Uo = R.DeviceTarget.dense(U), O.Ufunc.dense(U)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    P = random_perms(rng, L, nl)
    vec = rng.standard_normal(16 * nl)
    Z = np.stack([O.real_to_tangent(vec[16 * i:16 * i + 16], V[i]) for i in range(nl)])
    f, g = R.brickwall_unitary_target_and_projected_gradient(T, V, P)
    f_o, g_o = O.brickwall_unitary_target_and_projected_gradient(Uo, V, P, L)
    assert abs(f - f_o) < TOL * max(1, abs(f_o)) and reldiff(g, g_o) < TOL
    f, g, h = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
    f_o, g_o, h_o = O.brickwall_unitary_target_projected_hvp(Uo, V, Z, P, L)
    assert abs(f - f_o) < TOL * max(1, abs(f_o)) and reldiff(g, g_o) < TOL and reldiff(h, h_o) < TOL


def test_ising_workload_small_vs_oracle():
    """L=8 periodic brickwall d=5, Haar gates from the library RNG, U = expm(-iHt) of the Ising chain."""
    from scipy.linalg import expm
    from rqcopt_hpc_amd.workloads import ising_hamiltonian
    L, nl = 8, 5
    H = ising_hamiltonian(L, 1.0, 0.75, 0.6)
    U = expm(-1j * 0.25 * H)
    r = R.Rng(1)
    V = np.stack([r.haar_unitary() for _ in range(nl)])
    P = brickwall_perms(L, nl)
    r2 = R.Rng(2)
    vec = np.array([r2.randn() for _ in range(16 * nl)])
    Z = np.stack([R.real_to_tangent(vec[16 * i:16 * i + 16], V[i]) for i in range(nl)])
    T, Uo = R.DeviceTarget.dense(U), O.Ufunc.dense(U)
    f, g, h = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
    f_o, g_o, h_o = O.brickwall_unitary_target_projected_hvp(Uo, V, Z, P, L)
    assert abs(f - f_o) < TOL * max(1, abs(f_o)) and reldiff(g, g_o) < TOL and reldiff(h, h_o) < TOL


def test_ising_L8_d5_golden_and_rtr_trajectory():
    """L=8 Ising workload: f, projected gradient and HVP, then 10 RTR-HVP iterations against the
    oracle's trajectory (tests/golden/make_config0.py)."""
    from rqcopt_hpc_amd.workloads import expm_hermitian, ising_hamiltonian
    d = fixture("config0_ising_L8_d5")
    U = expm_hermitian(ising_hamiltonian(8, 1.0, 0.75, 0.6), 0.25)
    T = R.DeviceTarget.dense(U)
    f, g, h = R.brickwall_unitary_target_projected_hessian_vector_product(T, d["V"], d["Z"], d["P"])
    assert abs(f - d["f"]) < TOL * max(1, abs(d["f"]))
    assert reldiff(g, d["grad"]) < TOL and reldiff(h, d["hvp"]) < TOL
    f_iter, vopt = R.optimize_brickwall_circuit_hvp(T, d["V"], d["P"], 10)
    # accept/reject and radius branches could amplify rounding; none flips here
    assert np.max(np.abs(f_iter - d["f_iter"]) / np.maximum(1, np.abs(d["f_iter"]))) < 1e-9
    assert _max(vopt - d["vopt"]) < 1e-8


@pytest.mark.parametrize("radius_init,maxradius", [(2.0, 8.0), (1.0, 4.0)])
def test_mixed_rtr_takes_the_fp64_branches(radius_init, maxradius):
    """L=8 Ising workload, 12 RTR-HVP iterations with a large trust radius so that some steps are
    rejected (at iterations 0 and 2, or 0 and 11, with the oracle's evaluations).  A rejected step
    keeps x, so f_iter repeats; mixed precision must reject exactly the steps FP64 rejects: its
    ratio rho compares f values computed the same way (every mixed f carries the gate-rounding
    correction)."""
    from rqcopt_hpc_amd.workloads import expm_hermitian, ising_hamiltonian
    d = fixture("config0_ising_L8_d5")
    T = R.DeviceTarget.dense(expm_hermitian(ising_hamiltonian(8, 1.0, 0.75, 0.6), 0.25))

    def run():
        prm = R.rtr_default_params(16 * len(d["V"]))
        prm.radius_init, prm.maxradius = radius_init, maxradius
        return R.optimize_brickwall_circuit_hvp(T, d["V"], d["P"], 12, prm)

    f64, v64 = run()
    R.set_precision(R.PRECISION_MIXED)
    try:
        f32, v32 = run()
    finally:
        R.set_precision(R.PRECISION_FP64)
    rej64 = np.abs(np.diff(f64)) < 1e-10 * np.maximum(1, np.abs(f64[1:]))
    rej32 = np.abs(np.diff(f32)) < 1e-10 * np.maximum(1, np.abs(f32[1:]))
    print("rejected steps:", np.nonzero(rej64)[0], np.nonzero(rej32)[0], "max |df|", np.max(np.abs(f32 - f64)), flush=True)
    assert rej64.sum() >= 2
    assert np.array_equal(rej64, rej32)
    assert np.max(np.abs(f32 - f64)) / 256 < 1e-6
    T.close()


def test_column_blocking_and_determinism():
    rng = np.random.default_rng(8)
    L, nl = 8, 4
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    T = R.DeviceTarget.dense(U)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    Z = crandn(rng, (nl, 4, 4))
    P = brickwall_perms(L, nl)
    a = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
    b = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
    assert R.last_column_block() == 1 << L
    for x, y in zip(a, b):
        assert np.array_equal(x, y)  # bitwise reproducible
    N = 1 << L
    ng = nl * L // 2
    R.set_workspace_limit((2 * ng + 5) * N * 16 * 16)  # room for a block of at most 32 columns
    try:
        c = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
        assert 1 <= R.last_column_block() <= 32
    finally:
        R.set_workspace_limit(0)
    for x, y in zip(a, c):
        assert reldiff(x, y) < 1e-13


def test_mixed_precision_error_bounds():
    rng = np.random.default_rng(9)
    L, nl = 10, 5
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    T = R.DeviceTarget.dense(U)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    vec = rng.standard_normal(16 * nl)
    Z = np.stack([R.real_to_tangent(vec[16 * i:16 * i + 16], V[i]) for i in range(nl)])
    P = brickwall_perms(L, nl)
    f64 = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
    R.set_precision(R.PRECISION_MIXED)
    try:
        f32 = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
    finally:
        R.set_precision(R.PRECISION_FP64)
    N = 1 << L
    # stated bounds: |df| / N <= 1e-6, gradient and HVP within 1e-4 relative (FP32 sweeps, FP64 sums)
    assert abs(f32[0] - f64[0]) / N < 1e-6
    assert reldiff(f32[1], f64[1]) < 1e-4 and reldiff(f32[2], f64[2]) < 1e-4


def _round32(V):
    return (V.real.astype(np.float32) + 1j * V.imag.astype(np.float32)).astype(np.complex128)


@pytest.mark.parametrize("L,nl,engine", [(10, 5, "tile"), (8, 4, "gate")])
def test_mixed_gate_rounding_correction(L, nl, engine):
    """FP32 sweeps apply the gates rounded to complex float; the evaluations with a gradient add
    the first-order term sum dv (V - fl(V)).  Mixed runs on V and on fl(V) share every FP32
    operation (fl(fl(V)) = fl(V)), so their difference is that term alone; the FP64 path on the
    two gate sets gives its true value f(V) - f(fl(V)).  A dropped term, a wrong sign or a
    conjugated gradient fails the comparison by the size of the term itself."""
    rng = np.random.default_rng(12 + L)
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    T = R.DeviceTarget.dense(U)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    V32 = _round32(V)
    Z = crandn(rng, (nl, 4, 4))
    P = brickwall_perms(L, nl)
    if engine == "gate":
        R.set_sweep_engine(R.ENGINE_PER_GATE)
    try:
        f64, _ = R.brickwall_unitary_target_and_gradient(T, V, P)
        f64r, _ = R.brickwall_unitary_target_and_gradient(T, V32, P)
        R.set_precision(R.PRECISION_MIXED)
        try:
            fc, dvc = R.brickwall_unitary_target_and_gradient(T, V, P)
            fu, _ = R.brickwall_unitary_target_and_gradient(T, V32, P)
            hc = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)[0]
            hu = R.brickwall_unitary_target_hessian_vector_product(T, V32, Z, P)[0]
            ft = R.brickwall_unitary_target(T, V, P)
        finally:
            R.set_precision(R.PRECISION_FP64)
    finally:
        R.set_sweep_engine(R.ENGINE_TILE_PASS)
    exact = f64 - f64r
    print(f"L={L} {engine}: gate-rounding term {abs(exact):.3e}, corrected error {abs(fc - f64):.3e}", flush=True)
    # a mixed target evaluation runs the gradient sweep and applies the same correction
    assert ft == fc
    rnd = 1e-14 * max(1.0, abs(f64))                            # rounding of a difference of two f values
    assert abs(exact) > 1e3 * rnd                                # the term is there to be checked
    assert abs((fc - fu) - exact) < 1e-3 * abs(exact) + rnd
    assert abs((hc - hu) - exact) < 1e-3 * abs(exact) + rnd
    assert abs((fc - fu) - np.sum(dvc * (V - V32))) < rnd       # the term is sum dv (V - fl(V))
    assert abs(fc - f64) / (1 << L) < 1e-6
    T.close()


def test_device_target_brickwall_and_ufunc():
    rng = np.random.default_rng(10)
    L, nl = 6, 3
    Vt = np.stack([haar_unitary(rng) for _ in range(nl)])
    P = brickwall_perms(L, nl)
    T = R.DeviceTarget.brickwall(Vt, P)
    W = brickwall_matrix(Vt, P)
    assert _max(T.download() - W) < 1e-13
    # f(W = U) = -N and the Riemannian gradient vanishes at the optimum
    f, g = R.brickwall_unitary_target_and_projected_gradient(T, Vt, P)
    assert abs(f + (1 << L)) < 1e-11 and _max(g) < 1e-11
    T2 = R.DeviceTarget.from_ufunc(lambda x: W @ x, L)
    assert _max(T2.download() - W) < 1e-14
    psi = crandn(rng, 1 << L)
    out = np.empty_like(psi)
    r = R.lib().rqc_device_ufunc(ct.byref(R._sv(psi)), T.handle, ct.byref(R._sv(out)))
    assert r == 0 and reldiff(out, W @ psi) < 1e-13


def _hip():
    """The HIP runtime librqcopt_hip is linked against (already loaded in this process)."""
    h = ct.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    h.hipFree.argtypes = [ct.c_void_p]
    return h


def test_device_resident_statevectors():
    """Device pointers in struct statevector are used in place (no host staging)."""
    hip = _hip()
    rng = np.random.default_rng(11)
    L = 12
    G = crandn(rng, (4, 4))
    psi = crandn(rng, 1 << L)
    nbytes = psi.nbytes
    d_in, d_out = ct.c_void_p(), ct.c_void_p()
    assert hip.hipMalloc(ct.byref(d_in), nbytes) == 0 and hip.hipMalloc(ct.byref(d_out), nbytes) == 0
    try:
        assert hip.hipMemcpy(d_in, psi.ctypes.data, nbytes, 1) == 0  # host to device
        a = R.Statevector(d_in.value, L)
        b = R.Statevector(d_out.value, L)
        R.lib().apply_gate(R._p(R._cplx(G)), 3, 11, ct.byref(a), ct.byref(b))
        R.synchronize()
        out = np.empty_like(psi)
        assert hip.hipMemcpy(out.ctypes.data, d_out, nbytes, 2) == 0  # device to host
        assert reldiff(out, O.apply_gate(G, 3, 11, psi)) < 1e-14
        # a whole circuit with device-resident input and output
        gates = crandn(rng, (6, 4, 4))
        wires = random_wires(rng, L, 6).reshape(-1)
        g = R._gates(gates)
        r = R.lib().apply_quantum_circuit(R._p(g), 6, R._p(wires), ct.byref(a), ct.byref(b))
        assert r == 0
        assert hip.hipMemcpy(out.ctypes.data, d_out, nbytes, 2) == 0
        assert reldiff(out, O.apply_quantum_circuit(gates, wires.reshape(-1, 2), psi)) < 1e-13
    finally:
        hip.hipFree(d_in)
        hip.hipFree(d_out)


def test_rtr_optimiser_vs_oracle():
    rng = np.random.default_rng(12)
    L, nl = 6, 3
    P = brickwall_perms(L, nl)
    Vt = np.stack([haar_unitary(rng) for _ in range(nl)])
    U = brickwall_matrix(Vt, P)
    V0 = np.stack([haar_unitary(rng) for _ in range(nl)])
    f_iter, vopt = R.optimize_brickwall_circuit_hvp(R.DeviceTarget.dense(U), V0, P, 5)
    f_o, vopt_o = O.optimize_brickwall_circuit_hvp(O.Ufunc.dense(U), V0, P, L, 5)
    assert np.max(np.abs(f_iter - f_o) / np.maximum(1, np.abs(f_o))) < 1e-9
    assert _max(vopt - vopt_o) < 1e-8
    assert f_iter[-1] < f_iter[0]


@pytest.mark.parametrize("unitary", [True, False])
@pytest.mark.parametrize("L", [2, 3, 5, 7, 8, 11])
def test_sweep_engines_agree(L, unitary):
    """Tile-pass engine (several gates per HBM pass) vs per-gate engine vs oracle, random wires.
    Unitary gates run the tile-pass reverse sweep (uncompute with G^-1) from L = 5 on; general
    gates are handed to the per-gate engine for gradient and HVP."""
    rng = np.random.default_rng(200 + L)
    ng = 3 * L
    if unitary:
        gates = np.stack([haar_unitary(rng) for _ in range(ng)])
    else:
        gates = crandn(rng, (ng, 4, 4))
    dirs = crandn(rng, (ng, 4, 4))
    wires = random_wires(rng, L, ng)
    U = crandn(rng, (1 << L, 1 << L))
    T = R.DeviceTarget.dense(U)
    res, used = {}, {}
    for eng in (R.ENGINE_TILE_PASS, R.ENGINE_PER_GATE):
        R.set_sweep_engine(eng)
        try:
            r0 = R.circuit_unitary_target(T, gates, wires, L)
            u0 = R.last_sweep()
            r1 = R.circuit_unitary_target_and_gradient(T, gates, wires, L)
            u1 = R.last_sweep()
            r2 = R.circuit_unitary_target_hessian_vector_product(T, gates, dirs, wires, L)
            u2 = R.last_sweep()
            res[eng], used[eng] = (r0, r1, r2), (u0, u1, u2)
        finally:
            R.set_sweep_engine(R.ENGINE_TILE_PASS)
    tile_ok = L >= 5
    assert used[R.ENGINE_TILE_PASS][0][0] == (R.ENGINE_TILE_PASS if tile_ok else R.ENGINE_PER_GATE)
    for u in used[R.ENGINE_TILE_PASS][1:]:
        assert u[0] == (R.ENGINE_TILE_PASS if tile_ok and unitary else R.ENGINE_PER_GATE)
    assert all(u[0] == R.ENGINE_PER_GATE for u in used[R.ENGINE_PER_GATE])
    a, b = res[R.ENGINE_TILE_PASS], res[R.ENGINE_PER_GATE]
    assert reldiff(a[0], b[0]) < TOL
    assert reldiff(a[1][1], b[1][1]) < TOL
    for x, y in zip(a[2], b[2]):
        assert reldiff(x, y) < TOL
    if L <= 8:
        f_o, dg_o, hg_o = O.circuit_unitary_target_hvp(O.Ufunc.dense(U), gates, dirs, wires, L)
        assert reldiff(a[2][0], f_o) < TOL and reldiff(a[2][1], dg_o) < TOL and reldiff(a[2][2], hg_o) < TOL


def test_tile_pass_engine_column_blocks_and_mixed():
    rng = np.random.default_rng(300)
    L, nl = 12, 4
    P = brickwall_perms(L, nl)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    Z = crandn(rng, (nl, 4, 4))
    T = R.DeviceTarget.brickwall(np.stack([haar_unitary(rng) for _ in range(3)]), brickwall_perms(L, 3))
    full = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
    assert R.last_column_block() == 1 << L
    R.set_workspace_limit(64 << 20)
    try:
        blk = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
        assert R.last_column_block() < 1 << L
    finally:
        R.set_workspace_limit(0)
    for x, y in zip(full, blk):
        assert reldiff(x, y) < 1e-13
    R.set_sweep_engine(R.ENGINE_PER_GATE)
    try:
        ref = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
    finally:
        R.set_sweep_engine(R.ENGINE_TILE_PASS)
    for x, y in zip(full, ref):
        assert reldiff(x, y) < TOL
    R.set_precision(R.PRECISION_MIXED)
    try:
        mixed = R.brickwall_unitary_target_hessian_vector_product(T, V, Z, P)
    finally:
        R.set_precision(R.PRECISION_FP64)
    assert abs(mixed[0] - full[0]) / (1 << L) < 1e-6
    assert reldiff(mixed[1], full[1]) < 1e-4 and reldiff(mixed[2], full[2]) < 1e-4


# ------------------------------------------------------- full-size properties
@pytest.mark.parametrize("L,nl", [(12, 5), (14, 3)])
def test_large_size_invariants(L, nl):
    """At sizes beyond the oracle: f(W=U) = -N, zero Riemannian gradient at the optimum,
    HVP linear in the direction and symmetric."""
    rng = np.random.default_rng(L)
    P = brickwall_perms(L, nl)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    T = R.DeviceTarget.brickwall(V, P)
    N = 1 << L
    e1, e2 = rng.standard_normal(16 * nl), rng.standard_normal(16 * nl)
    Z1 = np.stack([R.real_to_tangent(e1[16 * i:16 * i + 16], V[i]) for i in range(nl)])
    Z2 = np.stack([R.real_to_tangent(e2[16 * i:16 * i + 16], V[i]) for i in range(nl)])
    f, g, h1 = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z1, P)
    assert abs(f + N) < 1e-12 * N
    assert _max(g) < 1e-12 * N
    _, _, h2 = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z2, P)
    _, _, h12 = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, 2.0 * Z1 - 0.5 * Z2, P)
    assert reldiff(h12, 2.0 * h1 - 0.5 * h2) < 1e-12
    assert abs(e2 @ h1 - e1 @ h2) < 1e-10 * max(1.0, abs(e1 @ h2))
    # at the optimum the projected Hessian is positive semidefinite
    assert e1 @ h1 >= -1e-9 * N
