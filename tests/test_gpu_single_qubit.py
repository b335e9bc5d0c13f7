"""GPU tests of the single-qubit gate path (include/rqcopt/rqc_gate1q.h; no reference counterpart,
the reference's two-qubit conventions of src/gate.c:10-188).

Statevector level: the HIP kernels give the oracle's single-qubit gate, backward map and gradient
(orc_apply_gate_1q*, themselves pinned by dense Kronecker products in test_oracle_pinning.py) to
1e-14, on every wire, for host and device buffers and interleaved arrays.

Circuits of one- and two-qubit gates: f, the gate gradients and the directional second
derivatives equal dense-matrix values (tests/helpers.py mixed_target_gradient; the HVP by central
differences of that dense gradient) to 1e-12 relative (the HVP to the differences' 1e-6), in both
sweep engines; mixed precision stays within its stated bounds.
"""
import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from helpers import crandn, haar_unitary, mixed_target_gradient, reldiff

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def gpu():
    R.require_gpu()
    yield
    R.set_precision(R.PRECISION_FP64)
    R.set_sweep_engine(R.ENGINE_TILE_PASS)


def test_single_qubit_gate_matches_oracle_on_every_wire():
    rng = np.random.default_rng(51)
    for L in (1, 4, 9):
        psi, dout = crandn(rng, 1 << L), crandn(rng, 1 << L)
        for i in range(L):
            V = crandn(rng, (2, 2))
            print(f"L={L} wire {i}", flush=True)
            out = R.apply_single_qubit_gate(V, i, psi)
            assert np.max(np.abs(out - O.apply_gate_1q(V, i, psi))) < 1e-14
            dpsi, dV = R.apply_single_qubit_gate_backward(V, i, psi, dout)
            dpsi_o, dV_o = O.apply_gate_1q_backward(V, i, psi, dout)
            assert np.max(np.abs(dpsi - dpsi_o)) < 1e-14
            assert np.max(np.abs(dV - dV_o)) < 1e-13 * max(1, np.max(np.abs(dV_o)))


def test_single_qubit_gate_on_arrays():
    rng = np.random.default_rng(52)
    L = 6
    for ns in (1, 5, 64):
        psi = crandn(rng, (1 << L, ns))
        for i in (0, 3, L - 1):
            V = haar_unitary(rng, 2)
            out = R.apply_single_qubit_gate_to_array(V, i, psi)
            ref = np.stack([O.apply_gate_1q(V, i, psi[:, k]) for k in range(ns)], axis=1)
            assert np.max(np.abs(out - ref)) < 1e-14


def _mixed_problem(rng, L):
    nq = [1, 2, 1, 1, 2, 2, 1, 2]
    wires = [(0, 0), (0, 1), (L - 1, L - 1), (2, 2), (L - 1, 0), (2, 3), (1, 1), (1, 2)]
    gates = [haar_unitary(rng, 2) if q == 1 else haar_unitary(rng) for q in nq]
    dirs = [crandn(rng, G.shape) for G in gates]
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    return U, gates, dirs, nq, wires


@pytest.mark.parametrize("engine", ["tile_pass", "per_gate"])
def test_mixed_circuit_target_gradient_hvp_match_dense(engine):
    rng = np.random.default_rng(53)
    L = 6
    U, gates, dirs, nq, wires = _mixed_problem(rng, L)
    R.set_sweep_engine(R.ENGINE_TILE_PASS if engine == "tile_pass" else R.ENGINE_PER_GATE)
    T = R.DeviceTarget.dense(U)
    f_d, g_d = mixed_target_gradient(U, gates, nq, wires, L)
    f = R.mixed_circuit_unitary_target(T, gates, nq, wires, L)
    assert abs(f - f_d) < 1e-12 * max(1, abs(f_d))
    f, g = R.mixed_circuit_unitary_target_and_gradient(T, gates, nq, wires, L)
    print(engine, "sweep engine used:", R.last_sweep(), flush=True)
    assert abs(f - f_d) < 1e-12 * max(1, abs(f_d))
    for a, b in zip(g, g_d):
        assert a.shape == b.shape and reldiff(a, b) < 1e-12
    f, g2, h = R.mixed_circuit_unitary_target_hessian_vector_product(T, gates, dirs, nq, wires, L)
    for a, b in zip(g2, g_d):
        assert reldiff(a, b) < 1e-12
    # directional derivative of the dense gradient along dirs (central differences)
    eps = 1e-6
    gp = mixed_target_gradient(U, [G + eps * D for G, D in zip(gates, dirs)], nq, wires, L)[1]
    gm = mixed_target_gradient(U, [G - eps * D for G, D in zip(gates, dirs)], nq, wires, L)[1]
    for k, (a, p, m) in enumerate(zip(h, gp, gm)):
        fd = (p - m) / (2 * eps)
        assert a.shape == fd.shape and reldiff(a, fd) < 1e-6, (k, reldiff(a, fd))
    T.close()


def test_mixed_circuit_mixed_precision_within_bounds():
    rng = np.random.default_rng(54)
    L = 8
    U, gates, dirs, nq, wires = _mixed_problem(rng, L)
    T = R.DeviceTarget.dense(U)
    f64, g64, h64 = R.mixed_circuit_unitary_target_hessian_vector_product(T, gates, dirs, nq, wires, L)
    R.set_precision(R.PRECISION_MIXED)
    f32, g32, h32 = R.mixed_circuit_unitary_target_hessian_vector_product(T, gates, dirs, nq, wires, L)
    # stated bounds: |df| / N <= 1e-6, gradient and HVP within 1e-4 relative
    assert abs(f32 - f64) / (1 << L) < 1e-6
    assert max(reldiff(a, b) for a, b in zip(g32, g64)) < 1e-4
    assert max(reldiff(a, b) for a, b in zip(h32, h64)) < 1e-4
    T.close()


def test_mixed_circuit_rejects_bad_gate_kinds():
    rng = np.random.default_rng(55)
    U, gates, dirs, nq, wires = _mixed_problem(rng, 4)
    T = R.DeviceTarget.dense(U)
    with pytest.raises((ValueError, R.RqcError)):
        R.mixed_circuit_unitary_target(T, gates, [3] + nq[1:], wires, 4)
    with pytest.raises(R.RqcError):
        R.mixed_circuit_unitary_target(T, gates, nq, [(7, 7)] + wires[1:], 4)
    T.close()
