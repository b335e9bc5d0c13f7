"""Workload generators (CPU): Hamiltonians, Trotter circuits, seeded inputs."""
import numpy as np

import oracle as O
from helpers import brickwall_matrix
from rqcopt_hpc_amd.workloads import (PAULI_X, PAULI_Z, brickwall_perms, config_inputs, expm_hermitian,
                                      ising_hamiltonian, ising_trotter_brickwall, xxz_hamiltonian)


def _kron_site(op, j, L):
    m = np.array([[1.0 + 0j]])
    for q in range(L):
        m = np.kron(m, op if q == j else np.eye(2))
    return m


def test_ising_hamiltonian_matches_kronecker_sum():
    L, J, h, g = 5, 1.0, 0.75, 0.6
    H = np.zeros((1 << L, 1 << L), np.complex128)
    for j in range(L):
        H += J * _kron_site(PAULI_Z, j, L) @ _kron_site(PAULI_Z, (j + 1) % L, L)
        H += h * _kron_site(PAULI_X, j, L) + g * _kron_site(PAULI_Z, j, L)
    assert np.max(np.abs(ising_hamiltonian(L, J, h, g) - H)) < 1e-14


def test_trotter_circuit_converges_to_exact_propagator():
    L, t = 6, 0.5
    U = expm_hermitian(ising_hamiltonian(L, 1.0, 0.75, 0.6), t)
    errs = []
    for steps in (1, 2, 4, 8):
        V, P = ising_trotter_brickwall(L, 1.0, 0.75, 0.6, t, steps)
        assert len(V) == 2 * steps + 1
        W = brickwall_matrix(V, P)
        assert np.max(np.abs(W.conj().T @ W - np.eye(1 << L))) < 1e-12
        errs.append(np.linalg.norm(W - U, 2))
    # Strang splitting is second order: error / 4 per doubling of the steps
    for a, b in zip(errs, errs[1:]):
        assert 3.0 < a / b < 5.0
    # the brickwall the GPU builds U from is what the oracle applies column by column
    V, P = ising_trotter_brickwall(L, 1.0, 0.75, 0.6, t, 4)
    e3 = np.zeros(1 << L, np.complex128)
    e3[3] = 1
    assert np.max(np.abs(O.apply_brickwall_unitary(V, P, e3) - brickwall_matrix(V, P)[:, 3])) < 1e-14


def test_xxz_hamiltonian_structure():
    L = 4
    rng = np.random.default_rng(0)
    fields = rng.uniform(-1, 1, L)
    H = xxz_hamiltonian(L, 0.5, fields)
    assert np.max(np.abs(H - H.conj().T)) < 1e-14
    # total magnetisation is conserved by XXZ with Z fields
    Mz = sum(_kron_site(PAULI_Z, j, L) for j in range(L))
    assert np.max(np.abs(H @ Mz - Mz @ H)) < 1e-12


def test_config_inputs_are_seeded_and_tangent():
    V1, P1, vec1, Z1 = config_inputs(8, 5, seed=3)
    V2, P2, vec2, Z2 = config_inputs(8, 5, seed=3)
    assert np.array_equal(V1, V2) and np.array_equal(Z1, Z2) and np.array_equal(P1, P2)
    assert np.array_equal(P1, brickwall_perms(8, 5))
    for V, Z, k in zip(V1, Z1, range(5)):
        assert np.max(np.abs(V.conj().T @ V - np.eye(4))) < 1e-14
        s = V.conj().T @ Z
        assert np.max(np.abs(s + s.conj().T)) < 1e-14           # Z in the tangent space at V
        assert np.max(np.abs(O.tangent_to_real(V, Z) - vec1[16 * k:16 * k + 16])) < 1e-14


def _circuit_matrix(gates, wires, L):
    from helpers import gate_matrix
    U = np.eye(1 << L, dtype=np.complex128)
    for G, (i, j) in zip(gates, wires):
        U = gate_matrix(G, int(i), int(j), L) @ U
    return U


def test_xxz_trotter_circuit_second_order_convergence():
    from rqcopt_hpc_amd.workloads import xxz_trotter_circuit
    L, t = 6, 0.4
    fields = np.linspace(-1, 1, L)
    U = expm_hermitian(xxz_hamiltonian(L, 0.5, fields), t)
    errs = [np.max(np.abs(_circuit_matrix(*xxz_trotter_circuit(L, 0.5, fields, t, s), L) - U)) for s in (2, 4, 8)]
    # second-order splitting: the error drops by about 4x per doubling of the steps
    assert errs[2] < errs[1] < errs[0] and errs[0] / errs[2] > 10 and errs[2] < 3e-3


def test_serving_chain_targets_match_dense_hamiltonians():
    from rqcopt_hpc_amd.serving import chain_bonds, chain_trotter_circuit, ising_term, xxz_term
    for L in (6, 7):
        bonds = chain_bonds(L)
        assert len(bonds) == (L if L % 2 == 0 else L - 1)
        fields = np.linspace(-0.8, 0.9, L)
        # dense H of the same chains: periodic for even L, open for odd L
        Hi = np.zeros((1 << L, 1 << L), np.complex128)
        Hx = np.zeros((1 << L, 1 << L), np.complex128)
        for j, q in bonds:
            Hi += _kron_site(PAULI_Z, j, L) @ _kron_site(PAULI_Z, q, L)
            for P, c in ((PAULI_X, 1.0), (np.array([[0, -1j], [1j, 0]]), 1.0), (PAULI_Z, 0.5)):
                Hx += c * _kron_site(P, j, L) @ _kron_site(P, q, L)
        for j in range(L):
            Hi += 0.75 * _kron_site(PAULI_X, j, L) + 0.6 * _kron_site(PAULI_Z, j, L)
            Hx += fields[j] * _kron_site(PAULI_Z, j, L)
        for H, term in ((Hi, ising_term(1.0, 0.75, 0.6)), (Hx, xxz_term(0.5, fields))):
            U = expm_hermitian(H, 0.3)
            e4 = np.max(np.abs(_circuit_matrix(*chain_trotter_circuit(L, term, 0.3, 4), L) - U))
            e16 = np.max(np.abs(_circuit_matrix(*chain_trotter_circuit(L, term, 0.3, 16), L) - U))
            assert e16 < e4 / 8 and e16 < 1e-3, (L, e4, e16)
        if L % 2 == 0:
            assert np.allclose(Hi, ising_hamiltonian(L, 1.0, 0.75, 0.6))


def test_serving_workload_layouts():
    from rqcopt_hpc_amd.serving import open_brickwall_wires, serving_workload
    solves, targets = serving_workload(5, nsolves=40)
    for s in solves:
        assert 8 <= s["L"] <= 14 and s["depth"] in (3, 5, 7, 9)
        g, w = targets[s["target"]]
        assert np.all(w < s["L"])
        if s["L"] % 2:
            assert not s["brickwall"] and len(s["gates"]) == len(s["wires"]) == s["depth"] * ((s["L"] - 1) // 2)
        else:
            assert s["brickwall"] and len(s["gates"]) == s["depth"]
        for G in s["gates"]:
            assert np.allclose(G.conj().T @ G, np.eye(4), atol=1e-12)
    w = open_brickwall_wires(7, 2)
    assert w.tolist() == [[0, 1], [2, 3], [4, 5], [1, 2], [3, 4], [5, 6]]
