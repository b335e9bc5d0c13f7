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
