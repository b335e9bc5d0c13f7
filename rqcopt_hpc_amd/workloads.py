"""Seeded inputs of the benchmark workloads (host-side input generation).

- periodic brickwall permutations (even layers: identity; odd layers: roll by
  one, giving the wrap pair (L-1, 0));
- Haar-random 4x4 gates from the library's seeded PCG generator;
- tangent directions vdirs = real_to_tangent(vec, V) with vec ~ N(0, 1);
- 1D Ising H = J sum Z_j Z_{j+1} + h sum X_j + g sum Z_j and Heisenberg-XXZ
  H = sum (X_j X_{j+1} + Y_j Y_{j+1} + Delta Z_j Z_{j+1}) + sum h_j Z_j, both
  periodic, as dense matrices for small L;
- the Strang-split Trotter circuit of the Ising chain as a brickwall of exact
  bond layers (one shared gate per layer), which the GPU turns into a dense
  device-resident U (DeviceTarget.brickwall).
Qubit 0 is the most significant bit of the amplitude index.
"""
import numpy as np

from . import Rng, real_to_tangent

PAULI_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
PAULI_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
PAULI_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
ID2 = np.eye(2, dtype=np.complex128)


def brickwall_perms(L, nlayers):
    ident = np.arange(L, dtype=np.int32)
    roll = np.roll(ident, 1).astype(np.int32)
    return np.stack([ident if i % 2 == 0 else roll for i in range(nlayers)])


def haar_gates(seed, nlayers):
    r = Rng(seed)
    return np.stack([r.haar_unitary() for _ in range(nlayers)])


def tangent_direction(seed, vlist):
    r = Rng(seed)
    vec = np.array([r.randn() for _ in range(16 * len(vlist))])
    Z = np.stack([real_to_tangent(vec[16 * i:16 * i + 16], vlist[i]) for i in range(len(vlist))])
    return vec, Z


def _site_op(op, j, L):
    out = np.array([[1.0 + 0j]])
    for q in range(L):
        out = np.kron(out, op if q == j else ID2)
    return out


def ising_hamiltonian(L, J, h, g):
    """Dense periodic Ising Hamiltonian (L <= 12)."""
    n = 1 << L
    idx = np.arange(n)
    bits = [((idx >> (L - 1 - q)) & 1) for q in range(L)]
    z = [1 - 2 * b for b in bits]
    diag = np.zeros(n)
    for j in range(L):
        diag += J * z[j] * z[(j + 1) % L] + g * z[j]
    H = np.diag(diag).astype(np.complex128)
    for j in range(L):
        H[idx ^ (1 << (L - 1 - j)), idx] += h
    return H


def xxz_hamiltonian(L, delta, fields):
    """Dense periodic Heisenberg-XXZ Hamiltonian with Z fields (L <= 12)."""
    n = 1 << L
    H = np.zeros((n, n), np.complex128)
    for j in range(L):
        k = (j + 1) % L
        for P, c in ((PAULI_X, 1.0), (PAULI_Y, 1.0), (PAULI_Z, delta)):
            H += c * (_site_op(P, j, L) @ _site_op(P, k, L))
        H += fields[j] * _site_op(PAULI_Z, j, L)
    return H


def expm_hermitian(H, t):
    """exp(-i t H) of a Hermitian matrix by eigendecomposition."""
    w, Q = np.linalg.eigh(H)
    return (Q * np.exp(-1j * t * w)[None, :]) @ Q.conj().T


def ising_bond(J, h, g):
    """Two-site term J ZZ + h/2 (XI + IX) + g/2 (ZI + IZ): every site belongs to one even and one odd bond."""
    return (J * np.kron(PAULI_Z, PAULI_Z) + 0.5 * h * (np.kron(PAULI_X, ID2) + np.kron(ID2, PAULI_X))
            + 0.5 * g * (np.kron(PAULI_Z, ID2) + np.kron(ID2, PAULI_Z)))


def ising_trotter_brickwall(L, J, h, g, t, steps):
    """Strang splitting [e^{-i tau/2 H_even} e^{-i tau H_odd} e^{-i tau/2 H_even}]^steps, adjacent half
    steps merged: 2*steps + 1 brickwall layers (even, odd, even, ..., even), tau = t / steps."""
    tau = t / steps
    hb = ising_bond(J, h, g)
    half, full = expm_hermitian(hb, tau / 2), expm_hermitian(hb, tau)
    layers = []
    for s in range(steps):
        layers.append(half if s == 0 else full)
        layers.append(full)
    layers.append(half)
    vlist = np.stack(layers)
    return vlist, brickwall_perms(L, len(layers))


def config_inputs(L, nlayers, seed):
    """Gates (Haar, seed), perms (periodic brickwall) and an HVP direction (seed + 1)."""
    V = haar_gates(seed, nlayers)
    P = brickwall_perms(L, nlayers)
    vec, Z = tangent_direction(seed + 1, V)
    return V, P, vec, Z
