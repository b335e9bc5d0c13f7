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
  device-resident U (DeviceTarget.brickwall);
- the same splitting of the XXZ chain with site-dependent fields as a general
  circuit of exact bond gates (DeviceTarget.circuit);
- the batched workload: independent problems with L in {8, 10, 12}, depth in
  {3, 5, 7}, XXZ targets (Delta = 0.5, fields uniform in [-1, 1], t uniform
  in [0.1, 1.0]) and Haar-random starting gates.
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


def xxz_bond(delta, hj, hk):
    """Two-site term XX + YY + Delta ZZ + hj/2 ZI + hk/2 IZ (every site lies in one even and one odd bond)."""
    return (np.kron(PAULI_X, PAULI_X) + np.kron(PAULI_Y, PAULI_Y) + delta * np.kron(PAULI_Z, PAULI_Z)
            + 0.5 * hj * np.kron(PAULI_Z, ID2) + 0.5 * hk * np.kron(ID2, PAULI_Z))


def xxz_trotter_circuit(L, delta, fields, t, steps):
    """Strang splitting of the periodic XXZ chain (even bonds (2k, 2k+1), odd bonds (2k+1, 2k+2 mod L)),
    adjacent even half steps merged; returns gates [ngates, 4, 4] and wires [ngates, 2]."""
    assert L % 2 == 0
    tau = t / steps
    gates, wires = [], []

    def layer(parity, dt):
        for k in range(L // 2):
            j = (2 * k + parity) % L
            q = (j + 1) % L
            gates.append(expm_hermitian(xxz_bond(delta, fields[j], fields[q]), dt))
            wires.append((j, q))

    for s in range(steps):
        layer(0, tau / 2 if s == 0 else tau)
        layer(1, tau)
    layer(0, tau / 2)
    return np.stack(gates), np.array(wires, dtype=np.int32)


def batch_config(seed, nproblems=64, qubits=(8, 10, 12), depths=(3, 5, 7), trotter_steps=4):
    """The batched workload's problems: dicts with L, depth, perms, t, fields, target (gates, wires), V0."""
    r = Rng(seed)
    probs = []
    for p in range(nproblems):
        L = qubits[int(r.randu() * len(qubits))]
        d = depths[int(r.randu() * len(depths))]
        t = 0.1 + 0.9 * r.randu()
        fields = np.array([2 * r.randu() - 1 for _ in range(L)])
        tg, tw = xxz_trotter_circuit(L, 0.5, fields, t, trotter_steps)
        V0 = np.stack([r.haar_unitary() for _ in range(d)])
        probs.append(dict(L=L, depth=d, perms=brickwall_perms(L, d), t=t, fields=fields, target_gates=tg,
                          target_wires=tw, V0=V0))
    return probs
