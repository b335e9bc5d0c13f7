"""The reference's Python interface (module rqcopt_hpc, reference pymodule/rqcopt_module.c):
argument checking on the CPU (exception types and messages of the reference), results on the
GPU against the oracle's optimisers."""
import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from rqcopt_hpc_amd import rqcopt_hpc as M
from helpers import brickwall_matrix, brickwall_perms, crandn, haar_unitary, random_wires


def _u(L, rng):
    return np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]


def test_argument_checks_follow_the_reference():
    rng = np.random.default_rng(1)
    U = _u(4, rng)
    G = [haar_unitary(rng) for _ in range(2)]
    with pytest.raises(ValueError, match="'nqubits' must be positive"):
        M.optimize_quantum_circuit(0, U, G, [(0, 1), (1, 2)], 3)
    with pytest.raises(ValueError, match="dimensions 2\\^nqubits x 2\\^nqubits"):
        M.optimize_quantum_circuit(3, U, G, [(0, 1), (1, 2)], 3)
    with pytest.raises(TypeError, match="complex double"):
        M.optimize_quantum_circuit(4, U.real.copy(), G, [(0, 1), (1, 2)], 3)
    with pytest.raises(SyntaxError, match="contiguous"):
        M.optimize_quantum_circuit(4, np.asfortranarray(U), G, [(0, 1), (1, 2)], 3)
    with pytest.raises(ValueError, match="cannot be empty"):
        M.optimize_quantum_circuit(4, U, [], [], 3)
    with pytest.raises(ValueError, match="same length"):
        M.optimize_quantum_circuit(4, U, G, [(0, 1)], 3)
    with pytest.raises(ValueError, match="outside of \\[0, nqubits\\)"):
        M.optimize_quantum_circuit(4, U, G, [(0, 1), (1, 4)], 3)
    with pytest.raises(ValueError, match="not distinct"):
        M.optimize_quantum_circuit(4, U, G, [(0, 1), (2, 2)], 3)
    with pytest.raises(ValueError, match="4 x 4"):
        M.optimize_quantum_circuit(4, U, [np.eye(2)] * 2, [(0, 1), (1, 2)], 3)
    with pytest.raises(ValueError, match="'niter' must be a positive integer"):
        M.optimize_quantum_circuit(4, U, G, [(0, 1), (1, 2)], 0)
    P = brickwall_perms(4, 2)
    with pytest.raises(ValueError, match="positive and even"):
        M.optimize_brickwall_circuit(3, _u(3, rng), G, P, 3)
    with pytest.raises(ValueError, match="same length"):
        M.optimize_brickwall_circuit(4, U, G, P[:1], 3)
    with pytest.raises(ValueError, match="not a valid permutation"):
        M.optimize_brickwall_circuit(4, U, G, [[0, 1, 2, 3], [0, 0, 1, 2]], 3)
    with pytest.raises(ValueError, match="sequence of length L = 4"):
        M.optimize_brickwall_circuit(4, U, G, [[0, 1, 2, 3], [0, 1, 2]], 3)
    with pytest.raises(ValueError, match="'nsamples' cannot be negative"):
        M.optimize_brickwall_circuit(4, U, G, P, 3, nsamples=-1)
    with pytest.raises(SyntaxError, match="must be a sequence"):
        M.optimize_brickwall_circuit(4, U, 5, P, 3)


@pytest.mark.gpu
def test_optimisers_match_the_oracle():
    R.require_gpu()
    rng = np.random.default_rng(2)
    L, nl = 6, 3
    P = brickwall_perms(L, nl)
    U = brickwall_matrix(np.stack([haar_unitary(rng) for _ in range(nl)]), P)
    V0 = [haar_unitary(rng) for _ in range(nl)]
    vopt, f_iter = M.optimize_brickwall_circuit(L, U, V0, P, 3)
    f_o, v_o = O.optimize_brickwall_circuit_hmat(O.Ufunc.dense(U), np.stack(V0), P, L, 3)
    assert np.max(np.abs(f_iter - f_o) / np.maximum(1, np.abs(f_o))) < 1e-9 and np.max(np.abs(vopt - v_o)) < 1e-8
    assert f_iter[-1] < f_iter[0]
    L = 5
    U = _u(L, rng)
    wires = random_wires(rng, L, 6)
    G0 = [haar_unitary(rng) for _ in range(6)]
    gopt, f_iter = M.optimize_quantum_circuit(L, U, G0, wires.tolist(), 3)
    f_o, g_o = O.optimize_quantum_circuit(O.Ufunc.dense(U), np.stack(G0), wires, L, 3)
    assert np.max(np.abs(f_iter - f_o) / np.maximum(1, np.abs(f_o))) < 1e-9 and np.max(np.abs(gopt - g_o)) < 1e-8
    assert f_iter[-1] < f_iter[0]
