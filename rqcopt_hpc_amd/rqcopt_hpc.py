"""The reference's Python interface (module `rqcopt_hpc`, reference
pymodule/rqcopt_module.c:34-253, 273-551, 570-600) on top of librqcopt_hip.

    optimize_quantum_circuit(nqubits, U, gates_start, wires, niter) -> (gates_opt, f_iter)
    optimize_brickwall_circuit(L, U, Vlist_start, perms, niter, nsamples=0, seed=42) -> (Vlist_opt, f_iter)

Arguments, checks, messages and exception types follow the reference module:
SyntaxError for arguments that are not sequences or not contiguous, ValueError
for sizes and indices out of range, TypeError for a U that is not complex
double.  The reference applies U on the host column by column; here U is copied
to the GPU once per call and stays resident while the optimiser runs.  The
brickwall optimiser uses the full Riemannian Hessian (as the reference module
does); the general-circuit optimiser uses Hessian-vector products
(reference src/circuit_opt.c).  The sampling variant (nsamples > 0) estimates
the trace from nsamples Haar-random states drawn from the reference's
generator seeded with `seed`, so the samples are the reference's.
"""
import ctypes as ct
from collections.abc import Sequence

import numpy as np

from . import DeviceTarget, Rng, _p, _perms, _target_args, _wires, lib, rtr_default_params


def _check_u(U, nqubits, name):
    if not isinstance(U, np.ndarray):
        raise SyntaxError("error parsing input; 'U' must be a numpy array")
    n = 1 << nqubits
    if U.ndim != 2:
        raise ValueError(f"'U' must have degree 2, received {U.ndim}")
    if U.shape != (n, n):
        raise ValueError(f"'U' must have dimensions 2^{name} x 2^{name}, received a {U.shape[0]} x {U.shape[1]} "
                         f"matrix and {name} = {nqubits}")
    if U.dtype != np.complex128:
        raise TypeError("'U' must have 'complex double' entries")
    if not U.flags["C_CONTIGUOUS"]:
        raise SyntaxError("'U' does not have contiguous C storage format")


def _read_gates(seq, name):
    gates = []
    for i in range(len(seq)):
        try:
            g = np.ascontiguousarray(seq[i], dtype=np.complex128)
        except (TypeError, ValueError):
            raise SyntaxError(f"cannot interpret {name}[{i}] as complex matrix") from None
        if g.ndim != 2:
            raise SyntaxError(f"cannot interpret {name}[{i}] as complex matrix")
        if g.shape != (4, 4):
            raise ValueError(f"{name}[{i}] must be a 4 x 4 matrix, received a {g.shape[0]} x {g.shape[1]} matrix")
        gates.append(g)
    return np.stack(gates)


def _read_int(x, what):
    try:
        return int(x)
    except (TypeError, ValueError):
        raise ValueError(f"cannot interpret '{what}' as integer") from None


def optimize_quantum_circuit(nqubits, U, gates_start, wires, niter):
    """Optimise the two-qubit gates of a circuit to approximate U (reference pymodule/rqcopt_module.c:34-253)."""
    if not isinstance(nqubits, int) or not isinstance(niter, int):
        raise SyntaxError("error parsing input; syntax: optimize_quantum_circuit(nqubits, U, gates_start, wires, niter)")
    if nqubits <= 0:
        raise ValueError(f"'nqubits' must be positive, received nqubits = {nqubits}")
    _check_u(U, nqubits, "nqubits")
    if not isinstance(gates_start, (Sequence, np.ndarray)):
        raise SyntaxError("'gates_start' must be a sequence")
    ngates = len(gates_start)
    if ngates <= 0:
        raise ValueError("sequence of initial quantum gates cannot be empty")
    if not isinstance(wires, (Sequence, np.ndarray)):
        raise SyntaxError("'wires' must be a sequence")
    if len(wires) != ngates:
        raise ValueError("'gates_start' and 'wires' sequences must have the same length")
    if niter <= 0:
        raise ValueError(f"'niter' must be a positive integer, received {niter}")
    gates = _read_gates(gates_start, "gates_start")
    w = np.empty((ngates, 2), np.int32)
    for i in range(ngates):
        if not isinstance(wires[i], (Sequence, np.ndarray)):
            raise SyntaxError(f"'wires[{i}]' must be a sequence")
        if len(wires[i]) != 2:
            raise ValueError(f"'wires[{i}]' must be a sequence of length 2")
        for j in range(2):
            w[i, j] = _read_int(wires[i][j], f"wires[{i}][{j}]")
            if not 0 <= w[i, j] < nqubits:
                raise ValueError(f"index 'wires[{i}][{j}] = {w[i, j]}' outside of [0, nqubits) index range")
        if w[i, 0] == w[i, 1]:
            raise ValueError(f"wire indices for gate {i} are not distinct, received index {w[i, 0]} twice")
    target = DeviceTarget.dense(U)
    try:
        params = rtr_default_params(16 * ngates)
        f_iter = np.empty(niter + 1)
        gopt = np.empty_like(gates)
        fp, ud = _target_args(target)
        wv = _wires(w)
        lib().optimize_quantum_circuit(fp, ud, _p(gates), ngates, nqubits, _p(wv), ct.byref(params), niter, _p(f_iter),
                                       _p(gopt))
    finally:
        target.close()
    return gopt, f_iter


def optimize_brickwall_circuit(L, U, Vlist_start, perms, niter, nsamples=0, seed=42):
    """Optimise the gates of a brickwall circuit to approximate U (reference pymodule/rqcopt_module.c:273-551)."""
    if not isinstance(L, int) or not isinstance(niter, int) or not isinstance(nsamples, int) or not isinstance(seed, int):
        raise SyntaxError("error parsing input; syntax: optimize_brickwall_circuit(L, U, Vlist_start, perms, niter, "
                          "nsamples=0, seed=42)")
    if L <= 0 or L % 2 != 0:
        raise ValueError(f"'L' must be positive and even, received L = {L}")
    _check_u(U, L, "L")
    if not isinstance(Vlist_start, (Sequence, np.ndarray)):
        raise SyntaxError("'Vlist_start' must be a sequence")
    nlayers = len(Vlist_start)
    if nlayers <= 0:
        raise ValueError("sequence of initial quantum gates cannot be empty")
    if not isinstance(perms, (Sequence, np.ndarray)):
        raise SyntaxError("'perms' must be a sequence")
    if len(perms) != nlayers:
        raise ValueError("'Vlist_start' and 'perms' sequences must have the same length (equal to the number of layers)")
    if niter <= 0:
        raise ValueError(f"'niter' must be a positive integer, received {niter}")
    if nsamples < 0:
        raise ValueError("'nsamples' cannot be negative")
    V = _read_gates(Vlist_start, "Vlist_start")
    P = np.empty((nlayers, L), np.int32)
    for i in range(nlayers):
        if not isinstance(perms[i], (Sequence, np.ndarray)):
            raise SyntaxError(f"'perms[{i}]' must be a sequence")
        if len(perms[i]) != L:
            raise ValueError(f"'perms[{i}]' must be a sequence of length L = {L}")
        for j in range(L):
            P[i, j] = _read_int(perms[i][j], f"perms[{i}][{j}]")
        if sorted(P[i].tolist()) != list(range(L)):
            raise ValueError(f"'perms[{i}]' is not a valid permutation")
    target = DeviceTarget.dense(U)
    try:
        params = rtr_default_params(16 * nlayers)
        f_iter = np.empty(niter + 1)
        vopt = np.empty_like(V)
        fp, ud = _target_args(target)
        prm, rows = _perms(P)
        if nsamples == 0:
            lib().optimize_brickwall_circuit_hmat(fp, ud, _p(V), nlayers, L, rows, ct.byref(params), niter, _p(f_iter),
                                                  _p(vopt))
        else:
            rng = Rng(seed)
            lib().optimize_brickwall_circuit_hmat_sampling.argtypes = [
                ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_long, ct.c_void_p,
                ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p]
            lib().optimize_brickwall_circuit_hmat_sampling(fp, ud, _p(V), nlayers, L, ct.cast(rows, ct.c_void_p), nsamples,
                                                           ct.addressof(rng.s), ct.addressof(params), niter, _p(f_iter),
                                                           _p(vopt))
    finally:
        target.close()
    return vopt, f_iter
