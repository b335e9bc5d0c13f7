"""Drop-in check on the GPU: the reference's own optimiser glue drives the library.

oracle/_ref/libdropin_opt.so is the reference's src/brickwall_opt.c and src/circuit_opt.c,
compiled unmodified against this library's headers and linked against librqcopt_hip.so (built in
the container that holds the reference checkout, oracle/Makefile target dropin; the reference
itself is not read here).  Its optimize_* functions call the library's target, gradient, HVP,
Hessian and trust-region entry points exactly as the reference calls its own, so their
trajectories must be those of the reference's example runs and of the oracle (1e-9 in f_iter,
1e-8 in the gates; the molecular run at its measured sensitivity, tests/helpers.py), and equal to
the library's own glue.
"""
import ctypes as ct
import os

import numpy as np
import pytest

import rqcopt_hpc_amd as R
from helpers import assert_molecular_trajectory, cplx, example_ising1d, example_molecular, fixture

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "oracle", "_ref", "libdropin_opt.so")


def _dropin():
    if not os.path.exists(SO):
        pytest.skip("oracle/_ref/libdropin_opt.so not built (needs the reference checkout at build time)")
    R.lib()
    return ct.CDLL(SO)


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(1, np.abs(np.asarray(b))))


def _brickwall(fn, T, V, P, niter):
    V = np.ascontiguousarray(V, np.complex128)
    perms, rows = R._perms(P)
    prm = R.rtr_default_params(16 * len(V))
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(V)
    fp, ud = R._target_args(T)
    fn(fp, ud, R._p(V), ct.c_int(len(V)), ct.c_int(perms.shape[1]), rows, ct.byref(prm), ct.c_int(niter), R._p(f_iter),
       R._p(vopt))
    return f_iter, vopt


def test_reference_brickwall_glue_hvp_on_library():
    from rqcopt_hpc_amd.workloads import expm_hermitian, ising_hamiltonian
    d = fixture("config0_ising_L8_d5")
    T = R.DeviceTarget.dense(expm_hermitian(ising_hamiltonian(8, 1.0, 0.75, 0.6), 0.25))
    f_iter, vopt = _brickwall(_dropin().optimize_brickwall_circuit_hvp, T, d["V"], d["P"], 10)
    print("reference glue (hvp) vs oracle trajectory", _rel(f_iter, d["f_iter"]), flush=True)
    assert _rel(f_iter, d["f_iter"]) < 1e-9 and np.max(np.abs(vopt - d["vopt"])) < 1e-8
    f_own, v_own = R.optimize_brickwall_circuit_hvp(T, d["V"], d["P"], 10)
    assert np.array_equal(f_iter, f_own) and np.array_equal(vopt, v_own)
    T.close()


def test_reference_brickwall_glue_hmat_on_library():
    U, V, P, run, _ = example_ising1d()
    T = R.DeviceTarget.dense(U)
    f_iter, vopt = _brickwall(_dropin().optimize_brickwall_circuit_hmat, T, V, P, 10)
    print("reference glue (hmat) vs reference example", _rel(f_iter, run["f_iter"]), flush=True)
    assert _rel(f_iter, run["f_iter"]) < 1e-9 and np.max(np.abs(vopt - cplx(run["Vlist"]))) < 1e-8
    T.close()


def test_reference_circuit_glue_on_library():
    U, G0, W, run = example_molecular()
    T = R.DeviceTarget.dense(U)
    G0 = np.ascontiguousarray(G0)
    w = np.ascontiguousarray(W, np.int32).reshape(-1)
    prm = R.rtr_default_params(16 * len(G0))
    f_iter = np.empty(11)
    gopt = np.empty_like(G0)
    fp, ud = R._target_args(T)
    _dropin().optimize_quantum_circuit(fp, ud, R._p(G0), ct.c_int(len(G0)), ct.c_int(5), R._p(w), ct.byref(prm),
                                       ct.c_int(10), R._p(f_iter), R._p(gopt))
    # the molecular run's late iterates move under rounding-level changes (tests/helpers.py)
    assert_molecular_trajectory(f_iter, gopt, run)
    f_own, g_own = R.optimize_quantum_circuit(T, G0, W, 5, 10)
    assert np.array_equal(f_iter, f_own) and np.array_equal(gopt, g_own)
    T.close()
