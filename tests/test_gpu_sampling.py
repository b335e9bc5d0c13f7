"""GPU tests of the statevector utilities (normalize, transpose, Haar states) and
the stochastic-trace targets (target.h *_sampling, brickwall_opt.h
optimize_brickwall_circuit_hmat_sampling).

The library draws the samples on the host from the caller's generator in the
reference's order, so with the same seed it sees the oracle's samples: FP64
results match the oracle to 1e-12 relative, for a device-resident U (seeds
from a GEMM) and for a host callback (seeds from the callback), and the
generator is left in the same state.
"""
import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from helpers import brickwall_perms, cplx, crandn, fixture, haar_unitary, perms_of, random_perms, random_wires, reldiff, stack

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def gpu():
    R.require_gpu()
    R.set_precision(R.PRECISION_FP64)
    yield


def test_statevector_utilities_vs_oracle_and_fixture():
    d = fixture("test_transpose_statevector_cplx")
    assert np.array_equal(R.transpose_statevector(cplx(d["psi"]), d["perm"]), cplx(d["chi"]))
    rng = np.random.default_rng(80)
    psi = crandn(rng, 1 << 7)
    assert reldiff(R.normalize_statevector(psi), O.normalize_statevector(psi)) < 1e-15
    assert np.array_equal(R.normalize_statevector(np.zeros(8, np.complex128)), np.zeros(8))
    perm = rng.permutation(7)
    assert np.array_equal(R.transpose_statevector(psi, perm), O.transpose_statevector(psi, perm))
    r1, r2 = R.Rng(33), O.Rng(33)
    for L in (3, 6):
        x1, x2 = R.haar_random_statevector(L, r1), O.haar_random_statevector(L, r2)
        assert np.max(np.abs(x1 - x2)) < 1e-15
    assert r1.uint64() == r2.uint64()


@pytest.mark.parametrize("nsamples", [1, 37, 200])
def test_sampled_targets_vs_oracle(nsamples):
    """Device-resident U (seeds by GEMM) and host callback; basis-column blocks of any width
    (a last block that is only partly filled)."""
    rng = np.random.default_rng(81 + nsamples)
    L, nl = 8, 3
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    Uo = O.Ufunc.dense(U)
    V = np.stack([haar_unitary(rng) for _ in range(nl)])
    P = random_perms(rng, L, nl)
    for T in (R.DeviceTarget.dense(U), R.HostUfunc(lambda x: U @ x)):
        r, ro = R.Rng(5), O.Rng(5)
        f = R.brickwall_unitary_target_sampling(T, V, P, nsamples, r)
        f_o = O.brickwall_unitary_target_sampling(Uo, V, P, L, nsamples, ro)
        assert abs(f - f_o) < TOL * max(1, abs(f_o))
        f, dv, h = R.brickwall_unitary_target_gradient_hessian_sampling(T, V, P, nsamples, r)
        f_o, dv_o, h_o = O.brickwall_unitary_target_gradient_hessian_sampling(Uo, V, P, L, nsamples, ro)
        assert abs(f - f_o) < TOL * max(1, abs(f_o)) and reldiff(dv, dv_o) < TOL and reldiff(h, h_o) < TOL
        f, g, H = R.brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(T, V, P, nsamples, r)
        f_o, g_o, H_o = O.brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(Uo, V, P, L, nsamples, ro)
        assert reldiff(g, g_o) < TOL and reldiff(H, H_o) < TOL and np.array_equal(H, H.T)
        assert r.uint64() == ro.uint64()
    wires = random_wires(rng, 7, 5)
    gates = np.stack([haar_unitary(rng) for _ in range(5)])
    U7 = np.linalg.qr(crandn(rng, (128, 128)))[0]
    f = R.circuit_unitary_target_sampling(R.DeviceTarget.dense(U7), gates, wires, 7, nsamples, R.Rng(6))
    f_o = O.circuit_unitary_target_sampling(O.Ufunc.dense(U7), gates, wires, 7, nsamples, O.Rng(6))
    assert abs(f - f_o) < TOL * max(1, abs(f_o))


def test_sampled_targets_reference_checks():
    """The reference's statistical checks with its fixtures, seeds and 10000 samples."""
    d = fixture("test_brickwall_unitary_target_sampling_cplx")
    V, P, U = stack(d, "V", 3), perms_of(d, 3), cplx(d["U"])
    T = R.DeviceTarget.dense(U)
    f = R.brickwall_unitary_target_sampling(T, V, P, 10000, R.Rng(677))
    f_ref = R.brickwall_unitary_target(T, V, P)
    assert abs(f - f_ref) / abs(f_ref) < 0.2
    d = fixture("test_brickwall_unitary_target_gradient_vector_hessian_matrix_sampling_cplx")
    V, P, U = stack(d, "V", 5), perms_of(d, 5), cplx(d["U"])
    T = R.DeviceTarget.dense(U)
    f, g, H = R.brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(T, V, P, 10000, R.Rng(401))
    f_ref, g_ref, H_ref = R.brickwall_unitary_target_gradient_vector_hessian_matrix(T, V, P)
    assert np.array_equal(H, H.T)
    assert abs(f - f_ref) / abs(f_ref) < 0.2
    assert np.max(np.abs(g - g_ref)) < 0.3 * max(1, np.max(np.abs(g_ref)))
    assert np.max(np.abs(H - H_ref)) < 0.4 * max(1, np.max(np.abs(H_ref)))


def test_sampling_optimiser_vs_oracle():
    rng = np.random.default_rng(82)
    L, nl = 6, 3
    P = brickwall_perms(L, nl)
    U = np.linalg.qr(crandn(rng, (1 << L, 1 << L)))[0]
    V0 = np.stack([haar_unitary(rng) for _ in range(nl)])
    f_iter, vopt = R.optimize_brickwall_circuit_hmat_sampling(R.DeviceTarget.dense(U), V0, P, 3, 64, R.Rng(7))
    f_o, v_o = O.optimize_brickwall_circuit_hmat_sampling(O.Ufunc.dense(U), V0, P, L, 3, 64, O.Rng(7))
    assert np.max(np.abs(f_iter - f_o) / np.maximum(1, np.abs(f_o))) < 1e-9 and np.max(np.abs(vopt - v_o)) < 1e-8
    from rqcopt_hpc_amd import rqcopt_hpc as M
    v2, f2 = M.optimize_brickwall_circuit(L, U, list(V0), P, 3, nsamples=64, seed=7)
    assert np.max(np.abs(f2 - f_o) / np.maximum(1, np.abs(f_o))) < 1e-9 and np.max(np.abs(v2 - v_o)) < 1e-8
