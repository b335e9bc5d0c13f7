"""ORACLE — test infrastructure only.

ctypes binding of the C restatement in oracle/oracle.c (built as
oracle/liboracle.so by `make -C oracle`).  Only tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py may import this package, and only as the
checker.  Nothing in rqcopt_hpc_amd/ imports it.

Arrays are numpy complex128 (C-contiguous).  Gates are (..., 4, 4) arrays;
perms are (nlayers, L) int32 arrays; wires are (ngates, 2) int32 arrays.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_c = ct.c_void_p
_i = ct.c_int
_l = ct.c_long
_d = ct.c_double

UFUNC = ct.CFUNCTYPE(ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p)


class CDouble2(ct.Structure):
    """double complex return value (two doubles in xmm0:xmm1 on SysV x86-64)."""
    _fields_ = [("re", ct.c_double), ("im", ct.c_double)]


class BrickwallOp(ct.Structure):
    _fields_ = [("vlist", ct.c_void_p), ("perms", ct.c_void_p), ("nlayers", ct.c_int)]


class OrcRng(ct.Structure):
    _fields_ = [("s0", ct.c_uint64), ("i0", ct.c_uint64), ("s1", ct.c_uint64), ("i1", ct.c_uint64)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ct.CDLL(path)
        L.orc_ufunc_dense.restype = ct.c_int
        L.orc_ufunc_scramble.restype = ct.c_int
        L.orc_randu.restype = _d
        L.orc_randn.restype = _d
        L.orc_crandn.restype = CDouble2
        L.orc_rand_uint64.restype = ct.c_uint64
        L.orc_rand_uint32.restype = ct.c_uint32
        L.orc_seed_rng_state.argtypes = [ct.c_uint64, ct.POINTER(OrcRng)]
        L.orc_truncated_cg_hvp.restype = ct.c_bool
        L.orc_inverse_matrix.restype = ct.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _cplx(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Ufunc:
    """Target operator for the oracle: a dense U, the reference test map, or a Python callable."""

    def __init__(self, kind, payload=None):
        self.kind = kind
        self.payload = payload
        self._keep = None

    @staticmethod
    def dense(U):
        return Ufunc("dense", _cplx(U))

    @staticmethod
    def scramble():
        return Ufunc("scramble")

    @staticmethod
    def brickwall(vlist, perms):
        """U applied matrix-free as a brickwall circuit (cheap columns for large L)."""
        vl, pm = _cplx(vlist), _i32(perms)
        u = Ufunc("brickwall", (vl, pm))
        u._keep = BrickwallOp(vl.ctypes.data, pm.ctypes.data, len(vl))
        return u

    def args(self):
        L = lib()
        if self.kind == "dense":
            fp = ct.cast(L.orc_ufunc_dense, ct.c_void_p)
            return fp, _p(self.payload)
        if self.kind == "scramble":
            return ct.cast(L.orc_ufunc_scramble, ct.c_void_p), None
        if self.kind == "brickwall":
            return ct.cast(L.orc_ufunc_brickwall, ct.c_void_p), ct.cast(ct.byref(self._keep), ct.c_void_p)
        raise ValueError(self.kind)


# ---------------------------------------------------------------- gate level
def apply_gate(G, i, j, psi):
    psi = _cplx(psi)
    out = np.empty_like(psi)
    L = int(np.log2(psi.size))
    lib().orc_apply_gate(_p(_cplx(G)), _i(i), _i(j), _i(L), _p(psi), _p(out))
    return out


def apply_gate_backward(G, i, j, psi, dpsi_out):
    psi = _cplx(psi)
    dpsi_out = _cplx(dpsi_out)
    dpsi = np.empty_like(psi)
    dG = np.empty((4, 4), np.complex128)
    L = int(np.log2(psi.size))
    lib().orc_apply_gate_backward(_p(_cplx(G)), _i(i), _i(j), _i(L), _p(psi), _p(dpsi_out), _p(dpsi), _p(dG))
    return dpsi, dG


def apply_gate_1q(V, i, psi):
    """Single-qubit gate V (2x2) on wire i (no reference counterpart; two-qubit conventions)."""
    psi = _cplx(psi)
    out = np.empty_like(psi)
    L = int(np.log2(psi.size))
    lib().orc_apply_gate_1q(_p(_cplx(V)), _i(i), _i(L), _p(psi), _p(out))
    return out


def apply_gate_1q_backward(V, i, psi, dpsi_out):
    psi = _cplx(psi)
    dpsi_out = _cplx(dpsi_out)
    dpsi = np.empty_like(psi)
    dV = np.empty((2, 2), np.complex128)
    L = int(np.log2(psi.size))
    lib().orc_apply_gate_1q_backward(_p(_cplx(V)), _i(i), _i(L), _p(psi), _p(dpsi_out), _p(dpsi), _p(dV))
    return dpsi, dV


def apply_gate_to_array(G, i, j, psi):
    """psi: (2^L, nstates) interleaved array."""
    psi = _cplx(psi)
    out = np.empty_like(psi)
    L = int(np.log2(psi.shape[0]))
    lib().orc_apply_gate_to_array(_p(_cplx(G)), _i(i), _i(j), _i(L), _i(psi.shape[1]), _p(psi), _p(out))
    return out


# ------------------------------------------------------------- circuit level
def apply_quantum_circuit(gates, wires, psi):
    gates, wires, psi = _cplx(gates), _i32(wires), _cplx(psi)
    out = np.empty_like(psi)
    L = int(np.log2(psi.size))
    lib().orc_apply_quantum_circuit(_p(gates), _i(len(gates)), _p(wires), _i(L), _p(psi), _p(out))
    return out


def quantum_circuit_forward(gates, wires, psi):
    gates, wires, psi = _cplx(gates), _i32(wires), _cplx(psi)
    L = int(np.log2(psi.size))
    cache = np.empty((len(gates), psi.size), np.complex128)
    out = np.empty_like(psi)
    lib().orc_quantum_circuit_forward(_p(gates), _i(len(gates)), _p(wires), _i(L), _p(psi), _p(cache), _p(out))
    return out, cache


def quantum_circuit_backward(gates, wires, cache, dpsi_out):
    gates, wires, cache, dpsi_out = _cplx(gates), _i32(wires), _cplx(cache), _cplx(dpsi_out)
    L = int(np.log2(dpsi_out.size))
    dpsi = np.empty_like(dpsi_out)
    dgates = np.empty((len(gates), 4, 4), np.complex128)
    lib().orc_quantum_circuit_backward(_p(gates), _i(len(gates)), _p(wires), _i(L), _p(cache), _p(dpsi_out),
                                       _p(dpsi), _p(dgates))
    return dpsi, dgates


def quantum_circuit_gates_hvp(gates, gatedirs, wires, psi, phi):
    gates, gatedirs, wires, psi, phi = _cplx(gates), _cplx(gatedirs), _i32(wires), _cplx(psi), _cplx(phi)
    L = int(np.log2(psi.size))
    out = np.empty_like(psi)
    dgates = np.empty((len(gates), 4, 4), np.complex128)
    hess = np.empty((len(gates), 4, 4), np.complex128)
    lib().orc_quantum_circuit_gates_hessian_vector_product(_p(gates), _p(gatedirs), _i(len(gates)), _p(wires), _i(L),
                                                           _p(psi), _p(phi), _p(out), _p(dgates), _p(hess))
    return out, dgates, hess


def apply_brickwall_unitary(vlist, perms, psi):
    vlist, perms, psi = _cplx(vlist), _i32(perms), _cplx(psi)
    L = int(np.log2(psi.size))
    out = np.empty_like(psi)
    lib().orc_apply_brickwall_unitary(_p(vlist), _i(len(vlist)), _p(perms), _i(L), _p(psi), _p(out))
    return out


def brickwall_unitary_forward(vlist, perms, psi):
    vlist, perms, psi = _cplx(vlist), _i32(perms), _cplx(psi)
    L = int(np.log2(psi.size))
    cache = np.empty((len(vlist) * (L // 2), psi.size), np.complex128)
    out = np.empty_like(psi)
    lib().orc_brickwall_unitary_forward(_p(vlist), _i(len(vlist)), _p(perms), _i(L), _p(psi), _p(cache), _p(out))
    return out, cache


def brickwall_unitary_backward(vlist, perms, cache, dpsi_out):
    vlist, perms, cache, dpsi_out = _cplx(vlist), _i32(perms), _cplx(cache), _cplx(dpsi_out)
    L = int(np.log2(dpsi_out.size))
    dpsi = np.empty_like(dpsi_out)
    dvlist = np.empty((len(vlist), 4, 4), np.complex128)
    lib().orc_brickwall_unitary_backward(_p(vlist), _i(len(vlist)), _p(perms), _i(L), _p(cache), _p(dpsi_out),
                                         _p(dpsi), _p(dvlist))
    return dpsi, dvlist


# ------------------------------------------------------------------- targets
def _fval():
    return np.zeros(1, np.complex128)


def circuit_unitary_target(uf, gates, wires, L):
    gates, wires = _cplx(gates), _i32(wires)
    f = _fval()
    fp, ud = uf.args()
    r = lib().orc_circuit_unitary_target(fp, ud, _p(gates), _i(len(gates)), _p(wires), _i(L), _p(f))
    assert r == 0
    return f[0]


def circuit_unitary_target_and_gradient(uf, gates, wires, L):
    gates, wires = _cplx(gates), _i32(wires)
    f = _fval()
    dg = np.empty((len(gates), 4, 4), np.complex128)
    fp, ud = uf.args()
    r = lib().orc_circuit_unitary_target_and_gradient(fp, ud, _p(gates), _i(len(gates)), _p(wires), _i(L), _p(f), _p(dg))
    assert r == 0
    return f[0], dg


def circuit_unitary_target_and_projected_gradient(uf, gates, wires, L):
    gates, wires = _cplx(gates), _i32(wires)
    f = _fval()
    g = np.empty(16 * len(gates))
    fp, ud = uf.args()
    r = lib().orc_circuit_unitary_target_and_projected_gradient(fp, ud, _p(gates), _i(len(gates)), _p(wires), _i(L),
                                                               _p(f), _p(g))
    assert r == 0
    return f[0], g


def circuit_unitary_target_hvp(uf, gates, gatedirs, wires, L):
    gates, gatedirs, wires = _cplx(gates), _cplx(gatedirs), _i32(wires)
    f = _fval()
    dg = np.empty((len(gates), 4, 4), np.complex128)
    h = np.empty((len(gates), 4, 4), np.complex128)
    fp, ud = uf.args()
    r = lib().orc_circuit_unitary_target_hessian_vector_product(fp, ud, _p(gates), _p(gatedirs), _i(len(gates)),
                                                               _p(wires), _i(L), _p(f), _p(dg), _p(h))
    assert r == 0
    return f[0], dg, h


def circuit_unitary_target_projected_hvp(uf, gates, gatedirs, wires, L):
    gates, gatedirs, wires = _cplx(gates), _cplx(gatedirs), _i32(wires)
    f = _fval()
    g = np.empty(16 * len(gates))
    h = np.empty(16 * len(gates))
    fp, ud = uf.args()
    r = lib().orc_circuit_unitary_target_projected_hessian_vector_product(fp, ud, _p(gates), _p(gatedirs),
                                                                         _i(len(gates)), _p(wires), _i(L), _p(f),
                                                                         _p(g), _p(h))
    assert r == 0
    return f[0], g, h


def brickwall_unitary_target(uf, vlist, perms, L):
    vlist, perms = _cplx(vlist), _i32(perms)
    f = _fval()
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target(fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms), _p(f))
    assert r == 0
    return f[0]


def brickwall_unitary_target_and_gradient(uf, vlist, perms, L):
    vlist, perms = _cplx(vlist), _i32(perms)
    f = _fval()
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_and_gradient(fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms), _p(f), _p(dv))
    assert r == 0
    return f[0], dv


def brickwall_unitary_target_and_projected_gradient(uf, vlist, perms, L):
    vlist, perms = _cplx(vlist), _i32(perms)
    f = _fval()
    g = np.empty(16 * len(vlist))
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_and_projected_gradient(fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms),
                                                                 _p(f), _p(g))
    assert r == 0
    return f[0], g


def brickwall_unitary_target_hvp(uf, vlist, vdirs, perms, L):
    vlist, vdirs, perms = _cplx(vlist), _cplx(vdirs), _i32(perms)
    f = _fval()
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    h = np.empty((len(vlist), 4, 4), np.complex128)
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_hessian_vector_product(fp, ud, _p(vlist), _p(vdirs), _i(len(vlist)),
                                                                 _p(perms), _i(L), _p(f), _p(dv), _p(h))
    assert r == 0
    return f[0], dv, h


def brickwall_hvp_columns(uf, vlist, vdirs, perms, L, b0, b1):
    vlist, vdirs, perms = _cplx(vlist), _cplx(vdirs), _i32(perms)
    f = _fval()
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    h = np.empty((len(vlist), 4, 4), np.complex128)
    fp, ud = uf.args()
    r = lib().orc_brickwall_hvp_columns(fp, ud, _p(vlist), _p(vdirs), _i(len(vlist)), _p(perms), _i(L), _l(b0), _l(b1),
                                        _p(f), _p(dv), _p(h))
    assert r == 0
    return f[0], dv, h


def brickwall_project_grad_hvp(vlist, vdirs, dvlist, hess_vdirs):
    """Projection step of the brickwall projected HVP applied to column sums (reference src/target.c:645-671)."""
    vlist, vdirs = _cplx(vlist), _cplx(vdirs)
    dv, h = np.array(dvlist, np.complex128), np.array(hess_vdirs, np.complex128)
    g = np.empty(16 * len(vlist))
    hv = np.empty(16 * len(vlist))
    lib().orc_brickwall_project_grad_hvp(_p(vlist), _p(vdirs), _i(len(vlist)), _p(dv), _p(h), _p(g), _p(hv))
    return g, hv


def brickwall_unitary_target_projected_hvp(uf, vlist, vdirs, perms, L):
    vlist, vdirs, perms = _cplx(vlist), _cplx(vdirs), _i32(perms)
    f = _fval()
    g = np.empty(16 * len(vlist))
    h = np.empty(16 * len(vlist))
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_projected_hessian_vector_product(fp, ud, _p(vlist), _p(vdirs),
                                                                           _i(len(vlist)), _p(perms), _i(L), _p(f),
                                                                           _p(g), _p(h))
    assert r == 0
    return f[0], g, h


def optimize_brickwall_circuit_hvp(uf, vlist_start, perms, L, niter):
    vlist_start, perms = _cplx(vlist_start), _i32(perms)
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(vlist_start)
    fp, ud = uf.args()
    lib().orc_optimize_brickwall_circuit_hvp(fp, ud, _p(vlist_start), _i(len(vlist_start)), _i(L), _p(perms),
                                             _i(niter), _p(f_iter), _p(vopt))
    return f_iter, vopt


def optimize_quantum_circuit(uf, gates_start, wires, L, niter):
    gates_start, wires = _cplx(gates_start), _i32(wires)
    f_iter = np.empty(niter + 1)
    gopt = np.empty_like(gates_start)
    fp, ud = uf.args()
    lib().orc_optimize_quantum_circuit(fp, ud, _p(gates_start), _i(len(gates_start)), _i(L), _p(wires), _i(niter),
                                       _p(f_iter), _p(gopt))
    return f_iter, gopt


# ------------------------------------------------------------------- statevector utilities, sampled targets
def normalize_statevector(psi):
    out = _cplx(psi).copy()
    lib().orc_normalize_statevector(_i(int(np.log2(out.size))), _p(out))
    return out


def transpose_statevector(psi, perm):
    psi = _cplx(psi)
    out = np.empty_like(psi)
    lib().orc_transpose_statevector(_i(int(np.log2(psi.size))), _p(psi), _p(_i32(perm)), _p(out))
    return out


def haar_random_statevector(L, rng):
    out = np.empty(1 << L, np.complex128)
    lib().orc_haar_random_statevector(_i(L), ct.byref(rng.s), _p(out))
    return out


def brickwall_unitary_target_sampling(uf, vlist, perms, L, nsamples, rng):
    vlist, perms = _cplx(vlist), _i32(perms)
    f = _fval()
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_sampling(fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms), _l(nsamples),
                                                    ct.byref(rng.s), _p(f))
    assert r == 0
    return f[0]


def circuit_unitary_target_sampling(uf, gates, wires, L, nsamples, rng):
    gates, wires = _cplx(gates), _i32(wires)
    f = _fval()
    fp, ud = uf.args()
    r = lib().orc_circuit_unitary_target_sampling(fp, ud, _p(gates), _i(len(gates)), _p(wires), _i(L), _l(nsamples),
                                                  ct.byref(rng.s), _p(f))
    assert r == 0
    return f[0]


def brickwall_unitary_target_gradient_hessian_sampling(uf, vlist, perms, L, nsamples, rng):
    vlist, perms = _cplx(vlist), _i32(perms)
    m = 16 * len(vlist)
    f = _fval()
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    hess = np.empty((m, m), np.complex128)
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_gradient_hessian_sampling(fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms),
                                                                    _l(nsamples), ct.byref(rng.s), _p(f), _p(dv), _p(hess))
    assert r == 0
    return f[0], dv, hess


def brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(uf, vlist, perms, L, nsamples, rng):
    vlist, perms = _cplx(vlist), _i32(perms)
    m = 16 * len(vlist)
    f = _fval()
    g = np.empty(m)
    H = np.empty((m, m))
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(
        fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms), _l(nsamples), ct.byref(rng.s), _p(f), _p(g), _p(H))
    assert r == 0
    return f[0], g, H


def optimize_brickwall_circuit_hmat_sampling(uf, vlist_start, perms, L, niter, nsamples, rng):
    vlist_start, perms = _cplx(vlist_start), _i32(perms)
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(vlist_start)
    fp, ud = uf.args()
    lib().orc_optimize_brickwall_circuit_hmat_sampling(fp, ud, _p(vlist_start), _i(len(vlist_start)), _i(L), _p(perms),
                                                       _l(nsamples), ct.byref(rng.s), _i(niter), _p(f_iter), _p(vopt))
    return f_iter, vopt


# ------------------------------------------------------------------- full-Hessian route
def apply_gate_placeholder(i, j, psi):
    psi = _cplx(psi)
    L = int(np.log2(psi.size))
    out = np.empty((psi.size, 16), np.complex128)
    lib().orc_apply_gate_placeholder(_i(i), _i(j), _i(L), _p(psi), _p(out))
    return out


def apply_gate_backward_array(i, j, psi, dpsi_out):
    psi, dpsi_out = _cplx(psi), _cplx(dpsi_out)
    L = int(np.log2(dpsi_out.size))
    ns = psi.shape[1]
    dg = np.empty((ns, 4, 4), np.complex128)
    lib().orc_apply_gate_backward_array(_i(i), _i(j), _i(L), _i(ns), _p(psi), _p(dpsi_out), _p(dg))
    return dg


def brickwall_unitary_backward_hessian(vlist, perms, cache, dpsi_out):
    vlist, perms, cache, dpsi_out = _cplx(vlist), _i32(perms), _cplx(cache), _cplx(dpsi_out)
    L = perms.shape[1]
    m = 16 * len(vlist)
    dpsi = np.empty_like(dpsi_out)
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    hess = np.empty((m, m), np.complex128)
    r = lib().orc_brickwall_unitary_backward_hessian(_p(vlist), _i(len(vlist)), _p(perms), _i(L), _p(cache),
                                                      _p(dpsi_out), _p(dpsi), _p(dv), _p(hess))
    assert r == 0
    return dpsi, dv, hess


def brickwall_unitary_target_gradient_hessian(uf, vlist, perms, L):
    vlist, perms = _cplx(vlist), _i32(perms)
    m = 16 * len(vlist)
    f = _fval()
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    hess = np.empty((m, m), np.complex128)
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_gradient_hessian(fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms), _p(f),
                                                           _p(dv), _p(hess))
    assert r == 0
    return f[0], dv, hess


def brickwall_unitary_target_gradient_vector_hessian_matrix(uf, vlist, perms, L):
    vlist, perms = _cplx(vlist), _i32(perms)
    m = 16 * len(vlist)
    f = _fval()
    g = np.empty(m)
    H = np.empty((m, m))
    fp, ud = uf.args()
    r = lib().orc_brickwall_unitary_target_gradient_vector_hessian_matrix(fp, ud, _p(vlist), _i(len(vlist)), _i(L),
                                                                         _p(perms), _p(f), _p(g), _p(H))
    assert r == 0
    return f[0], g, H


# ---------------------------------------------------------------- block-encoding targets
def _blockenc_call(name, hf, vlist, perms, L, *outs):
    vlist, perms = _cplx(vlist), _i32(perms)
    f = np.empty(1)
    fp, ud = hf.args()
    r = getattr(lib(), name)(fp, ud, _p(vlist), _i(len(vlist)), _i(L), _p(perms), _p(f), *[_p(o) for o in outs])
    assert r == 0, (name, r)
    return float(f[0])


def brickwall_blockenc_target(hf, vlist, perms, L):
    """f = ||P^dagger W P - H||_F^2 / 2; hf applies H to L/2-qubit states."""
    return _blockenc_call("orc_brickwall_blockenc_target", hf, vlist, perms, L)


def brickwall_blockenc_target_and_gradient(hf, vlist, perms, L):
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    return _blockenc_call("orc_brickwall_blockenc_target_and_gradient", hf, vlist, perms, L, dv), dv


def brickwall_blockenc_target_and_gradient_vector(hf, vlist, perms, L):
    g = np.empty(16 * len(vlist))
    return _blockenc_call("orc_brickwall_blockenc_target_and_gradient_vector", hf, vlist, perms, L, g), g


def brickwall_blockenc_target_gradient_hessian(hf, vlist, perms, L):
    m = 16 * len(vlist)
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    h1 = np.empty((m, m), np.complex128)
    h2 = np.empty((m, m), np.complex128)
    f = _blockenc_call("orc_brickwall_blockenc_target_gradient_hessian", hf, vlist, perms, L, dv, h1, h2)
    return f, dv, h1, h2


def brickwall_blockenc_target_gradient_vector_hessian_matrix(hf, vlist, perms, L):
    m = 16 * len(vlist)
    g = np.empty(m)
    H = np.empty((m, m))
    f = _blockenc_call("orc_brickwall_blockenc_target_gradient_vector_hessian_matrix", hf, vlist, perms, L, g, H)
    return f, g, H


def apply_brickwall_unitary_gate_placeholder(vlist, perms, l, cache):
    """Layer l's placeholder states carried through the later gates (2^L x 16)."""
    vlist, perms, cache = _cplx(vlist), _i32(perms), _cplx(cache)
    L = perms.shape[1]
    out = np.empty((1 << L, 16), np.complex128)
    r = lib().orc_apply_brickwall_unitary_gate_placeholder(_p(vlist), _i(len(vlist)), _p(perms), _i(L), _i(l), _p(cache),
                                                          _p(out))
    assert r == 0
    return out


def truncated_cg_hmat(grad, hess, radius, maxiter=None, abstol=1e-8, reltol=1e-6):
    grad = np.ascontiguousarray(grad, np.float64)
    hess = np.ascontiguousarray(hess, np.float64)
    n = grad.size
    z = np.empty(n)
    L = lib()
    L.orc_truncated_cg_hmat.restype = ct.c_bool
    on = L.orc_truncated_cg_hmat(_p(grad), _p(hess), _i(n), ct.c_double(radius), _i(2 * n if maxiter is None else maxiter),
                                 ct.c_double(abstol), ct.c_double(reltol), _p(z))
    return z, bool(on)


def optimize_brickwall_circuit_hmat(uf, vlist_start, perms, L, niter):
    vlist_start, perms = _cplx(vlist_start), _i32(perms)
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(vlist_start)
    fp, ud = uf.args()
    lib().orc_optimize_brickwall_circuit_hmat(fp, ud, _p(vlist_start), _i(len(vlist_start)), _i(L), _p(perms),
                                              _i(niter), _p(f_iter), _p(vopt))
    return f_iter, vopt


# ------------------------------------------------------------------- algebra
def _mat_fn(name, nin):
    def f(*args):
        ins = [_cplx(a) for a in args[:nin]]
        out = np.empty((4, 4), np.complex128)
        getattr(lib(), name)(*[_p(a) for a in ins], _p(out))
        return out
    return f


multiply = _mat_fn("orc_multiply", 2)
adjoint = _mat_fn("orc_adjoint", 1)
transpose = _mat_fn("orc_transpose", 1)
symm = _mat_fn("orc_symm", 1)
antisymm = _mat_fn("orc_antisymm", 1)
project_tangent = _mat_fn("orc_project_tangent", 2)
symmetric_triple_matrix_product = _mat_fn("orc_symmetric_triple_matrix_product", 3)
polar_factor = _mat_fn("orc_polar_factor", 1)


def inverse_matrix(a):
    out = np.empty((4, 4), np.complex128)
    r = lib().orc_inverse_matrix(_p(_cplx(a)), _p(out))
    return r, out


def real_to_antisymm(r):
    out = np.empty((4, 4), np.complex128)
    lib().orc_real_to_antisymm(_p(np.ascontiguousarray(r, np.float64)), _p(out))
    return out


def antisymm_to_real(w):
    out = np.empty(16)
    lib().orc_antisymm_to_real(_p(_cplx(w)), _p(out))
    return out


def real_to_tangent(r, v):
    out = np.empty((4, 4), np.complex128)
    lib().orc_real_to_tangent(_p(np.ascontiguousarray(r, np.float64)), _p(_cplx(v)), _p(out))
    return out


def tangent_to_real(v, z):
    out = np.empty(16)
    lib().orc_tangent_to_real(_p(_cplx(v)), _p(_cplx(z)), _p(out))
    return out


def retract(u, eta):
    out = np.empty((4, 4), np.complex128)
    lib().orc_retract(_p(_cplx(u)), _p(np.ascontiguousarray(eta, np.float64)), _p(out))
    return out


HVP = ct.CFUNCTYPE(None, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double))


def truncated_cg_hvp(grad, hvp_fn, radius, maxiter=None, abstol=1e-8, reltol=1e-6):
    grad = np.ascontiguousarray(grad, np.float64)
    n = grad.size
    if maxiter is None:
        maxiter = 2 * n

    def cb(x, fdata, vec, out):
        v = np.ctypeslib.as_array(vec, shape=(n,)).copy()
        o = np.ctypeslib.as_array(out, shape=(n,))
        o[:] = hvp_fn(v)

    cfun = HVP(cb)
    z = np.empty(n)
    lib().orc_truncated_cg_hvp.argtypes = [ct.c_void_p, ct.c_void_p, HVP, ct.c_void_p, ct.c_int, ct.c_double,
                                           ct.c_int, ct.c_double, ct.c_double, ct.c_void_p]
    on = lib().orc_truncated_cg_hvp(None, _p(grad), cfun, None, n, radius, maxiter, abstol, reltol, _p(z))
    return z, bool(on)


# ----------------------------------------------------------------------- RNG
class Rng:
    def __init__(self, seed):
        self.s = OrcRng()
        lib().orc_seed_rng_state(ct.c_uint64(seed), ct.byref(self.s))

    def uint64(self):
        return lib().orc_rand_uint64(ct.byref(self.s))

    def uint32(self):
        return lib().orc_rand_uint32(ct.byref(self.s))

    def randu(self):
        return lib().orc_randu(ct.byref(self.s))

    def randn(self):
        return lib().orc_randn(ct.byref(self.s))

    def crandn(self):
        r = lib().orc_crandn(ct.byref(self.s))
        return complex(r.re, r.im)
