"""rqcopt_hpc_amd — MI355X-native (gfx950) gate-sweep engine for Riemannian
quantum-circuit optimisation.

The product is the C-ABI shared library rqcopt_hpc_amd/librqcopt_hip.so
(HIP kernels + host driver), declared in include/rqcopt/*.h with the
reference's function names and argument orders (gate.h, statevector.h,
quantum_circuit.h, brickwall_circuit.h, target.h, matrix.h, rng.h,
trust_region.h, brickwall_opt.h, circuit_opt.h) plus rqc_device.h,
rqc_batch.h and rqc_server.h.  This module binds it with ctypes and offers a
numpy-level mirror of the same API: the gate, circuit and brickwall sweeps;
the unitary targets with gradients, HVPs and Hessians (exact and sampled);
the block-encoding targets; the trust-region optimisers; matrix and RNG
helpers; device-resident targets; batches (Batch) and the request server
(Server).  rqcopt_hpc_amd.rqcopt_hpc mirrors the reference's Python module.
Errors: a negative return code of the C ABI raises RqcError(code).  There is
no CPU fallback: without a HIP GPU the compute entry points raise.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librqcopt_hip.so")
_LIB = None

PRECISION_FP64 = 0
PRECISION_MIXED = 1
KERNEL_FORWARD, KERNEL_REVERSE, KERNEL_TARGET, KERNEL_OTHER = range(4)
ENGINE_TILE_PASS, ENGINE_PER_GATE = 0, 1
EVAL_NONE, EVAL_TARGET, EVAL_GRADIENT, EVAL_HVP = -1, 0, 1, 2
BATCH_GRAPH, BATCH_EAGER = 0, 1
BATCH_RTR_DEVICE, BATCH_RTR_HOST = 0, 1


class RqcError(RuntimeError):
    def __init__(self, where, code):
        super().__init__(f"{where} failed with code {code}")
        self.code = code


# --------------------------------------------------------------- ABI structs
class Statevector(ct.Structure):
    _fields_ = [("data", ct.c_void_p), ("nqubits", ct.c_int)]


class StatevectorArray(ct.Structure):
    _fields_ = [("data", ct.c_void_p), ("nqubits", ct.c_int), ("nstates", ct.c_int)]


class QuantumCircuitCache(ct.Structure):
    _fields_ = [("psi_list", ct.POINTER(Statevector)), ("nqubits", ct.c_int), ("ngates", ct.c_int)]


class Pcg32(ct.Structure):
    _fields_ = [("state", ct.c_uint64), ("inc", ct.c_uint64)]


class RngState(ct.Structure):
    _fields_ = [("gen", Pcg32 * 2)]


class TruncatedCgParams(ct.Structure):
    _fields_ = [("maxiter", ct.c_int), ("abstol", ct.c_double), ("reltol", ct.c_double)]


TARGET_FUNC = ct.CFUNCTYPE(ct.c_double, ct.c_void_p, ct.c_void_p)


class RtrParams(ct.Structure):
    _fields_ = [("tcg_params", TruncatedCgParams), ("rho_trust", ct.c_double), ("radius_init", ct.c_double),
                ("maxradius", ct.c_double), ("g_func", ct.c_void_p), ("g_data", ct.c_void_p),
                ("g_iter", ct.c_void_p)]


class CDouble2(ct.Structure):
    _fields_ = [("re", ct.c_double), ("im", ct.c_double)]


class CFloat2(ct.Structure):
    _fields_ = [("re", ct.c_float), ("im", ct.c_float)]


LINEAR_FUNC = ct.CFUNCTYPE(ct.c_int, ct.POINTER(Statevector), ct.c_void_p, ct.POINTER(Statevector))
HVP_FUNC = ct.CFUNCTYPE(None, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double))

# every function declared in include/rqcopt/*.h
EXPORTED = [
    # statevector.h
    "allocate_statevector", "free_statevector", "add_statevectors", "allocate_statevector_array",
    "free_statevector_array", "normalize_statevector", "transpose_statevector", "haar_random_statevector",
    # gate.h
    "apply_gate", "apply_gate_backward", "apply_gate_to_array", "apply_gate_backward_array", "apply_gate_placeholder",
    # quantum_circuit.h
    "apply_quantum_circuit", "allocate_quantum_circuit_cache", "free_quantum_circuit_cache",
    "quantum_circuit_forward", "quantum_circuit_backward", "quantum_circuit_gates_hessian_vector_product",
    # brickwall_circuit.h
    "apply_brickwall_unitary", "brickwall_unitary_forward", "brickwall_unitary_backward",
    "brickwall_unitary_backward_hessian", "apply_brickwall_unitary_gate_placeholder",
    # target.h
    "circuit_unitary_target", "circuit_unitary_target_and_gradient", "circuit_unitary_target_and_projected_gradient",
    "circuit_unitary_target_hessian_vector_product", "circuit_unitary_target_projected_hessian_vector_product",
    "brickwall_unitary_target", "brickwall_unitary_target_and_gradient",
    "brickwall_unitary_target_and_projected_gradient", "brickwall_unitary_target_hessian_vector_product",
    "brickwall_unitary_target_projected_hessian_vector_product", "brickwall_unitary_target_gradient_hessian",
    "brickwall_unitary_target_gradient_vector_hessian_matrix", "circuit_unitary_target_sampling",
    "brickwall_unitary_target_sampling", "brickwall_unitary_target_gradient_hessian_sampling",
    "brickwall_unitary_target_gradient_vector_hessian_matrix_sampling",
    "brickwall_blockenc_target", "brickwall_blockenc_target_and_gradient",
    "brickwall_blockenc_target_and_gradient_vector", "brickwall_blockenc_target_gradient_hessian",
    "brickwall_blockenc_target_gradient_vector_hessian_matrix",
    # matrix.h
    "zero_matrix", "identity_matrix", "scale_matrix", "conjugate_matrix", "add_matrix", "sub_matrix",
    "add_matrices", "sub_matrices", "transpose", "adjoint", "symm", "antisymm", "real_to_antisymm",
    "antisymm_to_real", "real_to_tangent", "tangent_to_real", "project_tangent", "multiply_matrices",
    "symmetric_triple_matrix_product", "inverse_matrix", "polar_factor", "retract",
    # rng.h
    "pcg32_srandom_r", "pcg32_random_r", "pcg32_boundedrand_r", "pcg32x2_srandom_r", "pcg32x2_random_r",
    "pcg32x2_boundedrand_r", "seed_rng_state", "rand_uint32", "rand_uint64", "rand_interval", "rand_choice", "randu",
    "randuf", "randn", "randnf", "crandn", "crandnf",
    # trust_region.h, brickwall_opt.h
    "truncated_cg_hvp", "riemannian_trust_region_optimize_hvp", "optimize_brickwall_circuit_hvp",
    "truncated_cg_hmat", "riemannian_trust_region_optimize_hmat", "optimize_brickwall_circuit_hmat",
    "optimize_brickwall_circuit_hmat_sampling",
    # circuit_opt.h
    "optimize_quantum_circuit",
    # rqc_device.h
    "rqc_device_ufunc", "rqc_target_create_dense", "rqc_target_create_from_ufunc", "rqc_target_create_brickwall",
    "rqc_target_destroy", "rqc_target_nqubits", "rqc_target_download", "rqc_set_precision", "rqc_get_precision",
    "rqc_set_workspace_limit", "rqc_release_workspace", "rqc_last_column_block", "rqc_profile_enable",
    "rqc_profile_reset", "rqc_profile_query", "rqc_profile_query_flops", "rqc_synchronize", "rqc_device_info", "rqc_measure_copy_bandwidth",
    "rqc_measure_stream_bandwidth", "rqc_set_sweep_engine", "rqc_last_sweep", "rqc_target_create_circuit",
    "rqc_brickwall_unitary_target_projected_hvp_noalloc", "rqc_noalloc_scratch_bytes",
    # rqc_batch.h
    "rqc_batch_create", "rqc_batch_destroy", "rqc_batch_set_exec", "rqc_batch_evaluate",
    "rqc_batch_optimize_brickwall_hvp", "rqc_batch_get_stats", "rqc_batch_problem_plan", "rqc_batch_problem_tiles",
    "rqc_set_autotune", "rqc_batch_set_rtr", "rqc_set_eval_autotune", "rqc_last_tiles",
    # rqc_gate1q.h
    "rqc_apply_single_qubit_gate", "rqc_apply_single_qubit_gate_backward", "rqc_apply_single_qubit_gate_to_array",
    "rqc_mixed_circuit_unitary_target", "rqc_mixed_circuit_unitary_target_and_gradient",
    "rqc_mixed_circuit_unitary_target_hessian_vector_product",
    # rqc_server.h
    "rqc_server_create", "rqc_server_destroy", "rqc_session_create_brickwall", "rqc_session_create_circuit",
    "rqc_session_ngroups", "rqc_session_destroy", "rqc_server_submit", "rqc_server_wait", "rqc_server_poll",
    "rqc_server_release", "rqc_server_get_stats", "rqc_server_optimize_brickwall_hvp",
]


def lib():
    """Load librqcopt_hip.so (built in-tree); raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m rqcopt_hpc_amd.build`")
        L = ct.CDLL(LIB_PATH)
        L.randu.restype = ct.c_double
        L.randn.restype = ct.c_double
        L.crandn.restype = CDouble2
        L.rand_uint64.restype = ct.c_uint64
        L.rand_uint32.restype = ct.c_uint32
        L.pcg32_random_r.restype = ct.c_uint32
        L.pcg32x2_random_r.restype = ct.c_uint64
        L.pcg32_boundedrand_r.restype = ct.c_uint32
        L.pcg32_boundedrand_r.argtypes = [ct.c_uint32, ct.c_void_p]
        L.pcg32x2_boundedrand_r.restype = ct.c_uint64
        L.pcg32x2_boundedrand_r.argtypes = [ct.c_uint64, ct.c_void_p]
        L.rand_interval.restype = ct.c_uint64
        L.rand_interval.argtypes = [ct.c_uint64, ct.c_void_p]
        L.rand_choice.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_void_p, ct.c_void_p]
        L.randuf.restype = ct.c_float
        L.randnf.restype = ct.c_float
        L.crandnf.restype = CFloat2
        L.seed_rng_state.argtypes = [ct.c_uint64, ct.POINTER(RngState)]
        L.truncated_cg_hvp.restype = ct.c_bool
        L.rqc_last_column_block.restype = ct.c_long
        L.rqc_noalloc_scratch_bytes.restype = ct.c_size_t
        L.rqc_set_workspace_limit.argtypes = [ct.c_size_t]
        L.rqc_profile_query.argtypes = [ct.c_int, ct.POINTER(ct.c_double), ct.POINTER(ct.c_long), ct.POINTER(ct.c_double)]
        L.rqc_profile_query_flops.argtypes = [ct.c_int, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
        _LIB = L
    return _LIB


def device_info():
    buf = ct.create_string_buffer(256)
    r = lib().rqc_device_info(buf, 256)
    return r, buf.value.decode()


def require_gpu():
    r, info = device_info()
    if r != 0:
        raise RuntimeError(f"librqcopt_hip needs a HIP GPU: {info}")
    return info


def _check(where, r):
    if r < 0:
        raise RqcError(where, r)
    return r


def _cplx(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class _ArrayPtr(ct.c_void_p):
    """A void* argument that holds a reference to the array it points into, so that a converted
    temporary (e.g. _p(_cplx(x)) for a float or list input) lives until the C call returns."""


def _p(a):
    p = _ArrayPtr(a.ctypes.data)
    p._keep = a
    return p


def _nq(n):
    L = int(n).bit_length() - 1
    if (1 << L) != n:
        raise ValueError(f"length {n} is not a power of two")
    return L


def _sv(a):
    a = _cplx(a)
    s = Statevector(a.ctypes.data, _nq(a.size))
    s._keep = a
    return s


def _gates(g):
    g = _cplx(g).reshape(-1, 4, 4)
    return np.ascontiguousarray(g)


def _perms(perms):
    perms = np.ascontiguousarray(perms, dtype=np.int32)
    rows = (ct.POINTER(ct.c_int) * perms.shape[0])(*[perms[i].ctypes.data_as(ct.POINTER(ct.c_int))
                                                     for i in range(perms.shape[0])])
    return perms, rows


def _wires(w):
    return np.ascontiguousarray(w, dtype=np.int32).reshape(-1)


# --------------------------------------------------------------- targets
class DeviceTarget:
    """U resident in device memory; pass it wherever a target is expected.

    Built from a dense matrix (DeviceTarget.dense), from a host callable
    assembled once (DeviceTarget.from_ufunc), or as a brickwall circuit formed
    on the GPU (DeviceTarget.brickwall, e.g. a Trotter product)."""

    def __init__(self, handle, nqubits):
        self.handle = ct.c_void_p(handle)
        self.nqubits = nqubits

    @staticmethod
    def dense(U):
        U = _cplx(U)
        L = _nq(U.shape[0])
        h = ct.c_void_p()
        _check("rqc_target_create_dense", lib().rqc_target_create_dense(L, _p(U), ct.byref(h)))
        return DeviceTarget(h.value, L)

    @staticmethod
    def from_ufunc(fn, nqubits):
        cb = HostUfunc(fn)
        h = ct.c_void_p()
        _check("rqc_target_create_from_ufunc",
               lib().rqc_target_create_from_ufunc(cb.cfunc, None, nqubits, ct.byref(h)))
        return DeviceTarget(h.value, nqubits)

    @staticmethod
    def brickwall(vlist, perms):
        vlist = _gates(vlist)
        perms, rows = _perms(perms)
        h = ct.c_void_p()
        _check("rqc_target_create_brickwall",
               lib().rqc_target_create_brickwall(perms.shape[1], _p(vlist), len(vlist), rows, ct.byref(h)))
        return DeviceTarget(h.value, perms.shape[1])

    @staticmethod
    def circuit(gates, wires, nqubits):
        """U = product of two-qubit gates (gates[k] on wires[k]) formed on the GPU."""
        gates, wires = _gates(gates), _wires(wires)
        h = ct.c_void_p()
        _check("rqc_target_create_circuit",
               lib().rqc_target_create_circuit(nqubits, _p(gates), len(gates), _p(wires), ct.byref(h)))
        return DeviceTarget(h.value, nqubits)

    def download(self):
        n = 1 << self.nqubits
        U = np.empty((n, n), np.complex128)
        _check("rqc_target_download", lib().rqc_target_download(self.handle, _p(U)))
        return U

    def ufunc_args(self):
        return ct.cast(lib().rqc_device_ufunc, ct.c_void_p), self.handle

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().rqc_target_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostUfunc:
    """A Python callable psi -> U psi exposed as a reference-style linear_func."""

    def __init__(self, fn):
        self.fn = fn

        def cb(psi, udata, out):
            n = 1 << psi.contents.nqubits
            x = np.ctypeslib.as_array(ct.cast(psi.contents.data, ct.POINTER(ct.c_double)), shape=(2 * n,))
            y = np.ctypeslib.as_array(ct.cast(out.contents.data, ct.POINTER(ct.c_double)), shape=(2 * n,))
            try:
                r = np.asarray(self.fn(x.view(np.complex128).copy()), dtype=np.complex128)
                y.view(np.complex128)[:] = r
                return 0
            except Exception:
                return -1

        self.cfunc = LINEAR_FUNC(cb)

    def ufunc_args(self):
        return ct.cast(self.cfunc, ct.c_void_p), None


def _target_args(target):
    if isinstance(target, (DeviceTarget, HostUfunc)):
        return target.ufunc_args()
    if callable(target):
        raise TypeError("wrap Python callables in HostUfunc (and keep a reference to it)")
    raise TypeError("target must be a DeviceTarget or a HostUfunc")


# --------------------------------------------------------------- gate level
def apply_gate(V, i, j, psi):
    psi = _cplx(psi)
    out = np.empty_like(psi)
    lib().apply_gate(_p(_mat4(V)), i, j, ct.byref(_sv(psi)), ct.byref(_sv(out)))
    return out


def apply_gate_backward(V, i, j, psi, dpsi_out):
    psi, dpsi_out = _cplx(psi), _cplx(dpsi_out)
    dpsi = np.empty_like(psi)
    dV = np.empty((4, 4), np.complex128)
    lib().apply_gate_backward(_p(_mat4(V)), i, j, ct.byref(_sv(psi)), ct.byref(_sv(dpsi_out)), ct.byref(_sv(dpsi)), _p(dV))
    return dpsi, dV


def apply_gate_to_array(V, i, j, psi):
    """psi: (2^L, nstates) interleaved statevector array."""
    psi = _cplx(psi)
    out = np.empty_like(psi)
    L = _nq(psi.shape[0])
    a = StatevectorArray(psi.ctypes.data, L, psi.shape[1])
    b = StatevectorArray(out.ctypes.data, L, psi.shape[1])
    lib().apply_gate_to_array(_p(_mat4(V)), i, j, ct.byref(a), ct.byref(b))
    return out


def add_statevectors(a, b):
    a, b = _cplx(a), _cplx(b)
    out = np.empty_like(a)
    lib().add_statevectors(ct.byref(_sv(a)), ct.byref(_sv(b)), ct.byref(_sv(out)))
    return out


# --------------------------------------------------------------- circuit level
def apply_quantum_circuit(gates, wires, psi):
    gates, w, psi = _gates(gates), _wires(wires), _cplx(psi)
    out = np.empty_like(psi)
    _check("apply_quantum_circuit", lib().apply_quantum_circuit(_p(gates), len(gates), _p(w), ct.byref(_sv(psi)),
                                                                ct.byref(_sv(out))))
    return out


class _Cache:
    """quantum_circuit_cache whose statevectors are numpy arrays."""

    def __init__(self, L, ngates, states=None):
        n = 1 << L
        self.states = np.zeros((ngates, n), np.complex128) if states is None else _cplx(states)
        self.svs = (Statevector * ngates)(*[Statevector(self.states[k].ctypes.data, L) for k in range(ngates)])
        self.c = QuantumCircuitCache(self.svs, L, ngates)


def quantum_circuit_forward(gates, wires, psi):
    gates, w, psi = _gates(gates), _wires(wires), _cplx(psi)
    cache = _Cache(_nq(psi.size), len(gates))
    out = np.empty_like(psi)
    _check("quantum_circuit_forward", lib().quantum_circuit_forward(_p(gates), len(gates), _p(w), ct.byref(_sv(psi)),
                                                                    ct.byref(cache.c), ct.byref(_sv(out))))
    return out, cache.states


def quantum_circuit_backward(gates, wires, cache_states, dpsi_out):
    gates, w, dpsi_out = _gates(gates), _wires(wires), _cplx(dpsi_out)
    cache = _Cache(_nq(dpsi_out.size), len(gates), cache_states)
    dpsi = np.empty_like(dpsi_out)
    dgates = np.empty((len(gates), 4, 4), np.complex128)
    _check("quantum_circuit_backward", lib().quantum_circuit_backward(_p(gates), len(gates), _p(w), ct.byref(cache.c),
                                                                      ct.byref(_sv(dpsi_out)), ct.byref(_sv(dpsi)),
                                                                      _p(dgates)))
    return dpsi, dgates


def quantum_circuit_gates_hessian_vector_product(gates, gatedirs, wires, psi, phi):
    gates, dirs, w = _gates(gates), _gates(gatedirs), _wires(wires)
    psi, phi = _cplx(psi), _cplx(phi)
    out = np.empty_like(psi)
    dg = np.empty((len(gates), 4, 4), np.complex128)
    hg = np.empty((len(gates), 4, 4), np.complex128)
    _check("quantum_circuit_gates_hessian_vector_product",
           lib().quantum_circuit_gates_hessian_vector_product(_p(gates), _p(dirs), len(gates), _p(w),
                                                              ct.byref(_sv(psi)), ct.byref(_sv(phi)),
                                                              ct.byref(_sv(out)), _p(dg), _p(hg)))
    return out, dg, hg


def apply_brickwall_unitary(vlist, perms, psi):
    vlist, psi = _gates(vlist), _cplx(psi)
    perms, rows = _perms(perms)
    out = np.empty_like(psi)
    _check("apply_brickwall_unitary", lib().apply_brickwall_unitary(_p(vlist), len(vlist), rows, ct.byref(_sv(psi)),
                                                                    ct.byref(_sv(out))))
    return out


def brickwall_unitary_forward(vlist, perms, psi):
    vlist, psi = _gates(vlist), _cplx(psi)
    perms, rows = _perms(perms)
    L = _nq(psi.size)
    cache = _Cache(L, len(vlist) * (L // 2))
    out = np.empty_like(psi)
    _check("brickwall_unitary_forward", lib().brickwall_unitary_forward(_p(vlist), len(vlist), rows,
                                                                        ct.byref(_sv(psi)), ct.byref(cache.c),
                                                                        ct.byref(_sv(out))))
    return out, cache.states


def brickwall_unitary_backward(vlist, perms, cache_states, dpsi_out):
    vlist, dpsi_out = _gates(vlist), _cplx(dpsi_out)
    perms, rows = _perms(perms)
    L = _nq(dpsi_out.size)
    cache = _Cache(L, len(vlist) * (L // 2), cache_states)
    dpsi = np.empty_like(dpsi_out)
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    _check("brickwall_unitary_backward", lib().brickwall_unitary_backward(_p(vlist), len(vlist), rows, ct.byref(cache.c),
                                                                          ct.byref(_sv(dpsi_out)), ct.byref(_sv(dpsi)),
                                                                          _p(dv)))
    return dpsi, dv


def normalize_statevector(psi):
    out = _cplx(psi).copy()
    lib().normalize_statevector(ct.byref(_sv(out)))
    return out


def transpose_statevector(psi, perm):
    psi = _cplx(psi)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    out = np.empty_like(psi)
    lib().transpose_statevector(ct.byref(_sv(psi)), _p(perm), ct.byref(_sv(out)))
    return out


def haar_random_statevector(nqubits, rng):
    """A Haar-random state drawn from rng (an Rng; advances it)."""
    out = np.empty(1 << nqubits, np.complex128)
    lib().haar_random_statevector(ct.byref(_sv(out)), ct.byref(rng.s))
    return out


def apply_gate_backward_array(V, i, j, psi, dpsi_out):
    """Gate gradients of every state of the (2^L, nstates) array psi: (nstates, 4, 4)."""
    psi, dpsi_out = _cplx(psi), _cplx(dpsi_out)
    L = _nq(psi.shape[0])
    dg = np.empty((psi.shape[1], 4, 4), np.complex128)
    a = StatevectorArray(psi.ctypes.data, L, psi.shape[1])
    lib().apply_gate_backward_array(_p(_mat4(V)), i, j, ct.byref(a), ct.byref(_sv(dpsi_out)), _p(dg))
    return dg


def apply_gate_placeholder(i, j, psi):
    """The 16 states E_rc psi (state 4r + c) as a (2^L, 16) array."""
    psi = _cplx(psi)
    L = _nq(psi.size)
    out = np.empty((psi.size, 16), np.complex128)
    lib().apply_gate_placeholder(i, j, ct.byref(_sv(psi)), ct.byref(StatevectorArray(out.ctypes.data, L, 16)))
    return out


def brickwall_unitary_backward_hessian(vlist, perms, cache_states, dpsi_out):
    vlist, dpsi_out = _gates(vlist), _cplx(dpsi_out)
    perms, rows = _perms(perms)
    L = _nq(dpsi_out.size)
    cache = _Cache(L, len(vlist) * (L // 2), cache_states)
    m = 16 * len(vlist)
    dpsi = np.empty_like(dpsi_out)
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    hess = np.empty((m, m), np.complex128)
    _check("brickwall_unitary_backward_hessian",
           lib().brickwall_unitary_backward_hessian(_p(vlist), len(vlist), rows, ct.byref(cache.c), ct.byref(_sv(dpsi_out)),
                                                    ct.byref(_sv(dpsi)), _p(dv), _p(hess)))
    return dpsi, dv, hess


def apply_brickwall_unitary_gate_placeholder(vlist, perms, l, cache_states):
    """Layer l's placeholder states carried through the later gates, summed: (2^L, 16)."""
    vlist = _gates(vlist)
    perms, rows = _perms(perms)
    cache_states = _cplx(cache_states)
    L = _nq(cache_states.shape[1])
    cache = _Cache(L, len(vlist) * (L // 2), cache_states)
    out = np.empty((1 << L, 16), np.complex128)
    _check("apply_brickwall_unitary_gate_placeholder",
           lib().apply_brickwall_unitary_gate_placeholder(_p(vlist), len(vlist), rows, l, ct.byref(cache.c),
                                                          ct.byref(StatevectorArray(out.ctypes.data, L, 16))))
    return out


# --------------------------------------------------------------- targets
def _f():
    return np.zeros(1, np.complex128)


def circuit_unitary_target(target, gates, wires, nqubits):
    gates, w, f = _gates(gates), _wires(wires), _f()
    fp, ud = _target_args(target)
    _check("circuit_unitary_target", lib().circuit_unitary_target(fp, ud, _p(gates), len(gates), _p(w), nqubits, _p(f)))
    return f[0]


def circuit_unitary_target_and_gradient(target, gates, wires, nqubits):
    gates, w, f = _gates(gates), _wires(wires), _f()
    dg = np.empty((len(gates), 4, 4), np.complex128)
    fp, ud = _target_args(target)
    _check("circuit_unitary_target_and_gradient",
           lib().circuit_unitary_target_and_gradient(fp, ud, _p(gates), len(gates), _p(w), nqubits, _p(f), _p(dg)))
    return f[0], dg


def circuit_unitary_target_and_projected_gradient(target, gates, wires, nqubits):
    gates, w, f = _gates(gates), _wires(wires), _f()
    g = np.empty(16 * len(gates))
    fp, ud = _target_args(target)
    _check("circuit_unitary_target_and_projected_gradient",
           lib().circuit_unitary_target_and_projected_gradient(fp, ud, _p(gates), len(gates), _p(w), nqubits, _p(f), _p(g)))
    return f[0], g


def circuit_unitary_target_hessian_vector_product(target, gates, gatedirs, wires, nqubits):
    gates, dirs, w, f = _gates(gates), _gates(gatedirs), _wires(wires), _f()
    dg = np.empty((len(gates), 4, 4), np.complex128)
    hg = np.empty((len(gates), 4, 4), np.complex128)
    fp, ud = _target_args(target)
    _check("circuit_unitary_target_hessian_vector_product",
           lib().circuit_unitary_target_hessian_vector_product(fp, ud, _p(gates), _p(dirs), len(gates), _p(w), nqubits,
                                                               _p(f), _p(dg), _p(hg)))
    return f[0], dg, hg


# --------------------------------------------------------------- single-qubit gates and mixed circuits
def _mat2(a):
    a = _cplx(a)
    if a.shape != (2, 2):
        raise ValueError(f"expected a 2x2 matrix, got shape {a.shape}")
    return a


def apply_single_qubit_gate(V, i, psi):
    """psi_out = V on wire i (2x2 V, row-major; qubit 0 the most significant bit)."""
    psi = _cplx(psi)
    out = np.empty_like(psi)
    _check("rqc_apply_single_qubit_gate", lib().rqc_apply_single_qubit_gate(_p(_mat2(V)), i, ct.byref(_sv(psi)),
                                                                            ct.byref(_sv(out))))
    return out


def apply_single_qubit_gate_backward(V, i, psi, dpsi_out):
    """(dpsi = V^T dpsi_out on wire i, dV = sum dpsi_out (x) psi)."""
    psi, dpsi_out = _cplx(psi), _cplx(dpsi_out)
    dpsi = np.empty_like(psi)
    dV = np.empty((2, 2), np.complex128)
    _check("rqc_apply_single_qubit_gate_backward",
           lib().rqc_apply_single_qubit_gate_backward(_p(_mat2(V)), i, ct.byref(_sv(psi)), ct.byref(_sv(dpsi_out)),
                                                      ct.byref(_sv(dpsi)), _p(dV)))
    return dpsi, dV


def apply_single_qubit_gate_to_array(V, i, psi):
    """psi: (2^L, nstates) interleaved statevector array."""
    psi = _cplx(psi)
    out = np.empty_like(psi)
    L = _nq(psi.shape[0])
    a = StatevectorArray(psi.ctypes.data, L, psi.shape[1])
    b = StatevectorArray(out.ctypes.data, L, psi.shape[1])
    _check("rqc_apply_single_qubit_gate_to_array",
           lib().rqc_apply_single_qubit_gate_to_array(_p(_mat2(V)), i, ct.byref(a), ct.byref(b)))
    return out


def _mixed(gates, nq, wires):
    nq = np.ascontiguousarray(nq, np.int32)
    if len(gates) != nq.size:
        raise ValueError("one qubit count per gate expected")
    packed = np.concatenate([(_mat2(G) if q == 1 else _mat4(G)).reshape(-1) for G, q in zip(gates, nq)])
    w = np.ascontiguousarray(np.asarray(wires).reshape(-1, 2), np.int32)
    return packed, nq, w


def _unpack_mixed(flat, nq):
    out, o = [], 0
    for q in nq:
        m = 2 if q == 1 else 4
        out.append(flat[o:o + m * m].reshape(m, m).copy())
        o += m * m
    return out


def mixed_circuit_unitary_target(target, gates, nq, wires, nqubits):
    """f = -Tr(U^dagger W) for a circuit of one-qubit (nq 1, 2x2, wires (i, any)) and two-qubit gates."""
    g, nq, w = _mixed(gates, nq, wires)
    f = _f()
    fp, ud = _target_args(target)
    _check("rqc_mixed_circuit_unitary_target",
           lib().rqc_mixed_circuit_unitary_target(fp, ud, _p(g), _p(nq), _p(w), len(nq), nqubits, _p(f)))
    return f[0]


def mixed_circuit_unitary_target_and_gradient(target, gates, nq, wires, nqubits):
    g, nq, w = _mixed(gates, nq, wires)
    f = _f()
    dg = np.empty_like(g)
    fp, ud = _target_args(target)
    _check("rqc_mixed_circuit_unitary_target_and_gradient",
           lib().rqc_mixed_circuit_unitary_target_and_gradient(fp, ud, _p(g), _p(nq), _p(w), len(nq), nqubits, _p(f),
                                                               _p(dg)))
    return f[0], _unpack_mixed(dg, nq)


def mixed_circuit_unitary_target_hessian_vector_product(target, gates, gatedirs, nq, wires, nqubits):
    g, nq, w = _mixed(gates, nq, wires)
    d, _, _ = _mixed(gatedirs, nq, wires)
    f = _f()
    dg = np.empty_like(g)
    hg = np.empty_like(g)
    fp, ud = _target_args(target)
    _check("rqc_mixed_circuit_unitary_target_hessian_vector_product",
           lib().rqc_mixed_circuit_unitary_target_hessian_vector_product(fp, ud, _p(g), _p(d), _p(nq), _p(w), len(nq),
                                                                         nqubits, _p(f), _p(dg), _p(hg)))
    return f[0], _unpack_mixed(dg, nq), _unpack_mixed(hg, nq)


def circuit_unitary_target_projected_hessian_vector_product(target, gates, gatedirs, wires, nqubits):
    gates, dirs, w, f = _gates(gates), _gates(gatedirs), _wires(wires), _f()
    g = np.empty(16 * len(gates))
    h = np.empty(16 * len(gates))
    fp, ud = _target_args(target)
    _check("circuit_unitary_target_projected_hessian_vector_product",
           lib().circuit_unitary_target_projected_hessian_vector_product(fp, ud, _p(gates), _p(dirs), len(gates), _p(w),
                                                                         nqubits, _p(f), _p(g), _p(h)))
    return f[0], g, h


def brickwall_unitary_target(target, vlist, perms, nqubits=None):
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    fp, ud = _target_args(target)
    _check("brickwall_unitary_target", lib().brickwall_unitary_target(fp, ud, _p(vlist), len(vlist), L, rows, _p(f)))
    return f[0]


def brickwall_unitary_target_and_gradient(target, vlist, perms, nqubits=None):
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    fp, ud = _target_args(target)
    _check("brickwall_unitary_target_and_gradient",
           lib().brickwall_unitary_target_and_gradient(fp, ud, _p(vlist), len(vlist), L, rows, _p(f), _p(dv)))
    return f[0], dv


def brickwall_unitary_target_and_projected_gradient(target, vlist, perms, nqubits=None):
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    g = np.empty(16 * len(vlist))
    fp, ud = _target_args(target)
    _check("brickwall_unitary_target_and_projected_gradient",
           lib().brickwall_unitary_target_and_projected_gradient(fp, ud, _p(vlist), len(vlist), L, rows, _p(f), _p(g)))
    return f[0], g


def brickwall_unitary_target_hessian_vector_product(target, vlist, vdirs, perms, nqubits=None):
    vlist, vdirs, f = _gates(vlist), _gates(vdirs), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    hz = np.empty((len(vlist), 4, 4), np.complex128)
    fp, ud = _target_args(target)
    _check("brickwall_unitary_target_hessian_vector_product",
           lib().brickwall_unitary_target_hessian_vector_product(fp, ud, _p(vlist), _p(vdirs), len(vlist), rows, L,
                                                                 _p(f), _p(dv), _p(hz)))
    return f[0], dv, hz


def brickwall_unitary_target_projected_hessian_vector_product(target, vlist, vdirs, perms, nqubits=None):
    vlist, vdirs, f = _gates(vlist), _gates(vdirs), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    g = np.empty(16 * len(vlist))
    h = np.empty(16 * len(vlist))
    fp, ud = _target_args(target)
    _check("brickwall_unitary_target_projected_hessian_vector_product",
           lib().brickwall_unitary_target_projected_hessian_vector_product(fp, ud, _p(vlist), _p(vdirs), len(vlist),
                                                                           rows, L, _p(f), _p(g), _p(h)))
    return f[0], g, h


def brickwall_unitary_target_projected_hvp_noalloc(target, vlist, vdirs, perms, nqubits=None):
    """The projected grad+HVP evaluation with the library's static device scratch as its whole
    workspace (no device allocation); `target` must be a DeviceTarget."""
    vlist, vdirs, f = _gates(vlist), _gates(vdirs), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    g = np.empty(16 * len(vlist))
    h = np.empty(16 * len(vlist))
    fp, ud = _target_args(target)
    _check("rqc_brickwall_unitary_target_projected_hvp_noalloc",
           lib().rqc_brickwall_unitary_target_projected_hvp_noalloc(fp, ud, _p(vlist), _p(vdirs), len(vlist), rows, L,
                                                                   _p(f), _p(g), _p(h)))
    return f[0], g, h


def noalloc_scratch_bytes():
    return int(lib().rqc_noalloc_scratch_bytes())


def brickwall_unitary_target_gradient_hessian(target, vlist, perms, nqubits=None):
    """f, dvlist and the complex Hessian hess[(i*16 + x), (j*16 + y)] (16 nlayers square)."""
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    m = 16 * len(vlist)
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    hess = np.empty((m, m), np.complex128)
    fp, ud = _target_args(target)
    _check("brickwall_unitary_target_gradient_hessian",
           lib().brickwall_unitary_target_gradient_hessian(fp, ud, _p(vlist), len(vlist), L, rows, _p(f), _p(dv), _p(hess)))
    return f[0], dv, hess


def brickwall_unitary_target_gradient_vector_hessian_matrix(target, vlist, perms, nqubits=None):
    """f, the projected gradient and the Riemannian Hessian matrix H (real, symmetric)."""
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    m = 16 * len(vlist)
    g = np.empty(m)
    H = np.empty((m, m))
    fp, ud = _target_args(target)
    _check("brickwall_unitary_target_gradient_vector_hessian_matrix",
           lib().brickwall_unitary_target_gradient_vector_hessian_matrix(fp, ud, _p(vlist), len(vlist), L, rows, _p(f), _p(g),
                                                                        _p(H)))
    return f[0], g, H


# ---------------------------------------------------------------- block-encoding targets
def _blockenc(name, hop, vlist, perms, nqubits, *outs):
    """Call a block-encoding entry point: hop applies H to states of nqubits/2 qubits."""
    vlist, f = _gates(vlist), np.zeros(1)
    perms, rows = _perms(perms)
    L = perms.shape[1] if nqubits is None else nqubits
    fp, ud = _target_args(hop)
    _check(name, getattr(lib(), name)(fp, ud, _p(vlist), len(vlist), L, rows, _p(f), *[_p(o) for o in outs]))
    return float(f[0])


def brickwall_blockenc_target(hop, vlist, perms, nqubits=None):
    """f = ||P^dagger W P - H||_F^2 / 2 of the brickwall circuit W (reference src/target.c:1100-1190)."""
    return _blockenc("brickwall_blockenc_target", hop, vlist, perms, nqubits)


def brickwall_blockenc_target_and_gradient(hop, vlist, perms, nqubits=None):
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    return _blockenc("brickwall_blockenc_target_and_gradient", hop, vlist, perms, nqubits, dv), dv


def brickwall_blockenc_target_and_gradient_vector(hop, vlist, perms, nqubits=None):
    g = np.empty(16 * len(vlist))
    return _blockenc("brickwall_blockenc_target_and_gradient_vector", hop, vlist, perms, nqubits, g), g


def brickwall_blockenc_target_gradient_hessian(hop, vlist, perms, nqubits=None):
    """f, dvlist, hess1 (complex symmetric) and hess2 (Hermitian), each 16 nlayers square."""
    m = 16 * len(vlist)
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    h1 = np.empty((m, m), np.complex128)
    h2 = np.empty((m, m), np.complex128)
    f = _blockenc("brickwall_blockenc_target_gradient_hessian", hop, vlist, perms, nqubits, dv, h1, h2)
    return f, dv, h1, h2


def brickwall_blockenc_target_gradient_vector_hessian_matrix(hop, vlist, perms, nqubits=None):
    """f, the projected gradient and the Riemannian Hessian matrix (real, symmetric)."""
    m = 16 * len(vlist)
    g = np.empty(m)
    H = np.empty((m, m))
    return _blockenc("brickwall_blockenc_target_gradient_vector_hessian_matrix", hop, vlist, perms, nqubits, g, H), g, H


def circuit_unitary_target_sampling(target, gates, wires, nqubits, nsamples, rng):
    gates, w, f = _gates(gates), _wires(wires), _f()
    fp, ud = _target_args(target)
    L = lib()
    L.circuit_unitary_target_sampling.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int,
                                                  ct.c_long, ct.c_void_p, ct.c_void_p]
    _check("circuit_unitary_target_sampling", L.circuit_unitary_target_sampling(
        fp, ud, _p(gates), len(gates), _p(w), nqubits, nsamples, ct.addressof(rng.s), _p(f)))
    return f[0]


def brickwall_unitary_target_sampling(target, vlist, perms, nsamples, rng, nqubits=None):
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    Lq = perms.shape[1] if nqubits is None else nqubits
    fp, ud = _target_args(target)
    L = lib()
    L.brickwall_unitary_target_sampling.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p,
                                                    ct.c_long, ct.c_void_p, ct.c_void_p]
    _check("brickwall_unitary_target_sampling", L.brickwall_unitary_target_sampling(
        fp, ud, _p(vlist), len(vlist), Lq, ct.cast(rows, ct.c_void_p), nsamples, ct.addressof(rng.s), _p(f)))
    return f[0]


def brickwall_unitary_target_gradient_hessian_sampling(target, vlist, perms, nsamples, rng, nqubits=None):
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    Lq = perms.shape[1] if nqubits is None else nqubits
    m = 16 * len(vlist)
    dv = np.empty((len(vlist), 4, 4), np.complex128)
    hess = np.empty((m, m), np.complex128)
    fp, ud = _target_args(target)
    L = lib()
    L.brickwall_unitary_target_gradient_hessian_sampling.argtypes = [
        ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_long, ct.c_void_p, ct.c_void_p,
        ct.c_void_p, ct.c_void_p]
    _check("brickwall_unitary_target_gradient_hessian_sampling", L.brickwall_unitary_target_gradient_hessian_sampling(
        fp, ud, _p(vlist), len(vlist), Lq, ct.cast(rows, ct.c_void_p), nsamples, ct.addressof(rng.s), _p(f), _p(dv),
        _p(hess)))
    return f[0], dv, hess


def brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(target, vlist, perms, nsamples, rng, nqubits=None):
    vlist, f = _gates(vlist), _f()
    perms, rows = _perms(perms)
    Lq = perms.shape[1] if nqubits is None else nqubits
    m = 16 * len(vlist)
    g = np.empty(m)
    H = np.empty((m, m))
    fp, ud = _target_args(target)
    L = lib()
    L.brickwall_unitary_target_gradient_vector_hessian_matrix_sampling.argtypes = [
        ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_long, ct.c_void_p, ct.c_void_p,
        ct.c_void_p, ct.c_void_p]
    _check("brickwall_unitary_target_gradient_vector_hessian_matrix_sampling",
           L.brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(
               fp, ud, _p(vlist), len(vlist), Lq, ct.cast(rows, ct.c_void_p), nsamples, ct.addressof(rng.s), _p(f), _p(g),
               _p(H)))
    return f[0], g, H


# --------------------------------------------------------------- optimiser
def rtr_default_params(n):
    """set_rtr_default_params (reference src/trust_region.h:68-76)."""
    return RtrParams(TruncatedCgParams(2 * n, 1e-8, 1e-6), 0.125, 0.01, 0.1, None, None, None)


def optimize_brickwall_circuit_hvp(target, vlist_start, perms, niter, params=None):
    vlist_start = _gates(vlist_start)
    perms, rows = _perms(perms)
    nl = len(vlist_start)
    params = rtr_default_params(16 * nl) if params is None else params
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(vlist_start)
    fp, ud = _target_args(target)
    lib().optimize_brickwall_circuit_hvp(fp, ud, _p(vlist_start), nl, perms.shape[1], rows, ct.byref(params), niter,
                                         _p(f_iter), _p(vopt))
    return f_iter, vopt


def optimize_brickwall_circuit_hmat(target, vlist_start, perms, niter, params=None):
    vlist_start = _gates(vlist_start)
    perms, rows = _perms(perms)
    nl = len(vlist_start)
    params = rtr_default_params(16 * nl) if params is None else params
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(vlist_start)
    fp, ud = _target_args(target)
    lib().optimize_brickwall_circuit_hmat(fp, ud, _p(vlist_start), nl, perms.shape[1], rows, ct.byref(params), niter,
                                          _p(f_iter), _p(vopt))
    return f_iter, vopt


def optimize_brickwall_circuit_hmat_sampling(target, vlist_start, perms, niter, nsamples, rng, params=None):
    vlist_start = _gates(vlist_start)
    perms, rows = _perms(perms)
    nl = len(vlist_start)
    params = rtr_default_params(16 * nl) if params is None else params
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(vlist_start)
    fp, ud = _target_args(target)
    L = lib()
    L.optimize_brickwall_circuit_hmat_sampling.argtypes = [
        ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_long, ct.c_void_p, ct.c_void_p,
        ct.c_int, ct.c_void_p, ct.c_void_p]
    L.optimize_brickwall_circuit_hmat_sampling(fp, ud, _p(vlist_start), nl, perms.shape[1], ct.cast(rows, ct.c_void_p),
                                               nsamples, ct.addressof(rng.s), ct.addressof(params), niter, _p(f_iter),
                                               _p(vopt))
    return f_iter, vopt


def optimize_quantum_circuit(target, gates_start, wires, nqubits, niter, params=None):
    gates_start, w = _gates(gates_start), _wires(wires)
    ng = len(gates_start)
    params = rtr_default_params(16 * ng) if params is None else params
    f_iter = np.empty(niter + 1)
    gopt = np.empty_like(gates_start)
    fp, ud = _target_args(target)
    lib().optimize_quantum_circuit(fp, ud, _p(gates_start), ng, nqubits, _p(w), ct.byref(params), niter, _p(f_iter),
                                   _p(gopt))
    return f_iter, gopt


def truncated_cg_hmat(grad, hess, radius, params=None):
    grad = np.ascontiguousarray(grad, np.float64)
    hess = np.ascontiguousarray(hess, np.float64)
    n = grad.size
    params = TruncatedCgParams(2 * n, 1e-8, 1e-6) if params is None else params
    z = np.empty(n)
    L = lib()
    L.truncated_cg_hmat.restype = ct.c_bool
    on = L.truncated_cg_hmat(_p(grad), _p(hess), n, ct.c_double(radius), ct.byref(params), _p(z))
    return z, bool(on)


def truncated_cg_hvp(grad, hvp_fn, radius, params=None):
    grad = np.ascontiguousarray(grad, np.float64)
    n = grad.size
    params = TruncatedCgParams(2 * n, 1e-8, 1e-6) if params is None else params

    def cb(x, fdata, vec, out):
        v = np.ctypeslib.as_array(vec, shape=(n,)).copy()
        np.ctypeslib.as_array(out, shape=(n,))[:] = hvp_fn(v)

    cfun = HVP_FUNC(cb)
    z = np.empty(n)
    L = lib()
    L.truncated_cg_hvp.argtypes = [ct.c_void_p, ct.c_void_p, HVP_FUNC, ct.c_void_p, ct.c_int, ct.c_double,
                                   ct.POINTER(TruncatedCgParams), ct.c_void_p]
    on = L.truncated_cg_hvp(None, _p(grad), cfun, None, n, radius, ct.byref(params), _p(z))
    return z, bool(on)


# --------------------------------------------------------------- batched optimisation (rqc_batch.h)
class _BatchProblem(ct.Structure):
    _fields_ = [("nqubits", ct.c_int), ("nlayers", ct.c_int), ("perms", ct.POINTER(ct.POINTER(ct.c_int))),
                ("target", ct.c_void_p)]


class BatchStats(ct.Structure):
    _fields_ = [("steps", ct.c_long), ("evaluations", ct.c_long), ("graph_launches", ct.c_long),
                ("device_ms", ct.c_double), ("workspace_bytes", ct.c_size_t)]


class Batch:
    """Many independent brickwall problems evaluated and optimised together on the GPU.

    problems: list of (perms [nlayers, L] int array, DeviceTarget).  Each step
    evaluates every problem whose kind is not EVAL_NONE (one graph launch per
    evaluation kind); optimize() runs one trust-region optimisation per problem
    in lock-step."""

    def __init__(self, problems, precision=PRECISION_FP64, nstreams=0, workspace_limit=0):
        self._keep = []
        arr = (_BatchProblem * len(problems))()
        self.nlayers = []
        for i, (perms, target) in enumerate(problems):
            perms, rows = _perms(perms)
            self._keep += [perms, rows, target]
            arr[i] = _BatchProblem(perms.shape[1], perms.shape[0], rows, target.handle.value)
            self.nlayers.append(perms.shape[0])
        self._keep.append(arr)
        self.n = len(problems)
        h = ct.c_void_p()
        L = lib()
        L.rqc_batch_create.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.c_int, ct.c_size_t, ct.c_void_p]
        _check("rqc_batch_create", L.rqc_batch_create(self.n, ct.cast(arr, ct.c_void_p), precision, nstreams,
                                                      workspace_limit, ct.byref(h)))
        self.handle = h

    def set_exec(self, mode):
        _check("rqc_batch_set_exec", lib().rqc_batch_set_exec(self.handle, mode))

    def set_rtr(self, mode):
        """BATCH_RTR_DEVICE (default): optimize() keeps the trust region's state and arithmetic on the
        GPU; BATCH_RTR_HOST: one host driver thread per problem."""
        _check("rqc_batch_set_rtr", lib().rqc_batch_set_rtr(self.handle, mode))

    @staticmethod
    def _ptrs(arrs):
        return (ct.c_void_p * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])

    def evaluate(self, kinds, vlists, vdirs=None):
        """Returns (f [n] complex, grads, hvps): lists of 16*nlayers arrays (None where not computed)."""
        kinds = np.ascontiguousarray(kinds, np.int32)
        vl = [_gates(v) for v in vlists]
        vd = [None if (vdirs is None or vdirs[i] is None) else _gates(vdirs[i]) for i in range(self.n)]
        f = np.zeros(self.n, np.complex128)
        grads = [np.full(16 * self.nlayers[i], np.nan) if kinds[i] >= EVAL_GRADIENT else None for i in range(self.n)]
        hvps = [np.full(16 * self.nlayers[i], np.nan) if kinds[i] == EVAL_HVP else None for i in range(self.n)]
        _check("rqc_batch_evaluate", lib().rqc_batch_evaluate(self.handle, _p(kinds), self._ptrs(vl), self._ptrs(vd),
                                                              _p(f), self._ptrs(grads), self._ptrs(hvps)))
        return f, grads, hvps

    def optimize(self, vlist_starts, niter, params=None):
        """Lock-step optimize_brickwall_circuit_hvp of every problem: (f_iters list, vlist_opt list).
        The trust region runs on the GPU unless set_rtr(BATCH_RTR_HOST) was called."""
        vl = [_gates(v) for v in vlist_starts]
        params = rtr_default_params(16 * self.nlayers[0]) if params is None else params
        f_iters = [np.empty(niter + 1) for _ in range(self.n)]
        vopt = [np.empty_like(v) for v in vl]
        _check("rqc_batch_optimize_brickwall_hvp",
               lib().rqc_batch_optimize_brickwall_hvp(self.handle, self._ptrs(vl), ct.byref(params), niter,
                                                      self._ptrs(f_iters), self._ptrs(vopt)))
        return f_iters, vopt

    def stats(self):
        st = BatchStats()
        _check("rqc_batch_get_stats", lib().rqc_batch_get_stats(self.handle, ct.byref(st)))
        return {k: getattr(st, k) for k, _ in BatchStats._fields_}

    def plan(self, p, kind):
        b, e, n = ct.c_long(), ct.c_int(), ct.c_int()
        _check("rqc_batch_problem_plan", lib().rqc_batch_problem_plan(self.handle, p, kind, ct.byref(b), ct.byref(e),
                                                                      ct.byref(n)))
        return b.value, e.value, n.value

    def tiles(self, p, kind=EVAL_HVP):
        """(log2 tile rows, log2 tile columns, most gates per pass) of problem p."""
        k, c, g = ct.c_int(), ct.c_int(), ct.c_int()
        _check("rqc_batch_problem_tiles", lib().rqc_batch_problem_tiles(self.handle, p, kind, ct.byref(k), ct.byref(c),
                                                                        ct.byref(g)))
        return k.value, c.value, g.value

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().rqc_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- request server (rqc_server.h)
class _ServerParams(ct.Structure):
    _fields_ = [("arena_bytes", ct.c_size_t), ("max_request_bytes", ct.c_size_t), ("nslots", ct.c_int)]


class _RequestInfo(ct.Structure):
    _fields_ = [("status", ct.c_int), ("slot", ct.c_int), ("queued_ms", ct.c_double), ("run_ms", ct.c_double),
                ("latency_ms", ct.c_double), ("start_ms", ct.c_double), ("end_ms", ct.c_double),
                ("workspace_bytes", ct.c_size_t), ("block", ct.c_long)]


class _ServerStats(ct.Structure):
    _fields_ = [("submitted", ct.c_long), ("completed", ct.c_long), ("failed", ct.c_long),
                ("peak_in_flight", ct.c_int), ("arena_bytes", ct.c_size_t), ("peak_arena_bytes", ct.c_size_t)]


class Session:
    """A circuit structure with its device target and precision, registered with a Server."""

    def __init__(self, server, handle, ngroups, target):
        self.server, self.handle, self.ngroups, self._target = server, handle, ngroups, target

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().rqc_session_destroy(self.handle)
            self.handle = None


class Pending:
    """A submitted request: its ticket and the output buffers the server writes on completion."""

    def __init__(self, ticket, f, grad, hvp, keep):
        self.ticket, self.f, self.grad, self.hvp, self._keep = ticket, f, grad, hvp, keep


class Server:
    """Continuous-batching request server: target / gradient / HVP requests of many sessions run
    concurrently on the GPU, admitted in arrival order as execution slots and arena memory free up."""

    def __init__(self, arena_bytes=0, max_request_bytes=0, nslots=0):
        prm = _ServerParams(arena_bytes, max_request_bytes, nslots)
        h = ct.c_void_p()
        L = lib()
        L.rqc_server_create.argtypes = [ct.POINTER(_ServerParams), ct.c_void_p]
        _check("rqc_server_create", L.rqc_server_create(ct.byref(prm), ct.byref(h)))
        self.handle = h

    def session_brickwall(self, perms, target, precision=PRECISION_FP64):
        perms, rows = _perms(perms)
        h = ct.c_void_p()
        _check("rqc_session_create_brickwall",
               lib().rqc_session_create_brickwall(self.handle, perms.shape[1], perms.shape[0], rows, target.handle,
                                                  precision, ct.byref(h)))
        return Session(self, h, perms.shape[0], target)

    def session_circuit(self, wires, nqubits, target, precision=PRECISION_FP64):
        wires = _wires(wires)
        h = ct.c_void_p()
        _check("rqc_session_create_circuit",
               lib().rqc_session_create_circuit(self.handle, nqubits, wires.size // 2, _p(wires), target.handle,
                                                precision, ct.byref(h)))
        return Session(self, h, wires.size // 2, target)

    def submit(self, session, kind, gates=None, dirs=None):
        g = None if gates is None else _gates(gates)
        d = None if dirs is None else _gates(dirs)
        f = np.zeros(1, np.complex128)
        grad = np.full(16 * session.ngroups, np.nan) if kind >= EVAL_GRADIENT else None
        hvp = np.full(16 * session.ngroups, np.nan) if kind == EVAL_HVP else None
        t = ct.c_long()
        L = lib()
        L.rqc_server_submit.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                        ct.c_void_p, ct.c_void_p, ct.c_void_p]
        _check("rqc_server_submit", L.rqc_server_submit(self.handle, session.handle, kind,
                                                        None if g is None else g.ctypes.data,
                                                        None if d is None else d.ctypes.data, f.ctypes.data,
                                                        None if grad is None else grad.ctypes.data,
                                                        None if hvp is None else hvp.ctypes.data, ct.byref(t)))
        return Pending(t.value, f, grad, hvp, (g, d))

    @staticmethod
    def _info(i):
        return {k: getattr(i, k) for k, _ in _RequestInfo._fields_}

    def wait(self, pending):
        """Block until done: (f, grad, hvp, info)."""
        info = _RequestInfo()
        L = lib()
        L.rqc_server_wait.argtypes = [ct.c_void_p, ct.c_long, ct.POINTER(_RequestInfo)]
        _check("rqc_server_wait", L.rqc_server_wait(self.handle, pending.ticket, ct.byref(info)))
        return pending.f[0], pending.grad, pending.hvp, self._info(info)

    def poll(self, pending):
        info = _RequestInfo()
        L = lib()
        L.rqc_server_poll.argtypes = [ct.c_void_p, ct.c_long, ct.POINTER(_RequestInfo)]
        r = _check("rqc_server_poll", L.rqc_server_poll(self.handle, pending.ticket, ct.byref(info)))
        return self._info(info) if r == 1 else None

    def release(self, pending):
        L = lib()
        L.rqc_server_release.argtypes = [ct.c_void_p, ct.c_long]
        _check("rqc_server_release", L.rqc_server_release(self.handle, pending.ticket))

    def optimize(self, session, vlist_start, niter, params=None, max_records=4096):
        """optimize_brickwall_circuit_hvp of a brickwall session on the calling thread, every target /
        gradient / HVP evaluation a request of this server (ctypes releases the GIL, so solves on
        several Python threads run concurrently): (f_iter, vlist_opt, request infos)."""
        v = _gates(vlist_start)
        params = rtr_default_params(16 * len(v)) if params is None else params
        f_iter = np.empty(niter + 1)
        vopt = np.empty_like(v)
        recs = (_RequestInfo * max_records)()
        nrec = ct.c_int()
        L = lib()
        L.rqc_server_optimize_brickwall_hvp.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int,
                                                        ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p]
        _check("rqc_server_optimize_brickwall_hvp",
               L.rqc_server_optimize_brickwall_hvp(self.handle, session.handle, v.ctypes.data, ct.byref(params), niter,
                                                   f_iter.ctypes.data, vopt.ctypes.data, recs, max_records,
                                                   ct.byref(nrec)))
        return f_iter, vopt, [self._info(recs[i]) for i in range(min(nrec.value, max_records))]

    def stats(self):
        st = _ServerStats()
        _check("rqc_server_get_stats", lib().rqc_server_get_stats(self.handle, ct.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().rqc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- algebra, RNG
def _mat4(a):
    a = _cplx(a)
    if a.shape != (4, 4):
        raise ValueError(f"expected a 4x4 matrix, got shape {a.shape}")
    return a


def _real16(r):
    r = np.ascontiguousarray(r, np.float64)
    if r.size != 16:
        raise ValueError(f"expected 16 real tangent coordinates, got {r.size}")
    return r


def _m1(name, a):
    out = np.empty((4, 4), np.complex128)
    getattr(lib(), name)(_p(_mat4(a)), _p(out))
    return out


def multiply_matrices(a, b):
    out = np.empty((4, 4), np.complex128)
    lib().multiply_matrices(_p(_mat4(a)), _p(_mat4(b)), _p(out))
    return out


def project_tangent(u, z):
    out = np.empty((4, 4), np.complex128)
    lib().project_tangent(_p(_mat4(u)), _p(_mat4(z)), _p(out))
    return out


def symmetric_triple_matrix_product(a, b, c):
    out = np.empty((4, 4), np.complex128)
    lib().symmetric_triple_matrix_product(_p(_mat4(a)), _p(_mat4(b)), _p(_mat4(c)), _p(out))
    return out


def polar_factor(a):
    return _m1("polar_factor", a)


def inverse_matrix(a):
    out = np.empty((4, 4), np.complex128)
    r = lib().inverse_matrix(_p(_mat4(a)), _p(out))
    return r, out


def real_to_tangent(r, v):
    out = np.empty((4, 4), np.complex128)
    lib().real_to_tangent(_p(_real16(r)), _p(_mat4(v)), _p(out))
    return out


def tangent_to_real(v, z):
    out = np.empty(16)
    lib().tangent_to_real(_p(_mat4(v)), _p(_mat4(z)), _p(out))
    return out


def retract(u, eta):
    out = np.empty((4, 4), np.complex128)
    lib().retract(_p(_mat4(u)), _p(_real16(eta)), _p(out))
    return out


class Rng:
    """seed_rng_state / rand_uint64 / randu / randn / crandn (reference src/rng.c)."""

    def __init__(self, seed):
        self.s = RngState()
        lib().seed_rng_state(ct.c_uint64(seed), ct.byref(self.s))

    def uint64(self):
        return lib().rand_uint64(ct.byref(self.s))

    def randu(self):
        return lib().randu(ct.byref(self.s))

    def randn(self):
        return lib().randn(ct.byref(self.s))

    def crandn(self):
        r = lib().crandn(ct.byref(self.s))
        return complex(r.re, r.im)

    def haar_unitary(self, n=4):
        """Haar-random unitary: QR of a complex Ginibre matrix drawn with crandn, phases of R's diagonal
        moved into Q (Mezzadri 2007)."""
        z = np.array([[self.crandn() for _ in range(n)] for _ in range(n)])
        q, r = np.linalg.qr(z)
        d = np.diag(r)
        return q * (d / np.abs(d))[None, :]


# --------------------------------------------------------------- device controls
def set_precision(mode):
    _check("rqc_set_precision", lib().rqc_set_precision(mode))


def set_autotune(on):
    _check("rqc_set_autotune", lib().rqc_set_autotune(1 if on else 0))


def set_eval_autotune(on, forget=False):
    """Tile autotuning of single evaluations: the first evaluation of each circuit structure times
    the tile candidates and keeps the fastest (forget=True also drops the choices made so far)."""
    _check("rqc_set_eval_autotune", lib().rqc_set_eval_autotune(2 if forget else (1 if on else 0)))


def last_tiles():
    """(log2 tile rows, log2 tile columns, most gates per pass) of the last evaluation."""
    k, c, g = ct.c_int(), ct.c_int(), ct.c_int()
    _check("rqc_last_tiles", lib().rqc_last_tiles(ct.byref(k), ct.byref(c), ct.byref(g)))
    return k.value, c.value, g.value


def set_sweep_engine(engine):
    _check("rqc_set_sweep_engine", lib().rqc_set_sweep_engine(engine))


def last_sweep():
    """(engine, passes) of the last target evaluation: ENGINE_TILE_PASS or ENGINE_PER_GATE and its
    number of HBM passes per sweep direction."""
    e, p = ct.c_int(), ct.c_int()
    _check("rqc_last_sweep", lib().rqc_last_sweep(ct.byref(e), ct.byref(p)))
    return e.value, p.value


def get_precision():
    return lib().rqc_get_precision()


def set_workspace_limit(nbytes):
    _check("rqc_set_workspace_limit", lib().rqc_set_workspace_limit(nbytes))


def release_workspace():
    lib().rqc_release_workspace()


def last_column_block():
    return lib().rqc_last_column_block()


def profile_enable(on=True):
    lib().rqc_profile_enable(1 if on else 0)


def profile_reset():
    lib().rqc_profile_reset()


def profile_query(kind):
    """(total ms, launches, algorithmic HBM bytes) of the timed launches of one kernel kind."""
    ms, n, b = ct.c_double(), ct.c_long(), ct.c_double()
    _check("rqc_profile_query", lib().rqc_profile_query(kind, ct.byref(ms), ct.byref(n), ct.byref(b)))
    return ms.value, n.value, b.value


def profile_query_flops(kind):
    """(algorithmic flops, issued flops) of the timed tile-pass launches of one kernel kind."""
    a, i = ct.c_double(), ct.c_double()
    _check("rqc_profile_query_flops", lib().rqc_profile_query_flops(kind, ct.byref(a), ct.byref(i)))
    return a.value, i.value


def measure_copy_bandwidth(nbytes=1 << 30, reps=10):
    """GB/s of a plain device-to-device streaming copy (read + write bytes)."""
    L = lib()
    L.rqc_measure_copy_bandwidth.argtypes = [ct.c_size_t, ct.c_int, ct.POINTER(ct.c_double)]
    g = ct.c_double()
    _check("rqc_measure_copy_bandwidth", L.rqc_measure_copy_bandwidth(nbytes, reps, ct.byref(g)))
    return g.value


def measure_stream_bandwidth(nbytes=1 << 30, reps=10):
    """GB/s of a device-to-device copy tuned for HBM bandwidth (read + write bytes): the
    streaming ceiling of the device."""
    L = lib()
    L.rqc_measure_stream_bandwidth.argtypes = [ct.c_size_t, ct.c_int, ct.POINTER(ct.c_double)]
    g = ct.c_double()
    _check("rqc_measure_stream_bandwidth", L.rqc_measure_stream_bandwidth(nbytes, reps, ct.byref(g)))
    return g.value


def synchronize():
    _check("rqc_synchronize", lib().rqc_synchronize())
