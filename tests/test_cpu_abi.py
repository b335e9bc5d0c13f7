"""CPU-only checks of the C-ABI library (no GPU needed).

- librqcopt_hip.so loads and exports every function declared in include/rqcopt/*.h;
- the headers compile as C99/C11 (a C caller, like the reference's own
  trust_region.c / brickwall_opt.c, can include them) and as C++;
- the host-side entry points (matrix.h, rng.h, the tCG of trust_region.h)
  agree with the oracle; the GPU entry points fail loudly without a GPU.
"""
import ctypes as ct
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
import rqcopt_hpc_amd as R
from helpers import crandn, fixture, haar_unitary

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "rqcopt")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INC)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INC, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][A-Za-z0-9_ \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
            name = m.group(1)
            line = text[m.start():text.find("\n", m.start() + 1)]
            if "typedef" in line or name in ("if", "for", "while", "return", "sizeof", "void", "int", "double",
                                              "bool", "char", "long"):
                continue
            names.add(name)
    # typedef'd function pointer types are not functions
    return sorted(n for n in names if n not in ("linear_func",))


def test_library_exports_every_declared_function():
    L = R.lib()
    decl = declared_functions()
    assert len(decl) >= 70
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(R.EXPORTED) == sorted(set(R.EXPORTED))
    assert set(decl) == set(R.EXPORTED)


C_CALLER = r"""
#include "rqcopt/gate.h"
#include "rqcopt/statevector.h"
#include "rqcopt/quantum_circuit.h"
#include "rqcopt/brickwall_circuit.h"
#include "rqcopt/target.h"
#include "rqcopt/matrix.h"
#include "rqcopt/rng.h"
#include "rqcopt/trust_region.h"
#include "rqcopt/brickwall_opt.h"
#include "rqcopt/rqc_device.h"
#include "rqcopt/rqc_batch.h"
#include "rqcopt/rqc_server.h"
#include "rqcopt/rqc_gate1q.h"
#include <stdio.h>
int main(void)
{
	struct rng_state rng;
	seed_rng_state(42, &rng);
	numeric z = crandn(&rng);
	struct mat4x4 a, b;
	identity_matrix(&a);
	scale_matrix(&a, 2.0);
	double eta[16] = {0};
	eta[1] = 0.1;
	retract(&a, eta, &b);
	struct rtr_params p;
	set_rtr_default_params(80, &p);
	printf("%.17g %.17g %.17g %d\n", creal(z), cimag(z), creal(b.data[0]), p.tcg_params.maxiter);
	/* take the address of every GPU entry point (link check only) */
	void* fns[] = { (void*)apply_gate, (void*)apply_gate_backward, (void*)apply_gate_to_array,
		(void*)apply_quantum_circuit, (void*)quantum_circuit_forward, (void*)quantum_circuit_backward,
		(void*)quantum_circuit_gates_hessian_vector_product, (void*)apply_brickwall_unitary,
		(void*)brickwall_unitary_forward, (void*)brickwall_unitary_backward,
		(void*)brickwall_unitary_target_projected_hessian_vector_product, (void*)rqc_device_ufunc,
		(void*)optimize_brickwall_circuit_hvp, (void*)riemannian_trust_region_optimize_hvp,
		(void*)rqc_batch_create, (void*)rqc_batch_evaluate, (void*)rqc_batch_optimize_brickwall_hvp,
		(void*)rqc_server_create, (void*)rqc_server_submit, (void*)rqc_server_wait };
	return fns[0] == 0;
}
"""


def test_headers_compile_as_c_and_link():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "caller.c")
        exe = os.path.join(d, "caller")
        open(src, "w").write(C_CALLER)
        subprocess.check_call(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                               src, "-o", exe, R.LIB_PATH, f"-Wl,-rpath,{os.path.dirname(R.LIB_PATH)}", "-lm"])
        out = subprocess.check_output([exe], text=True).split()
    r = O.Rng(42)
    z = r.crandn()
    assert float(out[0]) == z.real and float(out[1]) == z.imag
    assert int(out[3]) == 160
    u = O.retract(2 * np.eye(4), np.eye(1, 16, 1).ravel() * 0.1)
    assert abs(float(out[2]) - u[0, 0].real) < 1e-15


REF_SRC = "/root/reference/src"
DROPIN = [os.path.join(ROOT, "oracle", "_ref", n) for n in ("libdropin_opt.so", "libdropin_opt_refh.so")]


def test_reference_optimiser_glue_links_against_library():
    """The reference's own callers of the path (src/brickwall_opt.c and src/circuit_opt.c, compiled
    unmodified where they lie, once with this library's headers and once with the reference's) link
    against librqcopt_hip.so with --no-undefined, and every symbol they import is one the library
    declares.  Rebuilt here when the reference checkout is present (oracle/Makefile, target dropin)."""
    if os.path.isdir(REF_SRC):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "dropin"])
    elif not all(os.path.exists(p) for p in DROPIN):
        pytest.skip("reference checkout absent and the drop-in libraries were not built")
    for so in DROPIN:
        out = subprocess.check_output(["nm", "-D", "--undefined-only", so], text=True)
        imported = {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[0] == "U"}
        imported = {s for s in imported if "@" not in s}   # libc symbols carry a version
        assert imported and imported <= set(R.EXPORTED), sorted(imported - set(R.EXPORTED))
        d = ct.CDLL(so)
        for name in ("optimize_brickwall_circuit_hvp", "optimize_brickwall_circuit_hmat",
                     "optimize_brickwall_circuit_hmat_sampling", "optimize_quantum_circuit"):
            assert hasattr(d, name)


def test_headers_compile_as_cpp():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "caller.cpp")
        open(src, "w").write('#include "rqcopt/target.h"\n#include "rqcopt/brickwall_opt.h"\n'
                             '#include "rqcopt/rqc_device.h"\n#include "rqcopt/rqc_batch.h"\n#include "rqcopt/rqc_server.h"\nint main(){ struct mat4x4 a; zero_matrix(&a); return 0; }\n')
        subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", src,
                               "-o", os.path.join(d, "caller.o")])


def test_matrix_functions_match_oracle():
    rng = np.random.default_rng(3)
    for _ in range(10):
        a, b, c = crandn(rng, (4, 4)), crandn(rng, (4, 4)), crandn(rng, (4, 4))
        r = rng.standard_normal(16)
        u = haar_unitary(rng)
        assert np.array_equal(R.multiply_matrices(a, b), O.multiply(a, b))
        assert np.max(np.abs(R.project_tangent(u, b) - O.project_tangent(u, b))) < 1e-15
        assert np.max(np.abs(R.symmetric_triple_matrix_product(a, b, c) - O.symmetric_triple_matrix_product(a, b, c))) < 1e-13
        assert np.max(np.abs(R.real_to_tangent(r, u) - O.real_to_tangent(r, u))) < 1e-15
        assert np.max(np.abs(R.tangent_to_real(u, a) - O.tangent_to_real(u, a))) < 1e-14
        assert np.max(np.abs(R.polar_factor(a) - O.polar_factor(a))) < 1e-13
        assert np.max(np.abs(R.retract(u, 0.1 * r) - O.retract(u, 0.1 * r))) < 1e-14
        ri, inv = R.inverse_matrix(a)
        assert ri == 0 and np.max(np.abs(inv - O.inverse_matrix(a)[1])) < 1e-12
    assert R.inverse_matrix(np.zeros((4, 4)))[0] == -1


def test_algebra_wrappers_accept_float_and_list_inputs():
    """Inputs that are not complex128 arrays are converted to temporaries; the temporaries must stay
    alive until the C call has read them (two converted operands of one call must not share a buffer)."""
    a = np.arange(16.0).reshape(4, 4)
    b = 2.0 * np.eye(4)
    c = np.ones((4, 4))
    ref = O.multiply(a.astype(np.complex128), b.astype(np.complex128))
    for x, y in ((a, b), (a.tolist(), b.tolist()), (a.astype(np.float32), b), (a.T.copy().T, b)):
        assert np.array_equal(R.multiply_matrices(x, y), ref)
    ac, bc, cc = (m.astype(np.complex128) for m in (a, b, c))
    assert np.array_equal(R.symmetric_triple_matrix_product(a, b.tolist(), c),
                          R.symmetric_triple_matrix_product(ac, bc, cc))
    u = haar_unitary(np.random.default_rng(5))
    r = np.arange(16.0) / 16
    assert np.array_equal(R.project_tangent(u, a), R.project_tangent(u, ac))
    assert np.array_equal(R.real_to_tangent(r.tolist(), u), R.real_to_tangent(r, u))
    assert np.array_equal(R.real_to_tangent(r.astype(np.float32), u), R.real_to_tangent(r.astype(np.float32)
                                                                                       .astype(np.float64), u))
    assert np.array_equal(R.tangent_to_real(u, a.tolist()), R.tangent_to_real(u, ac))
    assert np.array_equal(R.retract(u, r.tolist()), R.retract(u, r))
    assert np.array_equal(R.polar_factor(a + np.eye(4)), R.polar_factor(ac + np.eye(4)))
    with pytest.raises(ValueError):
        R.multiply_matrices(np.eye(3), b)
    with pytest.raises(ValueError):
        R.real_to_tangent(np.zeros(15), u)


def test_rng_matches_oracle_bitwise():
    for seed in (0, 1, 42, 123456789):
        a, b = R.Rng(seed), O.Rng(seed)
        for _ in range(100):
            assert a.uint64() == b.uint64()
            assert a.randu() == b.randu()
            assert a.randn() == b.randn()
            assert a.crandn() == b.crandn()


def test_haar_unitary_generator():
    r = R.Rng(7)
    us = [r.haar_unitary() for _ in range(200)]
    for u in us[:10]:
        assert np.max(np.abs(u.conj().T @ u - np.eye(4))) < 1e-14
    # Haar measure: E|U_ij|^2 = 1/4, E U_ij = 0
    m = np.mean([np.abs(u) ** 2 for u in us], axis=0)
    assert np.max(np.abs(m - 0.25)) < 0.05
    assert np.max(np.abs(np.mean(us, axis=0))) < 0.1


def test_truncated_cg_matches_oracle():
    from helpers import fixture
    d = fixture("test_truncated_cg")
    H = d["hess"]
    z, on = R.truncated_cg_hvp(d["grad"], lambda v: H @ v, 1.5)
    assert np.max(np.abs(z - d["z"])) < 1e-14 and on == bool(d["on_boundary"])
    rng = np.random.default_rng(4)
    for radius in (1e-3, 0.1, 10.0):
        B = rng.standard_normal((16, 16))
        H = B + B.T
        g = rng.standard_normal(16)
        z1, on1 = R.truncated_cg_hvp(g, lambda v: H @ v, radius)
        z2, on2 = O.truncated_cg_hvp(g, lambda v: H @ v, radius)
        assert on1 == on2 and np.max(np.abs(z1 - z2)) < 1e-13


def test_gpu_entry_points_fail_loudly_without_gpu():
    r, info = R.device_info()
    if r == 0:
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        R.require_gpu()
    with pytest.raises(R.RqcError):
        R.set_precision(R.PRECISION_MIXED)


def test_truncated_cg_hmat_matches_fixture_and_oracle():
    d = fixture("test_truncated_cg")
    z, on = R.truncated_cg_hmat(d["grad"], d["hess"], 1.5)
    assert np.max(np.abs(z - d["z"])) < 1e-14 and on == bool(d["on_boundary"])
    rng = np.random.default_rng(40)
    for radius in (0.05, 0.5, 50.0):
        A = rng.standard_normal((20, 20))
        H = A + A.T
        g = rng.standard_normal(20)
        z1, on1 = R.truncated_cg_hmat(g, H, radius)
        z2, on2 = O.truncated_cg_hmat(g, H, radius)
        z3, on3 = R.truncated_cg_hvp(g, lambda v: H @ v, radius)
        assert on1 == on2 == on3 and np.max(np.abs(z1 - z2)) < 1e-12 and np.max(np.abs(z1 - z3)) < 1e-12


REF_SUBSET = os.path.join(os.path.dirname(O.__file__), "_ref", "libref_subset.so")


@pytest.mark.skipif(not os.path.exists(REF_SUBSET), reason="oracle/_ref not built (needs the reference checkout)")
def test_rng_helpers_match_reference_bitwise():
    """Bounded draws, Floyd sampling and the single-precision normals against the reference's own
    src/rng.c and src/pcg_basic.c compiled in oracle/_ref: same seeds, same sequences, bitwise."""
    P, Q = R.lib(), ct.CDLL(REF_SUBSET)
    Q.rand_interval.restype = ct.c_uint64
    Q.rand_interval.argtypes = [ct.c_uint64, ct.c_void_p]
    Q.rand_choice.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_void_p, ct.c_void_p]
    Q.pcg32_boundedrand_r.restype = ct.c_uint32
    Q.pcg32_boundedrand_r.argtypes = [ct.c_uint32, ct.c_void_p]
    Q.pcg32x2_boundedrand_r.restype = ct.c_uint64
    Q.pcg32x2_boundedrand_r.argtypes = [ct.c_uint64, ct.c_void_p]
    Q.randuf.restype = ct.c_float
    Q.randnf.restype = ct.c_float
    Q.crandnf.restype = R.CFloat2
    Q.seed_rng_state.argtypes = [ct.c_uint64, ct.POINTER(R.RngState)]
    for seed in (0, 3, 42, 2 ** 40 + 1):
        a, b = R.RngState(), R.RngState()
        P.seed_rng_state(seed, ct.byref(a))
        Q.seed_rng_state(seed, ct.byref(b))
        for bound in (1, 2, 3, 7, 1000, 2 ** 33 + 5, 2 ** 63 + 11):
            assert P.rand_interval(bound, ct.byref(a)) == Q.rand_interval(bound, ct.byref(b))
        for bound in (1, 5, 2 ** 31 + 1):
            assert (P.pcg32_boundedrand_r(bound, ct.byref(a.gen[0])) ==
                    Q.pcg32_boundedrand_r(bound, ct.byref(b.gen[0])))
        assert P.pcg32x2_boundedrand_r(12345, ct.byref(a)) == Q.pcg32x2_boundedrand_r(12345, ct.byref(b))
        for _ in range(40):
            assert P.randuf(ct.byref(a)) == Q.randuf(ct.byref(b))
            assert P.randnf(ct.byref(a)) == Q.randnf(ct.byref(b))
            x, y = P.crandnf(ct.byref(a)), Q.crandnf(ct.byref(b))
            assert (x.re, x.im) == (y.re, y.im)
        for bound, k in ((10, 10), (100, 7), (2 ** 40, 5), (1, 1)):
            u = np.zeros(k, np.uint64)
            v = np.zeros(k, np.uint64)
            P.rand_choice(bound, k, ct.byref(a), u.ctypes.data)
            Q.rand_choice(bound, k, ct.byref(b), v.ctypes.data)
            assert np.array_equal(u, v) and len(set(u.tolist())) == k and int(u.max()) < bound
