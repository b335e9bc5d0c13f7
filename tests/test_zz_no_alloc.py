"""The no-allocation projected grad+HVP entry point against the allocating one.

rqc_brickwall_unitary_target_projected_hvp_noalloc runs the same evaluation as
brickwall_unitary_target_projected_hessian_vector_product (reference src/target.c:631-678)
with the library's static 256 MiB device scratch as its whole workspace, so a large circuit is
swept in narrower column blocks.  Two calls in a row, the second on a smaller circuit, each
against the allocating entry point on the same inputs:
  FP64:  |df| / max(1, |f|) < 1e-12, max |dgrad| / max |grad| < 1e-12, the same for the HVP
         (only the order of the column-block sums differs);
  mixed: the same quantities within 2e-6 (complex FP32 sweeps, FP64 reductions, summed over
         other column blocks).
"""
import numpy as np
import pytest

import rqcopt_hpc_amd as R
from rqcopt_hpc_amd.workloads import config_inputs, ising_trotter_brickwall

pytestmark = pytest.mark.gpu


def _rel(a, b):
    (f, g, h), (f0, g0, h0) = a, b
    return (abs(f - f0) / max(1.0, abs(f0)), np.max(np.abs(g - g0)) / np.max(np.abs(g0)),
            np.max(np.abs(h - h0)) / np.max(np.abs(h0)))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_noalloc_matches_allocating_entry_point(prec, tol):
    R.require_gpu()
    assert R.noalloc_scratch_bytes() == 256 << 20
    R.set_precision(R.PRECISION_FP64 if prec == "fp64" else R.PRECISION_MIXED)
    try:
        # a circuit whose full column block exceeds the scratch, then a smaller one
        for L, depth, seed in ((12, 7, 3), (8, 5, 4)):
            print(f"case L={L} depth={depth} seed={seed} precision={prec}", flush=True)
            V, P, vec, Z = config_inputs(L, depth, seed)
            Vt, Pt = ising_trotter_brickwall(L, 1.0, 0.75, 0.6, 0.5, 2)
            T = R.DeviceTarget.brickwall(Vt, Pt)
            try:
                ref = R.brickwall_unitary_target_projected_hessian_vector_product(T, V, Z, P)
                B_ref = R.last_column_block()
                out = R.brickwall_unitary_target_projected_hvp_noalloc(T, V, Z, P)
                B = R.last_column_block()
            finally:
                T.close()
            e = _rel(out, ref)
            print(f"  column block {B} (allocating: {B_ref}); rel errors f {e[0]:.2e} grad {e[1]:.2e} hvp {e[2]:.2e}",
                  flush=True)
            if L == 12:
                assert B < (1 << L), "the L=12 plan should be column-blocked within the scratch"
            assert max(e) < tol, e
    finally:
        R.set_precision(R.PRECISION_FP64)


def test_noalloc_rejects_host_callback():
    R.require_gpu()
    L, depth = 6, 3
    V, P, vec, Z = config_inputs(L, depth, 1)
    U = np.eye(1 << L, dtype=np.complex128)
    cb = R.HostUfunc(lambda psi: U @ psi)
    with pytest.raises(R.RqcError):
        R.brickwall_unitary_target_projected_hvp_noalloc(cb, V, Z, P)
