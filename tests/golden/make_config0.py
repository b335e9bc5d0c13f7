"""Golden vectors of the L=8 Ising workload, written by the oracle.

Workload: L=8, periodic brickwall of depth 5, Haar-random gates from the seeded
PCG generator (seed 1), HVP direction vec ~ N(0,1) from seed 2 mapped to the
tangent space, target U = expm(-i t H) of the periodic Ising chain
H = J sum Z_j Z_{j+1} + h sum X_j + g sum Z_j with J=1, h=0.75, g=0.6, t=0.25
(formed densely in FP64; not stored, the test rebuilds it).  Stored: inputs, f, projected gradient and HVP, and
the 10-iteration RTR-HVP trajectory f_iter and final gates.

    python tests/golden/make_config0.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from rqcopt_hpc_amd.workloads import config_inputs, expm_hermitian, ising_hamiltonian  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures", "config0_ising_L8_d5.npz")


def main():
    L, d = 8, 5
    V, P, vec, Z = config_inputs(L, d, seed=1)
    H = ising_hamiltonian(L, 1.0, 0.75, 0.6)
    U = expm_hermitian(H, 0.25)
    uf = O.Ufunc.dense(U)
    f, g, h = O.brickwall_unitary_target_projected_hvp(uf, V, Z, P, L)
    f_iter, vopt = O.optimize_brickwall_circuit_hvp(uf, V, P, L, 10)
    np.savez_compressed(OUT, V=V, P=P, vec=vec, Z=Z, f=f, grad=g, hvp=h, f_iter=f_iter, vopt=vopt)
    print("f =", f, " f_iter =", f_iter)


if __name__ == "__main__":
    main()
