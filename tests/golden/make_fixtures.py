"""Convert the reference's own HDF5 test fixtures into committed .npz files.

Runs only where the reference checkout is present (it reads test/data/*.hdf5
there and links the image's libhdf5 from /opt/conda).  The GPU tests never run
it: they load the committed tests/golden/fixtures/*.npz instead.

Each .npz holds the datasets of one fixture under their original names
(complex arrays keep the trailing re/im axis the fixture uses, see
test/io_util.py:4-8 of the reference).

    python tests/golden/make_fixtures.py [name ...]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

REF_DATA = "/root/reference/test/data"
REF_EXAMPLES = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "fixtures")

# fixtures that pin the gate-sweep path (complex variants) plus the tCG one
FIXTURES = [
    "test_apply_gate_cplx",
    "test_apply_gate_backward_cplx",
    "test_apply_gate_to_array_cplx",
    "test_apply_quantum_circuit_cplx",
    "test_quantum_circuit_backward_cplx",
    "test_quantum_circuit_gates_hessian_vector_product_cplx",
    "test_apply_brickwall_unitary_cplx",
    "test_brickwall_unitary_backward_cplx",
    "test_circuit_unitary_target_cplx",
    "test_circuit_unitary_target_and_gradient_cplx",
    "test_circuit_unitary_target_hessian_vector_product_cplx",
    "test_circuit_unitary_target_projected_hessian_vector_product_cplx",
    "test_brickwall_unitary_target_cplx",
    "test_brickwall_unitary_target_and_gradient_cplx",
    "test_brickwall_unitary_target_and_projected_gradient_cplx",
    "test_brickwall_unitary_target_hessian_vector_product_cplx",
    "test_brickwall_unitary_target_projected_hessian_vector_product_cplx",
    "test_brickwall_unitary_target_gradient_hessian_cplx",
    "test_brickwall_unitary_target_gradient_vector_hessian_matrix_cplx",
    "test_truncated_cg",
    "test_polar_factor_cplx",
    "test_project_tangent_cplx",
    "test_real_to_tangent_cplx",
    "test_multiply_cplx",
    "test_transpose_statevector_cplx",
    "test_brickwall_unitary_target_sampling_cplx",
    "test_brickwall_unitary_target_gradient_hessian_sampling_cplx",
    "test_brickwall_unitary_target_gradient_vector_hessian_matrix_sampling_cplx",
    "test_brickwall_blockenc_target_cplx",
    "test_brickwall_blockenc_target_and_gradient_cplx",
    "test_brickwall_blockenc_target_and_gradient_vector_cplx",
    "test_brickwall_blockenc_target_gradient_hessian_cplx",
    "test_brickwall_blockenc_target_gradient_vector_hessian_matrix_cplx",
]

# run outputs of the reference's example programs (inputs and the 10-iteration trust-region
# trajectories they wrote), stored as example_<file name>; they pin the optimisers end to end:
#   ising1d_dynamics_opt.c           optimize_brickwall_circuit_hmat, L=6, 3 layers
#   ising1d_dynamics_opt_sampling.c  optimize_brickwall_circuit_hmat_sampling, 16 samples, seed 42
#   molecular_dynamics_opt.c         optimize_quantum_circuit (RTR-HVP), 5 qubits
EXAMPLES = [
    "ising1d/ising1d_dynamics_opt_n3_init",
    "ising1d/ising1d_dynamics_opt_n3",
    "ising1d/ising1d_dynamics_opt_sampling_n3",
    "molecular/molecular_dynamics_opt_init",
    "molecular/molecular_dynamics_opt",
]


def build_converter(workdir):
    exe = os.path.join(workdir, "h5_to_raw")
    subprocess.check_call([
        "gcc", "-O2", os.path.join(HERE, "h5_to_raw.c"),
        "-I/opt/conda/include", "-L/opt/conda/lib", "-lhdf5",
        "-Wl,-rpath,/opt/conda/lib", "-o", exe])
    return exe


def convert(exe, name, workdir, src=None):
    src = src or os.path.join(REF_DATA, name + ".hdf5")
    d = os.path.join(workdir, name)
    os.makedirs(d, exist_ok=True)
    out = subprocess.check_output([exe, src, d], text=True)
    arrays = {}
    for line in out.strip().splitlines():
        parts = line.split()
        key, dt, shape = parts[0], parts[1], tuple(int(x) for x in parts[2:])
        raw = np.fromfile(os.path.join(d, key + ".bin"), dtype="<" + dt)
        arrays[key] = raw.reshape(shape) if shape else raw.reshape(())
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
    return arrays


def main(only=None):
    """Convert every fixture, or only those named (test fixture names or example_<file name>)."""
    if not os.path.isdir(REF_DATA):
        sys.exit("reference fixtures not present; nothing to convert")
    os.makedirs(OUT, exist_ok=True)
    jobs = [(name, None) for name in FIXTURES]
    jobs += [("example_" + os.path.basename(e), os.path.join(REF_EXAMPLES, e + ".hdf5")) for e in EXAMPLES]
    if only:
        jobs = [j for j in jobs if j[0] in only]
    with tempfile.TemporaryDirectory() as work:
        exe = build_converter(work)
        for name, src in jobs:
            arrays = convert(exe, name, work, src)
            print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main(sys.argv[1:])
