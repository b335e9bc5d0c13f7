/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C (C99, double complex) restatement of the reference rqcopt_hpc
 * algorithms on the gate-sweep path, written to follow the reference's
 * arithmetic and summation order so that results agree to rounding.  It is the
 * checker for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and never as the thing measured or
 * shipped.  The product library (librqcopt_hip.so) never links or calls it.
 *
 * All names carry the prefix orc_ so that the oracle and the product library
 * can be loaded into one process without symbol clashes.
 *
 * Conventions (same as the reference):
 *   - complex entries are double complex (re, im);
 *   - amplitude index: qubit q <-> bit (L-1-q), qubit 0 is the most significant;
 *   - a 4x4 gate is row-major, basis |q_i q_j> with the first wire i the more
 *     significant index of the gate;
 *   - statevector arrays are interleaved: data[amp * nstates + k].
 *
 * Pinning: see tests/test_oracle_pinning.py (reference fixtures, the compiled
 * reference subset in oracle/_ref, and closed-form / brute-force checks).
 */
#ifndef RQC_ORACLE_H
#define RQC_ORACLE_H

#include <complex.h>
#include <stdint.h>
#include <stdbool.h>

typedef double complex ocplx;

/* per-column callback: psi_out = U psi  (cf. reference src/target.h:7) */
typedef int (*orc_linear_func)(int nqubits, const ocplx* psi, void* udata, ocplx* psi_out);

/* ---- gate level: reference src/gate.c ---- */
void orc_apply_gate(const ocplx* gate, int i, int j, int L, const ocplx* psi, ocplx* psi_out);
void orc_apply_gate_backward(const ocplx* gate, int i, int j, int L, const ocplx* psi,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dgate);
void orc_apply_gate_to_array(const ocplx* gate, int i, int j, int L, int nstates, const ocplx* psi, ocplx* psi_out);
/* single-qubit gate on wire i (no reference counterpart; the two-qubit conventions) */
void orc_apply_gate_1q(const ocplx* gate, int i, int L, const ocplx* psi, ocplx* psi_out);
void orc_apply_gate_1q_backward(const ocplx* gate, int i, int L, const ocplx* psi, const ocplx* dpsi_out,
	ocplx* dpsi, ocplx* dgate);

/* ---- circuit level: reference src/quantum_circuit.c ---- */
int orc_apply_quantum_circuit(const ocplx* gates, int ngates, const int* wires, int L, const ocplx* psi, ocplx* psi_out);
int orc_quantum_circuit_forward(const ocplx* gates, int ngates, const int* wires, int L, const ocplx* psi,
	ocplx* cache, ocplx* psi_out);
int orc_quantum_circuit_backward(const ocplx* gates, int ngates, const int* wires, int L, const ocplx* cache,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dgates);
int orc_quantum_circuit_gates_hessian_vector_product(const ocplx* gates, const ocplx* gatedirs, int ngates,
	const int* wires, int L, const ocplx* psi, const ocplx* phi, ocplx* psi_out, ocplx* dgates, ocplx* hess);

/* ---- brickwall level: reference src/brickwall_circuit.c; perms is nlayers x L ---- */
int orc_apply_brickwall_unitary(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* psi, ocplx* psi_out);
int orc_brickwall_unitary_forward(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* psi,
	ocplx* cache, ocplx* psi_out);
int orc_brickwall_unitary_backward(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* cache,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dvlist);

/* ---- targets: reference src/target.c ---- */
int orc_circuit_unitary_target(orc_linear_func ufunc, void* udata, const ocplx* gates, int ngates, const int* wires,
	int L, ocplx* fval);
int orc_circuit_unitary_target_and_gradient(orc_linear_func ufunc, void* udata, const ocplx* gates, int ngates,
	const int* wires, int L, ocplx* fval, ocplx* dgates);
int orc_circuit_unitary_target_and_projected_gradient(orc_linear_func ufunc, void* udata, const ocplx* gates,
	int ngates, const int* wires, int L, ocplx* fval, double* grad_vec);
int orc_circuit_unitary_target_hessian_vector_product(orc_linear_func ufunc, void* udata, const ocplx* gates,
	const ocplx* gatedirs, int ngates, const int* wires, int L, ocplx* fval, ocplx* dgates, ocplx* hess);
int orc_circuit_unitary_target_projected_hessian_vector_product(orc_linear_func ufunc, void* udata,
	const ocplx* gates, const ocplx* gatedirs, int ngates, const int* wires, int L, ocplx* fval,
	double* grad_vec, double* hvp_vec);

int orc_brickwall_unitary_target(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers, int L,
	const int* perms, ocplx* fval);
int orc_brickwall_unitary_target_and_gradient(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers,
	int L, const int* perms, ocplx* fval, ocplx* dvlist);
int orc_brickwall_unitary_target_and_projected_gradient(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	int nlayers, int L, const int* perms, ocplx* fval, double* grad_vec);
int orc_brickwall_unitary_target_hessian_vector_product(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	const ocplx* vdirs, int nlayers, const int* perms, int L, ocplx* fval, ocplx* dvlist, ocplx* hess_vdirs);
int orc_brickwall_unitary_target_projected_hessian_vector_product(orc_linear_func ufunc, void* udata,
	const ocplx* vlist, const ocplx* vdirs, int nlayers, const int* perms, int L, ocplx* fval,
	double* grad_vec, double* hvp_vec);

/* built-in ufuncs */
int orc_ufunc_dense(int nqubits, const ocplx* psi, void* udata, ocplx* psi_out);     /* udata = row-major N x N U */
int orc_ufunc_scramble(int nqubits, const ocplx* psi, void* udata, ocplx* psi_out);  /* reference test map */
/* U given matrix-free as a brickwall circuit (udata = struct orc_brickwall_op*) */
struct orc_brickwall_op { const ocplx* vlist; const int* perms; int nlayers; };
int orc_ufunc_brickwall(int nqubits, const ocplx* psi, void* udata, ocplx* psi_out);

/* column-range partial evaluation used for the multi-core CPU baseline:
   accumulates f, dvlist, hess_vdirs over basis columns [b0, b1) only */
int orc_brickwall_hvp_columns(orc_linear_func ufunc, void* udata, const ocplx* vlist, const ocplx* vdirs,
	int nlayers, const int* perms, int L, long b0, long b1, ocplx* fval, ocplx* dvlist, ocplx* hess_vdirs);

/* projection step of the brickwall projected HVP (reference src/target.c:645-671), applied to
   column sums dvlist / hess_vdirs (overwritten) that orc_brickwall_hvp_columns accumulated */
void orc_brickwall_project_grad_hvp(const ocplx* vlist, const ocplx* vdirs, int nlayers, ocplx* dvlist,
	ocplx* hess_vdirs, double* grad_vec, double* hvp_vec);

/* ---- 4x4 algebra: reference src/matrix.c ---- */
void orc_multiply(const ocplx* a, const ocplx* b, ocplx* c);
void orc_adjoint(const ocplx* a, ocplx* ah);
void orc_transpose(const ocplx* a, ocplx* at);
void orc_symm(const ocplx* w, ocplx* z);
void orc_antisymm(const ocplx* w, ocplx* z);
void orc_real_to_antisymm(const double* r, ocplx* w);
void orc_antisymm_to_real(const ocplx* w, double* r);
void orc_real_to_tangent(const double* r, const ocplx* v, ocplx* z);
void orc_tangent_to_real(const ocplx* v, const ocplx* z, double* r);
void orc_project_tangent(const ocplx* u, const ocplx* z, ocplx* p);
void orc_symmetric_triple_matrix_product(const ocplx* a, const ocplx* b, const ocplx* c, ocplx* ret);
int  orc_inverse_matrix(const ocplx* a, ocplx* ainv);
void orc_polar_factor(const ocplx* a, ocplx* u);
void orc_retract(const ocplx* u, const double* eta, ocplx* v);

/* ---- trust region: reference src/trust_region.c, src/brickwall_opt.c ---- */
typedef void (*orc_hvp_func)(const double* x, void* fdata, const double* vec, double* hvp);
bool orc_truncated_cg_hvp(const double* x, const double* grad, orc_hvp_func f_hvp, void* fdata, int n,
	double radius, int maxiter, double abstol, double reltol, double* z);
/* RTR-HVP brickwall optimiser with default parameters for n = 16 nlayers;
   f_iter has niter + 1 entries */
void orc_optimize_brickwall_circuit_hvp(orc_linear_func ufunc, void* udata, const ocplx* vlist_start, int nlayers,
	int L, const int* perms, int niter, double* f_iter, ocplx* vlist_opt);

/* RTR-HVP optimiser of a general circuit (reference src/circuit_opt.c:22-124), default parameters */
void orc_optimize_quantum_circuit(orc_linear_func ufunc, void* udata, const ocplx* gates_start, int ngates, int L,
	const int* wires, int niter, double* f_iter, ocplx* gates_opt);

/* ---- full-Hessian route: reference src/gate.c:269-460, src/brickwall_circuit.c:134-255,
        src/target.c:687-797,937-1040, src/trust_region.c:83-141,230-293, src/brickwall_opt.c ---- */
/* psi_out: 2^L x 16 interleaved (state 4r + c = E_rc psi on wires (i, j)) */
void orc_apply_gate_placeholder(int i, int j, int L, const ocplx* psi, ocplx* psi_out);
/* dgates[16 k + 4 r + c] = sum dpsi_out[r-amplitude] psi[c-amplitude, state k] */
void orc_apply_gate_backward_array(int i, int j, int L, int nstates, const ocplx* psi, const ocplx* dpsi_out, ocplx* dgates);
/* cache: nlayers*L/2 states (before each gate); hess: (16 nlayers)^2 */
int orc_brickwall_unitary_backward_hessian(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* cache,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dvlist, ocplx* hess);
int orc_brickwall_unitary_target_gradient_hessian(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers,
	int L, const int* perms, ocplx* fval, ocplx* dvlist, ocplx* hess);
int orc_brickwall_unitary_target_gradient_vector_hessian_matrix(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	int nlayers, int L, const int* perms, ocplx* fval, double* grad_vec, double* H);
bool orc_truncated_cg_hmat(const double* grad, const double* hess, int n, double radius, int maxiter, double abstol,
	double reltol, double* z);
/* RTR with the full Hessian, default parameters; f_iter has niter + 1 entries */
void orc_optimize_brickwall_circuit_hmat(orc_linear_func ufunc, void* udata, const ocplx* vlist_start, int nlayers,
	int L, const int* perms, int niter, double* f_iter, ocplx* vlist_opt);

/* ---- RNG: reference src/rng.c, src/pcg_basic.c ---- */
struct orc_pcg32 { uint64_t state, inc; };
struct orc_rng { struct orc_pcg32 gen[2]; };

/* ---- statevector utilities and sampled (stochastic-trace) targets: reference
        src/statevector.c:47-127, src/target.c:85-149,491-505,807-928,1050-1088,
        src/brickwall_opt.c:165-268 ---- */
void orc_normalize_statevector(int L, ocplx* psi);
void orc_transpose_statevector(int L, const ocplx* psi, const int* perm, ocplx* psi_trans);
void orc_haar_random_statevector(int L, struct orc_rng* rng, ocplx* psi);
int orc_circuit_unitary_target_sampling(orc_linear_func ufunc, void* udata, const ocplx* gates, int ngates, const int* wires,
	int L, long nsamples, struct orc_rng* rng, ocplx* fval);
int orc_brickwall_unitary_target_sampling(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers, int L,
	const int* perms, long nsamples, struct orc_rng* rng, ocplx* fval);
int orc_brickwall_unitary_target_gradient_hessian_sampling(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	int nlayers, int L, const int* perms, long nsamples, struct orc_rng* rng, ocplx* fval, ocplx* dvlist, ocplx* hess);
int orc_brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(orc_linear_func ufunc, void* udata,
	const ocplx* vlist, int nlayers, int L, const int* perms, long nsamples, struct orc_rng* rng, ocplx* fval,
	double* grad_vec, double* H);
void orc_optimize_brickwall_circuit_hmat_sampling(orc_linear_func ufunc, void* udata, const ocplx* vlist_start,
	int nlayers, int L, const int* perms, long nsamples, struct orc_rng* rng, int niter, double* f_iter, ocplx* vlist_opt);
void     orc_seed_rng_state(uint64_t seed, struct orc_rng* rng);
uint32_t orc_rand_uint32(struct orc_rng* rng);
uint64_t orc_rand_uint64(struct orc_rng* rng);
double   orc_randu(struct orc_rng* rng);
double   orc_randn(struct orc_rng* rng);
ocplx    orc_crandn(struct orc_rng* rng);

/* ---- block-encoding targets f = || P^dagger W P - H ||_F^2 / 2: reference
        src/target.c:1100-1720, src/brickwall_circuit.c:264-310; hfunc acts on
        L/2 qubits ---- */
int orc_brickwall_blockenc_target(orc_linear_func hfunc, void* hdata, const ocplx* vlist, int nlayers, int L,
	const int* perms, double* fval);
int orc_brickwall_blockenc_target_and_gradient(orc_linear_func hfunc, void* hdata, const ocplx* vlist, int nlayers,
	int L, const int* perms, double* fval, ocplx* dvlist);
int orc_brickwall_blockenc_target_and_gradient_vector(orc_linear_func hfunc, void* hdata, const ocplx* vlist,
	int nlayers, int L, const int* perms, double* fval, double* grad_vec);
/* psi_out: 2^L x 16 interleaved */
int orc_apply_brickwall_unitary_gate_placeholder(const ocplx* vlist, int nlayers, const int* perms, int L, int l,
	const ocplx* cache, ocplx* psi_out);
int orc_brickwall_blockenc_target_gradient_hessian(orc_linear_func hfunc, void* hdata, const ocplx* vlist, int nlayers,
	int L, const int* perms, double* fval, ocplx* dvlist, ocplx* hess1, ocplx* hess2);
int orc_brickwall_blockenc_target_gradient_vector_hessian_matrix(orc_linear_func hfunc, void* hdata, const ocplx* vlist,
	int nlayers, int L, const int* perms, double* fval, double* grad_vec, double* H);

#endif
