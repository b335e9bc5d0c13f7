/*
 * Targets f = -Tr(U^dagger C) of a circuit C, their gate gradients and
 * Hessian-vector products, evaluated on the GPU over all 2^nqubits basis
 * columns at once.  Replaces reference src/target.h:7-67 (the ten functions
 * below; the sampling, full-Hessian and block-encoding variants are not part
 * of this library).
 *
 * `ufunc` applies U to one state (reference src/target.h:7).  Two kinds are
 * accepted:
 *   - rqc_device_ufunc (see rqc_device.h) with an rqc_target descriptor as
 *     udata: U stays resident in GPU memory, no host callbacks;
 *   - any other host callback: it is called once per basis column (as in the
 *     reference) to assemble U, which is then uploaded for the GPU sweeps.
 * Return value: 0 on success; -1 allocation/device failure, -2 ufunc failure,
 * -3 invalid arguments.  Messages go to stderr.
 *
 * Note the argument order: the Hessian-vector functions take
 * (vlist, vdirs, nlayers, perms, nqubits, ...), the others
 * (vlist, nlayers, nqubits, perms, ...), as in the reference.
 */
#pragma once

#include "matrix.h"
#include "statevector.h"
#include "rng.h"

typedef int (*linear_func)(const struct statevector* RQC_RESTRICT psi, void* fdata, struct statevector* RQC_RESTRICT psi_out);

RQC_EXTERN_C_BEGIN

/* reference src/target.c:15-75 */
int circuit_unitary_target(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const int ngates, const int wires[], const int nqubits,
	numeric* fval);

/* reference src/target.c:158-247 */
int circuit_unitary_target_and_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const int ngates, const int wires[], const int nqubits,
	numeric* fval, struct mat4x4 dgates[]);

/* reference src/target.c:256-280 */
int circuit_unitary_target_and_projected_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const int ngates, const int wires[], const int nqubits,
	numeric* fval, double* grad_vec);

/* reference src/target.c:289-380 */
int circuit_unitary_target_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const struct mat4x4 gatedirs[], const int ngates, const int wires[], const int nqubits,
	numeric* fval, struct mat4x4 dgates[], struct mat4x4 hess_gatedirs[]);

/* reference src/target.c:389-436 */
int circuit_unitary_target_projected_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const struct mat4x4 gatedirs[], const int ngates, const int wires[], const int nqubits,
	numeric* fval, double* RQC_RESTRICT grad_vec, double* RQC_RESTRICT hvp_vec);

/* reference src/target.c:467-481 */
int brickwall_unitary_target(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	numeric* fval);

/* reference src/target.c:514-542 */
int brickwall_unitary_target_and_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	numeric* fval, struct mat4x4 dvlist[]);

/* reference src/target.c:551-575 */
int brickwall_unitary_target_and_projected_gradient(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	numeric* fval, double* grad_vec);

/* reference src/target.c:584-622 */
int brickwall_unitary_target_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const struct mat4x4 vdirs[], const int nlayers, const int* perms[], const int nqubits,
	numeric* fval, struct mat4x4 dvlist[], struct mat4x4 hess_vdirs[]);

/* reference src/target.c:631-678 — one "grad+HVP eval" */
int brickwall_unitary_target_projected_hessian_vector_product(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const struct mat4x4 vdirs[], const int nlayers, const int* perms[], const int nqubits,
	numeric* fval, double* RQC_RESTRICT grad_vec, double* RQC_RESTRICT hvp_vec);

/* Stochastic-trace targets: the trace over the 2^L basis columns is estimated
   from nsamples Haar-random states drawn from rng (in the reference's order:
   sample by sample, crandn entries then normalisation) and scaled by
   2^L / nsamples; rng advances as in the reference. */

/* reference src/target.c:85-149 */
int circuit_unitary_target_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 gates[], const int ngates, const int wires[], const int nqubits,
	const long nsamples, struct rng_state* rng, numeric* fval);

/* reference src/target.c:491-505 */
int brickwall_unitary_target_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	const long nsamples, struct rng_state* rng, numeric* fval);

/* reference src/target.c:807-928 */
int brickwall_unitary_target_gradient_hessian_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	const long nsamples, struct rng_state* rng,
	numeric* fval, struct mat4x4 dvlist[], numeric* hess);

/* reference src/target.c:1050-1088 */
int brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	const long nsamples, struct rng_state* rng,
	numeric* fval, double* grad_vec, double* H);

/* reference src/target.c:687-797: f, the gate gradients and the complex Hessian
   hess[((i*16 + x)*nlayers + j)*16 + y] ((16 nlayers)^2 numerics) */
int brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	numeric* fval, struct mat4x4 dvlist[], numeric* hess);

/* reference src/target.c:1003-1040 (projection :937-994): f, the projected gradient (16 nlayers reals)
   and the Riemannian Hessian matrix H ((16 nlayers)^2 reals, row-major, symmetric) */
int brickwall_unitary_target_gradient_vector_hessian_matrix(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	numeric* fval, double* grad_vec, double* H);

/* Block-encoding targets f = || P^dagger W P - H ||_F^2 / 2 (reference
   src/target.h:97-126, src/target.c:1100-1720): W is the brickwall circuit,
   P embeds a state of nqubits/2 qubits on the odd amplitude-index bits (bit i
   to bit 2i + 1, the even bits |0>), and hfunc applies H to states of
   nqubits/2 qubits (rqc_device_ufunc with a device target of nqubits/2
   qubits keeps H on the GPU).  nqubits must be even. */

/* reference src/target.c:1100-1190 */
int brickwall_blockenc_target(linear_func hfunc, void* hdata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	double* fval);

/* reference src/target.c:1202-1330 */
int brickwall_blockenc_target_and_gradient(linear_func hfunc, void* hdata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	double* fval, struct mat4x4 dvlist[]);

/* reference src/target.c:1341-1365 */
int brickwall_blockenc_target_and_gradient_vector(linear_func hfunc, void* hdata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	double* fval, double* grad_vec);

/* reference src/target.c:1380-1620: hess1 (second derivative in the gates with
   the residual held fixed, complex symmetric) and hess2 (in the gates and
   their conjugates, Hermitian), both (16 nlayers)^2 numerics in the layout of
   brickwall_unitary_target_gradient_hessian */
int brickwall_blockenc_target_gradient_hessian(linear_func hfunc, void* hdata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	double* fval, struct mat4x4 dvlist[], numeric* hess1, numeric* hess2);

/* reference src/target.c:1631-1720 */
int brickwall_blockenc_target_gradient_vector_hessian_matrix(linear_func hfunc, void* hdata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	double* fval, double* grad_vec, double* H);

RQC_EXTERN_C_END
