/*
 * Single-qubit gates (additive to the reference ABI, which has two-qubit gates only:
 * reference src/gate.h:7-13).  Conventions follow the two-qubit functions: qubit i is
 * bit L-1-i of the amplitude index, gates are row-major, the backward map applies the
 * transpose (not the adjoint) and the gate gradient is sum dpsi_out (x) psi without
 * conjugation (reference src/gate.c:83-188).  Statevectors may be host or HIP device
 * memory, as for the two-qubit functions.  Functions return 0, or < 0 on error (message
 * on stderr).
 */
#pragma once

#include "config.h"
#include "statevector.h"
#include "target.h"

RQC_EXTERN_C_BEGIN

/* psi_out = V_i psi, V a row-major 2x2 gate (4 numerics) on wire i */
int rqc_apply_single_qubit_gate(const numeric gate[4], int i, const struct statevector* RQC_RESTRICT psi,
	struct statevector* RQC_RESTRICT psi_out);
/* dpsi = V_i^T dpsi_out; dgate[2 r + c] = sum over the other wires of dpsi_out(r) psi(c) */
int rqc_apply_single_qubit_gate_backward(const numeric gate[4], int i, const struct statevector* RQC_RESTRICT psi,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, numeric dgate[4]);
/* psi_out = V_i psi for every state of an interleaved array (the layout of statevector_array) */
int rqc_apply_single_qubit_gate_to_array(const numeric gate[4], int i, const struct statevector_array* RQC_RESTRICT psi,
	struct statevector_array* RQC_RESTRICT psi_out);

/* Circuits of one- and two-qubit gates: gate k acts on nq[k] (1 or 2) wires, wires[2k]
   (and wires[2k + 1]); its entries are packed in gates (and gatedirs, dgates, hess_gatedirs)
   one gate after the other, 4 numerics for a one-qubit gate, 16 for a two-qubit gate.
   The targets, gradients and Hessian-vector products are those of
   circuit_unitary_target, circuit_unitary_target_and_gradient and
   circuit_unitary_target_hessian_vector_product (reference src/target.c:15-75,158-247,
   289-380) for such circuits.  On the GPU a one-qubit gate V on wire i runs as the
   two-qubit gate V (x) I on wires (i, (i + 1) mod nqubits) in the fused tile passes; its
   gradient and directional second derivative are the partial traces of the two-qubit
   gate's over the second wire. */
int rqc_mixed_circuit_unitary_target(linear_func ufunc, void* udata, const numeric* gates, const int nq[],
	const int wires[], int ngates, int nqubits, numeric* fval);
int rqc_mixed_circuit_unitary_target_and_gradient(linear_func ufunc, void* udata, const numeric* gates, const int nq[],
	const int wires[], int ngates, int nqubits, numeric* fval, numeric* dgates);
int rqc_mixed_circuit_unitary_target_hessian_vector_product(linear_func ufunc, void* udata, const numeric* gates,
	const numeric* gatedirs, const int nq[], const int wires[], int ngates, int nqubits, numeric* fval, numeric* dgates,
	numeric* hess_gatedirs);

RQC_EXTERN_C_END
