/*
 * Brickwall circuits: layer i applies the shared gate vlist[i] to the pairs
 * (perms[i][2k], perms[i][2k+1]), k = 0 .. nqubits/2-1; nqubits must be even.
 * Replaces reference src/brickwall_circuit.h:8-18 (the four functions below).
 */
#pragma once

#include "matrix.h"
#include "statevector.h"
#include "quantum_circuit.h"

RQC_EXTERN_C_BEGIN

/* reference src/brickwall_circuit.c:14-44 */
int apply_brickwall_unitary(const struct mat4x4 vlist[], const int nlayers, const int* perms[],
	const struct statevector* RQC_RESTRICT psi, struct statevector* RQC_RESTRICT psi_out);

/* reference src/brickwall_circuit.c:51-78; cache holds nlayers*nqubits/2 states */
int brickwall_unitary_forward(const struct mat4x4 vlist[], int nlayers, const int* perms[],
	const struct statevector* RQC_RESTRICT psi, struct quantum_circuit_cache* cache, struct statevector* RQC_RESTRICT psi_out);

/* reference src/brickwall_circuit.c:85-124; dvlist[i] sums the gradients of the pairs of layer i */
int brickwall_unitary_backward(const struct mat4x4 vlist[], int nlayers, const int* perms[], const struct quantum_circuit_cache* cache,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, struct mat4x4 dvlist[]);

/* reference src/brickwall_circuit.c:134-255: the backward pass with the gate gradients and the complex
   Hessian hess[((i*16 + x)*nlayers + j)*16 + y] = d^2 <dpsi_out, C psi> / dV_i[x] dV_j[y] of one input
   state (hess holds (16 nlayers)^2 numerics) */
int brickwall_unitary_backward_hessian(const struct mat4x4 vlist[], int nlayers, const int* perms[], const struct quantum_circuit_cache* cache,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, struct mat4x4 dvlist[], numeric* hess);

/* reference src/brickwall_circuit.c:264-310: the placeholder states of every gate of
   layer l (from the cached states), each carried through all later gates, summed:
   psi_out (16 states) holds the derivative of the circuit output by the 16 entries of
   layer l's gate */
int apply_brickwall_unitary_gate_placeholder(const struct mat4x4 vlist[], int nlayers, const int* perms[], int l,
	const struct quantum_circuit_cache* cache, struct statevector_array* RQC_RESTRICT psi_out);

RQC_EXTERN_C_END
