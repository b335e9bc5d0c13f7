/*
 * Two-qubit gate application and its backward pass, on the GPU.
 * Replaces reference src/gate.h:7-19 (apply_gate, apply_gate_backward,
 * apply_gate_to_array, apply_gate_backward_array, apply_gate_placeholder).
 *
 * Convention: qubit q is bit (nqubits-1-q) of the amplitude index (qubit 0 is
 * the most significant); the gate acts on |q_i q_j> with wire i the more
 * significant index of the 4x4 matrix; any i != j order is allowed.
 * The backward pass uses the holomorphic convention of the reference:
 * dpsi = gate^T dpsi_out, dgate = sum dpsi_out (x) psi (no conjugation).
 */
#pragma once

#include "matrix.h"
#include "statevector.h"

RQC_EXTERN_C_BEGIN

/* reference src/gate.c:10-76 */
void apply_gate(const struct mat4x4* gate, const int i, const int j, const struct statevector* RQC_RESTRICT psi,
	struct statevector* RQC_RESTRICT psi_out);

/* reference src/gate.c:83-188 */
void apply_gate_backward(const struct mat4x4* gate, const int i, const int j, const struct statevector* RQC_RESTRICT psi,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, struct mat4x4* dgate);

/* reference src/gate.c:195-262 */
void apply_gate_to_array(const struct mat4x4* gate, const int i, const int j, const struct statevector_array* RQC_RESTRICT psi,
	struct statevector_array* RQC_RESTRICT psi_out);

/* reference src/gate.c:269-366: dgates[k] = sum dpsi_out (x) psi[:, k] for every state k of the array
   (the gate enters only through its wires) */
void apply_gate_backward_array(const struct mat4x4* gate, const int i, const int j, const struct statevector_array* RQC_RESTRICT psi,
	const struct statevector* RQC_RESTRICT dpsi_out, struct mat4x4* dgates);

/* reference src/gate.c:374-460: the 16 states E_rc psi (state 4r + c) of a placeholder gate on (i, j);
   psi_out must hold 16 states */
void apply_gate_placeholder(const int i, const int j, const struct statevector* psi, struct statevector_array* psi_out);

RQC_EXTERN_C_END
