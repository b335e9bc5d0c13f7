/*
 * Sweeps of a general circuit of two-qubit gates on explicit wires, on the GPU.
 * Replaces reference src/quantum_circuit.h:7-36.
 * wires[2k], wires[2k+1] are the wires of gate k; gates run in list order.
 */
#pragma once

#include "matrix.h"
#include "statevector.h"

/* cache of intermediate states: psi_list[k] is the state before gate k
   (reference src/quantum_circuit.h:16-21) */
struct quantum_circuit_cache
{
	struct statevector* psi_list;
	int nqubits;
	int ngates;
};

RQC_EXTERN_C_BEGIN

/* reference src/quantum_circuit.c:13-35 */
int apply_quantum_circuit(const struct mat4x4 gates[], const int ngates, const int wires[],
	const struct statevector* RQC_RESTRICT psi, struct statevector* RQC_RESTRICT psi_out);

/* reference src/quantum_circuit.c:42-83 */
int allocate_quantum_circuit_cache(const int nqubits, const int ngates, struct quantum_circuit_cache* cache);
void free_quantum_circuit_cache(struct quantum_circuit_cache* cache);

/* reference src/quantum_circuit.c:90-108 */
int quantum_circuit_forward(const struct mat4x4 gates[], const int ngates, const int wires[],
	const struct statevector* RQC_RESTRICT psi, struct quantum_circuit_cache* cache, struct statevector* RQC_RESTRICT psi_out);

/* reference src/quantum_circuit.c:115-140 */
int quantum_circuit_backward(const struct mat4x4 gates[], const int ngates, const int wires[], const struct quantum_circuit_cache* cache,
	const struct statevector* RQC_RESTRICT dpsi_out, struct statevector* RQC_RESTRICT dpsi, struct mat4x4 dgates[]);

/* output state, gate gradients of <phi, C psi> and their directional derivative
   along gatedirs (reference src/quantum_circuit.c:147-236) */
int quantum_circuit_gates_hessian_vector_product(const struct mat4x4 gates[], const struct mat4x4 gatedirs[], const int ngates, const int wires[],
	const struct statevector* RQC_RESTRICT psi, const struct statevector* RQC_RESTRICT phi,
	struct statevector* RQC_RESTRICT psi_out, struct mat4x4 dgates[], struct mat4x4 hess_gatedirs[]);

RQC_EXTERN_C_END
