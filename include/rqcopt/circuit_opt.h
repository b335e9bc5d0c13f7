/*
 * Riemannian trust-region optimisation of a general circuit of two-qubit
 * gates (one independent gate per wire pair).  Replaces reference
 * src/circuit_opt.h:8 (optimize_quantum_circuit).  The driver is host code;
 * its target, gradient and HVP evaluations run on the GPU through target.h.
 */
#pragma once

#include "quantum_circuit.h"
#include "target.h"
#include "trust_region.h"

RQC_EXTERN_C_BEGIN

/* reference src/circuit_opt.c:104-124 (wrappers :22-79) */
void optimize_quantum_circuit(linear_func ufunc, void* udata, const struct mat4x4 gates_start[], const int ngates, const int nqubits,
	const int wires[], struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 gates_opt[]);

RQC_EXTERN_C_END
