/*
 * Statevectors and interleaved statevector arrays.
 * Replaces reference src/statevector.h:11-49 (all of it).
 *
 * Storage is host memory owned by the caller, as in the reference (64-byte
 * aligned allocations).  Every compute entry point of this library also
 * accepts `data` pointers that are HIP device allocations (detected per call)
 * and then works in place on the device without host copies.
 */
#pragma once

#include "config.h"
#include "rng.h"

/* reference src/statevector.h:11-15 */
struct statevector
{
	numeric* data;  /* 2^nqubits entries */
	int nqubits;
};

/* interleaved array of statevectors: data[amp * nstates + k] (reference src/statevector.h:39-44) */
struct statevector_array
{
	numeric* data;  /* 2^nqubits x nstates entries */
	int nqubits;
	int nstates;
};

RQC_EXTERN_C_BEGIN

/* reference src/statevector.c:12-45; returns 0 or -1 on allocation failure */
int allocate_statevector(int nqubits, struct statevector* psi);
void free_statevector(struct statevector* psi);

/* psi /= ||psi|| on the GPU, unless psi is zero (reference src/statevector.c:47-66) */
void normalize_statevector(struct statevector* RQC_RESTRICT psi);

/* permute the qubit wires: bit i of psi_trans's index is bit nqubits-1-perm[nqubits-1-i] of psi's
   (reference src/statevector.c:72-92), on the GPU */
void transpose_statevector(const struct statevector* RQC_RESTRICT psi, const int* perm, struct statevector* RQC_RESTRICT psi_trans);

/* crandn entries drawn in order from rng, then normalised (reference src/statevector.c:102-127);
   the generator advances exactly as the reference's does */
void haar_random_statevector(struct statevector* psi, struct rng_state* rng);

/* psi_sum = psi1 + psi2 on the GPU (reference src/statevector.c:132-143) */
void add_statevectors(const struct statevector* RQC_RESTRICT psi1, const struct statevector* RQC_RESTRICT psi2,
	struct statevector* RQC_RESTRICT psi_sum);

/* reference src/statevector.c:149-178 */
int allocate_statevector_array(int nqubits, int nstates, struct statevector_array* psi_array);
void free_statevector_array(struct statevector_array* psi_array);

RQC_EXTERN_C_END
