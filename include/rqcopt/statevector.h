/*
 * Statevectors and interleaved statevector arrays.
 * Replaces reference src/statevector.h:11-49.
 *
 * Storage is host memory owned by the caller, as in the reference.  The
 * allocators return page-locked (pinned) host memory when a GPU is present so
 * that uploads to the HIP kernels run at full PCIe rate; free_* releases it.
 * Every compute entry point of this library also accepts `data` pointers that
 * are HIP device allocations (detected per call) and then works in place on
 * the device without host copies.
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

/* psi_sum = psi1 + psi2 on the GPU (reference src/statevector.c:132-143) */
void add_statevectors(const struct statevector* RQC_RESTRICT psi1, const struct statevector* RQC_RESTRICT psi2,
	struct statevector* RQC_RESTRICT psi_sum);

/* reference src/statevector.c:149-178 */
int allocate_statevector_array(int nqubits, int nstates, struct statevector_array* psi_array);
void free_statevector_array(struct statevector_array* psi_array);

RQC_EXTERN_C_END
