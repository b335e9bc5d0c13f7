/*
 * Brickwall circuit optimisation by the RTR-HVP driver.
 * Replaces reference src/brickwall_opt.h:8-10 (optimize_brickwall_circuit_hvp).
 */
#pragma once

#include "brickwall_circuit.h"
#include "target.h"
#include "trust_region.h"
#include "rng.h"

RQC_EXTERN_C_BEGIN

/* reference src/brickwall_opt.c:127-150; f_iter has niter + 1 entries */
void optimize_brickwall_circuit_hvp(linear_func ufunc, void* udata,
	const struct mat4x4 vlist_start[], const int nlayers, const int nqubits, const int* perms[],
	struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 vlist_opt[]);

void optimize_brickwall_circuit_hmat(linear_func ufunc, void* udata,
	const struct mat4x4 vlist_start[], const int nlayers, const int nqubits, const int* perms[],
	struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 vlist_opt[]);

/* reference src/brickwall_opt.c:240-268: every evaluation draws fresh samples from rng */
void optimize_brickwall_circuit_hmat_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 vlist_start[], const int nlayers, const int nqubits, const int* perms[],
	const long nsamples, struct rng_state* rng,
	struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 vlist_opt[]);

RQC_EXTERN_C_END
