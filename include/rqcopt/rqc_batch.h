/*
 * Batched brickwall optimisation on one GPU (additive to the reference ABI; no
 * counterpart in the reference, whose optimisations run one at a time through
 * optimize_brickwall_circuit_hvp, reference src/brickwall_opt.c:127-150).
 *
 * A batch holds many independent problems of uneven qubit count and depth.
 * Every problem has its own device workspace, planned once; the evaluations
 * of one step (each problem's target, gradient or HVP) are enqueued as one
 * HIP graph per evaluation kind whose independent per-problem chains overlap
 * on the device.  Problems that sit out a step are skipped by device flags,
 * so the graphs are captured once and replayed for every step.  Results of a
 * problem do not depend on the other problems of the batch (same plan, same
 * grids, fixed-order reductions).
 */
#pragma once

#include "config.h"
#include "matrix.h"
#include "rqc_device.h"
#include "trust_region.h"

struct rqc_batch;  /* opaque */

/* one problem: a brickwall circuit (even nqubits >= 6) and its device target */
struct rqc_batch_problem
{
	int nqubits;
	int nlayers;
	const int* const* perms;          /* nlayers permutations of the nqubits wires */
	const struct rqc_target* target;  /* dense U resident in device memory */
};

enum rqc_eval_kind
{
	RQC_EVAL_NONE     = -1,  /* the problem sits out this step */
	RQC_EVAL_TARGET   = 0,   /* f = -Tr(U^dagger C) */
	RQC_EVAL_GRADIENT = 1,   /* f and the projected gradient (16 * nlayers reals) */
	RQC_EVAL_HVP      = 2,   /* f, projected gradient and projected Hessian-vector product */
};

enum rqc_batch_exec
{
	RQC_BATCH_GRAPH = 0,  /* one captured HIP graph launch per evaluation kind and step (default) */
	RQC_BATCH_EAGER = 1,  /* direct launches of the active problems on the batch's streams */
};

enum rqc_batch_rtr
{
	RQC_BATCH_RTR_DEVICE = 0,  /* trust region on the GPU: tCG, projection, retraction and radius updates
	                              in FP64 device kernels between graph launches (default) */
	RQC_BATCH_RTR_HOST   = 1,  /* one host driver thread per problem (the reference's driver), in lock-step */
};

struct rqc_batch_stats
{
	long steps;            /* evaluation steps run */
	long evaluations;      /* problem evaluations run */
	long graph_launches;   /* HIP graph launches */
	double device_ms;      /* summed wall time of the steps (upload to results on the host) */
	size_t workspace_bytes;/* device memory held by the batch */
};

RQC_EXTERN_C_BEGIN

/* precision: RQC_PRECISION_FP64 or RQC_PRECISION_MIXED; nstreams: parallel
   branches of the step graphs (0 = automatic); workspace_limit: bytes of
   device memory for all problems together (0 = 85% of the free memory) */
int rqc_batch_create(int nproblems, const struct rqc_batch_problem problems[], int precision, int nstreams,
	size_t workspace_limit, struct rqc_batch** batch);
void rqc_batch_destroy(struct rqc_batch* batch);
int rqc_batch_set_exec(struct rqc_batch* batch, int exec);

/* one step: problem p is evaluated in kinds[p] at vlists[p] (nlayers gates),
   with directions vdirs[p] for RQC_EVAL_HVP (else ignored, may be NULL);
   writes fvals[p], grads[p] and hvps[p] (16 * nlayers doubles; NULL to skip) */
int rqc_batch_evaluate(struct rqc_batch* batch, const int kinds[], const struct mat4x4* const vlists[],
	const struct mat4x4* const vdirs[], numeric fvals[], double* const grads[], double* const hvps[]);

/* one Riemannian trust-region optimisation per problem (the algorithm of
   optimize_brickwall_circuit_hvp, reference src/brickwall_opt.c:127-150 over
   src/trust_region.c:159-217,306-368), all run in lock-step: each step evaluates
   every problem that has a pending target, gradient or HVP request.  params
   applies to every problem (its g_func is ignored); f_iters[p] receives
   niter + 1 values, vlist_opt[p] the optimised gates.
   RQC_BATCH_RTR_DEVICE (default): the driver's state (gates, gradient, tCG vectors,
   radius) stays on the GPU; FP64 kernels run the tCG recurrences, the projection to
   tangent directions, the polar retraction and the acceptance test, and refill the
   gate tables, between launches of the captured evaluation graphs, so a tCG step needs
   no host round trip (the host reads one active-problem count per chunk of steps).
   RQC_BATCH_RTR_HOST: one host driver thread per problem. */
int rqc_batch_optimize_brickwall_hvp(struct rqc_batch* batch, const struct mat4x4* const vlist_start[],
	const struct rtr_params* params, int niter, double* const f_iters[], struct mat4x4* const vlist_opt[]);
int rqc_batch_set_rtr(struct rqc_batch* batch, int mode);

int rqc_batch_get_stats(const struct rqc_batch* batch, struct rqc_batch_stats* stats);
/* column block, sweep engine and passes of problem p for an evaluation kind */
int rqc_batch_problem_plan(const struct rqc_batch* batch, int p, int kind, long* block, int* engine, int* passes);
/* tile shape (2^rows x 2^cols amplitudes) and the most gates per pass of problem p: chosen by the
   autotuner at rqc_batch_create (rqc_set_autotune), which times whole HVP steps of the batch for
   each candidate once per batch structure and precision */
int rqc_batch_problem_tiles(const struct rqc_batch* batch, int p, int kind, int* tile_rows_log2, int* tile_cols_log2,
	int* gates_per_pass);

RQC_EXTERN_C_END
