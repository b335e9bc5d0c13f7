/*
 * Request server on one GPU (additive to the reference ABI; no counterpart in
 * the reference, which evaluates one target at a time on the caller's thread,
 * reference src/target.c:389-436).
 *
 * Target, gradient and HVP evaluations from many concurrent optimisations
 * arrive as a stream of requests.  Each request belongs to a session: a
 * circuit structure, its device-resident target and a precision, whose
 * evaluation plans are made once.  The server owns one device arena; a
 * request is admitted in arrival order when an execution slot (a worker with
 * its own HIP stream) is free and its workspace fits into the free part of
 * the arena.  Admitted requests run concurrently on their streams, and a
 * finished request frees its slot and memory at once for the next one
 * (continuous batching).  A request's column block is fixed by the per-request
 * workspace cap, never by the memory free at the time, so its result does not
 * depend on what else is in flight.
 */
#pragma once

#include <stddef.h>
#include "config.h"
#include "matrix.h"
#include "rqc_device.h"
#include "rqc_batch.h"
#include "trust_region.h"

struct rqc_server;   /* opaque */
struct rqc_session;  /* opaque */

struct rqc_server_params
{
	size_t arena_bytes;        /* device memory for request workspaces (0 = 80% of the free memory) */
	size_t max_request_bytes;  /* state-array cap of one request; fixes its column block (0 = arena / 4) */
	int nslots;                /* requests in flight on the device (0 = 8) */
};

struct rqc_request_info
{
	int status;                /* 0 done, < 0 error code, 1 queued or running */
	int slot;                  /* execution slot that ran it */
	double queued_ms;          /* submission to admission */
	double run_ms;             /* admission to results on the host */
	double latency_ms;         /* submission to results on the host */
	double start_ms, end_ms;   /* admission and completion, since the server was created */
	size_t workspace_bytes;    /* arena bytes the request held */
	long block;                /* column block of its sweeps */
};

struct rqc_server_stats
{
	long submitted;
	long completed;
	long failed;
	int peak_in_flight;        /* most requests running at once */
	size_t arena_bytes;
	size_t peak_arena_bytes;   /* most arena bytes held at once */
};

RQC_EXTERN_C_BEGIN

int rqc_server_create(const struct rqc_server_params* params, struct rqc_server** server);
/* waits for every request, then releases the server (sessions must be destroyed first or are released here) */
void rqc_server_destroy(struct rqc_server* server);

/* a brickwall session (even nqubits; one gate per layer) or a general-circuit
   session (one gate per wire pair); precision RQC_PRECISION_FP64 or _MIXED */
int rqc_session_create_brickwall(struct rqc_server* server, int nqubits, int nlayers, const int* perms[],
	const struct rqc_target* target, int precision, struct rqc_session** session);
int rqc_session_create_circuit(struct rqc_server* server, int nqubits, int ngates, const int wires[],
	const struct rqc_target* target, int precision, struct rqc_session** session);
/* number of gate groups (layers of a brickwall, gates of a circuit) */
int rqc_session_ngroups(const struct rqc_session* session);
void rqc_session_destroy(struct rqc_session* session);

/* queue one evaluation (kind: RQC_EVAL_TARGET, _GRADIENT or _HVP).  gates: one
   per group, or NULL to reuse the session's gates of its previous request;
   dirs: one per group (HVP only).  Inputs are copied before the call returns;
   fval, grad and hvp (16 reals per group, projected; NULL to skip) are written
   when the request completes. */
int rqc_server_submit(struct rqc_server* server, struct rqc_session* session, int kind, const struct mat4x4 gates[],
	const struct mat4x4 dirs[], numeric* fval, double* grad, double* hvp, long* ticket);
/* block until the request is done (returns its status) */
int rqc_server_wait(struct rqc_server* server, long ticket, struct rqc_request_info* info);
/* 1 if done, 0 if not yet; info filled when done */
int rqc_server_poll(struct rqc_server* server, long ticket, struct rqc_request_info* info);
/* forget a finished request's record (records are otherwise kept until the server is destroyed) */
int rqc_server_release(struct rqc_server* server, long ticket);
int rqc_server_get_stats(const struct rqc_server* server, struct rqc_server_stats* stats);

/* one trust-region optimisation (the algorithm of optimize_brickwall_circuit_hvp, reference
   src/brickwall_opt.c:127-150, or for a general-circuit session of optimize_quantum_circuit,
   src/circuit_opt.c:104-124: one gate per group) of a session on the calling thread, whose target,
   gradient and HVP evaluations are requests of this server: solves on several host threads
   share the device like any other requests.  f_iter receives niter + 1 values, vlist_opt the
   optimised gates; records (up to max_records, may be NULL) receive the info of every request
   in issue order and *nrecords their number. */
int rqc_server_optimize_brickwall_hvp(struct rqc_server* server, struct rqc_session* session,
	const struct mat4x4 vlist_start[], const struct rtr_params* params, int niter, double* f_iter,
	struct mat4x4 vlist_opt[], struct rqc_request_info* records, int max_records, int* nrecords);

RQC_EXTERN_C_END
