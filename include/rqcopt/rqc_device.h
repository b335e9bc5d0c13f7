/*
 * GPU-specific additions to the C ABI (MI355X / gfx950).  None of these
 * replaces a reference function: they are additive and do not change the
 * signatures of gate.h, statevector.h, quantum_circuit.h, brickwall_circuit.h
 * or target.h.
 *
 * Device-resident targets: the reference passes U as a host callback that is
 * invoked once per basis column (reference src/target.h:7).  A callback cannot
 * run inside a GPU sweep, so this library exports a sentinel callback,
 * rqc_device_ufunc.  Passing it as `ufunc` with an rqc_target* as `udata`
 * keeps U resident in HBM across calls; the trust-region driver and any
 * other caller of target.h use it unchanged.
 */
#pragma once

#include <stddef.h>
#include "config.h"
#include "matrix.h"
#include "statevector.h"
#include "target.h"

struct rqc_target;  /* opaque: a dense 2^L x 2^L complex128 U in device memory */

enum rqc_precision
{
	RQC_PRECISION_FP64  = 0,  /* complex FP64 sweeps and reductions */
	RQC_PRECISION_MIXED = 1,  /* complex FP32 sweeps, FP64 reductions, FP64 projection */
};

enum rqc_sweep_engine
{
	RQC_ENGINE_TILE_PASS = 0,  /* several gates per HBM pass through LDS tiles (default) */
	RQC_ENGINE_PER_GATE  = 1,  /* one gate per pass over the column block */
};

enum rqc_kernel_kind
{
	RQC_KERNEL_FORWARD = 0,   /* forward sweep of one gate over the column block */
	RQC_KERNEL_REVERSE = 1,   /* reverse (adjoint) sweep of one gate with gradient / HVP reductions */
	RQC_KERNEL_TARGET  = 2,   /* -conj(U) seed and trace reduction */
	RQC_KERNEL_OTHER   = 3,   /* fills, partial-sum reductions, projection */
	RQC_KERNEL_KINDS   = 4,
};

RQC_EXTERN_C_BEGIN

/* Sentinel linear_func.  Called directly it computes psi_out = U psi with U
   taken from the rqc_target in fdata (on the GPU). */
int rqc_device_ufunc(const struct statevector* RQC_RESTRICT psi, void* fdata, struct statevector* RQC_RESTRICT psi_out);

/* U given densely (row-major, 2^nqubits x 2^nqubits; host or device pointer): copied to device memory.
   Device targets hold 1 .. 15 qubits (U up to 17 GB); larger ones return -3 (untested sizes). */
int rqc_target_create_dense(int nqubits, const numeric* U, struct rqc_target** target);
/* U assembled once from a host callback (2^nqubits calls) and kept on the device */
int rqc_target_create_from_ufunc(linear_func ufunc, void* udata, int nqubits, struct rqc_target** target);
/* U = brickwall circuit (e.g. a Trotter product of exact bond layers), formed on the GPU */
int rqc_target_create_brickwall(int nqubits, const struct mat4x4 vlist[], int nlayers, const int* perms[], struct rqc_target** target);
/* U = a general circuit of two-qubit gates (gates[k] on wires[2k], wires[2k+1]), formed on the GPU */
int rqc_target_create_circuit(int nqubits, const struct mat4x4 gates[], int ngates, const int wires[], struct rqc_target** target);
void rqc_target_destroy(struct rqc_target* target);
int rqc_target_nqubits(const struct rqc_target* target);
/* copy U back to a host buffer of 4^nqubits numerics */
int rqc_target_download(const struct rqc_target* target, numeric* U);

/* precision of the target / gradient / HVP sweeps (process-wide; default FP64) */
int rqc_set_precision(int mode);
int rqc_get_precision(void);

/* sweep engine of the target / gradient / HVP functions (process-wide); also
   selectable with the environment variable RQC_SWEEP_ENGINE=pergate */
int rqc_set_sweep_engine(int engine);

/* tile-shape autotuning of batches (process-wide; default on, or the
   environment variable RQC_AUTOTUNE=0 switches it off): the tile height and
   the gates per pass are timed on whole steps of the batch, once per batch
   structure and precision */
int rqc_set_autotune(int on);

/* tile autotuning of single evaluations (process-wide; default off, or the environment
   variable RQC_EVAL_AUTOTUNE=1): the first target / gradient / HVP evaluation of a circuit
   structure times the tile height (2^10 or 2^9 FP64 elements, 2^11 or 2^10 FP32, which fixes the
   fused qubits K per pass) and the most gates per pass (8 or 6) on the evaluation itself and keeps
   the fastest for that structure; on = 2 also forgets the choices made so far */
int rqc_set_eval_autotune(int on);
/* tile shape (2^rows x 2^cols amplitudes) and most gates per pass of the last evaluation */
int rqc_last_tiles(int* tile_rows_log2, int* tile_cols_log2, int* gates_per_pass);

/* device workspace: upper bound in bytes (0 = automatic, about 85% of free HBM);
   the workspace is kept between calls and released here */
int rqc_set_workspace_limit(size_t bytes);
void rqc_release_workspace(void);
/* column-block width used by the last target evaluation */
long rqc_last_column_block(void);
/* sweep engine (RQC_ENGINE_*) of the last target evaluation and its number of
   HBM passes per sweep direction (the gate count for the per-gate engine);
   the tile-pass engine hands gradient / HVP sweeps whose gates are not unitary
   to 1e-6, and circuits too small for one tile, to the per-gate engine */
int rqc_last_sweep(int* engine, int* passes);

/* kernel timing with HIP events on the library's stream: when enabled, every
   launch is bracketed by events and its duration is added per kind, together
   with the launch's algorithmic HBM bytes (each state array it must read or
   write counted once; any output pointer may be NULL) */
void rqc_profile_enable(int on);
void rqc_profile_reset(void);
int rqc_profile_query(int kind, double* total_ms, long* launches, double* alg_bytes);
/* algorithmic flops of the timed tile-pass launches of one kind (per amplitude and gate of the
   tiles a launch processes: forward 96, reverse 192 for the adjoint update and the outer
   products, plain sweeps a third of these; a reverse launch that carries a forward pass adds
   that pass's), and the flops the matrix cores issue for them: the reverse passes' state
   recovery besides, in the forms the kernels use (FP64 dual: forward 72, reverse 240, with three
   real products per complex gate product; FP32 the real 16 x 16 form: forward 128, reverse 384).
   No reference counterpart: the library's own measurement. */
int rqc_profile_query_flops(int kind, double* alg_flops, double* issued_flops);

/* device-to-device copy of `bytes` bytes by a plain streaming kernel (reads
   `bytes`, writes `bytes`), timed with HIP events over `reps` launches; gives the
   measured copy bandwidth in GB/s (read + write bytes per second).  Used to
   calibrate HBM byte counters and the roofline peak on the running device. */
int rqc_measure_copy_bandwidth(size_t bytes, int reps, double* gbps);

/* the same measurement with a copy kernel tuned for HBM bandwidth (four 16-byte elements per
   thread in flight, non-temporal loads and stores): the device's streaming ceiling, against
   which the tile passes' HBM rates are compared */
int rqc_measure_stream_bandwidth(size_t bytes, int reps, double* gbps);

/* No counterpart in the reference: the same evaluation as
   brickwall_unitary_target_projected_hessian_vector_product (target.h), for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  It allocates and frees no device memory: its whole workspace is a
   static device array of rqc_noalloc_scratch_bytes() bytes (256 MiB), held by each GPU the
   library runs on until the process ends, and the column block is the widest that fits it.
   U must be a device target (ufunc = rqc_device_ufunc, udata = rqc_target*).  A circuit whose
   single-column plan exceeds the scratch returns -3 with a message.  Calls are serialised by
   the library and finish before they return; the scratch keeps the last call's contents, and
   anything that overlapped it on one GPU outside the library would race. */
int rqc_brickwall_unitary_target_projected_hvp_noalloc(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const struct mat4x4 vdirs[], const int nlayers, const int* perms[], const int nqubits,
	numeric* fval, double* RQC_RESTRICT grad_vec, double* RQC_RESTRICT hvp_vec);
size_t rqc_noalloc_scratch_bytes(void);

/* blocks until all work of the library's stream has finished */
int rqc_synchronize(void);
/* "gfx950 <name> <CUs> <HBM bytes>" into buf; returns 0 when a device is present */
int rqc_device_info(char* buf, int len);

RQC_EXTERN_C_END
