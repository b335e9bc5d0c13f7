// Planned target evaluations of librqcopt_hip: the interface between the
// sweep engine (rqc_engine.hip) and the runtimes that run many evaluations at
// once (batched optimisation and the request server, rqc_serving.hip).
//
// An evaluation is split into three steps so that independent evaluations can
// be placed in their own workspaces, enqueued on their own streams and
// captured in HIP graphs:
//   plan_eval         depends on the circuit structure only (qubits, wires,
//                     mode, precision, column block); fixes the sweep engine,
//                     the passes, the grids and the workspace layout
//   fill_eval_tables  host side, per evaluation: gate matrices (and inverses,
//                     transposes, directions) in the layout the kernels read
//   enqueue_eval      kernel launches only (plus the optional table upload),
//                     on one slot's stream, into one workspace
// Results are written to a fixed region of the workspace: f, then dv and hz
// (16 complex per group), then the projected grad and hvp (16 doubles per group).
#pragma once

#include <memory>
#include <vector>

#include "rqc_internal.hpp"

namespace rqc {

enum class Mode { Target, Grad, Hvp };

// a sequential circuit of two-qubit gates
struct Circuit
{
	int L = 0;
	int ngates = 0;
	std::vector<zc> gates;  // ngates * 16
	std::vector<zc> dirs;   // ngates * 16 or empty
	std::vector<int> wi, wj;
};

// gates (and optional directions) on the given wires (2 * ngates entries)
Circuit make_circuit(const struct mat4x4* gates, const struct mat4x4* dirs, int ngates, const int* wires, int L);
// brickwall expansion: layer i's gate on the pairs (perms[i][2k], perms[i][2k + 1])
Circuit make_brickwall(const struct mat4x4* vlist, const struct mat4x4* vdirs, int nlayers, const int* const* perms, int L);
// every gate unitary to 1e-6 (the tile-pass reverse sweep recovers states through G^-1)
bool gates_near_unitary(const Circuit& C);
// the `unitary` argument of plan_eval for an evaluation of kind `mode`: a plain FP64 target
// evaluation has no reverse sweep and needs no check; every other kind, and a mixed-precision
// target evaluation (which runs the gradient sweep, see plan_eval), needs gates_near_unitary
bool sweep_unitary(const Circuit& C, Mode mode, bool fp32);

struct TilePlan;  // tile-pass specific part of a plan (passes, pass arguments, grids)

// tile height (log2 of the elements of a tile) and gates per pass; 0 = default
struct TileChoice
{
	int logE = 0;
	int gmax = 0;
};

struct EvalPlan
{
	Mode mode = Mode::Target;
	bool fp32 = false;     // complex FP32 sweeps (mixed precision)
	bool project = false;  // Riemannian projection of gradient and HVP
	bool tiles = false;    // tile-pass engine (else one gate per pass)
	int L = 0, G = 0, per = 1, ng = 0;
	long N = 0, B = 0;     // 2^L and the column-block width
	bool sampled = false;  // columns are given states (stochastic trace), not the basis
	long ncols = 0;        // columns swept (N, or the number of samples)
	long nblk = 0;         // column blocks (the last one zero-padded when sampled)
	int nst = 0;           // state arrays of N x B amplitudes in the workspace
	int passes = 0;        // HBM passes per sweep direction
	unsigned gtr = 0;      // trace-reduction grid
	unsigned grev_pg = 0;  // per-gate reverse grid
	size_t n_acc = 0, n_part = 0;
	// workspace layout: byte offsets from a 256-byte aligned base
	size_t off_acc = 0, off_part = 0, off_tab = 0, off_res = 0, off_states = 0, bytes = 0;
	size_t mats_bytes = 0, terms_bytes = 0, tab_bytes = 0;
	size_t rmats_off = 0;  // table offset of G^-T and W^T per gate (FP64; tile-pass gradient / HVP plans)
	int tile_K = 0, tile_logC = 0, gmax = 0;   // tile shape and gates per pass (tile-pass plans)
	std::shared_ptr<const TilePlan> tp;

	size_t res_bytes() const { return sizeof(zc) * (1 + 32 * (size_t)ng) + sizeof(double) * 32 * (size_t)ng; }
	double state_bytes() const { return (double)nst * N * B * (fp32 ? 8.0 : 16.0); }
};

// In mixed precision (fp32) a target evaluation is planned as a gradient evaluation: the FP32
// sweeps apply the gates rounded to complex float, and f is corrected to first order for that
// rounding with the gate gradient (k_finalize), so that every mixed-precision f, with or without
// a gradient, is computed the same way (a trust-region ratio compares f values of both kinds).
// `unitary`: the gates pass gates_near_unitary (the tile-pass engine is used for
// reverse sweeps only then; see sweep_unitary); `budget`: bytes for the state arrays
// (0 = automatic, 85% of the free device memory plus `held`); `ncols` > 0: the
// sweep's columns are that many given states instead of the 2^L basis columns
// (the column block is then at most the power of two covering them)
EvalPlan plan_eval(const Context& dev, const Circuit& C, Mode mode, int per, bool fp32, bool project, bool unitary,
	size_t budget, size_t held, const TileChoice* choice = nullptr, long ncols = 0);
// host image of the evaluation's table region (pl.tab_bytes bytes at `host`)
void fill_eval_tables(const EvalPlan& pl, const Circuit& C, char* host);

// Device-side refill of the gate-dependent part of a tile-pass plan's table region, for
// circuits whose gates are their groups' gates (brickwall layers: every gate of group l
// is gates[l]); the position terms stay those of a host fill made once.  All pointers
// are device addresses.
struct DeviceTableDesc
{
	char* tab = nullptr;            // the plan's table region (pl.tab_bytes)
	const int* slot_group = nullptr;    // group of each execution slot (table_slot_groups)
	const zc* gates = nullptr;      // ng * 16: the groups' gates
	const zc* dirs = nullptr;       // ng * 16: their directions (HVP plans)
	const int* skip = nullptr;      // nothing is written when *skip != 0 (may be NULL)
	int* bad = nullptr;             // set to 1 when a gate is singular
	size_t mats_bytes = 0, terms_bytes = 0, rmats_off = 0;
	int G = 0, ng = 0, mode = 0, fp32 = 0;
};
// group of every execution slot of a tile-pass plan
std::vector<int> table_slot_groups(const EvalPlan& pl);
// one workgroup per descriptor (descriptors in device memory)
void enqueue_fill_tables(hipStream_t s, const DeviceTableDesc* d_descs, int n);
// Block-encoding objective f = ||P^dagger W P - H||_F^2 / 2 in place of the
// trace (sampled plans over the 2^(L/2) columns P e_b, with P putting bit i of
// an index of L/2 qubits on bit 2i + 1 and the even bits |0>): the residual
// kernel reads H, adds |d|^2 / 2 to f and seeds the adjoint with conj(d) on the
// embedded rows.  In HVP sweeps it also stores the projected tangent
// P^dagger dpsi of the sweep's direction at J (2^L entries, J[b 2^(L/2) + a]).
struct BlockEnc
{
	const zc* H = nullptr;  // device, 2^(L/2) x 2^(L/2), row-major
	zc* J = nullptr;        // device, or NULL
};

// enqueue the evaluation on slot s into workspace `work` (pl.bytes).  The table
// region is at `tab` and the result region at `res` (NULL: their places in the
// workspace); the table image is uploaded first when host_tab is given.  With
// `skip`, every kernel returns at once when *skip != 0 at its start (a device
// flag, read when the kernel runs; tile-pass plans only).  Sampled plans take
// the input columns X (nblk blocks of N x B states in the sweep's precision,
// the last block zero-padded) and the trace seeds W = U X (same layout, FP64),
// or, with `be`, the block-encoding residual in place of the trace (W unused).
void enqueue_eval(const Context& dev, Slot& s, const EvalPlan& pl, const Circuit& C, const zc* dU, char* work,
	char* tab, char* res, const char* host_tab, const int* skip, const void* X = nullptr, const zc* W = nullptr,
	const BlockEnc* be = nullptr);

struct EvalResult
{
	zc f = {0, 0};
	std::vector<zc> dv, hz;          // ngroups * 16
	std::vector<double> grad, hvp;   // ngroups * 16
};
// decode the result region (pl.res_bytes() bytes copied from work + pl.off_res)
void unpack_result(const EvalPlan& pl, const char* res, EvalResult& R);

// the device-resident dense target of rqc_device_ufunc (checked; L < 0 skips the qubit check)
const zc* device_target_matrix(void* udata, int L);

// error reporting of the int-returning entry points
int report_error(const char* where, const std::exception& e);

}  // namespace rqc
