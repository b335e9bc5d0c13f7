// Internal runtime of librqcopt_hip: device context, stream, workspace,
// kernel timing and error reporting.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "zc.hpp"
#include "rqcopt/rqc_device.h"

namespace rqc {

// error thrown inside the library and turned into a return code (or, for the
// void entry points of the reference ABI, a message and abort) at the boundary
struct Error : std::runtime_error
{
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define RQC_HIP(call)                                                                      \
	do {                                                                                   \
		hipError_t e_ = (call);                                                            \
		if (e_ != hipSuccess) {                                                            \
			throw ::rqc::Error(-1, std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
		}                                                                                  \
	} while (0)

// grow-only device allocation
struct DeviceBuffer
{
	void* ptr = nullptr;
	size_t bytes = 0;
	void reserve(size_t n);
	void release();
	template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

// An execution slot: one HIP stream and the bookkeeping of the evaluations
// enqueued on it.  Evaluations on different slots share nothing on the device
// and may run concurrently.
struct Slot
{
	hipStream_t stream = nullptr;
	long last_block = 0;   // column-block width of the last target evaluation
	int last_engine = -1;  // sweep engine of the last target evaluation
	int last_passes = 0;   // HBM passes per sweep direction of the last target evaluation
	int last_K = 0, last_logC = 0, last_gmax = 0;   // its tile shape and most gates per pass

	// kernel timing (HIP events on this slot's stream)
	bool profile = false;
	struct Ev { hipEvent_t a, b; int kind; double bytes, flops, issued; };
	std::vector<Ev> events;
	size_t events_used = 0;
};

// the device, its settings and the main slot, which serves the reference ABI
struct Context : Slot
{
	int device = -1;
	int cu_count = 0;
	size_t hbm_bytes = 0;
	std::string name;
	int precision = RQC_PRECISION_FP64;
	int engine = RQC_ENGINE_TILE_PASS;
	size_t lds_per_block = 65536;
	size_t ws_limit = 0;  // 0 = automatic
	bool autotune = true;  // tile autotuning of batched problems
	bool eval_autotune = false;  // tile autotuning of single evaluations (first call per structure)

	DeviceBuffer work;     // workspace of target evaluations (states, partials, tables, results)
	DeviceBuffer small;    // accumulators of the statevector-level API
	DeviceBuffer vecs;     // uploaded statevectors of the statevector-level API
	DeviceBuffer ubuf;     // U assembled from a host callback
	DeviceBuffer samples;  // sample states and trace seeds of stochastic-trace evaluations
	char* scratch = nullptr;   // device address of the static no-allocation scratch array (kScratchBytes)

	double prof_ms[RQC_KERNEL_KINDS] = {0, 0, 0, 0};
	long prof_n[RQC_KERNEL_KINDS] = {0, 0, 0, 0};
	double prof_bytes[RQC_KERNEL_KINDS] = {0, 0, 0, 0};   // algorithmic HBM bytes of the timed launches
	double prof_flops[RQC_KERNEL_KINDS] = {0, 0, 0, 0};   // their algorithmic flops (tile passes)
	double prof_issued[RQC_KERNEL_KINDS] = {0, 0, 0, 0};  // and the flops they issue (with the recovery)

	std::recursive_mutex mu;
};

// the process-wide context on the current HIP device; throws if no GPU is present
Context& ctx();

// size of the static device array that the no-allocation entry point uses as its workspace
constexpr size_t kScratchBytes = (size_t)256 << 20;

// bracket a kernel launch with timing events when the slot profiles; `bytes` is the
// launch's algorithmic HBM traffic (arrays it must read and write once), summed per kind
struct LaunchTimer
{
	Slot& s;
	Slot::Ev* ev = nullptr;
	LaunchTimer(Slot& sl, int kind, double bytes = 0, double flops = 0, double issued = 0);
	~LaunchTimer();
};

// flush the main slot's pending timing events into the per-kind totals
void profile_flush(Context& c);

// true if p is a HIP device allocation (or managed memory usable by kernels)
bool is_device_pointer(const void* p);

// report an error from a void entry point of the reference ABI and stop
[[noreturn]] void fatal(const char* where, const std::exception& e);

}  // namespace rqc
