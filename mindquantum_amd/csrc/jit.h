// jit.h - per-circuit code generation for the tile passes.
//
// The interpreted pass kernel (k_tile_pass) reads every op descriptor and switches on its kind
// and register positions; across that switch the register allocator copies the whole register
// state on many paths.  For a plan that runs often the same passes are emitted as HIP source
// with every op kind, position, angle slot and tap index a compile-time constant, compiled with
// hipRTC on host threads (one program per pass, in parallel) and loaded as modules.  Launches
// use the interpreted kernel until the compiled passes are ready.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "program.h"
#include "schedule.h"

namespace mqa {

// Name of the generated kernel: mqa_jit_tile_pass_fwd / mqa_jit_tile_pass_adj.
const char* jit_kernel_name(bool adjoint);

// HIP source of pass `pass` of a plan (dtype 0: complex64, 1: complex128).  It includes
// "device_ops.h" and defines one extern "C" kernel (jit_kernel_name) taking TilePassArgs.
// Forms of a compiled pass: kTile, one workgroup per (tile, sample) with the tiles in LDS (every
// plan); kSingleBuffer (back-sweeps), psi and lambda stay in registers between register groups and
// share one LDS tile buffer (device_ops.h).
enum class JitForm { kTile = 0, kSingleBuffer = 2 };
std::string jit_pass_source(const Plan& pl, int pass, int dtype, JitForm form = JitForm::kTile);

// CoefForm of every angle slot a plan's ops read (index = slot; program.h): the layout of the
// coefficient table its compiled passes take.  Empty when no op has an angle.
std::vector<uint8_t> jit_coef_forms(const Plan& pl);

// Whether a plan's compiled passes may take the single-buffer form.
inline bool jit_can_single_buffer(const Plan& pl, int dtype) {
    (void)dtype;
    if (!pl.adjoint) return false;
    for (const DevOp& op : pl.dev_ops)  // dense 3- / 4-qubit blocks re-read their inputs from the tile
        if (op.kind == OPK_M3 || op.kind == OPK_M4) return false;
    return true;
}

// Wall time spent compiling generated passes in this process.
double jit_compile_seconds_total();

// Compile one generated source for gfx950.  Returns false (log filled) on failure.
bool jit_compile(const std::string& src, std::string& code, std::string& log);

class JitPlan {
  public:
    // Copies the plan and starts compiling every pass on a background thread.
    JitPlan(const Plan& pl, int dtype, JitForm form = JitForm::kTile);
    ~JitPlan();
    JitPlan(const JitPlan&) = delete;
    JitPlan& operator=(const JitPlan&) = delete;

    bool ready() const { return state_.load() == kReady; }
    bool failed() const { return state_.load() == kFailed; }
    void wait();                       // block until compiled (or failed)
    std::string error() const;
    double compile_seconds() const { return seconds_; }
    JitForm form() const { return form_; }

    // Launch pass `pass` on the current device and stream (modules load on first use; a load
    // failure unloads what loaded and marks the plan failed).
    hipError_t launch(int pass, const TilePassArgs& a, uint64_t batch, size_t lds, hipStream_t s);

  private:
    enum { kCompiling = 0, kReady = 1, kFailed = 2 };
    void compile_all();

    Plan plan_;
    int dtype_;
    JitForm form_;
    std::atomic<int> state_{kCompiling};
    std::vector<std::string> code_;
    std::string error_;
    double seconds_ = 0;
    std::thread worker_;
    mutable std::mutex mu_;
    std::map<int, std::vector<hipModule_t>> modules_;    // per device
    std::map<int, std::vector<hipFunction_t>> functions_;
};

}  // namespace mqa
