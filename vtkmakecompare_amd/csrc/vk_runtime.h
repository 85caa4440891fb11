// vk_runtime.h -- host-side runtime shared by the entry points: device
// selection, the library stream, the residency registry that maps a
// caller's host array to its device mirror, workspace and error state.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/vtkgpu.h"
#include "vk_common.h"

namespace vk {

struct Runtime {
  bool initialised = false;
  int device = 0;
  hipStream_t stream = nullptr;
  // three more streams (default priority) and the events that fork work
  // from `stream` onto them and join it back: the fused steps solve the
  // velocity components' ADI sweeps on these while the temperature's run
  // on `stream`
  hipStream_t aux[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[3] = {nullptr, nullptr, nullptr};
  // set while kernels are being timed: the work meant for the auxiliary
  // streams is issued on the forking stream instead, one kernel after
  // another, so that each kernel's events time it alone
  bool serial = false;
};
Runtime& rt();

// Records an error message and returns the code.
int fail(int code, const std::string& msg);
// the code of the last fail(): what resolve() refused an array with
int last_status();
// counts entry-point calls (every C-ABI entry point bumps it on entry):
// state derived from a call's arguments can be reused within one call
unsigned long long& call_epoch();
// Checks a HIP status; records and returns VK_ERR_HIP on failure.
int hip_check(hipError_t e, const char* what);

// Forks work from `s` onto the auxiliary streams in `mask` (bit c for
// aux[c]): they start after everything enqueued on `s` so far.  join(c)
// makes `s` continue only after aux[c]'s work; the destructor joins every
// stream not yet joined, so `s` waits for all of them on every path out of
// the caller.
struct AuxFork {
  hipStream_t s;
  bool pending[3] = {false, false, false};
  int err = 0;
  AuxFork(hipStream_t s_, unsigned mask) : s(s_) {
    if (rt().serial) return;
    err = hip_check(hipEventRecord(rt().ev_fork, s), "fork");
    for (int c = 0; c < 3 && !err; c++) {
      if (!(mask & (1u << c))) continue;
      err = hip_check(hipStreamWaitEvent(rt().aux[c], rt().ev_fork, 0), "fork wait");
      pending[c] = !err;
    }
  }
  hipStream_t aux(int c) const { return rt().serial ? s : rt().aux[c]; }
  int join(int c) {
    if (!pending[c]) return 0;
    pending[c] = false;
    int e = hip_check(hipEventRecord(rt().ev_join[c], rt().aux[c]), "join");
    if (!e) e = hip_check(hipStreamWaitEvent(s, rt().ev_join[c], 0), "join wait");
    return e;
  }
  ~AuxFork() {
    for (int c = 0; c < 3; c++) (void)join(c);
  }
};


// Resolves a caller pointer to device memory holding at least `bytes`:
// a registered host array (or an address inside one) maps to its mirror,
// device memory is used as is.  Returns nullptr and records the error
// otherwise.
void* resolve(const void* p, size_t bytes, const char* what);
// bytes from p to the end of its device allocation, when the runtime knows
bool device_extent(const void* p, size_t* bytes);
// whether p is the base of a registered host array (vk_register)
bool is_registered(const void* p);
// whether p points into device memory (false before vk_init)
bool is_device(const void* p);

// Registers a function that vk_finalize calls, after every stream has
// drained and before the streams are destroyed, to free device memory a
// translation unit keeps in caches of its own.  A hook is registered once
// per process and stays registered across finalize/init cycles.
void on_finalize(void (*hook)());

// Device memory for every buffer the library allocates or hands out
// (vk_malloc, registered arrays' mirrors, workspace, tables, checkpoints).
// With VK_GUARD set (and not "0") in the environment when vk_init runs,
// each allocation lies between two zones of 64 KiB of a fixed byte
// pattern; vk_guard_check reports any zone that a kernel has written.
// dev_free also takes memory that dev_malloc did not allocate.
hipError_t dev_malloc(void** p, size_t bytes, const char* what);
hipError_t dev_free(void* p);

// Grow-only device workspace, one per slot.
enum WorkSlot {
  kWorkPhi = 0,
  kWorkSpec,
  kWorkSpec2,
  kWorkGrad,
  kWorkMisc,
  kWorkRecip,
  kWorkQueue,
  kWorkPhi2,
  kWorkRecipF,
  kWorkRhs,
  kWorkFlags,
  kWorkDiag,
  kWorkPhi3,
  kWorkDivMax,
  kWorkLineBatch,
  kWorkLblQ,
  kWorkLblRes,
  kWorkSlots
};
// `grown` (optional) reports whether this call allocated the slot anew.
void* workspace(WorkSlot slot, size_t bytes, bool* grown = nullptr);

// Per-kernel timing with HIP events on the library stream, active between
// vk_profile_begin and vk_profile_end (a no-op otherwise).  A ProfScope
// brackets one kernel launch (or one launch group) with a named pair of
// events; vk_profile_entry reports the summed duration per name.
struct ProfScope {
  int slot = -1;
  hipStream_t s = nullptr;
  ProfScope(const char* name, hipStream_t stream);
  ~ProfScope();
};

}  // namespace vk

#define VK_REQUIRE_INIT()                                                   \
  do {                                                                      \
    if (!vk::rt().initialised)                                              \
      return vk::fail(VK_ERR_NOT_INITIALISED, "vk_init has not been called"); \
    ++vk::call_epoch();                                                     \
  } while (0)

#define VK_RESOLVE(var, type, ptr, bytes)                                   \
  type* var = static_cast<type*>(vk::resolve((const void*)(ptr), (bytes), #ptr)); \
  if (!var) return vk::last_status()

#define VK_LAUNCH_CHECK(what)                                               \
  do {                                                                      \
    int _rc = vk::hip_check(hipGetLastError(), what);                       \
    if (_rc) return _rc;                                                    \
  } while (0)
