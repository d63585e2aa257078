// Stepping without any allocation (no counterpart in the reference): lbm_enqueue_steps with the
// force reduction's scratch (FP64 per-wave partials and the completion counter, which every other
// entry point keeps in buffers of the handle) taken from one file-scope device array per GPU. For
// callers that may not allocate device memory at all while they step: code under HIP graph capture,
// tight serving loops, harnesses that forbid allocation.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/lbm.h"
#include "solver_internal.hpp"

using namespace lbm;

namespace {

constexpr size_t SCRATCH_PARTIALS = 32768;  // force partials of 3 doubles: 768 KiB
constexpr int MAX_DEVICES = 64;

// SCRATCH_PARTIALS x 3 FP64 partials, then the u32 completion counter (in the last 256 bytes).
// Every GPU the library runs on holds its own copy until the process ends.
static __device__ __align__(256) double scratch[3 * SCRATCH_PARTIALS + 32];

std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};  // device address of each GPU's copy of scratch, once looked up

__global__ void clear_counter_kernel(uint32_t* counter) { *counter = 0u; }

// The address of scratch on a device, looked up once per device with that device current. In host
// code the name `scratch` is a stand-in, not a device address: only this pointer is passed on.
int scratch_on(int device, double** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(LBM_ERR_INVALID, "lbm_enqueue_steps_noalloc: bad device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(LBM_ERR_DEVICE, std::string("lbm_enqueue_steps_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<double*>(p);
  }
  *out = g_ptr[device];
  return LBM_OK;
}

}  // namespace

extern "C" int lbm_enqueue_steps_noalloc(lbm_solver* s, int64_t n) {
  if (!s || n < 0) return fail(LBM_ERR_INVALID, "bad arguments");
  if (int rc = check_usable(s)) return rc;
  const size_t need = force_partials_needed(s);
  if (need > SCRATCH_PARTIALS)
    return fail(LBM_ERR_UNSUPPORTED, "lbm_enqueue_steps_noalloc: this tunnel's force reduction needs " +
                                         std::to_string(need) + " partials, the static scratch holds " +
                                         std::to_string(SCRATCH_PARTIALS) + "; use lbm_enqueue_steps");
  double* ptr = nullptr;
  if (int rc = scratch_on(s->device, &ptr)) return rc;
  uint32_t* counter = reinterpret_cast<uint32_t*>(ptr + 3 * SCRATCH_PARTIALS);
  // The array keeps whatever the last call on this GPU left; the counter must start at zero.
  clear_counter_kernel<<<1, 1, 0, s->stream>>>(counter);
  for (int64_t k = 0; k < n; k++)
    if (int rc = enqueue_step(s, nullptr, -1, ptr, counter)) return rc;
  HIP_TRY(hipGetLastError());
  return LBM_OK;
}
