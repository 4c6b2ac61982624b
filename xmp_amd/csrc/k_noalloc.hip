// k_noalloc.hip -- xmpIntegersPowmNoAllocAsync: modular exponentiation for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// No reference counterpart: the reference's xmpIntegersPowm grows the
// handle's scratch with cudaMalloc as needed (src/xmp.cu:105-129,
// powm_operators.h:308-328).  Here the workspace -- R^2 mod m, the division
// workspace of long moduli, the window table and, for an in-place call, the
// output -- lives in one file-scope device array of kScratchBytes.  Each
// device that runs this entry point holds its own copy of the array from the
// first call until the process ends.
//
// The array's name is only a host-side stand-in for the device copy: its
// device address is taken once per device with hipGetSymbolAddress (with
// that device current) and only that address is handed to kernels.
//
// Every call on one device shares that device's array: calls that overlap
// on one device (two streams, two host threads) race and must be ordered by
// the caller.  Nothing in the array has to start at zero: every kernel writes
// what it later reads (R^2 mod m, then the table slots, then the output).
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/xmp_ext.h"
#include "internal.hpp"
#include "powm.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;   // 256 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kScratchBytes / sizeof(uint32_t)];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// the device address of this device's copy of `scratch` (the device is current)
xmpError_t device_scratch(int device, void **ptr) {
  if (device < 0 || device >= kMaxDevices) return xmpErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      return xmpErrorCuda;
    }
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  return xmpErrorSuccess;
}

}  // namespace

extern "C" {

xmpError_t XMPAPI xmpIntegersPowmNoAllocAsync(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t base,
                                              const xmpIntegers_t exp, const xmpIntegers_t mod, uint32_t count) {
  XMP_TRACE("xmpIntegersPowmNoAllocAsync");
  XMP_CHECK_NE(handle, nullptr);
  XMP_CHECK_NE(out, nullptr);
  XMP_CHECK_NE(base, nullptr);
  XMP_CHECK_NE(exp, nullptr);
  XMP_CHECK_NE(mod, nullptr);
  void *ptr = nullptr;
  {
    XMP_SET_DEVICE(handle);
    xmpError_t e = device_scratch(handle->device, &ptr);
    if (e != xmpErrorSuccess) return e;
  }
  const xmp::FixedWorkspace ws{(uint8_t *)ptr, kScratchBytes};
  return xmp::powm_run(handle, out, base, exp, mod, count, &ws);
}

}  // extern "C"
