// internal.hpp -- host-side object model of the MI355X big-integer library.
//
// Replaces the reference's src/include/xmp_internal.h:27-253.  Differences by
// design:
//  * one device layout per integer array ("interleaved"): limb j of instance i
//    lives at limbs[j*stride + i] with stride = round_up(count, 64), so a
//    wave64 reading limb j of 64 consecutive instances touches one 256-byte
//    segment.  The reference keeps a compact and a strided copy plus a format
//    state machine (xmp_internal.h:194-239); here import/export convert
//    directly, so memory per array is halved and no lazy transpose exists.
//  * the handle carries the tuner table used to pick Montgomery layouts.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "../../include/xmp.h"

namespace xmp {

// A roctx range over one API call, so that `rocprofv3 --marker-trace` shows
// each call of the library by name beside the kernels it launched.  Without
// a profiler attached a range costs one call into the roctx stub.
struct TraceRange {
  explicit TraceRange(const char *name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};
#define XMP_TRACE(name) ::xmp::TraceRange xmp_trace_range_(name)

constexpr uint32_t kMaxIndexArrays = 4;   // reference xmp_internal.h:59
constexpr uint32_t kStrideAlign = 64;     // instances per wave64

inline uint32_t round_up(uint32_t n, uint32_t d) { return (n + d - 1) / d * d; }
inline uint32_t div_round_up(uint32_t n, uint32_t d) { return (n + d - 1) / d; }

struct TunerTable;  // tuner.cpp

}  // namespace xmp

struct _xmpExecutionPolicy_t {
  size_t scratch_size_limit = 0;
  uint32_t *indices[xmp::kMaxIndexArrays] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t indices_count[xmp::kMaxIndexArrays] = {0, 0, 0, 0};
  xmpAlgorithm_t algorithm = xmpAlgorithmDefault;
};

struct _xmpHandle_t {
  hipStream_t stream = nullptr;
  int32_t device = 0;
  xmpAllocFunc ha = nullptr, da = nullptr;
  xmpFreeFunc hf = nullptr, df = nullptr;
  size_t scratchSize = 0;
  void *scratch = nullptr;
  size_t tmpOutSize = 0;
  void *tmpOut = nullptr;
  size_t memorySize = 0;
  uint32_t arch = 0;      // gfx major*100+minor*10+stepping (950 on MI355X)
  uint32_t cuCount = 0;   // compute units
  xmpExecutionPolicy_t policy = nullptr;
};

struct _xmpIntegers_t {
  uint32_t *limbs = nullptr;   // interleaved, nlimbs * stride words
  uint32_t stride = 0;
  uint32_t nlimbs = 0;
  uint32_t precision = 0;
  uint32_t count = 0;
  int32_t device = 0;
};

namespace xmp {

extern const _xmpExecutionPolicy_t kDefaultPolicy;

// RAII device pin (reference xmp_internal.h:241-253).
class DeviceSetter {
 public:
  explicit DeviceSetter(int device) {
    (void)hipGetDevice(&old_);
    if (old_ != device) (void)hipSetDevice(device);
  }
  ~DeviceSetter() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != old_) (void)hipSetDevice(old_);
  }

 private:
  int old_ = 0;
};

inline bool hip_ok() { return hipPeekAtLastError() == hipSuccess; }

#define XMP_CHECK_NE(p, q) \
  if ((p) == (q)) return xmpErrorInvalidParameter;
#define XMP_SET_DEVICE(h)                       \
  xmp::DeviceSetter __device_setter((h)->device); \
  if (!xmp::hip_ok()) return xmpErrorCuda;
#define XMP_CHECK_HIP() \
  if (!xmp::hip_ok()) return xmpErrorCuda;

// true when `p` is device-accessible memory; other pointers are treated as
// host memory (reference cudaPointerGetAttributes test, src/xmp.cu:522-531)
bool is_device_pointer(const void *p);

// grow-only scratch owned by the handle (reference src/xmp.cu:105-129)
xmpError_t set_necessary_scratch(xmpHandle_t h, size_t bytes);
// grow-only temporary output for in-place ops (reference src/xmp.cu:131-148)
xmpError_t set_necessary_tmp_out(xmpHandle_t h, size_t bytes);

// Input/output descriptor of one operand for a kernel launch: base pointer,
// limb count, stride, array count and optional index array.
struct Operand {
  const uint32_t *data;
  uint32_t len;      // limbs
  uint32_t stride;
  uint32_t count;    // instances in the array (inputs read instance % count)
  const uint32_t *idx;
  uint32_t idx_count;
};

struct OutOperand {
  uint32_t *data;
  uint32_t len;
  uint32_t stride;
  const uint32_t *idx;   // out index = idx[i] (no modulo, reference implementations.h:61)
};

inline Operand make_in(xmpIntegers_t x, xmpExecutionPolicy_t p, int slot) {
  Operand o{x->limbs, x->nlimbs, x->stride, x->count, nullptr, 0};
  if (p && slot >= 0 && p->indices[slot]) {
    o.idx = p->indices[slot];
    o.idx_count = p->indices_count[slot];
  }
  return o;
}

inline OutOperand make_out(xmpIntegers_t x, uint32_t *buffer, xmpExecutionPolicy_t p, int slot) {
  OutOperand o{buffer ? buffer : x->limbs, x->nlimbs, x->stride, nullptr};
  if (p && slot >= 0 && p->indices[slot]) o.idx = p->indices[slot];
  return o;
}

}  // namespace xmp
