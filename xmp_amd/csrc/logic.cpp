// logic.cpp -- bitwise entry points: ior, and, xor, not, shift, popcount, compare.
//
// Validation order and error codes follow the reference (src/logic.cu:26-605).
// Deliberate differences, all on paths where the reference misbehaves:
//  * index slots follow the argument order (ior/and/xor: 0 = c, 1 = a, 2 = b;
//    not/shf: 0 = c, 1 = a; popc: 0 = a; cmp: 0 = a, 1 = b).  The reference
//    assigns slot 1 (slot 0 for cmp) to whichever input has the larger
//    precision (logic.cu:56-64, 244-252), so the two agree whenever
//    a.precision >= b.precision;
//  * compare reads the selected instance of the longer operand
//    (implementations.h:795 reads instance 0's high limbs);
//  * in-place operations with an output index array leave the unselected
//    output instances unchanged (logic.cu copies the whole scratch back);
//  * shift_count == 0 is rejected with xmpErrorInvalidParameter (the
//    reference takes thread % 0).
#include <algorithm>

#include "internal.hpp"
#include "kernels.hpp"

using namespace xmp;

namespace {

inline size_t out_bytes(xmpIntegers_t x) { return sizeof(uint32_t) * (size_t)x->nlimbs * x->stride; }

xmpError_t sync(xmpHandle_t h, xmpError_t e) {
  if (e != xmpErrorSuccess) return e;
  XMP_SET_DEVICE(h);
  (void)hipStreamSynchronize(h->stream);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

// In-place output buffer: `buf` (the handle's scratch or tmpOut, already sized)
// starts as a copy of c when an output index array leaves some instances untouched.
xmpError_t stage_inplace(xmpHandle_t h, xmpIntegers_t c, uint32_t *buf) {
  if (h->policy->indices[0] &&
      hipMemcpyAsync(buf, c->limbs, out_bytes(c), hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
    return xmpErrorCuda;
  return xmpErrorSuccess;
}

xmpError_t copy_back(xmpHandle_t h, xmpIntegers_t c, const uint32_t *buf) {
  if (hipMemcpyAsync(c->limbs, buf, out_bytes(c), hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
    return xmpErrorCuda;
  return xmpErrorSuccess;
}

// ior / and / xor / not (reference logic.cu:211-539)
xmpError_t bitwise_async(int op, xmpHandle_t h, xmpIntegers_t c, xmpIntegers_t a, xmpIntegers_t b, uint32_t count) {
  xmpExecutionPolicy_t p = h->policy;
  const bool unary = op == kBitNot;
  if (c->device != h->device || a->device != h->device || (!unary && b->device != h->device))
    return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  XMP_SET_DEVICE(h);
  if (c->count < count) return xmpErrorInvalidCount;
  if (p->indices[0] && p->indices_count[0] < count) return xmpErrorInvalidCount;
  if (c->precision < (unary ? a->precision : std::max(a->precision, b->precision))) return xmpErrorInvalidPrecision;
  const bool inplace = c == a || (!unary && c == b);
  uint32_t *buf = c->limbs;
  if (inplace) {
    xmpError_t e = set_necessary_scratch(h, out_bytes(c));
    if (e != xmpErrorSuccess) return e;
    buf = (uint32_t *)h->scratch;
    if ((e = stage_inplace(h, c, buf)) != xmpErrorSuccess) return e;
  }
  const Operand A = make_in(a, p, 1);
  const Operand B = unary ? A : make_in(b, p, 2);
  if (launch_bitwise(op, make_out(c, buf, p, 0), A, B, count, h->stream) != hipSuccess) return xmpErrorCuda;
  if (inplace) {
    xmpError_t e = copy_back(h, c, buf);
    if (e != xmpErrorSuccess) return e;
  }
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

}  // namespace

extern "C" {

// c = sign(a - b) per instance: -1, 0 or +1 (reference logic.cu:26-116).
// c may be host or device memory; host results land when the stream completes.
xmpError_t XMPAPI xmpIntegersCmpAsync(xmpHandle_t handle, int32_t *c, const xmpIntegers_t a, const xmpIntegers_t b,
                                      uint32_t count) {
  XMP_TRACE("xmpIntegersCmpAsync");
  XMP_CHECK_NE(c, nullptr);
  if (a->device != handle->device || b->device != handle->device) return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  XMP_SET_DEVICE(handle);
  const bool host = !is_device_pointer(c);
  int32_t *dst = c;
  if (host) {
    xmpError_t e = set_necessary_scratch(handle, sizeof(int32_t) * (size_t)count);
    if (e != xmpErrorSuccess) return e;
    dst = (int32_t *)handle->scratch;
  }
  xmpExecutionPolicy_t p = handle->policy;
  if (launch_cmp(dst, make_in(a, p, 0), make_in(b, p, 1), count, handle->stream) != hipSuccess) return xmpErrorCuda;
  if (host && hipMemcpyAsync(c, dst, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToHost, handle->stream) !=
                  hipSuccess)
    return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}
xmpError_t XMPAPI xmpIntegersCmp(xmpHandle_t handle, int32_t *c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersCmpAsync(handle, c, a, b, count));
}

// c = a << shift[i % shift_count] (truncated) or a >> -shift[...] (reference logic.cu:118-208).
// shift may be host or device memory.
xmpError_t XMPAPI xmpIntegersShfAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const int32_t *shift, const uint32_t shift_count, uint32_t count) {
  XMP_TRACE("xmpIntegersShfAsync");
  XMP_CHECK_NE(shift, nullptr);
  xmpExecutionPolicy_t p = handle->policy;
  if (c->device != handle->device || a->device != handle->device) return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  XMP_SET_DEVICE(handle);
  if (c->count < count) return xmpErrorInvalidCount;
  if (p->indices[0] && p->indices_count[0] < count) return xmpErrorInvalidCount;
  if (shift_count == 0) return xmpErrorInvalidParameter;
  const uint32_t scount = std::min(shift_count, count);
  const bool host = !is_device_pointer(shift);
  const bool inplace = c == a;
  const int32_t *src = shift;
  if (host) {
    xmpError_t e = set_necessary_scratch(handle, sizeof(int32_t) * (size_t)scount);
    if (e != xmpErrorSuccess) return e;
    src = (const int32_t *)handle->scratch;
    if (hipMemcpyAsync(handle->scratch, shift, sizeof(int32_t) * (size_t)scount, hipMemcpyHostToDevice,
                       handle->stream) != hipSuccess)
      return xmpErrorCuda;
  }
  uint32_t *buf = c->limbs;
  if (inplace) {
    xmpError_t e = set_necessary_tmp_out(handle, out_bytes(c));
    if (e != xmpErrorSuccess) return e;
    buf = (uint32_t *)handle->tmpOut;
    if ((e = stage_inplace(handle, c, buf)) != xmpErrorSuccess) return e;
  }
  if (launch_shift(make_out(c, buf, p, 0), make_in(a, p, 1), src, scount, count, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  if (inplace) {
    xmpError_t e = copy_back(handle, c, buf);
    if (e != xmpErrorSuccess) return e;
  }
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}
xmpError_t XMPAPI xmpIntegersShf(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const int32_t *shift,
                                 const uint32_t shift_count, uint32_t count) {
  return sync(handle, xmpIntegersShfAsync(handle, c, a, shift, shift_count, count));
}

xmpError_t XMPAPI xmpIntegersIorAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersIorAsync");
  return bitwise_async(kBitIor, handle, c, a, b, count);
}
xmpError_t XMPAPI xmpIntegersIor(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersIorAsync(handle, c, a, b, count));
}
xmpError_t XMPAPI xmpIntegersAndAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersAndAsync");
  return bitwise_async(kBitAnd, handle, c, a, b, count);
}
xmpError_t XMPAPI xmpIntegersAnd(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersAndAsync(handle, c, a, b, count));
}
xmpError_t XMPAPI xmpIntegersXorAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersXorAsync");
  return bitwise_async(kBitXor, handle, c, a, b, count);
}
xmpError_t XMPAPI xmpIntegersXor(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersXorAsync(handle, c, a, b, count));
}
xmpError_t XMPAPI xmpIntegersNotAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, uint32_t count) {
  XMP_TRACE("xmpIntegersNotAsync");
  return bitwise_async(kBitNot, handle, c, a, nullptr, count);
}
xmpError_t XMPAPI xmpIntegersNot(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, uint32_t count) {
  return sync(handle, xmpIntegersNotAsync(handle, c, a, count));
}

// c[i] = number of one bits of instance i (reference logic.cu:541-605).
// c may be host or device memory.
xmpError_t XMPAPI xmpIntegersPopcAsync(xmpHandle_t handle, uint32_t *c, const xmpIntegers_t a, uint32_t count) {
  XMP_TRACE("xmpIntegersPopcAsync");
  XMP_CHECK_NE(c, nullptr);
  if (a->device != handle->device) return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  if (a->count < count) return xmpErrorInvalidCount;
  XMP_SET_DEVICE(handle);
  const bool host = !is_device_pointer(c);
  uint32_t *dst = c;
  if (host) {
    xmpError_t e = set_necessary_scratch(handle, sizeof(uint32_t) * (size_t)count);
    if (e != xmpErrorSuccess) return e;
    dst = (uint32_t *)handle->scratch;
  }
  if (launch_popc(dst, make_in(a, handle->policy, 0), count, handle->stream) != hipSuccess) return xmpErrorCuda;
  if (host && hipMemcpyAsync(c, dst, sizeof(uint32_t) * (size_t)count, hipMemcpyDeviceToHost, handle->stream) !=
                  hipSuccess)
    return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}
xmpError_t XMPAPI xmpIntegersPopc(xmpHandle_t handle, uint32_t *c, const xmpIntegers_t a, uint32_t count) {
  return sync(handle, xmpIntegersPopcAsync(handle, c, a, count));
}

}  // extern "C"
