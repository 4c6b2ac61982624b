// api.cpp -- handle, execution policy, integers and transfer entry points.
//
// Behaviour (validation order, error codes, ownership) follows the
// reference's src/xmp.cu:28-813; the storage model is this library's own
// single interleaved layout (internal.hpp).
#include <algorithm>
#include <cstdlib>
#include <vector>
#include <cstring>
#include <string>

#include "internal.hpp"
#include "kernels.hpp"

namespace xmp {

const _xmpExecutionPolicy_t kDefaultPolicy{};

static void *device_malloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
static void device_free(void *p) {
  if (p) (void)hipFree(p);
}

xmpError_t set_necessary_scratch(xmpHandle_t h, size_t bytes) {
  size_t max_size = h->policy ? h->policy->scratch_size_limit : 0;
  if (max_size == 0) max_size = h->memorySize / 4 * 3;   // reference xmp.cu:110-115
  if (bytes > max_size) return xmpErrorIncreaseScratchLimit;
  if (h->scratchSize < bytes) {
    if (h->scratch) h->df(h->scratch);
    h->scratch = h->da(bytes);
    h->scratchSize = h->scratch ? bytes : 0;
  }
  if (h->scratch == nullptr && bytes > 0) return xmpErrorInvalidCudaMalloc;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

// The reference checks handle->scratch here instead of tmpOut (xmp.cu:144,
// a bug); this checks the buffer it allocated.
xmpError_t set_necessary_tmp_out(xmpHandle_t h, size_t bytes) {
  if (h->tmpOutSize < bytes) {
    if (h->tmpOut) h->df(h->tmpOut);
    h->tmpOut = h->da(bytes);
    h->tmpOutSize = h->tmpOut ? bytes : 0;
  }
  if (h->tmpOut == nullptr && bytes > 0) return xmpErrorInvalidCudaMalloc;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

// true when `p` is device-accessible memory (device or managed allocation)
bool is_device_pointer(const void *p) {
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();   // unregistered host memory: clear and treat as host
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

}  // namespace xmp

using namespace xmp;

extern "C" {

xmpError_t XMPAPI xmpHandleCreate(xmpHandle_t *handle) {
  return xmpHandleCreateWithMemoryFunctions(handle, nullptr, nullptr, nullptr, nullptr);
}

xmpError_t XMPAPI xmpHandleCreateWithMemoryFunctions(xmpHandle_t *handle, xmpAllocFunc ha, xmpFreeFunc hf,
                                                     xmpAllocFunc da, xmpFreeFunc df) {
  XMP_CHECK_NE(handle, nullptr);
  if (!ha) ha = malloc;
  if (!hf) hf = free;
  if (!da) da = device_malloc;
  if (!df) df = device_free;
  void *mem = ha(sizeof(_xmpHandle_t));
  *handle = nullptr;
  if (!mem) return xmpErrorInvalidMalloc;
  xmpHandle_t h = new (mem) _xmpHandle_t();
  h->ha = ha;
  h->hf = hf;
  h->da = da;
  h->df = df;
  h->policy = const_cast<xmpExecutionPolicy_t>(&kDefaultPolicy);
  *handle = h;
  if (hipGetDevice(&h->device) != hipSuccess) return xmpErrorInvalidDevice;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) return xmpErrorInvalidDevice;
  // the code objects are built for gfx950 (MI355X) only
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return xmpErrorInvalidDevice;
  h->memorySize = prop.totalGlobalMem;
  h->arch = 950;
  h->cuCount = (uint32_t)prop.multiProcessorCount;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpHandleDestroy(xmpHandle_t handle) {
  XMP_CHECK_NE(handle, nullptr);
  {
    XMP_SET_DEVICE(handle);
    if (handle->scratch) handle->df(handle->scratch);
    if (handle->tmpOut) handle->df(handle->tmpOut);
  }
  xmpFreeFunc hf = handle->hf;
  handle->~_xmpHandle_t();
  hf(handle);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpHandleSetStream(xmpHandle_t handle, hipStream_t stream) {
  XMP_CHECK_NE(handle, nullptr);
  handle->stream = stream;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpHandleGetStream(xmpHandle_t handle, hipStream_t *stream) {
  XMP_CHECK_NE(stream, nullptr);
  *stream = handle->stream;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpHandleSetCudaStream(xmpHandle_t handle, hipStream_t stream) {
  return xmpHandleSetStream(handle, stream);
}
xmpError_t XMPAPI xmpHandleGetCudaStream(xmpHandle_t handle, hipStream_t *stream) {
  return xmpHandleGetStream(handle, stream);
}

xmpError_t XMPAPI xmpHandleGetDevice(xmpHandle_t handle, int32_t *device) {
  XMP_CHECK_NE(device, nullptr);
  *device = handle->device;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpHandleSetExecutionPolicy(xmpHandle_t handle, xmpExecutionPolicy_t policy) {
  handle->policy = policy ? policy : const_cast<xmpExecutionPolicy_t>(&kDefaultPolicy);
  return xmpErrorSuccess;
}

// The reference writes *df under an hf guard (xmp.cu:155); each pointer is
// guarded by itself here.
xmpError_t XMPAPI xmpHandleGetMemoryFunctions(xmpHandle_t handle, xmpAllocFunc *ha, xmpFreeFunc *hf,
                                              xmpAllocFunc *da, xmpFreeFunc *df) {
  if (ha) *ha = handle->ha;
  if (hf) *hf = handle->hf;
  if (da) *da = handle->da;
  if (df) *df = handle->df;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpExecutionPolicyCreate(xmpHandle_t handle, xmpExecutionPolicy_t *policy) {
  XMP_CHECK_NE(policy, nullptr);
  XMP_SET_DEVICE(handle);
  void *mem = handle->ha(sizeof(_xmpExecutionPolicy_t));
  if (!mem) return xmpErrorInvalidMalloc;
  *policy = new (mem) _xmpExecutionPolicy_t();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpExecutionPolicyDestroy(xmpHandle_t handle, xmpExecutionPolicy_t policy) {
  XMP_CHECK_NE(policy, nullptr);
  XMP_SET_DEVICE(handle);
  for (uint32_t i = 0; i < kMaxIndexArrays; i++)
    if (policy->indices[i]) handle->df(policy->indices[i]);
  if (handle->policy == policy) handle->policy = const_cast<xmpExecutionPolicy_t>(&kDefaultPolicy);
  policy->~_xmpExecutionPolicy_t();
  handle->hf(policy);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

// reference xmp.cu:238-280: indices are copied into device memory owned by the policy
xmpError_t XMPAPI xmpExecutionPolicySetIndicesAsync(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                                    uint32_t which_integer, uint32_t *indices, uint32_t count) {
  XMP_TRACE("xmpExecutionPolicySetIndicesAsync");
  XMP_SET_DEVICE(handle);
  if (which_integer >= kMaxIndexArrays) return xmpErrorInvalidParameter;
  uint32_t **pind = &policy->indices[which_integer];
  uint32_t *pcount = &policy->indices_count[which_integer];
  if (indices == nullptr) {
    if (*pind) {
      handle->df(*pind);
      *pind = nullptr;
    }
    *pcount = 0;
    return xmpErrorSuccess;
  }
  if (count == 0) return xmpErrorInvalidCount;
  if (*pind && *pcount != count) {
    handle->df(*pind);
    *pind = nullptr;
  }
  if (!*pind) {
    *pind = (uint32_t *)handle->da(sizeof(uint32_t) * count);
    if (!*pind) return xmpErrorInvalidCudaMalloc;
  }
  *pcount = count;
  if (hipMemcpyAsync(*pind, indices, sizeof(uint32_t) * count, hipMemcpyDefault, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpExecutionPolicySetIndices(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                               uint32_t which_integer, uint32_t *indices, uint32_t count) {
  xmpError_t e = xmpExecutionPolicySetIndicesAsync(handle, policy, which_integer, indices, count);
  (void)hipStreamSynchronize(handle->stream);
  return e;
}

xmpError_t XMPAPI xmpExecutionPolicySetParameter(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                                 xmpExecutionPolicyParam_t param, xmpExecutionPolicyValue_t val) {
  (void)handle;
  switch (param) {
    case xmpAlgorithm:
      if ((uint32_t)val.algorithm > (uint32_t)xmpAlgorithmDistributedMP) return xmpErrorInvalidParameter;
      policy->algorithm = val.algorithm;
      return xmpErrorSuccess;
    case xmpScratchSizeLimit:
      policy->scratch_size_limit = val.size;
      return xmpErrorSuccess;
    default:
      return xmpErrorInvalidParameter;
  }
}

xmpError_t XMPAPI xmpExecutionPolicyGetParameter(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                                 xmpExecutionPolicyParam_t param, xmpExecutionPolicyValue_t &val) {
  (void)handle;
  switch (param) {
    case xmpAlgorithm:
      val.size = 0;
      val.algorithm = policy->algorithm;
      return xmpErrorSuccess;
    case xmpScratchSizeLimit:
      val.size = policy->scratch_size_limit;
      return xmpErrorSuccess;
    default:
      return xmpErrorInvalidParameter;
  }
}

// reference xmp.cu:317-354; one interleaved buffer, stride rounded to 64 instances
xmpError_t XMPAPI xmpIntegersCreate(xmpHandle_t handle, xmpIntegers_t *x, uint32_t precision, uint32_t count) {
  XMP_CHECK_NE(x, nullptr);
  XMP_SET_DEVICE(handle);
  *x = nullptr;
  if (count == 0 || count > 0xfffffffcu) return xmpErrorInvalidCount;
  if (precision == 0 || precision % 32 != 0) return xmpErrorUnsupported;
  void *mem = handle->ha(sizeof(_xmpIntegers_t));
  if (!mem) return xmpErrorInvalidMalloc;
  xmpIntegers_t v = new (mem) _xmpIntegers_t();
  v->count = count;
  v->precision = precision;
  v->nlimbs = precision / 32;
  v->device = handle->device;
  v->stride = round_up(count, kStrideAlign);
  v->limbs = (uint32_t *)handle->da(sizeof(uint32_t) * (size_t)v->nlimbs * v->stride);
  *x = v;
  if (!v->limbs) return xmpErrorInvalidCudaMalloc;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersDestroy(xmpHandle_t handle, xmpIntegers_t x) {
  XMP_CHECK_NE(x, nullptr);
  XMP_SET_DEVICE(handle);
  if (x->limbs) handle->df(x->limbs);
  x->~_xmpIntegers_t();
  handle->hf(x);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersGetPrecision(xmpHandle_t handle, xmpIntegers_t x, uint32_t *precision) {
  (void)handle;
  XMP_CHECK_NE(precision, nullptr);
  *precision = x->precision;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersGetCount(xmpHandle_t handle, xmpIntegers_t x, uint32_t *count) {
  (void)handle;
  XMP_CHECK_NE(count, nullptr);
  *count = x->count;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

static xmpError_t import_internal(xmpHandle_t handle, xmpIntegers_t out, uint32_t words, int32_t order, size_t size,
                                  int32_t endian, uint32_t nails, void *in, uint32_t count, bool async) {
  XMP_CHECK_NE(in, nullptr);
  xmpExecutionPolicy_t policy = handle->policy;
  if (out->device != handle->device) return xmpErrorInvalidDevice;
  XMP_SET_DEVICE(handle);
  if (endian == 0) endian = -1;   // native order on little-endian hosts
  if (size != 1 && size != 2 && size != 4 && size != 8) return xmpErrorInvalidParameter;
  if (count == 0 || count > out->count || words == 0 || (order != 1 && order != -1) || (endian != 1 && endian != -1))
    return xmpErrorInvalidParameter;
  if ((uint64_t)words * size * 8 != out->precision) return xmpErrorInvalidPrecision;
  if (nails >= size * 8) return xmpErrorInvalidParameter;
  const uint32_t *idx = policy->indices[0];
  uint32_t idx_count = policy->indices_count[0];
  // with gather indices the source holds max(index)+1 rows; without, count rows
  size_t rows = count;
  if (idx) {
    std::vector<uint32_t> h(idx_count);
    if (hipMemcpy(h.data(), idx, sizeof(uint32_t) * idx_count, hipMemcpyDeviceToHost) != hipSuccess) return xmpErrorCuda;
    rows = 0;
    for (uint32_t k = 0; k < count; k++) rows = std::max<size_t>(rows, (size_t)h[k % idx_count] + 1);
  }
  size_t bytes = rows * words * size;
  const uint8_t *src = (const uint8_t *)in;
  if (!is_device_pointer(in)) {
    xmpError_t e = set_necessary_scratch(handle, bytes);
    if (e != xmpErrorSuccess) return e;
    if (hipMemcpyAsync(handle->scratch, in, bytes, hipMemcpyHostToDevice, handle->stream) != hipSuccess)
      return xmpErrorCuda;
    src = (const uint8_t *)handle->scratch;
  }
  if (launch_import(out->limbs, out->stride, out->nlimbs, src, count, words, order, endian, nails, (uint32_t)size,
                    idx, idx_count, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  if (!async) (void)hipStreamSynchronize(handle->stream);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersImport(xmpHandle_t handle, xmpIntegers_t out, uint32_t words, int32_t order, size_t size,
                                    int32_t endian, uint32_t nails, void *in, uint32_t count) {
  XMP_TRACE("xmpIntegersImport");
  return import_internal(handle, out, words, order, size, endian, nails, in, count, false);
}
xmpError_t XMPAPI xmpIntegersImportAsync(xmpHandle_t handle, xmpIntegers_t out, uint32_t words, int32_t order,
                                         size_t size, int32_t endian, uint32_t nails, void *in, uint32_t count) {
  XMP_TRACE("xmpIntegersImportAsync");
  return import_internal(handle, out, words, order, size, endian, nails, in, count, true);
}

static xmpError_t export_internal(xmpHandle_t handle, void *out, uint32_t *words, int32_t order, size_t size,
                                  int32_t endian, uint32_t nails, xmpIntegers_t in, uint32_t count, bool async) {
  XMP_CHECK_NE(out, nullptr);
  xmpExecutionPolicy_t policy = handle->policy;
  if (in->device != handle->device) return xmpErrorInvalidDevice;
  XMP_SET_DEVICE(handle);
  if (endian == 0) endian = -1;
  if (size != 1 && size != 2 && size != 4 && size != 8) return xmpErrorInvalidParameter;
  if (count == 0 || count > in->count || (order != 1 && order != -1) || (endian != 1 && endian != -1))
    return xmpErrorInvalidParameter;
  if ((in->nlimbs * 4) % size != 0) return xmpErrorInvalidParameter;
  if (nails >= size * 8) return xmpErrorInvalidParameter;
  uint32_t w = in->nlimbs * 4 / (uint32_t)size;
  const uint32_t *idx = policy->indices[0];
  uint32_t idx_count = policy->indices_count[0];
  size_t rows = count;
  if (idx) {
    std::vector<uint32_t> h(idx_count);
    if (hipMemcpy(h.data(), idx, sizeof(uint32_t) * idx_count, hipMemcpyDeviceToHost) != hipSuccess) return xmpErrorCuda;
    rows = 0;
    for (uint32_t k = 0; k < count; k++) rows = std::max<size_t>(rows, (size_t)h[k % idx_count] + 1);
  }
  size_t bytes = rows * in->nlimbs * 4;
  bool host = !is_device_pointer(out);
  uint8_t *dst = (uint8_t *)out;
  if (host) {
    xmpError_t e = set_necessary_scratch(handle, bytes);
    if (e != xmpErrorSuccess) return e;
    dst = (uint8_t *)handle->scratch;
    // scattered export leaves unselected rows of the caller's buffer untouched
    if (idx && hipMemcpyAsync(dst, out, bytes, hipMemcpyHostToDevice, handle->stream) != hipSuccess)
      return xmpErrorCuda;
  }
  if (launch_export(dst, in->limbs, in->stride, in->nlimbs, count, w, order, endian, nails, (uint32_t)size, idx,
                    idx_count, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  if (host && hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  if (words) *words = w;
  if (!async) (void)hipStreamSynchronize(handle->stream);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersExport(xmpHandle_t handle, void *out, uint32_t *words, int32_t order, size_t size,
                                    int32_t endian, uint32_t nails, xmpIntegers_t in, uint32_t count) {
  XMP_TRACE("xmpIntegersExport");
  return export_internal(handle, out, words, order, size, endian, nails, in, count, false);
}
xmpError_t XMPAPI xmpIntegersExportAsync(xmpHandle_t handle, void *out, uint32_t *words, int32_t order, size_t size,
                                         int32_t endian, uint32_t nails, xmpIntegers_t in, uint32_t count) {
  XMP_TRACE("xmpIntegersExportAsync");
  return export_internal(handle, out, words, order, size, endian, nails, in, count, true);
}

// reference xmp.cu:734-813; the reorder runs on the handle's stream (the
// reference launches it on the default stream, xmp.cu:806).
xmpError_t XMPAPI xmpIntegersSetAsync(xmpHandle_t handle, xmpIntegers_t out, xmpIntegers_t in, uint32_t count) {
  XMP_TRACE("xmpIntegersSetAsync");
  if (in->device != handle->device && out->device != handle->device) return xmpErrorInvalidDevice;
  XMP_SET_DEVICE(handle);
  if (in->precision != out->precision) return xmpErrorInvalidPrecision;
  if (out->count < count || in->count < count) return xmpErrorInvalidCount;
  if (count == 0) return xmpErrorSuccess;
  xmpExecutionPolicy_t policy = handle->policy;
  const uint32_t *oidx = policy->indices[0], *iidx = policy->indices[1];
  if (!oidx && !iidx && in->count == out->count) {
    if (out != in &&
        hipMemcpyAsync(out->limbs, in->limbs, sizeof(uint32_t) * (size_t)in->nlimbs * in->stride,
                       hipMemcpyDeviceToDevice, handle->stream) != hipSuccess)
      return xmpErrorCuda;
  } else {
    const uint32_t *src = in->limbs;
    if (in == out) {   // in-place permutation goes through scratch
      size_t bytes = sizeof(uint32_t) * (size_t)in->nlimbs * in->stride;
      xmpError_t e = set_necessary_scratch(handle, bytes);
      if (e != xmpErrorSuccess) return e;
      if (hipMemcpyAsync(handle->scratch, in->limbs, bytes, hipMemcpyDeviceToDevice, handle->stream) != hipSuccess)
        return xmpErrorCuda;
      src = (const uint32_t *)handle->scratch;
    }
    if (launch_set(out->limbs, out->stride, out->nlimbs, src, in->stride, in->nlimbs, in->count, count, oidx,
                   policy->indices_count[0], iidx, policy->indices_count[1], handle->stream) != hipSuccess)
      return xmpErrorCuda;
  }
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersSet(xmpHandle_t handle, xmpIntegers_t out, xmpIntegers_t in, uint32_t count) {
  xmpError_t e = xmpIntegersSetAsync(handle, out, in, count);
  if (e != xmpErrorSuccess) return e;
  XMP_SET_DEVICE(handle);
  return hipStreamSynchronize(handle->stream) == hipSuccess ? xmpErrorSuccess : xmpErrorCuda;
}

}  // extern "C"
