/*
 * xmp.h -- public C API of the MI355X batched big-integer library.
 *
 * Drop-in for the reference's src/include/xmp.h:40-226: same function names,
 * parameter order, enum ordinals and semantics.  The only type change is
 * cudaStream_t -> hipStream_t (xmp.h:125-138 in the reference);
 * xmpErrorCuda keeps its ordinal and means "HIP runtime error".
 *
 * Integers are arrays of `count` unsigned integers of `precision` bits
 * (a multiple of 32), stored on one GPU as 32-bit little-endian limbs.
 *
 * Beyond the reference the library exports the batch, serving and tuning
 * entry points declared in xmp_ext.h.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#define XMPAPI __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* reference xmp.h:40-54 (ordinals preserved) */
typedef enum {
  xmpErrorSuccess,
  xmpErrorInvalidDispatchTable,
  xmpErrorInvalidParameter,
  xmpErrorInvalidMalloc,
  xmpErrorInvalidCudaMalloc,
  xmpErrorInvalidCount,
  xmpErrorInvalidDevice,
  xmpErrorInvalidFormat,
  xmpErrorInvalidPrecision,
  xmpErrorIncreaseScratchLimit,
  xmpErrorUnsupported,
  xmpErrorUnspecified,
  xmpErrorCuda,
} xmpError_t;

/* reference xmp.h:57-60 */
typedef enum {
  xmpAlgorithm,
  xmpScratchSizeLimit
} xmpExecutionPolicyParam_t;

/* reference xmp.h:62-67.  On this library: Default lets the tuner choose;
 * RegMP forces one lane per instance; DistributedMP forces a multi-lane
 * layout; DigitMP is accepted and treated as Default (there is no
 * digit-serial path; see xmp_ext.h). */
typedef enum {
  xmpAlgorithmDefault,
  xmpAlgorithmRegMP,
  xmpAlgorithmDigitMP,
  xmpAlgorithmDistributedMP,
} xmpAlgorithm_t;

/* reference xmp.h:69-74.  In C++ the union keeps its converting
 * constructors; in C it is a plain union.  Either way it is 8 bytes and is
 * passed like a size_t. */
#ifdef __cplusplus
union xmpExecutionPolicyValue_t {
  xmpAlgorithm_t algorithm;
  size_t size;
  xmpExecutionPolicyValue_t() { size = 0; }
  xmpExecutionPolicyValue_t(xmpAlgorithm_t a) { size = 0; algorithm = a; }
  xmpExecutionPolicyValue_t(size_t s) { size = s; }
};
#else
typedef union xmpExecutionPolicyValue_t {
  xmpAlgorithm_t algorithm;
  size_t size;
} xmpExecutionPolicyValue_t;
#endif

/* reference xmp.h:76-108 (inline, not an exported symbol) */
static inline const char *xmpGetErrorString(xmpError_t error) {
  switch (error) {
    case xmpErrorSuccess: return "xmpErrorSuccess";
    case xmpErrorInvalidDispatchTable: return "xmpErrorInvalidDispatchTable";
    case xmpErrorInvalidParameter: return "xmpErrorInvalidParameter";
    case xmpErrorInvalidMalloc: return "xmpErrorInvalidMalloc";
    case xmpErrorInvalidCudaMalloc: return "xmpErrorInvalidCudaMalloc";
    case xmpErrorInvalidCount: return "xmpErrorInvalidCount";
    case xmpErrorInvalidDevice: return "xmpErrorInvalidDevice";
    case xmpErrorInvalidFormat: return "xmpErrorInvalidFormat";
    case xmpErrorInvalidPrecision: return "xmpErrorPrecision";
    case xmpErrorIncreaseScratchLimit: return "xmpIncreaseScratchLimit";
    case xmpErrorUnsupported: return "xmpErrorUnsupported";
    case xmpErrorUnspecified: return "xmpErrorUnspecified";
    case xmpErrorCuda: return "xmpErrorCuda";
    default: return "xmpErrorUnknown";
  }
}

/* reference xmp.h:111-116 */
typedef struct _xmpHandle_t *xmpHandle_t;
typedef struct _xmpIntegers_t *xmpIntegers_t;
typedef struct _xmpExecutionPolicy_t *xmpExecutionPolicy_t;

typedef void *(*xmpAllocFunc)(size_t);
typedef void (*xmpFreeFunc)(void *);

/* ---- handle: reference xmp.h:119-141, src/xmp.cu:39-200 ---- */
xmpError_t XMPAPI xmpHandleCreate(xmpHandle_t *handle);
xmpError_t XMPAPI xmpHandleCreateWithMemoryFunctions(xmpHandle_t *handle, xmpAllocFunc ha, xmpFreeFunc hf,
                                                     xmpAllocFunc da, xmpFreeFunc df);
xmpError_t XMPAPI xmpHandleDestroy(xmpHandle_t handle);
xmpError_t XMPAPI xmpHandleSetStream(xmpHandle_t handle, hipStream_t stream);
xmpError_t XMPAPI xmpHandleGetStream(xmpHandle_t handle, hipStream_t *stream);
xmpError_t XMPAPI xmpHandleSetExecutionPolicy(xmpHandle_t handle, xmpExecutionPolicy_t policy);
xmpError_t XMPAPI xmpHandleGetMemoryFunctions(xmpHandle_t handle, xmpAllocFunc *ha, xmpFreeFunc *hf,
                                              xmpAllocFunc *da, xmpFreeFunc *df);
/* declared but never defined by the reference (xmp.h:135-138); defined here
 * as aliases of Set/GetStream */
xmpError_t XMPAPI xmpHandleSetCudaStream(xmpHandle_t handle, hipStream_t stream);
xmpError_t XMPAPI xmpHandleGetCudaStream(xmpHandle_t handle, hipStream_t *stream);
xmpError_t XMPAPI xmpHandleGetDevice(xmpHandle_t handle, int32_t *device);

/* ---- execution policy: reference xmp.h:144-152, src/xmp.cu:203-314 ---- */
xmpError_t XMPAPI xmpExecutionPolicyCreate(xmpHandle_t handle, xmpExecutionPolicy_t *policy);
xmpError_t XMPAPI xmpExecutionPolicyDestroy(xmpHandle_t handle, xmpExecutionPolicy_t policy);
xmpError_t XMPAPI xmpExecutionPolicySetIndices(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                               uint32_t which_integer, uint32_t *indices, uint32_t count);
xmpError_t XMPAPI xmpExecutionPolicySetIndicesAsync(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                                    uint32_t which_integer, uint32_t *indices, uint32_t count);
xmpError_t XMPAPI xmpExecutionPolicySetParameter(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                                 xmpExecutionPolicyParam_t param, xmpExecutionPolicyValue_t val);
#ifdef __cplusplus
/* reference xmp.h:152 takes a C++ reference; the ABI is a pointer */
xmpError_t XMPAPI xmpExecutionPolicyGetParameter(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                                 xmpExecutionPolicyParam_t param, xmpExecutionPolicyValue_t &val);
#else
xmpError_t XMPAPI xmpExecutionPolicyGetParameter(xmpHandle_t handle, xmpExecutionPolicy_t policy,
                                                 xmpExecutionPolicyParam_t param, xmpExecutionPolicyValue_t *val);
#endif

/* ---- integers and transfer: reference xmp.h:156-173, src/xmp.cu:317-813 ---- */
xmpError_t XMPAPI xmpIntegersCreate(xmpHandle_t handle, xmpIntegers_t *x, uint32_t precision, uint32_t count);
xmpError_t XMPAPI xmpIntegersDestroy(xmpHandle_t handle, xmpIntegers_t x);
xmpError_t XMPAPI xmpIntegersGetPrecision(xmpHandle_t handle, xmpIntegers_t x, uint32_t *precision);
xmpError_t XMPAPI xmpIntegersGetCount(xmpHandle_t handle, xmpIntegers_t x, uint32_t *count);
xmpError_t XMPAPI xmpIntegersImport(xmpHandle_t handle, xmpIntegers_t out, uint32_t words, int32_t order, size_t size,
                                    int32_t endian, uint32_t nails, void *in, uint32_t count);
xmpError_t XMPAPI xmpIntegersImportAsync(xmpHandle_t handle, xmpIntegers_t out, uint32_t words, int32_t order,
                                         size_t size, int32_t endian, uint32_t nails, void *in, uint32_t count);
xmpError_t XMPAPI xmpIntegersExport(xmpHandle_t handle, void *out, uint32_t *words, int32_t order, size_t size,
                                    int32_t endian, uint32_t nails, xmpIntegers_t in, uint32_t count);
xmpError_t XMPAPI xmpIntegersExportAsync(xmpHandle_t handle, void *out, uint32_t *words, int32_t order, size_t size,
                                         int32_t endian, uint32_t nails, xmpIntegers_t in, uint32_t count);
xmpError_t XMPAPI xmpIntegersSet(xmpHandle_t handle, xmpIntegers_t out, xmpIntegers_t in, uint32_t count);
xmpError_t XMPAPI xmpIntegersSetAsync(xmpHandle_t handle, xmpIntegers_t out, xmpIntegers_t in, uint32_t count);

/* ---- math: reference xmp.h:180-199 ----
 * add_sub.cu:26-256, mul.cu:131-303, div.cu:26-517, powm.cu:240-307 */
xmpError_t XMPAPI xmpIntegersAdd(xmpHandle_t handle, xmpIntegers_t s, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersAddAsync(xmpHandle_t handle, xmpIntegers_t s, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersSub(xmpHandle_t handle, xmpIntegers_t d, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersSubAsync(xmpHandle_t handle, xmpIntegers_t d, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersMul(xmpHandle_t handle, xmpIntegers_t p, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersMulAsync(xmpHandle_t handle, xmpIntegers_t p, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
/* exported but undeclared by the reference (src/mul.cu:26) */
xmpError_t XMPAPI xmpIntegersSqrAsync(xmpHandle_t handle, xmpIntegers_t p, const xmpIntegers_t a, uint32_t count);
xmpError_t XMPAPI xmpIntegersDiv(xmpHandle_t handle, xmpIntegers_t q, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersDivAsync(xmpHandle_t handle, xmpIntegers_t q, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersMod(xmpHandle_t handle, xmpIntegers_t m, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersModAsync(xmpHandle_t handle, xmpIntegers_t m, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersDivMod(xmpHandle_t handle, xmpIntegers_t q, xmpIntegers_t m, const xmpIntegers_t a,
                                    const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersDivModAsync(xmpHandle_t handle, xmpIntegers_t q, xmpIntegers_t m,
                                         const xmpIntegers_t a, const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersPowm(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t base,
                                  const xmpIntegers_t exp, const xmpIntegers_t mod, uint32_t count);
xmpError_t XMPAPI xmpIntegersPowmAsync(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t base,
                                       const xmpIntegers_t exp, const xmpIntegers_t mod, uint32_t count);

/* ---- bitwise: reference xmp.h:206-226, src/logic.cu.  Index slots follow
 * the argument order (0 = output, then the inputs); Cmp and Popc write one
 * int32/uint32 per instance to a host or device array. ---- */
xmpError_t XMPAPI xmpIntegersShf(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const int32_t *shift,
                                 const uint32_t shift_count, uint32_t count);
xmpError_t XMPAPI xmpIntegersShfAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const int32_t *shift, const uint32_t shift_count, uint32_t count);
xmpError_t XMPAPI xmpIntegersIor(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersIorAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersAnd(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersAndAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersXor(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersXorAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count);
xmpError_t XMPAPI xmpIntegersNot(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, uint32_t count);
xmpError_t XMPAPI xmpIntegersNotAsync(xmpHandle_t handle, xmpIntegers_t c, const xmpIntegers_t a, uint32_t count);
xmpError_t XMPAPI xmpIntegersPopc(xmpHandle_t handle, uint32_t *c, const xmpIntegers_t a, uint32_t count);
xmpError_t XMPAPI xmpIntegersPopcAsync(xmpHandle_t handle, uint32_t *c, const xmpIntegers_t a, uint32_t count);
xmpError_t XMPAPI xmpIntegersCmp(xmpHandle_t handle, int32_t *c, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count);
xmpError_t XMPAPI xmpIntegersCmpAsync(xmpHandle_t handle, int32_t *c, const xmpIntegers_t a, const xmpIntegers_t b,
                                      uint32_t count);

#ifdef __cplusplus
}
#endif
