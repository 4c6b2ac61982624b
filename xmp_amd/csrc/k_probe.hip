// k_probe.hip -- issue-rate probe for v_mad_u64_u32 on this device.
//
// Streams of independent 32x32+64 multiply-adds (eight chains per lane) run
// for a fixed count on every CU; the result, in multiply-adds per clock per
// CU, is the measured ceiling the powm roofline is read against (64 would be
// one per lane per clock).  No reference counterpart.
#include <hip/hip_runtime.h>

#include "../../include/xmp_ext.h"
#include "internal.hpp"

namespace {

constexpr int kIters = 32768;

__global__ void __launch_bounds__(64) mad_probe_kernel(uint64_t *out, uint32_t seed) {
  uint64_t acc[8];
  const uint32_t a = threadIdx.x ^ seed, b = blockIdx.x + seed;
#pragma unroll
  for (int i = 0; i < 8; i++) acc[i] = (uint64_t)i * seed;
  // The multiply-adds are volatile asm: written in C, a * (b + i) is loop
  // invariant and the compiler reduces each chain to a plain 64-bit add.
  for (int it = 0; it < kIters; it++) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint64_t carry;
      asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc[i]), "=s"(carry) : "v"(a), "v"(b + i));
    }
  }
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r ^= acc[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

}  // namespace

extern "C" xmpError_t XMPAPI xmpProbeMadRate(xmpHandle_t handle, uint32_t waves_per_simd, double *mads_per_clk_per_cu,
                                             double *seconds) {
  XMP_CHECK_NE(mads_per_clk_per_cu, nullptr);
  if (waves_per_simd == 0 || waves_per_simd > 8) return xmpErrorInvalidParameter;
  XMP_SET_DEVICE(handle);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, handle->device) != hipSuccess) return xmpErrorCuda;
  const uint32_t blocks = handle->cuCount * 4 * waves_per_simd;
  uint64_t *out = (uint64_t *)handle->da(sizeof(uint64_t) * blocks * 64);
  if (!out) return xmpErrorInvalidCudaMalloc;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(mad_probe_kernel, dim3(blocks), dim3(64), 0, handle->stream, out, 1u);
  (void)hipEventRecord(e0, handle->stream);
  hipLaunchKernelGGL(mad_probe_kernel, dim3(blocks), dim3(64), 0, handle->stream, out, 3u);
  (void)hipEventRecord(e1, handle->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  handle->df(out);
  const double mads = (double)blocks * 64 * kIters * 8;
  const double clk_hz = (double)prop.clockRate * 1e3;
  *mads_per_clk_per_cu = mads / (ms * 1e-3) / handle->cuCount / clk_hz;
  if (seconds) *seconds = ms * 1e-3;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}
