// devmp.hpp -- device helpers shared by the arithmetic kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace xmp {

// Input instance selection (reference docs/xmp.docx para 15-17,
// src/include/implementations.h:57-60): idx[i % idx_count] % count, or
// i % count without an index array (broadcast from count == 1 works).
__device__ __forceinline__ uint32_t in_index(const Operand &o, uint32_t i) {
  uint32_t k = o.idx ? o.idx[i % o.idx_count] : i;
  return k % o.count;
}

// Output instance selection: idx[i] without modulo (implementations.h:61).
__device__ __forceinline__ uint32_t out_index(const OutOperand &o, uint32_t i) { return o.idx ? o.idx[i] : i; }

__device__ __forceinline__ uint32_t ld_limb(const Operand &o, uint32_t inst, uint32_t j) {
  return j < o.len ? o.data[(size_t)j * o.stride + inst] : 0u;
}

// 32-bit add with carry in/out (lowers to v_add_co_u32 / v_addc_co_u32)
__device__ __forceinline__ uint32_t addc(uint32_t a, uint32_t b, uint32_t cin, uint32_t &cout) {
  uint64_t s = (uint64_t)a + b + cin;
  cout = (uint32_t)(s >> 32);
  return (uint32_t)s;
}

// 32-bit subtract with borrow in/out (v_sub_co_u32 / v_subb_co_u32)
__device__ __forceinline__ uint32_t subb(uint32_t a, uint32_t b, uint32_t bin, uint32_t &bout) {
  uint64_t d = (uint64_t)a - b - bin;
  bout = (uint32_t)(d >> 63);
  return (uint32_t)d;
}

// funnel shift: low 32 bits of ((hi:lo) >> s), s in [0,31]  (v_alignbit_b32)
__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
  return __builtin_amdgcn_alignbit(hi, lo, s);
}

}  // namespace xmp
