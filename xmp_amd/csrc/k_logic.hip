// k_logic.hip -- bitwise operations on the interleaved layout: ior, and, xor,
// not, shift, popcount and compare.
//
// Reference: src/logic.cu (entry points) and the strided_*_kernel family of
// src/include/implementations.h:766-1067, one thread per instance walking its
// limbs.  Semantics kept: inputs are zero-extended to the output length, not
// fills the limbs above its input with ones, a positive shift is a left shift
// truncated to the output, a negative one a right shift, compare returns
// -1/0/+1.  These are HBM streams: without index arrays each lane takes four
// consecutive instances, so every limb is one non-temporal dwordx4 per
// operand, as in the add/sub kernels.  Compare reads from the top limb down
// and stops as soon as every lane of the wave has decided.
#include "devmp.hpp"
#include "kernels.hpp"

namespace xmp {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int OP>
__device__ __forceinline__ uint32_t bitop(uint32_t a, uint32_t b) {
  if constexpr (OP == kBitIor) return a | b;
  else if constexpr (OP == kBitAnd) return a & b;
  else if constexpr (OP == kBitXor) return a ^ b;
  else return ~a;   // not: b is unused
}

template <int OP>
__device__ __forceinline__ u32x4 bitop4(u32x4 a, u32x4 b) {
  if constexpr (OP == kBitIor) return a | b;
  else if constexpr (OP == kBitAnd) return a & b;
  else if constexpr (OP == kBitXor) return a ^ b;
  else return ~a;
}

__device__ __forceinline__ u32x4 ld4(const uint32_t *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
}
__device__ __forceinline__ void st4(uint32_t *p, u32x4 v) { __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(p)); }

// Limb j of the result for inputs a (a_len limbs) and b (b_len limbs): above
// an input's length its limbs read as 0, so not gives all ones there.
template <int OP>
__global__ void __launch_bounds__(64) bitwise_vec4_kernel(uint32_t *__restrict__ c, uint32_t c_len, uint32_t c_stride,
                                                          const uint32_t *__restrict__ a, uint32_t a_len,
                                                          uint32_t a_stride, const uint32_t *__restrict__ b,
                                                          uint32_t b_len, uint32_t b_stride, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nvec = count / 4;
  const u32x4 zero = {0, 0, 0, 0};
  if (t < nvec) {
    const uint32_t i = 4 * t;
    uint32_t j = 0;
    for (; j + 8 <= c_len; j += 8) {
      u32x4 av[8], bv[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        av[k] = (j + k < a_len) ? ld4(a + (size_t)(j + k) * a_stride + i) : zero;
        bv[k] = (OP != kBitNot && j + k < b_len) ? ld4(b + (size_t)(j + k) * b_stride + i) : zero;
      }
#pragma unroll
      for (int k = 0; k < 8; k++) st4(c + (size_t)(j + k) * c_stride + i, bitop4<OP>(av[k], bv[k]));
    }
    for (; j < c_len; j++) {
      u32x4 av = (j < a_len) ? ld4(a + (size_t)j * a_stride + i) : zero;
      u32x4 bv = (OP != kBitNot && j < b_len) ? ld4(b + (size_t)j * b_stride + i) : zero;
      st4(c + (size_t)j * c_stride + i, bitop4<OP>(av, bv));
    }
  } else if (t < nvec + (count & 3)) {
    const uint32_t i = 4 * nvec + (t - nvec);
    for (uint32_t j = 0; j < c_len; j++) {
      uint32_t av = j < a_len ? a[(size_t)j * a_stride + i] : 0u;
      uint32_t bv = (OP != kBitNot && j < b_len) ? b[(size_t)j * b_stride + i] : 0u;
      c[(size_t)j * c_stride + i] = bitop<OP>(av, bv);
    }
  }
}

// General path: one instance per lane, index arrays and broadcast.
template <int OP>
__global__ void __launch_bounds__(256) bitwise_general_kernel(OutOperand c, Operand a, Operand b, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint32_t ai = in_index(a, t), ci = out_index(c, t);
  const uint32_t bi = OP != kBitNot ? in_index(b, t) : 0u;
  for (uint32_t j = 0; j < c.len; j++) {
    const uint32_t bv = OP != kBitNot ? ld_limb(b, bi, j) : 0u;
    c.data[(size_t)j * c.stride + ci] = bitop<OP>(ld_limb(a, ai, j), bv);
  }
}

// c = a * 2^s truncated to c.len limbs (s >= 0) or floor(a / 2^-s) (s < 0),
// s = shift[i % shift_count].  Output limb i holds bits [32 i - s, 32 i - s + 32)
// of a: with q = floor(-s / 32) and r = -s mod 32 that is the funnel of input
// limbs q + i + 1 and q + i by r, limbs outside [0, a.len) reading as 0.
__global__ void __launch_bounds__(256) shift_kernel(OutOperand c, Operand a, const int32_t *shift,
                                                    uint32_t shift_count, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint32_t ai = in_index(a, t), ci = out_index(c, t);
  const int64_t down = -(int64_t)shift[t % shift_count];    // right shift in bits (may be negative)
  const int64_t q = down >= 0 ? down / 32 : -((-down + 31) / 32);   // floor(down / 32)
  const uint32_t r = (uint32_t)(down - 32 * q);                   // in [0, 31]
  auto limb = [&](int64_t j) -> uint32_t {
    return (j >= 0 && j < (int64_t)a.len) ? a.data[(size_t)j * a.stride + ai] : 0u;
  };
  uint32_t lo = limb(q);
  for (uint32_t i = 0; i < c.len; i++) {
    const uint32_t hi = limb(q + i + 1);
    c.data[(size_t)i * c.stride + ci] = funnel(hi, lo, r);
    lo = hi;
  }
}

__global__ void __launch_bounds__(256) popc_kernel(uint32_t *out, Operand a, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint32_t ai = in_index(a, t);
  uint32_t n = 0;
  for (uint32_t j = 0; j < a.len; j++) n += __builtin_popcount(a.data[(size_t)j * a.stride + ai]);
  out[t] = n;
}

// Scans limbs from the top (the longer operand's length) down.  A lane is
// decided at its first differing limb; the wave leaves the loop once every
// lane is decided, so random inputs read about one limb per operand.
__global__ void __launch_bounds__(256) cmp_kernel(int32_t *out, Operand a, Operand b, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < count;
  const uint32_t ai = active ? in_index(a, t) : 0u, bi = active ? in_index(b, t) : 0u;
  int32_t res = 0;
  const uint32_t n = a.len > b.len ? a.len : b.len;
  for (uint32_t j = n; j-- > 0;) {
    if (active && res == 0) {
      const uint32_t x = ld_limb(a, ai, j), y = ld_limb(b, bi, j);
      res = x > y ? 1 : (x < y ? -1 : 0);
    }
    if (__builtin_amdgcn_ballot_w64(active && res == 0) == 0) break;
  }
  if (active) out[t] = res;
}

template <int OP>
hipError_t launch_bitwise_op(OutOperand c, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  const bool fast = !c.idx && !a.idx && a.count >= count && (OP == kBitNot || (!b.idx && b.count >= count));
  if (fast) {
    const uint32_t threads = count / 4 + (count & 3);
    hipLaunchKernelGGL((bitwise_vec4_kernel<OP>), dim3(div_round_up(threads, 64)), dim3(64), 0, stream, c.data, c.len,
                       c.stride, a.data, a.len, a.stride, b.data, b.len, b.stride, count);
  } else {
    hipLaunchKernelGGL((bitwise_general_kernel<OP>), dim3(div_round_up(count, 256)), dim3(256), 0, stream, c, a, b,
                       count);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_bitwise(int op, OutOperand c, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  switch (op) {
    case kBitIor: return launch_bitwise_op<kBitIor>(c, a, b, count, stream);
    case kBitAnd: return launch_bitwise_op<kBitAnd>(c, a, b, count, stream);
    case kBitXor: return launch_bitwise_op<kBitXor>(c, a, b, count, stream);
    case kBitNot: return launch_bitwise_op<kBitNot>(c, a, a, count, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_shift(OutOperand c, Operand a, const int32_t *shift, uint32_t shift_count, uint32_t count,
                        hipStream_t stream) {
  hipLaunchKernelGGL(shift_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, c, a, shift, shift_count,
                     count);
  return hipGetLastError();
}

hipError_t launch_popc(uint32_t *out, Operand a, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL(popc_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, out, a, count);
  return hipGetLastError();
}

hipError_t launch_cmp(int32_t *out, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL(cmp_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, out, a, b, count);
  return hipGetLastError();
}

}  // namespace xmp
