// k_arith.hip -- elementwise add, sub and mul on the interleaved layout.
//
// add/sub (reference src/add_sub.cu:35-256, implementations.h:36-248,
// digitmp/add.h, digitmp/sub.h): s = (a + b) mod 2^P(s), d = (a - b) mod
// 2^P(d); inputs are zero-extended to the output length, the carry of an add
// lands in the next limb, a negative difference wraps.  These are HBM-bound:
// the fast path gives each lane four consecutive instances so that every limb
// is one 16-byte load per operand (dwordx4) and the carry chains run in
// registers (v_add_co_u32/v_addc_co_u32).
//
// mul (reference src/mul.cu:140-303, digitmp/mul.h:24-55): p = a*b mod
// 2^P(p).  Operands are split into 28-bit digits held in registers and the
// product is formed column by column (product scanning) in 64-bit
// accumulators: each partial product is one v_mad_u64_u32 and carries are
// resolved once per column, because up to 256 products of 56 bits fit in 64
// bits.
#include "devmp.hpp"
#include "kernels.hpp"

namespace xmp {

namespace {

template <bool SUB>
__device__ __forceinline__ void addsub_limb(uint32_t &r, uint32_t a, uint32_t b, uint32_t &c) {
  if (SUB)
    r = subb(a, b, c, c);
  else
    r = addc(a, b, c, c);
}

// Fast path: no index arrays and both inputs hold at least `count` instances.
template <bool SUB>
__global__ void __launch_bounds__(256) addsub_vec4_kernel(uint32_t *__restrict__ s, uint32_t s_len, uint32_t s_stride,
                                                          const uint32_t *__restrict__ a, uint32_t a_len,
                                                          uint32_t a_stride, const uint32_t *__restrict__ b,
                                                          uint32_t b_len, uint32_t b_stride, uint32_t count) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nvec = count / 4;
  if (t < nvec) {
    uint32_t i = 4 * t;
    uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    uint32_t j = 0;
    for (; j + 8 <= s_len; j += 8) {
      uint4 av[8], bv[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        av[k] = (j + k < a_len) ? *reinterpret_cast<const uint4 *>(a + (size_t)(j + k) * a_stride + i)
                                : make_uint4(0, 0, 0, 0);
        bv[k] = (j + k < b_len) ? *reinterpret_cast<const uint4 *>(b + (size_t)(j + k) * b_stride + i)
                                : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        uint4 r;
        addsub_limb<SUB>(r.x, av[k].x, bv[k].x, c0);
        addsub_limb<SUB>(r.y, av[k].y, bv[k].y, c1);
        addsub_limb<SUB>(r.z, av[k].z, bv[k].z, c2);
        addsub_limb<SUB>(r.w, av[k].w, bv[k].w, c3);
        *reinterpret_cast<uint4 *>(s + (size_t)(j + k) * s_stride + i) = r;
      }
    }
    for (; j < s_len; j++) {
      uint4 av = (j < a_len) ? *reinterpret_cast<const uint4 *>(a + (size_t)j * a_stride + i) : make_uint4(0, 0, 0, 0);
      uint4 bv = (j < b_len) ? *reinterpret_cast<const uint4 *>(b + (size_t)j * b_stride + i) : make_uint4(0, 0, 0, 0);
      uint4 r;
      addsub_limb<SUB>(r.x, av.x, bv.x, c0);
      addsub_limb<SUB>(r.y, av.y, bv.y, c1);
      addsub_limb<SUB>(r.z, av.z, bv.z, c2);
      addsub_limb<SUB>(r.w, av.w, bv.w, c3);
      *reinterpret_cast<uint4 *>(s + (size_t)j * s_stride + i) = r;
    }
  } else if (t < nvec + (count & 3)) {
    uint32_t i = 4 * nvec + (t - nvec);
    uint32_t c = 0;
    for (uint32_t j = 0; j < s_len; j++) {
      uint32_t av = j < a_len ? a[(size_t)j * a_stride + i] : 0u;
      uint32_t bv = j < b_len ? b[(size_t)j * b_stride + i] : 0u;
      uint32_t r;
      addsub_limb<SUB>(r, av, bv, c);
      s[(size_t)j * s_stride + i] = r;
    }
  }
}

// General path: one instance per lane with index arrays and broadcast.
template <bool SUB>
__global__ void __launch_bounds__(256) addsub_general_kernel(OutOperand s, Operand a, Operand b, uint32_t count) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t ai = in_index(a, t), bi = in_index(b, t), si = out_index(s, t);
  uint32_t c = 0;
  for (uint32_t j = 0; j < s.len; j++) {
    uint32_t r;
    addsub_limb<SUB>(r, ld_limb(a, ai, j), ld_limb(b, bi, j), c);
    s.data[(size_t)j * s.stride + si] = r;
  }
}

// ------------------------------------------------------------------ mul
constexpr int kDigit = 28;
constexpr uint32_t kDigitMask = (1u << kDigit) - 1;

// Load NL 32-bit limbs of one instance and split them into ND 28-bit digits.
template <int NL, int ND>
__device__ __forceinline__ void load_digits(uint32_t (&d)[ND], const Operand &o, uint32_t inst) {
  uint32_t l[NL + 1];
#pragma unroll
  for (int k = 0; k < NL; k++) l[k] = ld_limb(o, inst, k);
  l[NL] = 0;
#pragma unroll
  for (int k = 0; k < ND; k++) {
    constexpr int dummy = 0;
    (void)dummy;
    const int bit = kDigit * k;
    const int w = bit >> 5, sh = bit & 31;
    uint32_t lo = w < NL + 1 ? l[w] : 0u;
    uint32_t hi = w + 1 < NL + 1 ? l[w + 1] : 0u;
    d[k] = (sh ? ((lo >> sh) | (hi << (32 - sh))) : lo) & kDigitMask;
  }
}

// Product scanning over ND_A x ND_B digits; output converted back to 32-bit
// limbs and truncated to p.len limbs.
template <int NLA, int NLB>
__global__ void __launch_bounds__(256) mul_digits_kernel(OutOperand p, Operand a, Operand b, uint32_t count) {
  constexpr int NDA = (32 * NLA + kDigit - 1) / kDigit;
  constexpr int NDB = (32 * NLB + kDigit - 1) / kDigit;
  constexpr int NDP = NDA + NDB;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t ai = in_index(a, t), bi = in_index(b, t), pi = out_index(p, t);
  uint32_t da[NDA], db[NDB];
  load_digits<NLA, NDA>(da, a, ai);
  load_digits<NLB, NDB>(db, b, bi);
  uint64_t acc = 0;
  // bit buffer for digits -> limbs repacking
  uint64_t bits = 0;
  int nbits = 0;
  uint32_t limb = 0;
#pragma unroll
  for (int k = 0; k < NDP; k++) {
#pragma unroll
    for (int i = 0; i < NDA; i++) {
      const int j = k - i;
      if (j >= 0 && j < NDB) acc += (uint64_t)da[i] * db[j];
    }
    uint32_t digit = (uint32_t)acc & kDigitMask;
    acc >>= kDigit;
    bits |= (uint64_t)digit << nbits;
    nbits += kDigit;
    if (nbits >= 32) {
      if (limb < p.len) p.data[(size_t)limb * p.stride + pi] = (uint32_t)bits;
      limb++;
      bits >>= 32;
      nbits -= 32;
    }
  }
  if (nbits > 0) {
    if (limb < p.len) p.data[(size_t)limb * p.stride + pi] = (uint32_t)bits;
    limb++;
  }
  for (uint32_t k = limb; k < p.len; k++) p.data[(size_t)k * p.stride + pi] = 0u;
}

// Generic fallback for any sizes: schoolbook with 32-bit limbs, the row
// accumulator kept in the output array itself (each lane owns its column).
__global__ void __launch_bounds__(256) mul_general_kernel(OutOperand p, Operand a, Operand b, uint32_t count) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t ai = in_index(a, t), bi = in_index(b, t), pi = out_index(p, t);
  uint32_t *out = p.data + pi;
  for (uint32_t k = 0; k < p.len; k++) out[(size_t)k * p.stride] = 0u;
  uint32_t alen = min(a.len, p.len);
  for (uint32_t i = 0; i < alen; i++) {
    uint32_t ax = ld_limb(a, ai, i);
    uint32_t carry = 0;
    uint32_t jmax = min(b.len, p.len - i);
    for (uint32_t j = 0; j < jmax; j++) {
      uint64_t x = (uint64_t)ax * ld_limb(b, bi, j) + out[(size_t)(i + j) * p.stride] + carry;
      out[(size_t)(i + j) * p.stride] = (uint32_t)x;
      carry = (uint32_t)(x >> 32);
    }
    if (i + jmax < p.len) {
      uint32_t c2;
      uint32_t *o = &out[(size_t)(i + jmax) * p.stride];
      *o = addc(*o, carry, 0u, c2);
      for (uint32_t k = i + jmax + 1; c2 && k < p.len; k++) {
        uint32_t *q = &out[(size_t)k * p.stride];
        *q = addc(*q, 0u, c2, c2);
      }
    }
  }
}

template <int NLA, int NLB>
hipError_t launch_mul_digits(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL((mul_digits_kernel<NLA, NLB>), dim3(div_round_up(count, 256)), dim3(256), 0, stream, p, a, b,
                     count);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_add(OutOperand s, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  if (!s.idx && !a.idx && !b.idx && a.count >= count && b.count >= count) {
    uint32_t threads = count / 4 + (count & 3);
    hipLaunchKernelGGL(addsub_vec4_kernel<false>, dim3(div_round_up(threads, 256)), dim3(256), 0, stream, s.data,
                       s.len, s.stride, a.data, a.len, a.stride, b.data, b.len, b.stride, count);
  } else {
    hipLaunchKernelGGL(addsub_general_kernel<false>, dim3(div_round_up(count, 256)), dim3(256), 0, stream, s, a, b,
                       count);
  }
  return hipGetLastError();
}

hipError_t launch_sub(OutOperand d, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  if (!d.idx && !a.idx && !b.idx && a.count >= count && b.count >= count) {
    uint32_t threads = count / 4 + (count & 3);
    hipLaunchKernelGGL(addsub_vec4_kernel<true>, dim3(div_round_up(threads, 256)), dim3(256), 0, stream, d.data,
                       d.len, d.stride, a.data, a.len, a.stride, b.data, b.len, b.stride, count);
  } else {
    hipLaunchKernelGGL(addsub_general_kernel<true>, dim3(div_round_up(count, 256)), dim3(256), 0, stream, d, a, b,
                       count);
  }
  return hipGetLastError();
}

hipError_t launch_mul(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  // register-resident digit kernels for the common square shapes
  if (a.len == b.len) {
    switch (a.len) {
      case 4: return launch_mul_digits<4, 4>(p, a, b, count, stream);
      case 8: return launch_mul_digits<8, 8>(p, a, b, count, stream);
      case 16: return launch_mul_digits<16, 16>(p, a, b, count, stream);
      case 32: return launch_mul_digits<32, 32>(p, a, b, count, stream);
      case 64: return launch_mul_digits<64, 64>(p, a, b, count, stream);
      default: break;
    }
  }
  hipLaunchKernelGGL(mul_general_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, p, a, b, count);
  return hipGetLastError();
}

}  // namespace xmp
