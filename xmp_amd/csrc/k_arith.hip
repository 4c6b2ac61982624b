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
// 2^P(p).  The longer operand is split into 28-bit digits held in registers
// and the product is formed row by row in 64-bit "lazy" accumulators: each
// partial product is one v_mad_u64_u32 and carries are resolved once per
// retired digit, because up to 256 products of 56 bits fit in 64 bits.
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
// Operands are streamed once, so loads and stores are non-temporal (they do
// not displace cached data); when both inputs are at least as long as the
// output (FULL) the limb loop carries no zero-extension tests.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ u32x4 ld4(const uint32_t *p) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
  else return *reinterpret_cast<const u32x4 *>(p);
}

__device__ __forceinline__ void st4(uint32_t *p, u32x4 v) { __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(p)); }

template <bool SUB>
__device__ __forceinline__ u32x4 addsub4(u32x4 a, u32x4 b, uint32_t (&c)[4]) {
  uint32_t r0, r1, r2, r3;
  addsub_limb<SUB>(r0, a.x, b.x, c[0]);
  addsub_limb<SUB>(r1, a.y, b.y, c[1]);
  addsub_limb<SUB>(r2, a.z, b.z, c[2]);
  addsub_limb<SUB>(r3, a.w, b.w, c[3]);
  return u32x4{r0, r1, r2, r3};
}

template <bool SUB, bool FULL>
__global__ void __launch_bounds__(64) addsub_vec4_kernel(uint32_t *__restrict__ s, uint32_t s_len, uint32_t s_stride,
                                                         const uint32_t *__restrict__ a, uint32_t a_len,
                                                         uint32_t a_stride, const uint32_t *__restrict__ b,
                                                         uint32_t b_len, uint32_t b_stride, uint32_t count) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nvec = count / 4;
  if (t < nvec) {
    uint32_t i = 4 * t;
    uint32_t c[4] = {0, 0, 0, 0};
    uint32_t j = 0;
    const u32x4 zero = {0, 0, 0, 0};
    for (; j + 8 <= s_len; j += 8) {
      u32x4 av[8], bv[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (FULL) {
          av[k] = ld4<true>(a + (size_t)(j + k) * a_stride + i);
          bv[k] = ld4<true>(b + (size_t)(j + k) * b_stride + i);
        } else {
          av[k] = (j + k < a_len) ? ld4<true>(a + (size_t)(j + k) * a_stride + i) : zero;
          bv[k] = (j + k < b_len) ? ld4<true>(b + (size_t)(j + k) * b_stride + i) : zero;
        }
      }
#pragma unroll
      for (int k = 0; k < 8; k++) st4(s + (size_t)(j + k) * s_stride + i, addsub4<SUB>(av[k], bv[k], c));
    }
    for (; j < s_len; j++) {
      u32x4 av = (FULL || j < a_len) ? ld4<true>(a + (size_t)j * a_stride + i) : zero;
      u32x4 bv = (FULL || j < b_len) ? ld4<true>(b + (size_t)j * b_stride + i) : zero;
      st4(s + (size_t)j * s_stride + i, addsub4<SUB>(av, bv, c));
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

template <bool SUB>
void launch_addsub_vec4(const OutOperand &s, const Operand &a, const Operand &b, uint32_t count, hipStream_t stream) {
  const uint32_t threads = count / 4 + (count & 3);
  const dim3 grid(div_round_up(threads, 64)), block(64);
  if (a.len >= s.len && b.len >= s.len)
    hipLaunchKernelGGL((addsub_vec4_kernel<SUB, true>), grid, block, 0, stream, s.data, s.len, s.stride, a.data,
                       a.len, a.stride, b.data, b.len, b.stride, count);
  else
    hipLaunchKernelGGL((addsub_vec4_kernel<SUB, false>), grid, block, 0, stream, s.data, s.len, s.stride, a.data,
                       a.len, a.stride, b.data, b.len, b.stride, count);
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
    const int bit = kDigit * k;
    const int w = bit >> 5, sh = bit & 31;
    uint32_t lo = w < NL + 1 ? l[w] : 0u;
    uint32_t hi = w + 1 < NL + 1 ? l[w + 1] : 0u;
    d[k] = (sh ? ((lo >> sh) | (hi << (32 - sh))) : lo) & kDigitMask;
  }
}

// Row scanning (operand scanning) over radix-2^28 digits: the multiplicand's
// NDA digits and NDA lazy 64-bit accumulators stay in registers; the
// multiplier is streamed from HBM one digit per row through a bit buffer, and
// each row retires one finished product digit, so the accumulator window
// never grows.  Output digits are repacked into 32-bit limbs on the fly and
// truncated to p.len limbs.
template <int NLA>
__global__ void __launch_bounds__(64, (NLA >= 64 ? 1 : 2)) mul_rows_kernel(OutOperand p, Operand a, Operand b, uint32_t count) {
  constexpr int NDA = (32 * NLA + kDigit - 1) / kDigit;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint32_t ai = in_index(a, t), bi = in_index(b, t), pi = out_index(p, t);
  uint32_t da[NDA];
  load_digits<NLA, NDA>(da, a, ai);
  uint64_t acc[NDA];
#pragma unroll
  for (int k = 0; k < NDA; k++) acc[k] = 0;
  uint64_t ibuf = 0;   // multiplier bits not yet consumed
  uint32_t inb = 0, next = 0;
  uint64_t obuf = 0;   // product bits not yet stored
  uint32_t onb = 0, limb = 0;
  const uint32_t ndb = (32u * b.len + kDigit - 1) / kDigit;
  auto emit = [&](uint32_t digit) {
    obuf |= (uint64_t)digit << onb;
    onb += kDigit;
    if (onb >= 32) {
      if (limb < p.len) p.data[(size_t)limb * p.stride + pi] = (uint32_t)obuf;
      limb++;
      obuf >>= 32;
      onb -= 32;
    }
  };
  for (uint32_t i = 0; i < ndb; i++) {
    if (inb < (uint32_t)kDigit) {
      ibuf |= (uint64_t)ld_limb(b, bi, next++) << inb;
      inb += 32;
    }
    const uint32_t d = (uint32_t)ibuf & kDigitMask;
    ibuf >>= kDigit;
    inb -= kDigit;
    uint64_t f = acc[0] + (uint64_t)da[0] * d;
#pragma unroll
    for (int j = 1; j < NDA; j++) acc[j - 1] = acc[j] + (uint64_t)da[j] * d;
    acc[NDA - 1] = 0;
    acc[0] += f >> kDigit;
    emit((uint32_t)f & kDigitMask);
  }
  uint64_t c = 0;
#pragma unroll
  for (int k = 0; k < NDA; k++) {
    uint64_t x = acc[k] + c;
    emit((uint32_t)x & kDigitMask);
    c = x >> kDigit;
  }
  if (onb > 0) {
    if (limb < p.len) p.data[(size_t)limb * p.stride + pi] = (uint32_t)obuf;
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

template <int NLA>
hipError_t launch_mul_rows(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL((mul_rows_kernel<NLA>), dim3(div_round_up(count, 64)), dim3(64), 0, stream, p, a, b, count);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_add(OutOperand s, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  if (!s.idx && !a.idx && !b.idx && a.count >= count && b.count >= count) {
    launch_addsub_vec4<false>(s, a, b, count, stream);
  } else {
    hipLaunchKernelGGL(addsub_general_kernel<false>, dim3(div_round_up(count, 256)), dim3(256), 0, stream, s, a, b,
                       count);
  }
  return hipGetLastError();
}

hipError_t launch_sub(OutOperand d, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  if (!d.idx && !a.idx && !b.idx && a.count >= count && b.count >= count) {
    launch_addsub_vec4<true>(d, a, b, count, stream);
  } else {
    hipLaunchKernelGGL(addsub_general_kernel<true>, dim3(div_round_up(count, 256)), dim3(256), 0, stream, d, a, b,
                       count);
  }
  return hipGetLastError();
}

hipError_t launch_mul(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  // register-resident row kernels: the longer operand (a) in registers, any multiplier length
  if (a.len <= 4) return launch_mul_rows<4>(p, a, b, count, stream);
  if (a.len <= 8) return launch_mul_rows<8>(p, a, b, count, stream);
  if (a.len <= 16) return launch_mul_rows<16>(p, a, b, count, stream);
  if (a.len <= 32) return launch_mul_rows<32>(p, a, b, count, stream);
  if (a.len <= 48) return launch_mul_rows<48>(p, a, b, count, stream);
  if (a.len <= 64) return launch_mul_rows<64>(p, a, b, count, stream);
  hipLaunchKernelGGL(mul_general_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, p, a, b, count);
  return hipGetLastError();
}

}  // namespace xmp
