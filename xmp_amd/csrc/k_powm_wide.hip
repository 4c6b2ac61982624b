// k_powm_wide.hip -- modular exponentiation for moduli wider than any lane
// layout holds (above 33,278 bits): one 256- or 512-thread workgroup per
// instance.
//
// Reference: the digit-serial fallback the reference uses above 8192 bits
// (src/include/powm/digitized.h, powm_implementations.h:154-286), with the
// same fixed-window schedule as every other powm here (powm/powm.h:94-98).
//
// The instance's ND = T*C radix-2^24 digits are spread over the T threads
// (four or eight waves) of a workgroup, C per thread, with the CIOS
// accumulator as 64-bit lazy sums in registers (2*ND products of < 2^48 plus
// carries stay below 2^64 for ND <= 16384).  Per CIOS step every thread reads the step's multiplier digit
// from LDS, thread 0 publishes the quotient digit of the next step, and every
// thread hands the low digit of its retired accumulator to the thread below,
// through LDS with one workgroup barrier per step (the quotient and the
// hand-over are double-buffered, so a step's reads never meet the next step's
// writes).  R = 2^(24*ND) = 3 * 2^(11 + log2 C), so R^2 mod m comes from one
// short division of 2^(rbits + 3) and log2(T*C) + 3 Montgomery squarings.
//
// Tiers (T x C): 256 x 8, 256 x 16, 256 x 32 and 512 x 32, i.e. moduli up to
// 49,150, 98,302, 196,606 and 393,214 bits.  The 512-thread tier holds 16,384
// digits in 64 KiB of LDS with 8 waves on one CU (two per SIMD).  The layout
// trades speed for reach: each step pays a workgroup barrier, and a modulus
// just above a tier pays the next tier's digit count.
#include <algorithm>

#include "devmp.hpp"
#include "lanes.hpp"
#include "powm.hpp"

namespace xmp {

namespace {

constexpr int kWideD = 24;
constexpr uint32_t kWideMask = (1u << kWideD) - 1;

template <int C, int T>
struct Wide {
  static constexpr int kWideThreads = T;
  static constexpr int ND = kWideThreads * C;
  static_assert(C >= 2, "a thread's lowest accumulator must not depend on its neighbour");
  static_assert(2.0 * ND * (double)kWideMask * kWideMask + 4.0 * (1ull << 40) < 18446744073709551615.0,
                "lazy accumulator bound");
  uint32_t m[C];
  uint32_t np0;
  int tid;
  uint32_t *b;      // ND multiplier digits (LDS)
  uint32_t *xchg;   // 2 x kWideThreads hand-over digits (LDS)
  uint32_t *qbuf;   // 2 quotient digits (LDS)
  uint64_t *cbuf;   // kWideThreads carries (LDS)

  // r = a * b / R (mod m), r < 2m, normalised; b is the LDS multiplier.
  // Every thread of the workgroup calls it (it contains barriers).
  __device__ void mul(uint32_t (&r)[C], const uint32_t (&a)[C]) const {
    uint64_t t[C];
#pragma unroll
    for (int k = 0; k < C; k++) t[k] = 0;
    if (tid == 0) qbuf[0] = ((uint32_t)((uint64_t)a[0] * b[0]) * np0) & kWideMask;
    __syncthreads();
    for (int i = 0; i < ND; i++) {
      const uint32_t bi = b[i];
      const uint32_t q = qbuf[i & 1];
      uint64_t f = t[0] + (uint64_t)a[0] * bi + (uint64_t)q * m[0];
#pragma unroll
      for (int k = 1; k < C; k++) t[k - 1] = t[k] + (uint64_t)a[k] * bi + (uint64_t)q * m[k];
      t[0] += f >> kWideD;
      xchg[(i & 1) * kWideThreads + tid] = (uint32_t)f & kWideMask;
      // the next quotient digit: thread 0's new lowest accumulator does not
      // depend on the value it receives from thread 1 (C >= 2)
      if (tid == 0 && i + 1 < ND) qbuf[(i + 1) & 1] = ((uint32_t)(t[0] + (uint64_t)a[0] * b[i + 1]) * np0) & kWideMask;
      __syncthreads();
      t[C - 1] = tid + 1 < kWideThreads ? xchg[(i & 1) * kWideThreads + tid + 1] : 0u;
    }
    normalize(r, t);
  }

  // digits < 2^24: a local carry pass, then carries move up one thread per
  // round until no thread has one (the value is < 2m < R, so nothing leaves
  // the top thread; its carry is dropped to bound the loop).
  __device__ void normalize(uint32_t (&r)[C], const uint64_t (&t)[C]) const {
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < C; k++) {
      const uint64_t x = t[k] + c;
      r[k] = (uint32_t)x & kWideMask;
      c = x >> kWideD;
    }
    for (int round = 0; round < kWideThreads; round++) {
      cbuf[tid] = tid + 1 < kWideThreads ? c : 0ull;
      __syncthreads();
      uint64_t cin = tid > 0 ? cbuf[tid - 1] : 0ull;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < C; k++) {
        const uint64_t x = (uint64_t)r[k] + cin;
        r[k] = (uint32_t)x & kWideMask;
        cin = x >> kWideD;
      }
      c = cin;
      if (!__syncthreads_or(tid + 1 < kWideThreads && c != 0)) break;
    }
  }

  __device__ void stage(const uint32_t (&a)[C]) const {
#pragma unroll
    for (int k = 0; k < C; k++) b[tid * C + k] = a[k];
    __syncthreads();
  }
};

// Thread t holds digits t*C .. t*C+C-1; tables store digit t*C+k of entry e
// at ((e * C + k) * T + t), so a table access is coalesced over threads.
template <int C, int T>
__global__ void __launch_bounds__(T) powm_wide_kernel(PowmArgs args, uint32_t squarings) {
  using W = Wide<C, T>;
  constexpr int kWideThreads = T;
  constexpr int ND = W::ND;
  __shared__ uint32_t s_b[ND];
  __shared__ uint32_t s_x[2 * kWideThreads];
  __shared__ uint32_t s_q[2];
  __shared__ uint64_t s_c[kWideThreads];
  W w;
  w.tid = threadIdx.x;
  w.b = s_b;
  w.xchg = s_x;
  w.qbuf = s_q;
  w.cbuf = s_c;
  const int tid = threadIdx.x;
  const uint32_t li = blockIdx.x;   // the grid is exactly args.count workgroups
  const uint32_t gi = args.start + li;
  const uint32_t mi = in_index(args.mod, gi);
#pragma unroll
  for (int k = 0; k < C; k++)
    w.m[k] = limb_digit_guarded<kWideD>(args.mod.data, args.mod.stride, mi, args.mod.len, (uint32_t)(tid * C + k));
  {
    const uint32_t m0 = ld_limb(args.mod, mi, 0);
    uint32_t inv = m0;
#pragma unroll
    for (int it = 0; it < 4; it++) inv = inv * (2u - m0 * inv);
    w.np0 = (0u - inv) & kWideMask;
  }
  uint32_t a[C];

  if (squarings) {
    // R^2 set-up: args.base holds 2^t R mod m (one per modulus); square it
    for (int k = 0; k < C; k++)
      a[k] = limb_digit_guarded<kWideD>(args.base.data, args.base.stride, mi, args.base.len, (uint32_t)(tid * C + k));
    for (uint32_t s = 0; s < squarings; s++) {
      w.stage(a);
      w.mul(a, a);
      __syncthreads();
    }
  } else {
    const uint32_t bi = in_index(args.base, gi), ei = in_index(args.exp, gi);
    const uint32_t wbits = args.window_bits;
    const uint32_t entries = 1u << wbits;
    uint32_t *tab = args.table + (size_t)li * entries * ND;
    auto tab_at = [&](uint32_t e, int k) -> uint32_t * { return tab + ((size_t)e * C + k) * kWideThreads + tid; };
    // slot 0 = R mod m, slot 1 = base R mod m, slot e = slot e-1 * slot 1
#pragma unroll
    for (int k = 0; k < C; k++)
      s_b[tid * C + k] = limb_digit_guarded<kWideD>(args.r2, args.r2_stride, mi, args.mod.len, (uint32_t)(tid * C + k));
#pragma unroll
    for (int k = 0; k < C; k++) a[k] = (tid == 0 && k == 0) ? 1u : 0u;
    __syncthreads();
    w.mul(a, a);
#pragma unroll
    for (int k = 0; k < C; k++) *tab_at(0, k) = a[k];
#pragma unroll
    for (int k = 0; k < C; k++)
      a[k] = limb_digit_guarded<kWideD>(args.base.data, args.base.stride, bi, args.base.len, (uint32_t)(tid * C + k));
    w.mul(a, a);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C; k++) *tab_at(1, k) = a[k];
    w.stage(a);
    for (uint32_t e = 2; e < entries; e++) {
      w.mul(a, a);
#pragma unroll
      for (int k = 0; k < C; k++) *tab_at(e, k) = a[k];
    }
    __syncthreads();
    const uint32_t bits = args.exp.len * 32;
    uint32_t pos = bits - ((bits % wbits) ? (bits % wbits) : wbits);
    auto window = [&](uint32_t p, uint32_t len) {
      const uint32_t wd = p >> 5, sh = p & 31;
      return funnel(ld_limb(args.exp, ei, wd + 1), ld_limb(args.exp, ei, wd), sh) & ((1u << len) - 1);
    };
    uint32_t idx = window(pos, bits - pos);
#pragma unroll
    for (int k = 0; k < C; k++) a[k] = *tab_at(idx, k);
    while (pos > 0) {
      pos -= wbits;
      idx = window(pos, wbits);
      for (uint32_t s = 0; s < wbits; s++) {
        w.stage(a);
        w.mul(a, a);
        __syncthreads();
      }
#pragma unroll
      for (int k = 0; k < C; k++) s_b[tid * C + k] = *tab_at(idx, k);
      __syncthreads();
      w.mul(a, a);
      __syncthreads();
    }
    // leave Montgomery form (a * 1 / R <= m), then r == m -> 0, even m -> 0
#pragma unroll
    for (int k = 0; k < C; k++) s_b[tid * C + k] = (tid == 0 && k == 0) ? 1u : 0u;
    __syncthreads();
    w.mul(a, a);
    bool ne = false;
#pragma unroll
    for (int k = 0; k < C; k++) ne |= a[k] != w.m[k];
    ne = __syncthreads_or(ne) != 0;
    const bool odd = (ld_limb(args.mod, mi, 0) & 1u) != 0;
#pragma unroll
    for (int k = 0; k < C; k++)
      if (!(ne && odd)) a[k] = 0;
  }
  __syncthreads();
  w.stage(a);
  uint32_t *out = args.out.data;
  const uint32_t oi = squarings ? mi : out_index(args.out, gi);
  for (uint32_t wd = (uint32_t)tid; wd < args.out.len; wd += kWideThreads) {
    const uint32_t bit = 32u * wd, d0 = bit / kWideD, s0 = bit % kWideD;
    const uint32_t x0 = d0 < (uint32_t)ND ? s_b[d0] : 0u;
    const uint32_t x1 = d0 + 1 < (uint32_t)ND ? s_b[d0 + 1] : 0u;
    const uint32_t x2 = d0 + 2 < (uint32_t)ND ? s_b[d0 + 2] : 0u;
    uint64_t lo = ((uint64_t)x1 << kWideD) | x0;
    uint32_t v = (uint32_t)(lo >> s0);
    const uint32_t sh2 = 2u * kWideD - s0;
    if (sh2 < 32) v |= x2 << sh2;
    out[(size_t)wd * args.out.stride + oi] = v;
  }
}

// ---- the memory tier: moduli above 393,214 bits --------------------------
// The same CIOS steps with the instance's digits in global memory instead of
// registers and LDS: 1024 threads, C radix-2^20 digits each (any C >= 20, so
// R = 2^(20*1024*C) lies within 20,480 bits above the modulus), digit j =
// tid*C + k stored at k*1024 + tid so that every access is coalesced.  An
// instance's block (multiplicand, modulus, multiplier, the 64-bit
// accumulator and its window table) is a few MB and stays in L2 while its
// workgroup runs.  Radix 2^20 keeps the lazy sums below 2^64 for up to
// 4,194,304 digits (2*ND*(2^20-1)^2 plus carries), so moduli up to
// 83,886,078 bits.  Each step costs a workgroup barrier and C rounds of L2
// traffic per thread: this is the reach of the reference's digit-serial
// fallback, not a fast layout.
constexpr int kMemD = 20;
constexpr uint32_t kMemMask = (1u << kMemD) - 1;
constexpr int kMemThreads = 1024;
constexpr uint32_t kMemMinC = 20;
constexpr uint32_t kMemMaxC = 4096;
static_assert(2.0 * kMemThreads * kMemMaxC * (double)kMemMask * kMemMask + 4.0 * (1ull << 44) <
                  18446744073709551615.0,
              "memory-tier lazy accumulator bound");

struct WideMem {
  uint32_t C, ND;
  int tid;
  const uint32_t *m;   // modulus digits
  uint32_t *b;         // multiplier digits
  uint64_t *t;         // accumulator
  uint32_t *xchg;      // 2 x kMemThreads hand-over digits (LDS)
  uint32_t *qbuf;      // 2 quotient digits (LDS)
  uint64_t *cbuf;      // kMemThreads carries (LDS)
  uint32_t np0;

  __device__ __forceinline__ uint32_t at(uint32_t j) const { return (j % C) * kMemThreads + j / C; }

  // r = a * b / R (mod m), r < 2m, normalised; a and r (which may be the same
  // array) and the multiplier b are in global memory.  Every thread calls it.
  __device__ void mul(uint32_t *r, const uint32_t *a) const {
    for (uint32_t k = 0; k < C; k++) t[k * kMemThreads + tid] = 0;
    if (tid == 0) qbuf[0] = ((uint32_t)((uint64_t)a[0] * b[0]) * np0) & kMemMask;
    __syncthreads();
    for (uint32_t i = 0; i < ND; i++) {
      const uint32_t bi = b[at(i)];
      const uint32_t q = qbuf[i & 1];
      const uint64_t f = t[tid] + (uint64_t)a[tid] * bi + (uint64_t)q * m[tid];
      // position k moves to k-1; position 1 also takes the retired digit's carry
      uint64_t t0 = 0;
#pragma unroll 4
      for (uint32_t k = 1; k < C; k++) {
        const uint32_t x = k * kMemThreads + tid;
        uint64_t v = t[x] + (uint64_t)a[x] * bi + (uint64_t)q * m[x];
        if (k == 1) {
          v += f >> kMemD;
          t0 = v;
        }
        t[x - kMemThreads] = v;
      }
      xchg[(i & 1) * kMemThreads + tid] = (uint32_t)f & kMemMask;
      // thread 0's new lowest position does not depend on thread 1's hand-over (C >= 2)
      if (tid == 0 && i + 1 < ND) qbuf[(i + 1) & 1] = ((uint32_t)(t0 + (uint64_t)a[0] * b[at(i + 1)]) * np0) & kMemMask;
      __syncthreads();
      t[(C - 1) * kMemThreads + tid] = tid + 1 < kMemThreads ? xchg[(i & 1) * kMemThreads + tid + 1] : 0u;
    }
    normalize(r);
  }

  // a local carry pass, then carries move up one thread per round until no
  // thread has one (the value is < 2m < R: nothing leaves the top thread)
  __device__ void normalize(uint32_t *r) const {
    uint64_t c = 0;
    for (uint32_t k = 0; k < C; k++) {
      const uint32_t x = k * kMemThreads + tid;
      const uint64_t v = t[x] + c;
      r[x] = (uint32_t)v & kMemMask;
      c = v >> kMemD;
    }
    for (int round = 0; round < kMemThreads; round++) {
      cbuf[tid] = tid + 1 < kMemThreads ? c : 0ull;
      __syncthreads();
      uint64_t cin = tid > 0 ? cbuf[tid - 1] : 0ull;
      __syncthreads();
      for (uint32_t k = 0; k < C && cin; k++) {
        const uint32_t x = k * kMemThreads + tid;
        const uint64_t v = (uint64_t)r[x] + cin;
        r[x] = (uint32_t)v & kMemMask;
        cin = v >> kMemD;
      }
      c = cin;
      if (!__syncthreads_or(tid + 1 < kMemThreads && c != 0)) break;
    }
  }

  // the multiplier := a (this thread's digits), visible to every thread after the barrier
  __device__ void stage(const uint32_t *a) const {
    for (uint32_t k = 0; k < C; k++) b[k * kMemThreads + tid] = a[k * kMemThreads + tid];
    __syncthreads();
  }
};

// words of one instance's block: the window table, then the multiplicand,
// modulus and multiplier (ND each) and the accumulator (2 ND)
__host__ __device__ inline size_t mem_block_words(uint32_t nd, uint32_t entries) {
  return ((size_t)entries + 5) * nd;
}

__global__ void __launch_bounds__(kMemThreads) powm_mem_kernel(PowmArgs args, uint32_t squarings, uint32_t C) {
  __shared__ uint32_t s_x[2 * kMemThreads];
  __shared__ uint32_t s_q[2];
  __shared__ uint64_t s_c[kMemThreads];
  const int tid = threadIdx.x;
  const uint32_t ND = kMemThreads * C;
  const uint32_t li = blockIdx.x;   // the grid is exactly args.count workgroups
  const uint32_t gi = args.start + li;
  const uint32_t mi = in_index(args.mod, gi);
  const uint32_t entries = squarings ? 0u : 1u << args.window_bits;
  uint32_t *blk = args.table + (size_t)li * mem_block_words(ND, entries);
  uint32_t *tab = blk;
  uint32_t *a = blk + (size_t)entries * ND;
  uint32_t *mm = a + ND;
  WideMem w;
  w.C = C;
  w.ND = ND;
  w.tid = tid;
  w.m = mm;
  w.b = mm + ND;
  w.t = reinterpret_cast<uint64_t *>(mm + 2 * (size_t)ND);
  w.xchg = s_x;
  w.qbuf = s_q;
  w.cbuf = s_c;
  for (uint32_t k = 0; k < C; k++)
    mm[k * kMemThreads + tid] =
        limb_digit_guarded<kMemD>(args.mod.data, args.mod.stride, mi, args.mod.len, (uint32_t)tid * C + k);
  {
    const uint32_t m0 = ld_limb(args.mod, mi, 0);
    uint32_t inv = m0;
#pragma unroll
    for (int it = 0; it < 4; it++) inv = inv * (2u - m0 * inv);
    w.np0 = (0u - inv) & kMemMask;
  }
  auto load = [&](uint32_t *dst, const uint32_t *src, uint32_t stride, uint32_t inst, uint32_t len) {
    for (uint32_t k = 0; k < C; k++)
      dst[k * kMemThreads + tid] = limb_digit_guarded<kMemD>(src, stride, inst, len, (uint32_t)tid * C + k);
  };
  auto unit = [&](uint32_t *dst) {   // the integer 1
    for (uint32_t k = 0; k < C; k++) dst[k * kMemThreads + tid] = (tid == 0 && k == 0) ? 1u : 0u;
  };
  auto copy = [&](uint32_t *dst, const uint32_t *src) {
    for (uint32_t k = 0; k < C; k++) dst[k * kMemThreads + tid] = src[k * kMemThreads + tid];
  };
  __syncthreads();

  if (squarings) {
    // R^2 set-up: args.base holds 2^t R mod m (one per modulus); square it
    load(a, args.base.data, args.base.stride, mi, args.base.len);
    for (uint32_t s = 0; s < squarings; s++) {
      w.stage(a);
      w.mul(a, a);
      __syncthreads();
    }
  } else {
    const uint32_t bi = in_index(args.base, gi), ei = in_index(args.exp, gi);
    const uint32_t wbits = args.window_bits;
    // slot 0 = R mod m = 1 * (R^2 mod m) / R, slot 1 = base R mod m, slot e = slot e-1 * slot 1
    load(w.b, args.r2, args.r2_stride, mi, args.mod.len);
    unit(a);
    __syncthreads();
    w.mul(tab, a);
    load(a, args.base.data, args.base.stride, bi, args.base.len);
    __syncthreads();
    w.mul(a, a);
    __syncthreads();
    copy(tab + ND, a);
    w.stage(a);
    for (uint32_t e = 2; e < entries; e++) {
      w.mul(tab + (size_t)e * ND, tab + (size_t)(e - 1) * ND);
      __syncthreads();
    }
    const uint32_t bits = args.exp.len * 32;
    uint32_t pos = bits - ((bits % wbits) ? (bits % wbits) : wbits);
    auto window = [&](uint32_t p, uint32_t len) {
      const uint32_t wd = p >> 5, sh = p & 31;
      return funnel(ld_limb(args.exp, ei, wd + 1), ld_limb(args.exp, ei, wd), sh) & ((1u << len) - 1);
    };
    uint32_t idx = window(pos, bits - pos);
    copy(a, tab + (size_t)idx * ND);
    __syncthreads();
    while (pos > 0) {
      pos -= wbits;
      idx = window(pos, wbits);
      for (uint32_t s = 0; s < wbits; s++) {
        w.stage(a);
        w.mul(a, a);
        __syncthreads();
      }
      copy(w.b, tab + (size_t)idx * ND);
      __syncthreads();
      w.mul(a, a);
      __syncthreads();
    }
    // leave Montgomery form (a * 1 / R <= m), then r == m -> 0, even m -> 0
    unit(w.b);
    __syncthreads();
    w.mul(a, a);
    __syncthreads();
    bool ne = false;
    for (uint32_t k = 0; k < C; k++) ne |= a[k * kMemThreads + tid] != mm[k * kMemThreads + tid];
    ne = __syncthreads_or(ne) != 0;
    const bool odd = (ld_limb(args.mod, mi, 0) & 1u) != 0;
    if (!(ne && odd))
      for (uint32_t k = 0; k < C; k++) a[k * kMemThreads + tid] = 0u;
  }
  __syncthreads();
  // digits to 32-bit limbs: limb wd takes digits floor(32 wd / 20) .. + 2
  uint32_t *out = args.out.data;
  const uint32_t oi = squarings ? mi : out_index(args.out, gi);
  auto digit = [&](uint32_t j) -> uint32_t { return j < ND ? a[(j % C) * kMemThreads + j / C] : 0u; };
  for (uint32_t wd = (uint32_t)tid; wd < args.out.len; wd += kMemThreads) {
    const uint32_t bit = 32u * wd, d0 = bit / kMemD, s0 = bit % kMemD;
    const uint64_t lo = ((uint64_t)digit(d0 + 2) << (2 * kMemD)) | ((uint64_t)digit(d0 + 1) << kMemD) | digit(d0);
    out[(size_t)wd * args.out.stride + oi] = (uint32_t)(lo >> s0);
  }
}

template <int C, int T>
hipError_t launch_wide(const PowmArgs &args, uint32_t squarings, hipStream_t stream) {
  hipLaunchKernelGGL((powm_wide_kernel<C, T>), dim3(args.count), dim3(T), 0, stream, args, squarings);
  return hipGetLastError();
}

// a tier is named by its digit count in units of 256 digits: 8, 16, 32 (256
// threads with 8, 16, 32 digits each) and 64 (512 threads with 32) hold the
// digits in registers; 4C for C >= 20 is the memory tier with C digits per
// thread of 1024
bool memory_tier(uint32_t tier) { return tier >= 4 * kMemMinC; }

hipError_t launch_tier(uint32_t tier, const PowmArgs &args, uint32_t squarings, hipStream_t stream) {
  if (memory_tier(tier)) {
    if (tier % 4 || tier / 4 > kMemMaxC || !args.table) return hipErrorInvalidValue;
    hipLaunchKernelGGL(powm_mem_kernel, dim3(args.count), dim3(kMemThreads), 0, stream, args, squarings, tier / 4);
    return hipGetLastError();
  }
  switch (tier) {
    case 8: return launch_wide<8, 256>(args, squarings, stream);
    case 16: return launch_wide<16, 256>(args, squarings, stream);
    case 32: return launch_wide<32, 256>(args, squarings, stream);
    case 64: return launch_wide<32, 512>(args, squarings, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

uint32_t powm_wide_digits_per_thread(uint32_t precision) {
  for (uint32_t tier : {8u, 16u, 32u, 64u})
    if (kWideD * 256u * tier >= precision + 2) return tier;
  // the memory tier: the fewest radix-2^20 digits per thread of 1024 that hold the modulus
  const uint64_t c = ((uint64_t)precision + 2 + (uint64_t)kMemD * kMemThreads - 1) / ((uint64_t)kMemD * kMemThreads);
  if (c > kMemMaxC) return 0;
  return 4u * (uint32_t)std::max<uint64_t>(c, kMemMinC);
}

uint32_t powm_wide_rbits(uint32_t tier) {
  return memory_tier(tier) ? (uint32_t)kMemD * kMemThreads * (tier / 4) : kWideD * 256u * tier;
}

size_t powm_wide_table_words(uint32_t tier, uint32_t count, uint32_t window_bits) {
  if (memory_tier(tier)) return (size_t)count * mem_block_words(kMemThreads * (tier / 4), 1u << window_bits);
  return (size_t)count * ((size_t)1 << window_bits) * 256u * tier;
}

size_t powm_wide_r2_scratch_words(uint32_t tier, uint32_t count) {
  return memory_tier(tier) ? (size_t)count * mem_block_words(kMemThreads * (tier / 4), 0) : 0;
}

hipError_t launch_powm_wide(uint32_t tier, const PowmArgs &args, hipStream_t stream) {
  if (args.count == 0) return hipErrorInvalidValue;
  return launch_tier(tier, args, 0, stream);
}

hipError_t launch_wide_r2(uint32_t c, uint32_t *r2, uint32_t r2_stride, const uint32_t *m, uint32_t m_len,
                          uint32_t m_stride, uint32_t count, uint32_t *ws, uint32_t *tier_ws, hipStream_t stream) {
  const uint32_t rbits = powm_wide_rbits(c);
  const uint32_t j = (uint32_t)__builtin_ctz(rbits), t = rbits >> j;
  if (!ws || count == 0 || (memory_tier(c) && !tier_ws)) return hipErrorInvalidValue;
  // Y = 2^(rbits+t) mod m into r2 (t = 3 for the register tiers, 5 times C's
  // odd part for the memory tier: a quotient of a few hundred limbs at most),
  // then j squarings into the workspace, then the canonical value back into r2
  hipError_t e = launch_pow2_mod(r2, m_len, r2_stride, m, m_len, m_stride, count, rbits + t, ws, stream);
  if (e != hipSuccess) return e;
  const uint32_t ws_stride = round_up(count, 256);
  PowmArgs a{};
  a.mod = Operand{m, m_len, m_stride, count, nullptr, 0};
  a.base = Operand{r2, m_len, r2_stride, count, nullptr, 0};
  a.out = OutOperand{ws, m_len + 1, ws_stride, nullptr};
  a.count = count;
  a.table = tier_ws;   // the memory tier's per-instance blocks
  if ((e = launch_tier(c, a, j, stream)) != hipSuccess) return e;
  return launch_r2_canonical(r2, r2_stride, ws, ws_stride, m, m_len, m_stride, count, stream);
}

}  // namespace xmp
