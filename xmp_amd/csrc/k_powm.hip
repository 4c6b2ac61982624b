// k_powm.hip -- batched modular exponentiation, out = base^exp mod m.
//
// Reference: src/powm.cu:240-307 (entry), src/include/powm/powm.h:39-127
// (fixed-window left-to-right exponentiation over every exponent bit with a
// full 2^w table whose slot 0 is the Montgomery one), warp_distributed.h and
// warpmp/mulmod_*.h (an instance spread over W lanes of a 32-wide warp, CIOS
// Montgomery products with 16-bit XMAD limb products, shuffle broadcasts and
// a ballot-based carry resolution).
//
// This kernel is designed for gfx950's VALU, where v_mad_u64_u32 issues at
// (nearly) full rate but every carry instruction costs as much as the
// multiply:
//
//  * operands are held as radix-2^D digits (D = 28, or 27 above 127 digits)
//    and the CIOS accumulator as 64-bit "lazy" sums: each partial product is a
//    single v_mad_u64_u32 into its 64-bit accumulator and no carry is
//    propagated inside the multiplication (2*ND products of < 2^(2D) each fit
//    in 64 bits);
//  * R = 2^(D*ND) >= 4m, so every product of inputs < 2m is < 2m again
//    (Walter's bound) and no conditional subtraction is needed until the
//    final conversion;
//  * an instance spans L lanes of the wave (L = 1, 2, 4, 8), each holding K
//    digits of the multiplicand, the modulus and the accumulator in
//    registers.  Per CIOS step the quotient digit is broadcast from the
//    group's first lane and one 28-bit digit moves one lane down, both with
//    DPP moves;
//  * the multiplier (the current value when squaring, a table entry when
//    multiplying) is staged in LDS and read four digits at a time with an
//    LDS broadcast;
//  * the 2^w window table lives in HBM, one contiguous block per instance.
//
// The exponent is scanned over all of its precision's bits with a fixed
// window (the reference's constant-time schedule, powm.h:94-98), and the
// result is reduced to the canonical value < m (GMP semantics).
#include "devmp.hpp"
#include "powm.hpp"

namespace xmp {

namespace {

// Cross-lane moves inside an instance's group of L lanes (groups are aligned
// within 16-lane DPP rows), done with DPP moves on the VALU instead of LDS
// permutes:
//   bcast0    -- the group's first lane's value to every lane of the group
//   from_next -- lane l receives lane l+1's value, the group's last lane 0
//   from_prev -- lane l receives lane l-1's value, the group's first lane 0
// DPP controls: quad_perm [0,0,2,2] = 0xA0, [0,0,0,0] = 0x00, row_shl:1 =
// 0x101, row_shr:1 = 0x111, row_shr:4 = 0x114, row_newbcast:0 = 0x150 (lane 0
// of each row to the whole row).
template <int L>
struct Group {
  int l;   // lane within the instance group
  __device__ __forceinline__ explicit Group(int lane) : l(lane & (L - 1)) {}
  __device__ __forceinline__ uint32_t bcast0(uint32_t x) const {
    if constexpr (L == 1) {
      return x;
    } else if constexpr (L == 2) {
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xA0, 0xF, 0xF, false);
    } else if constexpr (L == 4) {
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x00, 0xF, 0xF, false);
    } else if constexpr (L == 8) {
      uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x00, 0xF, 0xF, false);
      uint32_t z = (uint32_t)__builtin_amdgcn_mov_dpp((int)y, 0x114, 0xF, 0xF, true);
      return (l & 4) ? z : y;
    } else {
      static_assert(L == 16, "groups are at most one DPP row");
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x150, 0xF, 0xF, false);
    }
  }
  __device__ __forceinline__ uint32_t from_next(uint32_t x) const {
    if constexpr (L == 1) {
      return 0;
    } else {
      uint32_t v = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x101, 0xF, 0xF, true);
      if constexpr (L == 16) return v;   // the row's last lane already reads 0
      return l == L - 1 ? 0u : v;
    }
  }
  __device__ __forceinline__ uint64_t from_prev(uint64_t x) const {
    if constexpr (L == 1) {
      return 0;
    } else {
      uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)x, 0x111, 0xF, 0xF, true);
      uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(x >> 32), 0x111, 0xF, 0xF, true);
      uint64_t v = ((uint64_t)hi << 32) | lo;
      if constexpr (L == 16) return v;   // the row's first lane already reads 0
      return l == 0 ? 0ull : v;
    }
  }
  __device__ __forceinline__ bool any(bool p) const {
    if (L == 1) return p;
    uint32_t v = p ? 1u : 0u;
#pragma unroll
    for (int o = 1; o < L; o <<= 1) v |= (uint32_t)__shfl_xor((int)v, o, L);
    return v != 0;
  }
};

template <int L, int K, int D>
struct Mont {
  static constexpr int ND = L * K;
  static constexpr uint32_t MASK = (1u << D) - 1;
  Group<L> g;
  uint32_t m[K];
  uint32_t np0;

  __device__ __forceinline__ explicit Mont(int lane) : g(lane) {}

  // One CIOS step: t = (t + a*b_i + q*m) / 2^D with q making the lowest digit
  // vanish.  Lane l's first accumulator f (position l*K) is split as
  // f = f_lo + 2^D f_hi: f_lo moves to the previous lane's top position and
  // f_hi stays at this lane's new position 0.  For the group's first lane
  // f_lo = 0, so the code is the same on every lane and only a 28-bit value
  // crosses lanes.
  __device__ __forceinline__ void step(uint64_t (&t)[K], const uint32_t (&a)[K], uint32_t bi) const {
    uint64_t f = t[0] + (uint64_t)a[0] * bi;
    uint32_t q = ((uint32_t)f * np0) & MASK;
    q = g.bcast0(q);
    f += (uint64_t)q * m[0];
    const uint32_t up = g.from_next((uint32_t)f & MASK);
#pragma unroll
    for (int j = 1; j < K; j++) t[j - 1] = t[j] + (uint64_t)a[j] * bi + (uint64_t)q * m[j];
    t[K - 1] = up;
    t[0] += f >> D;
  }

  // r = a * b / R (mod m), r < 2m, normalised digits.  b: group's ND digits in LDS.
  __device__ __forceinline__ void mul(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t *b) const {
    uint64_t t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = 0;
    constexpr int ND4 = ND / 4 * 4;
    for (int i = 0; i < ND4; i += 4) {
      uint4 bb = *reinterpret_cast<const uint4 *>(b + i);
      step(t, a, bb.x);
      step(t, a, bb.y);
      step(t, a, bb.z);
      step(t, a, bb.w);
    }
#pragma unroll
    for (int i = ND4; i < ND; i++) step(t, a, b[i]);
    normalize(r, t);
  }

  __device__ __forceinline__ void normalize(uint32_t (&r)[K], const uint64_t (&t)[K]) const {
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      uint64_t x = t[k] + c;
      r[k] = (uint32_t)x & MASK;
      c = x >> D;
    }
#pragma unroll
    for (int round = 1; round < L; round++) {
      c = g.from_prev(c);
#pragma unroll
      for (int k = 0; k < K; k++) {
        uint64_t x = (uint64_t)r[k] + c;
        r[k] = (uint32_t)x & MASK;
        c = x >> D;
      }
    }
  }
};

// digit k of an integer given by 32-bit limbs (interleaved) as radix-2^D
template <int D>
__device__ __forceinline__ uint32_t limb_digit(const uint32_t *data, uint32_t stride, uint32_t inst, uint32_t len,
                                               uint32_t digit) {
  uint32_t bit = (uint32_t)D * digit;
  uint32_t w = bit >> 5, s = bit & 31;
  uint32_t lo = w < len ? data[(size_t)w * stride + inst] : 0u;
  uint32_t hi = w + 1 < len ? data[(size_t)(w + 1) * stride + inst] : 0u;
  return funnel(hi, lo, s) & ((1u << D) - 1);
}

__device__ __forceinline__ uint32_t exp_bits(const Operand &e, uint32_t inst, uint32_t pos, uint32_t len) {
  uint32_t w = pos >> 5, s = pos & 31;
  uint32_t lo = ld_limb(e, inst, w);
  uint32_t hi = ld_limb(e, inst, w + 1);
  return funnel(hi, lo, s) & ((1u << len) - 1);
}

// One tile: the 64/L instances [tile*64/L, tile*64/L + 64/L) of a launch,
// processed by one wave.  `bbuf` is this wave's LDS staging area
// (G * NDP words, 16-byte aligned).
template <int L, int K, int D>
__device__ __forceinline__ void powm_tile(const PowmArgs &args, uint32_t tile, uint32_t *bbuf) {
  using M = Mont<L, K, D>;
  constexpr int ND = M::ND;
  constexpr int NDP = (ND + 3) / 4 * 4;
  constexpr int G = 64 / L;   // instances per wave

  const int lane = threadIdx.x;
  const int grp = lane / L;
  M mont(lane);
  const int l = mont.g.l;
  const uint32_t local = tile * G + grp;
  const bool active = local < args.count;
  const uint32_t li = active ? local : args.count - 1;   // clamp: inactive lanes recompute a valid instance
  const uint32_t gi = args.start + li;
  const uint32_t bi = in_index(args.base, gi), ei = in_index(args.exp, gi), mi = in_index(args.mod, gi);
  uint32_t *b = bbuf + grp * NDP;

  // modulus as digits; R^2 mod m straight into the LDS multiplier buffer
  uint32_t a[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    uint32_t d = (uint32_t)(l * K + k);
    mont.m[k] = limb_digit<D>(args.mod.data, args.mod.stride, mi, args.mod.len, d);
    b[l * K + k] = limb_digit<D>(args.r2, args.r2_stride, mi, args.mod.len, d);
  }
  {
    // np0 = -m^-1 mod 2^D (Newton, as reference regmp/np0.h:24-33)
    uint32_t m0 = ld_limb(args.mod, mi, 0);
    uint32_t inv = m0;
#pragma unroll
    for (int it = 0; it < 4; it++) inv = inv * (2u - m0 * inv);
    mont.np0 = (0u - inv) & M::MASK;
  }

  const uint32_t wbits = args.window_bits;
  const uint32_t T = 1u << wbits;
  uint32_t *tab = args.table + (size_t)local * T * ND;   // table sized for whole waves
#define tab_at(e, k) (tab + ((size_t)(e) * K + (k)) * L + l)

  // slot 0 = R mod m = 1 * (R^2 mod m) / R
#pragma unroll
  for (int k = 0; k < K; k++) a[k] = (l == 0 && k == 0) ? 1u : 0u;
  __syncthreads();
  mont.mul(a, a, b);
#pragma unroll
  for (int k = 0; k < K; k++) *tab_at(0, k) = a[k];
  // slot 1 = base * R mod m
#pragma unroll
  for (int k = 0; k < K; k++)
    a[k] = limb_digit<D>(args.base.data, args.base.stride, bi, args.base.len, (uint32_t)(l * K + k));
  mont.mul(a, a, b);
  __syncthreads();
  // slots 2..T-1: slot[e] = slot[e-1] * slot[1]
#pragma unroll
  for (int k = 0; k < K; k++) {
    *tab_at(1, k) = a[k];
    b[l * K + k] = a[k];
  }
  __syncthreads();
  for (uint32_t e = 2; e < T; e++) {
    mont.mul(a, a, b);
#pragma unroll
    for (int k = 0; k < K; k++) *tab_at(e, k) = a[k];
  }
  __syncthreads();

  // exponent scan, top window first
  const uint32_t bits = args.exp.len * 32;
  uint32_t pos = bits - ((bits % wbits) ? (bits % wbits) : wbits);
  uint32_t idx = exp_bits(args.exp, ei, pos, bits - pos);
#pragma unroll
  for (int k = 0; k < K; k++) a[k] = *tab_at(idx, k);
  while (pos > 0) {
    pos -= wbits;
    idx = exp_bits(args.exp, ei, pos, wbits);
    for (uint32_t s = 0; s < wbits; s++) {
#pragma unroll
      for (int k = 0; k < K; k++) b[l * K + k] = a[k];
      __syncthreads();
      mont.mul(a, a, b);
      __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; k++) b[l * K + k] = *tab_at(idx, k);
    __syncthreads();
    mont.mul(a, a, b);
    __syncthreads();
  }
#undef tab_at
  // leave Montgomery form: a * 1 / R, which is <= m
#pragma unroll
  for (int k = 0; k < K; k++) b[l * K + k] = (l == 0 && k == 0) ? 1u : 0u;
  __syncthreads();
  uint32_t r[K];
  mont.mul(r, a, b);
  __syncthreads();
  // canonical: r == m  ->  0
  bool ne = false;
#pragma unroll
  for (int k = 0; k < K; k++) ne |= (r[k] != mont.m[k]);
  ne = mont.g.any(ne);
  // an even modulus is outside Montgomery's domain: defined result 0
  const bool odd = (ld_limb(args.mod, mi, 0) & 1u) != 0;
#pragma unroll
  for (int k = 0; k < K; k++) b[l * K + k] = (ne && odd) ? r[k] : 0u;
  __syncthreads();
  if (active) {
    const uint32_t oi = out_index(args.out, gi);
    for (uint32_t w = (uint32_t)l; w < args.out.len; w += L) {
      uint32_t bit = 32u * w;
      uint32_t d0 = bit / D, s0 = bit % D;
      uint32_t x0 = d0 < (uint32_t)ND ? b[d0] : 0u;
      uint32_t x1 = d0 + 1 < (uint32_t)ND ? b[d0 + 1] : 0u;
      uint32_t x2 = d0 + 2 < (uint32_t)ND ? b[d0 + 2] : 0u;
      uint64_t lo = ((uint64_t)x1 << D) | x0;
      uint32_t v = (uint32_t)(lo >> s0);
      uint32_t sh2 = 2u * D - s0;
      if (sh2 < 32) v |= x2 << sh2;
      args.out.data[(size_t)w * args.out.stride + oi] = v;
    }
  }
}

template <int L, int K, int D>
constexpr int lds_words() {
  return (64 / L) * ((L * K + 3) / 4 * 4);
}

template <int L, int K, int D>
__global__ void __launch_bounds__(64, 2) powm_kernel(PowmArgs args) {
  __shared__ __attribute__((aligned(16))) uint32_t bbuf[lds_words<L, K, D>()];
  powm_tile<L, K, D>(args, blockIdx.x, bbuf);
}

struct VariantEntry {
  PowmVariant v;
  void (*kernel)(PowmArgs);
};

#define XMP_VARIANT(L, K, D) {{L, K, D}, powm_kernel<L, K, D>}
const VariantEntry kVariants[] = {
    // <=1024-bit: one lane per instance
    XMP_VARIANT(1, 12, 28),  // 336-bit R : up to 320-bit moduli
    XMP_VARIANT(1, 20, 28),  // 560 : 512-bit
    XMP_VARIANT(1, 28, 28),  // 784 : 768-bit
    XMP_VARIANT(1, 37, 28),  // 1036 : 1024-bit
    XMP_VARIANT(2, 19, 28),  // 1064 : 1024-bit on two lanes
    XMP_VARIANT(2, 28, 28),  // 1568 : 1536-bit
    XMP_VARIANT(2, 37, 28),  // 2072 : 2048-bit
    XMP_VARIANT(4, 19, 28),  // 2128 : 2048-bit on four lanes
    XMP_VARIANT(4, 28, 28),  // 3136 : 3072-bit
    XMP_VARIANT(2, 55, 28),  // 3080 : 3072-bit on two lanes
    XMP_VARIANT(4, 38, 27),  // 4104 : 4096-bit
    XMP_VARIANT(8, 19, 27),  // 4104 : 4096-bit on eight lanes
    XMP_VARIANT(8, 29, 27),  // 6264 : 6144-bit
    XMP_VARIANT(8, 38, 27),  // 8208 : 8192-bit
    XMP_VARIANT(16, 19, 27), // 8208 : 8192-bit on sixteen lanes (one DPP row)
};
#undef XMP_VARIANT
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

// ---- persistent mixed-size kernel --------------------------------------
// Every wave pulls (class, tile) pairs from a device queue ordered by
// decreasing cost and runs the tile with its class's layout, so tiles of all
// precision classes share one launch and no SIMD idles at a class boundary.
// Each variant's tile is its own (non-inlined) function: it gets its own
// register allocation, and the call costs nothing next to a tile's runtime.
template <int I>
__device__ __attribute__((noinline)) void run_variant(const PowmArgs &args, uint32_t tile, uint32_t *lds);

#define XMP_RUN(I, L, K, D)                                                                          \
  template <>                                                                                        \
  __device__ __attribute__((noinline)) void run_variant<I>(const PowmArgs &args, uint32_t tile, uint32_t *lds) { \
    powm_tile<L, K, D>(args, tile, lds);                                                             \
  }
XMP_RUN(0, 1, 12, 28)
XMP_RUN(1, 1, 20, 28)
XMP_RUN(2, 1, 28, 28)
XMP_RUN(3, 1, 37, 28)
XMP_RUN(4, 2, 19, 28)
XMP_RUN(5, 2, 28, 28)
XMP_RUN(6, 2, 37, 28)
XMP_RUN(7, 4, 19, 28)
XMP_RUN(8, 4, 28, 28)
XMP_RUN(9, 2, 55, 28)
XMP_RUN(10, 4, 38, 27)
XMP_RUN(11, 8, 19, 27)
XMP_RUN(12, 8, 29, 27)
XMP_RUN(13, 8, 38, 27)
XMP_RUN(14, 16, 19, 27)
#undef XMP_RUN
static_assert(kNumVariants == 15, "run_variant<> must list every compiled variant in kVariants order");
static_assert(9 < kNumVariants, "variant 9 is the unbatched 2x55 layout");

__global__ void __launch_bounds__(64, 2) powm_batch_kernel(BatchArgs ba) {
  extern __shared__ __attribute__((aligned(16))) uint32_t dyn_lds[];
  while (true) {
    uint32_t idx = 0;
    if (threadIdx.x == 0) idx = atomicAdd(ba.counter, 1u);
    idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)__shfl((int)idx, 0));
    if (idx >= ba.ntiles) break;
    const uint32_t cls = (uint32_t)__builtin_amdgcn_readfirstlane((int)ba.tile_class[idx]);
    const uint32_t tile = (uint32_t)__builtin_amdgcn_readfirstlane((int)ba.tile_index[idx]);
    const PowmArgs &args = ba.classes[cls];
    switch (__builtin_amdgcn_readfirstlane(ba.class_variant[cls])) {
      case 0: run_variant<0>(args, tile, dyn_lds); break;
      case 1: run_variant<1>(args, tile, dyn_lds); break;
      case 2: run_variant<2>(args, tile, dyn_lds); break;
      case 3: run_variant<3>(args, tile, dyn_lds); break;
      case 4: run_variant<4>(args, tile, dyn_lds); break;
      case 5: run_variant<5>(args, tile, dyn_lds); break;
      case 6: run_variant<6>(args, tile, dyn_lds); break;
      case 7: run_variant<7>(args, tile, dyn_lds); break;
      case 8: run_variant<8>(args, tile, dyn_lds); break;
      case 10: run_variant<10>(args, tile, dyn_lds); break;
      case 11: run_variant<11>(args, tile, dyn_lds); break;
      case 12: run_variant<12>(args, tile, dyn_lds); break;
      case 13: run_variant<13>(args, tile, dyn_lds); break;
      case 14: run_variant<14>(args, tile, dyn_lds); break;
      default: break;
    }
    __syncthreads();
  }
}

__global__ void queue_reset_kernel(uint32_t *counter) {
  if (threadIdx.x == 0) counter[threadIdx.x] = 0u;
}

}  // namespace

int powm_variant_count() { return kNumVariants; }
PowmVariant powm_variant(int i) { return kVariants[i].v; }

int powm_find_variant(int lanes, int digits, int radix) {
  for (int i = 0; i < kNumVariants; i++)
    if (kVariants[i].v.lanes == lanes && kVariants[i].v.digits == digits && kVariants[i].v.radix == radix) return i;
  return -1;
}

bool variant_fits(const PowmVariant &v, uint32_t precision) {
  const uint32_t nd = (uint32_t)(v.lanes * v.digits);
  // lazy accumulators: 2*ND products of < 2^(2D) plus a carry must fit in 64 bits
  const uint32_t max_nd = v.radix == 28 ? 127u : 511u;
  return nd <= max_nd && variant_rbits(v) >= precision + 2;
}

size_t powm_table_words(const PowmVariant &v, uint32_t count, uint32_t window_bits) {
  return (size_t)count * ((size_t)1 << window_bits) * (size_t)(v.lanes * v.digits);
}

// the 2-lane x 55-digit layout spills as a called function: not batched
bool powm_variant_batchable(int variant) { return variant >= 0 && variant < kNumVariants && variant != 9; }

size_t powm_variant_lds_bytes(int variant) {
  const PowmVariant &v = kVariants[variant].v;
  return sizeof(uint32_t) * (size_t)(64 / v.lanes) * (size_t)((v.lanes * v.digits + 3) / 4 * 4);
}

hipError_t launch_powm_batch(const BatchArgs &ba, uint32_t grid, size_t lds_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, ba.counter);
  hipLaunchKernelGGL(powm_batch_kernel, dim3(grid), dim3(64), lds_bytes, stream, ba);
  return hipGetLastError();
}

hipError_t launch_powm_variant(int variant, const PowmArgs &args, hipStream_t stream) {
  if (variant < 0 || variant >= kNumVariants || args.count == 0) return hipErrorInvalidValue;
  const PowmVariant &v = kVariants[variant].v;
  const uint32_t per_block = 64u / (uint32_t)v.lanes;
  hipLaunchKernelGGL(kVariants[variant].kernel, dim3(div_round_up(args.count, per_block)), dim3(64), 0, stream, args);
  return hipGetLastError();
}

}  // namespace xmp
