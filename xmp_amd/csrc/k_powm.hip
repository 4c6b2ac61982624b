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
//  * operands are held as radix-2^D digits (D = 28; the 4x37 and 8x37 groups
//    run carry passes inside a product to keep D = 28 above 127 digits; the
//    other wide layouts use D = 27 up to 511 digits and D = 26 up to 2047)
//    and the CIOS accumulator as 64-bit "lazy" sums: each partial product is a
//    single v_mad_u64_u32 into its 64-bit accumulator and no carry is
//    propagated inside the multiplication (2*ND products of < 2^(2D) each fit
//    in 64 bits);
//  * R = 2^(D*K*nl) >= 4m, nl <= L the lanes the modulus' digits occupy, so
//    every product of inputs < 2m is < 2m again (Walter's bound) and no
//    conditional subtraction is needed until the final conversion; a
//    product's steps run over those nl*K digits only;
//  * an instance spans L lanes of the wave (L = 1 ... 64), each holding K
//    digits of the multiplicand, the modulus and the accumulator in
//    registers.  Per CIOS step the quotient digit is broadcast from the
//    group's first lane and one 28-bit digit moves one lane down, both as
//    DPP operands of the instructions that mask them;
//  * the multiplier (the current value when squaring, a table entry when
//    multiplying) is staged in LDS and read with LDS broadcasts;
//  * squarings compute most products of two different digits once and
//    double them: within one lane as a triangle, across a lane group by a
//    tournament on register slots that gives every lane the same share
//    (K + K/2 + 1 multiply-adds per step instead of 2K);
//  * the 2^w window table lives in HBM, one contiguous block per instance.
//
// The exponent is scanned over all of its precision's bits with a fixed
// window (the reference's constant-time schedule, powm.h:94-98), and the
// result is reduced to the canonical value < m (GMP semantics).
#include <cmath>
#include <utility>

#include "devmp.hpp"
#include "lanes.hpp"
#include "powm.hpp"

namespace xmp {

namespace {



// Every powm workgroup is a single wave64, and the LDS operations of one wave
// complete in program order, so the staging through LDS needs no workgroup
// barrier: a wavefront-scope fence keeps the compiler from moving LDS
// accesses across the hand-over, and costs no instruction.  (Inside the
// persistent kernel's non-inlined tile functions the compiler cannot see the
// workgroup size and would otherwise emit a barrier with a full wait.)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// low 32 bits of x * y, issued as v_mad_u64_u32 (the multiply the kernels
// are built on) instead of v_mul_lo_u32
__device__ __forceinline__ uint32_t mul_lo_mad(uint32_t x, uint32_t y) {
  uint64_t r, c;
  asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(c) : "v"(x), "v"(y));
  return (uint32_t)r;
}

// np0 = -m0^-1 mod 2^D for odd m0 (Newton's iteration doubles the correct
// low bits, from 3 for inv = m0: reference regmp/np0.h:24-33)
template <int D>
__device__ __forceinline__ uint32_t neg_inverse(uint32_t m0) {
  uint32_t inv = m0;
#pragma unroll
  for (int it = 0; it < 4; it++) inv = inv * (2u - m0 * inv);
  return (0u - inv) & ((1u << D) - 1);
}

template <int L, int K, int D>
struct Mont {
  static constexpr int ND = L * K;
  static constexpr uint32_t MASK = (1u << D) - 1;
  Group<L> g;
  uint32_t m[K];
  uint32_t np0;
  // lanes whose digits the modulus occupies (powm_active_lanes): the steps
  // of a product run over their nl*K digits, R = 2^(D*K*nl)
  uint32_t nl = L;
  // the digit mask held in a register: an AND with a register operand can
  // take a DPP move of its other operand as a modifier, so the quotient's
  // broadcast and the digit hand-over each fold into their mask instruction
  uint32_t mask;

  __device__ __forceinline__ explicit Mont(int lane) : g(lane) {
    uint32_t v = MASK;
    asm volatile("v_mov_b32 %0, %1" : "=v"(mask) : "s"(v));
  }

  // One CIOS step: t = (t + a*b_i + q*m) / 2^D with q making the lowest digit
  // vanish.  Lane l's first accumulator f (position l*K) is split as
  // f = f_lo + 2^D f_hi: f_lo moves to the previous lane's top position and
  // f_hi stays at this lane's new position 0.  For the group's first lane
  // f_lo = 0, so the code is the same on every lane and only a 28-bit value
  // crosses lanes.
  template <bool QF = false>
  __device__ __forceinline__ void step(uint64_t (&t)[K], const uint32_t (&a)[K], uint32_t bi) const {
    uint64_t f = t[0] + (uint64_t)a[0] * bi;
    const uint32_t q = quotient<QF>(f);
    f += (uint64_t)q * m[0];
    const uint32_t up = g.next_digit((uint32_t)f, mask);
#pragma unroll
    for (int j = 1; j < K; j++) t[j - 1] = t[j] + (uint64_t)a[j] * bi + (uint64_t)q * m[j];
    t[K - 1] = up;
    t[0] += f >> D;
  }

  // The step's quotient digit, the group's first lane's q = f * np0 mod 2^D,
  // on every lane of the group.  Quotient-free (QF): m holds a multiple
  // m~ = m * np0 of the modulus, which is -1 mod 2^D, so q = f mod 2^D.
  template <bool QF>
  __device__ __forceinline__ uint32_t quotient(uint64_t f) const {
    if constexpr (QF) return g.bcast0_and((uint32_t)f, mask);
    return g.bcast0_and(mul_lo_mad((uint32_t)f, np0), mask);
  }

  // One CIOS step of a sum of two products: t = (t + a*b_i + c*d_i + q*m) / 2^D.
  // A position collects three products of < 2^(2D) per step, so `steps`
  // such steps need 3*steps*2^(2D) plus a carry below 2^64 (checked by the
  // launcher).
  __device__ __forceinline__ void step2(uint64_t (&t)[K], const uint32_t (&a)[K], uint32_t bi,
                                        const uint32_t (&c)[K], uint32_t di) const {
    uint64_t f = t[0] + (uint64_t)a[0] * bi + (uint64_t)c[0] * di;
    const uint32_t q = quotient<false>(f);
    f += (uint64_t)q * m[0];
    const uint32_t up = g.next_digit((uint32_t)f, mask);
#pragma unroll
    for (int j = 1; j < K; j++)
      t[j - 1] = t[j] + (uint64_t)a[j] * bi + (uint64_t)c[j] * di + (uint64_t)q * m[j];
    t[K - 1] = up;
    t[0] += f >> D;
  }

  // r = a * b / R (mod m), r < 2m, normalised digits.  b: group's ND digits in
  // LDS, of which the first nl*K are the multiplier's (R = 2^(D*K*nl)).  With
  // every lane's digits the step count is a constant, as the compiler best
  // schedules it; otherwise a run-time count.
  template <bool QF = false>
  __device__ __forceinline__ void mul(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t *b) const {
    if (L == 1 || nl == (uint32_t)L)
      mul_steps<QF>(r, a, b, ND);
    else
      mul_steps<QF>(r, a, b, (int)nl * K);
  }

  template <bool QF>
  __device__ __forceinline__ void mul_steps(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t *b,
                                            const int ns) const {
    uint64_t t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = 0;
    const int ns4 = ns / 4 * 4;
    for (int i = 0; i < ns4; i += 4) {
      uint4 bb = *reinterpret_cast<const uint4 *>(b + i);
      step<QF>(t, a, bb.x);
      step<QF>(t, a, bb.y);
      step<QF>(t, a, bb.z);
      step<QF>(t, a, bb.w);
      if constexpr (kCarryPasses)
        if ((i + 4) % 64 == 0) carry_pass(t);
    }
#pragma unroll 3
    for (int i = ns4; i < ns; i++) step<QF>(t, a, b[i]);
    normalize(r, t);
  }

  __device__ __forceinline__ void normalize(uint32_t (&r)[K], const uint64_t (&t)[K]) const {
    normalize_group<L, K, D>(r, t, g);
  }

  // Radix-2^28 layouts of more than 127 digits (4x37, 8x37) would overflow
  // the 64-bit lazy sums over a whole product (2*ND products of < 2^56 per
  // position), so they run carry passes within it: every position keeps its
  // low D bits and hands the rest to the next position (the previous lane's
  // top carry crosses by one DPP move; the group's top position has none,
  // since the running value stays below 3m < R).  The value is unchanged and
  // every sum drops below 2^D + 2^(64-D).  After that a position can take 85
  // steps of up to three products' worth (a squaring step adds a doubled
  // product and q*m_j: below 3*2^56), so a pass follows every 64 steps of a
  // product and every floor(85/K) lanes of a squaring.
  static constexpr bool kCarryPasses = D == 28 && ND > 127;
  static constexpr int kPassLanes = 85 / K > 0 ? 85 / K : 1;

  __device__ __forceinline__ void carry_pass(uint64_t (&t)[K]) const {
    const uint64_t below = g.from_prev(t[K - 1] >> D);
#pragma unroll
    for (int k = K - 1; k >= 1; k--) t[k] = (t[k] & MASK) + (t[k - 1] >> D);
    t[0] = (t[0] & MASK) + below;
  }

  // r = a^2 / R (mod m) for one-lane instances: the CIOS steps with the
  // multiplier digits taken from the lane's own registers and the symmetric
  // partial products folded, so step i adds a_i^2 at position i and 2 a_i a_j
  // for j > i.  A position's products all arrive no later than the step that
  // retires it (pair (i, j) lands at i + j >= i), so the quotient digits are
  // those of the plain product.  ND(ND+1)/2 + ND^2 multiply-adds instead of
  // 2 ND^2.  Bound per accumulator: (ND/2+1) doubled products < 2^(2D+1) plus
  // ND products < 2^(2D), the same 2^(2D)(2ND+2) as the plain product.
  template <bool QF = false>
  __device__ __forceinline__ void sqr(uint32_t (&r)[K], const uint32_t (&a)[K]) const {
    static_assert(L == 1, "symmetric squaring needs the whole instance in one lane");
    uint64_t t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = 0;
    // one instantiation per step, so every register index is a constant
    sqr_steps<QF>(t, a, std::make_integer_sequence<int, K>());
    normalize(r, t);
  }

  template <bool QF, int... I>
  __device__ __forceinline__ void sqr_steps(uint64_t (&t)[K], const uint32_t (&a)[K],
                                            std::integer_sequence<int, I...>) const {
    (sqr_step<QF, I>(t, a), ...);
  }

  template <bool QF, int I>
  __device__ __forceinline__ void sqr_step(uint64_t (&t)[K], const uint32_t (&a)[K]) const {
    const uint32_t ai = a[I], ai2 = a[I] << 1;
    uint64_t f = t[0];
    if constexpr (I == 0) f += (uint64_t)ai * ai;
    const uint32_t q = (QF ? (uint32_t)f : mul_lo_mad((uint32_t)f, np0)) & MASK;
    f += (uint64_t)q * m[0];
#pragma unroll
    for (int j = 1; j < K; j++) {
      uint64_t x = t[j] + (uint64_t)q * m[j];
      if (j == I) x += (uint64_t)ai * a[j];
      else if (j > I) x += (uint64_t)ai2 * a[j];
      t[j - 1] = x;
    }
    t[K - 1] = 0;
    t[0] += f >> D;
  }

  // Symmetric squaring on a lane group: r = a^2 / R (mod m), with a's digits
  // also staged in LDS at b (the step multipliers).  Step i = li*K + KI
  // multiplies digit i with the lanes' own digits; each pair x != y must
  // contribute 2 a_x a_y in total, and is split over the steps of x and y so
  // that every lane does the same number of multiply-adds per step:
  //  - pairs in different register slots (x mod K != y mod K): at the step of
  //    x when slot(x) beats slot(y) in a regular tournament on the K slots
  //    (slot s beats s+1 .. s+(K-1)/2 cyclically; for even K, s also beats
  //    s+K/2 when s < K/2), as 2 a_x a_y;
  //  - pairs in one slot on two lanes: at both steps, as a_x a_y each time,
  //    and the square a_x^2 at step x, so that slot multiplies by a_i on
  //    every lane.
  // A pair (x, y) lands at position x + y >= max(x, y), never below the
  // position its step retires, so every position holds the same total as in
  // the plain product when it is retired: quotients, carries and the result
  // are those of mul(r, a, b), bit for bit, with the same accumulator bound.
  // Per step K + floor(K/2) + 1 multiply-adds per lane instead of 2K.
  template <bool QF = false>
  __device__ __forceinline__ void sqr_group(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t *b) const {
    uint64_t t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = 0;
    uint32_t a2[K];
#pragma unroll
    for (int k = 0; k < K; k++) a2[k] = kOwnDoubled ? a[k] << 1 : 0u;
    // one lane's K steps are unrolled; the loop over lanes is kept (unrolling
    // it too only raises register pressure) and stops after the nl lanes
    // that hold digits
    const int nlr = (int)nl;
#pragma unroll 1
    for (int li = 0; li < nlr; li++) {
      // fresh copies behind an empty asm each iteration, so the compiler
      // cannot hoist the digits' 64-bit extensions out of the loop (that
      // doubles the registers they take and spills).  Groups of two to eight
      // lanes with up to 38 digits fit their registers without it under the
      // ILP scheduler (no spill at all for 2x37, whose launch is then
      // 1.2-1.8 % faster; 4x38 and 8x38 run 0.5 % faster); 2x55 keeps it,
      // and layouts of up to 10 digits per lane have registers to spare.
      constexpr bool kCopyBarrier = K > 10 && !((L == 2 || L == 4 || L == 8) && K <= 38);
      uint32_t ac[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        ac[k] = a[k];
        if constexpr (kCopyBarrier) asm("" : "+v"(ac[k]));
      }
      gsqr_steps<QF>(t, ac, a2, b + li * K, std::make_integer_sequence<int, K>());
      if constexpr (kCarryPasses)
        if ((li + 1) % kPassLanes == 0 && li + 1 < nlr) carry_pass(t);
    }
    normalize(r, t);
  }

  // Layouts of up to 10 digits per lane keep their own digits doubled in K
  // more registers, so a step multiplies 2a_j by b_i instead of doubling b_i
  // first: K shifts per squaring instead of one per step (a step of 16x3 is
  // 14 VALU instructions).  With more digits the step is long enough that
  // the shift matters less than the registers.
  static constexpr bool kOwnDoubled = K <= 10;

  static constexpr bool beats(int x, int y) {
    const int d = (y - x + K) % K;
    return 2 * d < K || (2 * d == K && x < K / 2);
  }

  template <bool QF, int... I>
  __device__ __forceinline__ void gsqr_steps(uint64_t (&t)[K], const uint32_t (&a)[K], const uint32_t (&a2)[K],
                                             const uint32_t *bl, std::integer_sequence<int, I...>) const {
    (gsqr_step<QF, I>(t, a, a2, bl), ...);
  }

  // 2 a_j b_i, from the doubled own digit or the doubled multiplier digit
  __device__ __forceinline__ uint64_t twice(const uint32_t (&a)[K], const uint32_t (&a2)[K], int j, uint32_t bi,
                                            uint32_t b2) const {
    return kOwnDoubled ? (uint64_t)a2[j] * bi : (uint64_t)a[j] * b2;
  }

  template <bool QF, int KI>
  __device__ __forceinline__ void gsqr_step(uint64_t (&t)[K], const uint32_t (&a)[K], const uint32_t (&a2)[K],
                                            const uint32_t *bl) const {
    const uint32_t bi = bl[KI];
    const uint32_t b2 = kOwnDoubled ? 0u : bi << 1;
    uint64_t f = t[0];
    if constexpr (KI == 0) f += (uint64_t)a[0] * bi;
    else if constexpr (beats(KI, 0)) f += twice(a, a2, 0, bi, b2);
    const uint32_t q = quotient<QF>(f);
    f += (uint64_t)q * m[0];
    const uint32_t up = g.next_digit((uint32_t)f, mask);
    gsqr_cols<KI>(t, a, a2, q, bi, b2, std::make_integer_sequence<int, K - 1>());
    t[K - 1] = up;
    t[0] += f >> D;
  }

  template <int KI, int... J>
  __device__ __forceinline__ void gsqr_cols(uint64_t (&t)[K], const uint32_t (&a)[K], const uint32_t (&a2)[K],
                                            uint32_t q, uint32_t bi, uint32_t b2,
                                            std::integer_sequence<int, J...>) const {
    (gsqr_col<KI, J + 1>(t, a, a2, q, bi, b2), ...);
  }

  template <int KI, int J>
  __device__ __forceinline__ void gsqr_col(uint64_t (&t)[K], const uint32_t (&a)[K], const uint32_t (&a2)[K],
                                           uint32_t q, uint32_t bi, uint32_t b2) const {
    uint64_t x = t[J] + (uint64_t)q * m[J];
    if constexpr (J == KI) x += (uint64_t)a[J] * bi;
    else if constexpr (beats(KI, J)) x += twice(a, a2, J, bi, b2);
    t[J - 1] = x;
  }
};


// An instance's ND normalised radix-2^D digits, staged in LDS at b, written
// as `len` interleaved 32-bit limbs; lane l of the group writes limbs l, l+L, ...
template <int L, int D>
__device__ __forceinline__ void store_limbs(const uint32_t *b, int ND, uint32_t *out, uint32_t stride, uint32_t inst,
                                            uint32_t len, int l) {
  for (uint32_t w = (uint32_t)l; w < len; w += L) {
    uint32_t bit = 32u * w;
    uint32_t d0 = bit / D, s0 = bit % D;
    uint32_t x0 = d0 < (uint32_t)ND ? b[d0] : 0u;
    uint32_t x1 = d0 + 1 < (uint32_t)ND ? b[d0 + 1] : 0u;
    uint32_t x2 = d0 + 2 < (uint32_t)ND ? b[d0 + 2] : 0u;
    uint64_t lo = ((uint64_t)x1 << D) | x0;
    uint32_t v = (uint32_t)(lo >> s0);
    uint32_t sh2 = 2u * D - s0;
    if (sh2 < 32) v |= x2 << sh2;
    out[(size_t)w * stride + inst] = v;
  }
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
template <int L, int K, int D, bool QF>
__device__ __forceinline__ void powm_tile_mode(const PowmArgs &args, uint32_t tile, uint32_t *bbuf) {
  using M = Mont<L, K, D>;
  constexpr int ND = M::ND;
  constexpr int NDP = (ND + 3) / 4 * 4;
  constexpr int G = 64 / L;   // instances per wave

  const int lane = threadIdx.x;
  const int grp = lane / L;
  M mont(lane);
  mont.nl = powm_active_lanes(L, K, D, args.mod.len);
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
    mont.m[k] = limb_digit_guarded<D>(args.mod.data, args.mod.stride, mi, args.mod.len, d);
    b[l * K + k] = limb_digit_guarded<D>(args.r2, args.r2_stride, mi, args.mod.len, d);
  }
  mont.np0 = neg_inverse<D>(ld_limb(args.mod, mi, 0));
  if constexpr (QF) {
    // m~ = m * np0 < 2^(32*len + D), which the active lanes hold
    uint64_t t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = (uint64_t)mont.m[k] * mont.np0;
    normalize_group<L, K, D>(mont.m, t, mont.g);
  }

  const uint32_t wbits = args.window_bits;
  const uint32_t T = 1u << wbits;
  uint32_t *tab = args.table + (size_t)local * T * ND;   // table sized for whole waves
#define tab_at(e, k) (tab + ((size_t)(e) * K + (k)) * L + l)

  // slot 0 = R mod m = 1 * (R^2 mod m) / R
#pragma unroll
  for (int k = 0; k < K; k++) a[k] = (l == 0 && k == 0) ? 1u : 0u;
  wave_sync();
  mont.template mul<QF>(a, a, b);
#pragma unroll
  for (int k = 0; k < K; k++) *tab_at(0, k) = a[k];
  // slot 1 = base * R mod m
#pragma unroll
  for (int k = 0; k < K; k++)
    a[k] = limb_digit_guarded<D>(args.base.data, args.base.stride, bi, args.base.len, (uint32_t)(l * K + k));
  mont.template mul<QF>(a, a, b);
  wave_sync();
  // slots 2..T-1: slot[e] = slot[e-1] * slot[1]
#pragma unroll
  for (int k = 0; k < K; k++) {
    *tab_at(1, k) = a[k];
    b[l * K + k] = a[k];
  }
  wave_sync();
  for (uint32_t e = 2; e < T; e++) {
    mont.template mul<QF>(a, a, b);
#pragma unroll
    for (int k = 0; k < K; k++) *tab_at(e, k) = a[k];
  }
  wave_sync();

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
      if constexpr (L == 1) {
        mont.template sqr<QF>(a, a);
      } else {
#pragma unroll
        for (int k = 0; k < K; k++) b[l * K + k] = a[k];
        wave_sync();
        mont.template sqr_group<QF>(a, a, b);
        wave_sync();
      }
    }
#pragma unroll
    for (int k = 0; k < K; k++) b[l * K + k] = *tab_at(idx, k);
    wave_sync();
    mont.template mul<QF>(a, a, b);
    wave_sync();
  }
#undef tab_at
  // leave Montgomery form: a * 1 / R, which is <= m.  Quotient-free, a is
  // below 2 m~ and congruent to the result mod m, and this one product
  // reduces by m itself: (a + Q m) / R < 2 m~ / R + m <= m + 1/2.
  if constexpr (QF) {
    // the load addresses are recomputed here (an opaque copy of the index),
    // not kept live from the prologue through the exponentiation
    uint32_t mr = mi, lr = (uint32_t)(l * K), sr = args.mod.stride;
    asm volatile("" : "+v"(mr), "+v"(lr), "+s"(sr));
#pragma unroll
    for (int k = 0; k < K; k++)
      mont.m[k] = limb_digit_guarded<D>(args.mod.data, sr, mr, args.mod.len, lr + (uint32_t)k);
  }
#pragma unroll
  for (int k = 0; k < K; k++) b[l * K + k] = (l == 0 && k == 0) ? 1u : 0u;
  wave_sync();
  uint32_t r[K];
  mont.mul(r, a, b);
  wave_sync();
  // canonical: r == m  ->  0
  bool ne = false;
#pragma unroll
  for (int k = 0; k < K; k++) ne |= (r[k] != mont.m[k]);
  ne = mont.g.any(ne);
  // an even modulus is outside Montgomery's domain: defined result 0
  const bool odd = (ld_limb(args.mod, mi, 0) & 1u) != 0;
#pragma unroll
  for (int k = 0; k < K; k++) b[l * K + k] = (ne && odd) ? r[k] : 0u;
  wave_sync();
  if (active) store_limbs<L, D>(b, ND, args.out.data, args.out.stride, out_index(args.out, gi), args.out.len, l);
}

// the persistent batch's tiles take the quotient of every step
template <int L, int K, int D>
__device__ __forceinline__ void powm_tile(const PowmArgs &args, uint32_t tile, uint32_t *bbuf) {
  powm_tile_mode<L, K, D, false>(args, tile, bbuf);
}

template <int L, int K, int D>
constexpr int lds_words() {
  return (64 / L) * ((L * K + 3) / 4 * 4);
}

// One wave per tile.  The quotient-free form (QF) is a kernel of its own, so
// that each form gets its own register allocation; the launcher takes it
// when the active lanes have D bits to spare (powm_quotient_free): one
// multiply-add fewer per step, the reduction running by m~ = m * np0, one
// digit longer than m, at the same R.
template <int L, int K, int D, bool QF = false>
__global__ void __launch_bounds__(64, 2) powm_kernel(PowmArgs args) {
  __shared__ __attribute__((aligned(16))) uint32_t bbuf[lds_words<L, K, D>()];
  powm_tile_mode<L, K, D, QF>(args, blockIdx.x, bbuf);
}

#ifdef XMP_POWM_LATENCY_UNIT
}  // namespace

// The few-digit latency layouts, compiled in a unit of their own
// (k_powm_latency.hip) with the default instruction scheduler, which issues
// their short steps 1-3 % faster than the ILP-oriented one used for this file.
template <int L, int K>
void launch_latency_kernel(dim3 grid, const PowmArgs &args, hipStream_t stream) {
  if (powm_quotient_free(L, K, 28, args.mod.len))
    hipLaunchKernelGGL((powm_kernel<L, K, 28, true>), grid, dim3(64), 0, stream, args);
  else
    hipLaunchKernelGGL((powm_kernel<L, K, 28>), grid, dim3(64), 0, stream, args);
}

hipError_t launch_powm_latency(int lanes, int digits, const PowmArgs &args, hipStream_t stream) {
  if (args.count == 0) return hipErrorInvalidValue;
  const dim3 grid(div_round_up(args.count, 64u / (uint32_t)lanes));
  if (lanes == 4 && digits == 10)
    launch_latency_kernel<4, 10>(grid, args, stream);
  else if (lanes == 8 && digits == 5)
    launch_latency_kernel<8, 5>(grid, args, stream);
  else if (lanes == 8 && digits == 10)
    launch_latency_kernel<8, 10>(grid, args, stream);
  else if (lanes == 16 && digits == 5)
    launch_latency_kernel<16, 5>(grid, args, stream);
  else if (lanes == 16 && digits == 3)
    launch_latency_kernel<16, 3>(grid, args, stream);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace xmp
#else

struct VariantEntry {
  PowmVariant v;
  void (*kernel)(PowmArgs);      // null: launched by launch_powm_latency
  // The quotient-free form, compiled for the layouts whose classes have the
  // D spare bits and where it measured faster: 4-7 % for 1024- and 2048-bit
  // moduli on 16x3, 16x5, 8x5 and 4x10, 1.5 % for 8192-bit ones on 8x37.
  // 4x37's form spilled 72 bytes per lane against 32 and ran no faster, so
  // 4096-bit moduli keep the quotient.
  void (*qf_kernel)(PowmArgs);
};

#define XMP_VARIANT(L, K, D) {{L, K, D}, powm_kernel<L, K, D>, nullptr}
#define XMP_VARIANT_QF(L, K, D) {{L, K, D}, powm_kernel<L, K, D>, powm_kernel<L, K, D, true>}
#define XMP_LATENCY_VARIANT(L, K, D) {{L, K, D}, nullptr, nullptr}
const VariantEntry kVariants[] = {
    // <=1024-bit: one lane per instance
    XMP_VARIANT(1, 12, 28),  // 336-bit R : up to 320-bit moduli
    XMP_VARIANT_QF(1, 20, 28),  // 560 : 512-bit
    XMP_VARIANT(1, 28, 28),  // 784 : 768-bit
    XMP_VARIANT(1, 37, 28),  // 1036 : 1024-bit
    XMP_VARIANT_QF(2, 19, 28),  // 1064 : 1024-bit on two lanes
    XMP_VARIANT_QF(2, 28, 28),  // 1568 : 1536-bit
    XMP_VARIANT(2, 37, 28),  // 2072 : 2048-bit
    XMP_VARIANT_QF(4, 19, 28),  // 2128 : 2048-bit on four lanes
    XMP_VARIANT_QF(4, 28, 28),  // 3136 : 3072-bit
    XMP_VARIANT(2, 55, 28),  // 3080 : 3072-bit on two lanes
    XMP_VARIANT(4, 38, 27),  // 4104 : 4096-bit
    XMP_VARIANT(8, 19, 27),  // 4104 : 4096-bit on eight lanes
    XMP_VARIANT_QF(8, 29, 27),  // 6264 : 6144-bit
    XMP_VARIANT(8, 38, 27),  // 8208 : 8192-bit
    XMP_VARIANT(16, 19, 27), // 8208 : 8192-bit on sixteen lanes (one DPP row)
    // small moduli, one lane per instance
    XMP_VARIANT(1, 5, 28),   // 140 : 128-bit
    XMP_VARIANT(1, 10, 28),  // 280 : 256-bit
    // above 8192 bits: half a wave / a whole wave per instance, radix 2^26
    XMP_VARIANT_QF(32, 20, 26), // 16640 : 16384-bit
    XMP_VARIANT_QF(64, 20, 26), // 33280 : 32768-bit
    // low-latency layouts (fewer digits per lane, so one wave finishes sooner)
    XMP_LATENCY_VARIANT(4, 10, 28),  // 1120 : 1024-bit on four lanes
    XMP_LATENCY_VARIANT(8, 5, 28),   // 1120 : 1024-bit on eight lanes
    XMP_LATENCY_VARIANT(8, 10, 28),  // 2240 : 2048-bit on eight lanes
    XMP_LATENCY_VARIANT(16, 5, 28),  // 2240 : 2048-bit on sixteen lanes
    XMP_LATENCY_VARIANT(16, 3, 28),  // 1344 : 1024-bit on sixteen lanes
    // radix 2^28 above 127 digits, with carry passes inside each product
    XMP_VARIANT(4, 37, 28),  // 4144 : 4096-bit
    XMP_VARIANT_QF(8, 37, 28),  // 8288 : 8192-bit
};
#undef XMP_VARIANT
#undef XMP_VARIANT_QF
#undef XMP_LATENCY_VARIANT
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

// ---- R^2 mod m for long moduli by Montgomery squarings ------------------
// With R = 2^rbits = 2^(t * 2^j) (t odd), Montgomery squaring maps
// x*R mod m to x^2*R mod m, so j squarings of Y = 2^t*R mod m give
// 2^rbits*R = R^2 mod m.  Y is one division of 2^(rbits+t), whose quotient
// has only about t + (rbits - precision) bits: for 8192-bit moduli 17 limbs
// instead of the 257 of a direct 2^(2 rbits) mod m.
struct R2Args {
  Operand mod;
  const uint32_t *y;   // 2^(rbits+t) mod m, mod.len limbs, interleaved by modulus
  uint32_t y_stride;
  uint32_t *out;       // the squared value (< 2m), out_len = mod.len + 1 limbs
  uint32_t out_stride, out_len;
  uint32_t count, squarings;
};

template <int L, int K, int D>
__global__ void __launch_bounds__(64, 2) r2_square_kernel(R2Args args) {
  using M = Mont<L, K, D>;
  constexpr int ND = M::ND;
  constexpr int NDP = (ND + 3) / 4 * 4;
  constexpr int G = 64 / L;
  __shared__ __attribute__((aligned(16))) uint32_t bbuf[lds_words<L, K, D>()];
  const int lane = threadIdx.x;
  const int grp = lane / L;
  M mont(lane);
  mont.nl = powm_active_lanes(L, K, D, args.mod.len);
  const int l = mont.g.l;
  const uint32_t local = blockIdx.x * G + grp;
  const bool active = local < args.count;
  const uint32_t mi = active ? local : args.count - 1;
  uint32_t *b = bbuf + grp * NDP;
  uint32_t a[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint32_t d = (uint32_t)(l * K + k);
    mont.m[k] = limb_digit_guarded<D>(args.mod.data, args.mod.stride, mi, args.mod.len, d);
    a[k] = limb_digit_guarded<D>(args.y, args.y_stride, mi, args.mod.len, d);
  }
  mont.np0 = neg_inverse<D>(ld_limb(args.mod, mi, 0));
  for (uint32_t s = 0; s < args.squarings; s++) {
#pragma unroll
    for (int k = 0; k < K; k++) b[l * K + k] = a[k];
    wave_sync();
    mont.sqr_group(a, a, b);
    wave_sync();
  }
#pragma unroll
  for (int k = 0; k < K; k++) b[l * K + k] = a[k];
  wave_sync();
  if (active) store_limbs<L, D>(b, ND, args.out, args.out_stride, mi, args.out_len, l);
}

// r2 = v - m if v >= m else v, one lane per modulus (v < 2m has len+1 limbs)
__global__ void __launch_bounds__(256) r2_canonical_kernel(uint32_t *r2, uint32_t r2_stride, const uint32_t *v,
                                                           uint32_t v_stride, const uint32_t *m, uint32_t m_len,
                                                           uint32_t m_stride, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t br = 0;
  for (uint32_t k = 0; k <= m_len; k++) {
    const uint32_t mk = k < m_len ? m[(size_t)k * m_stride + t] : 0u;
    (void)__builtin_subc(v[(size_t)k * v_stride + t], mk, br, &br);
  }
  const bool keep = br != 0;   // v < m
  br = 0;
  for (uint32_t k = 0; k < m_len; k++) {
    const uint32_t vk = v[(size_t)k * v_stride + t];
    const uint32_t d = __builtin_subc(vk, m[(size_t)k * m_stride + t], br, &br);
    r2[(size_t)k * r2_stride + t] = keep ? vk : d;
  }
}

struct R2Entry {
  PowmVariant v;
  void (*kernel)(R2Args);
};
#define XMP_R2(L, K, D) {{L, K, D}, r2_square_kernel<L, K, D>}
// the layouts whose moduli can exceed the register division (more than
// kMaxRegDivisorLimbs limbs)
const R2Entry kR2Variants[] = {
    XMP_R2(8, 29, 27), XMP_R2(8, 38, 27), XMP_R2(16, 19, 27), XMP_R2(32, 20, 26), XMP_R2(64, 20, 26),
    XMP_R2(8, 37, 28),
};
#undef XMP_R2

// ---- CRT recombination ---------------------------------------------------
// out = (mp*cp + mq*cq) mod n for each request (the server's kind-1 tail,
// reference samples/crt_rsa/rsa.cc:205-219) as one Montgomery reduction:
// with cp2 = cp*R mod n and cq2 = cq*R mod n resident (R = 2^(D*steps)),
// REDC(mp*cp2 + mq*cq2) = (mp*cp + mq*cq) mod n up to one subtraction of n,
// since the sum is below 2^(h+1) n for h-bit mp and mq and R >= 2^(h+1).
// Each step takes one digit of mp and of mq (staged in LDS) against the
// lane group's digits of cp2 and cq2, with one quotient digit for both.
template <int L, int K, int D>
__global__ void __launch_bounds__(64, 2) crt_combine_kernel(CrtArgs args) {
  using M = Mont<L, K, D>;
  constexpr int ND = M::ND;
  constexpr int NDP = (ND + 3) / 4 * 4;
  constexpr int G = 64 / L;
  __shared__ __attribute__((aligned(16))) uint32_t buf[2 * G * NDP];
  const int lane = threadIdx.x;
  const int grp = lane / L;
  M mont(lane);
  const int l = mont.g.l;
  const uint32_t local = blockIdx.x * G + grp;
  const bool active = local < args.count;
  const uint32_t i = active ? local : args.count - 1;   // clamp: inactive lanes recompute a valid request
  uint32_t *b = buf + grp * NDP;         // mp's digits, then the result's
  uint32_t *d = buf + (G + grp) * NDP;   // mq's digits, then n's
  const uint32_t ni = in_index(args.n, i), pi = in_index(args.cp2, i), qi = in_index(args.cq2, i);
  const uint32_t xi = in_index(args.mp, i), yi = in_index(args.mq, i);
  uint32_t a[K], c[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint32_t dk = (uint32_t)(l * K + k);
    mont.m[k] = limb_digit_guarded<D>(args.n.data, args.n.stride, ni, args.n.len, dk);
    a[k] = limb_digit_guarded<D>(args.cp2.data, args.cp2.stride, pi, args.cp2.len, dk);
    c[k] = limb_digit_guarded<D>(args.cq2.data, args.cq2.stride, qi, args.cq2.len, dk);
    b[dk] = limb_digit_guarded<D>(args.mp.data, args.mp.stride, xi, args.mp.len, dk);
    d[dk] = limb_digit_guarded<D>(args.mq.data, args.mq.stride, yi, args.mq.len, dk);
  }
  mont.np0 = neg_inverse<D>(ld_limb(args.n, ni, 0));
  wave_sync();
  uint64_t t[K];
#pragma unroll
  for (int k = 0; k < K; k++) t[k] = 0;
  for (uint32_t s = 0; s < args.steps; s++) mont.step2(t, a, b[s], c, d[s]);
  uint32_t r[K];
  mont.normalize(r, t);
  wave_sync();
#pragma unroll
  for (int k = 0; k < K; k++) {
    b[l * K + k] = r[k];
    d[l * K + k] = mont.m[k];
  }
  wave_sync();
  if (l == 0) {
    // r < 2n: r - n into d (digit-serial borrow), kept where it did not borrow out
    uint32_t br = 0;
    for (int j = 0; j < ND; j++) {
      const uint32_t x = b[j] + (M::MASK + 1u) - d[j] - br;
      d[j] = x & M::MASK;
      br = 1u - (x >> D);
    }
    if (!br)
      for (int j = 0; j < ND; j++) b[j] = d[j];
  }
  wave_sync();
  if (active) store_limbs<L, D>(b, ND, args.out.data, args.out.stride, out_index(args.out, i), args.out.len, l);
}

struct CrtEntry {
  PowmVariant v;
  void (*kernel)(CrtArgs);
};
#define XMP_CRT(L, K, D) {{L, K, D}, crt_combine_kernel<L, K, D>}
// moduli n below 2^1064, 2^2128 and 2^4104; 19 digits per lane keep the three
// operands and the accumulator in registers (37 spill)
const CrtEntry kCrtVariants[] = {XMP_CRT(2, 19, 28), XMP_CRT(4, 19, 28), XMP_CRT(8, 19, 27)};
#undef XMP_CRT

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
XMP_RUN(15, 1, 5, 28)
XMP_RUN(16, 1, 10, 28)
XMP_RUN(17, 32, 20, 26)
XMP_RUN(18, 64, 20, 26)
XMP_RUN(24, 4, 37, 28)
#undef XMP_RUN
// Variant 9 (2x55), the low-latency layouts 19..23 and 25 (8x37) are not part
// of the persistent kernel: as called functions 2x55 and 8x37 spill far more
// than their siblings (8x37 672 bytes per lane against 500 for 8x38, and a
// mixed batch with its 8192-bit class on 8x37 took 1.07 s against 1.01 s).
static_assert(kNumVariants == 26, "run_variant<> must list every batched variant in kVariants order");
constexpr bool batched_variant(int i) { return i >= 0 && (i < 19 || i == 24) && i != 9; }

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
      case 15: run_variant<15>(args, tile, dyn_lds); break;
      case 16: run_variant<16>(args, tile, dyn_lds); break;
      case 17: run_variant<17>(args, tile, dyn_lds); break;
      case 18: run_variant<18>(args, tile, dyn_lds); break;
      case 24: run_variant<24>(args, tile, dyn_lds); break;
      default: break;
    }
    wave_sync();
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

// lazy accumulators: 2*ND products of < 2^(2D) plus a carry must fit in 64 bits
// (without carry passes)
constexpr uint32_t max_digits(int radix) { return radix == 28 ? 127u : radix == 27 ? 511u : radix == 26 ? 2047u : 0u; }
static_assert(2.0 * 127 * 268435455.0 * 268435455.0 + 68719476736.0 < 18446744073709551616.0, "radix 28 bound");
static_assert(2.0 * 511 * 134217727.0 * 134217727.0 + 137438953472.0 < 18446744073709551616.0, "radix 27 bound");
static_assert(2.0 * 2047 * 67108863.0 * 67108863.0 + 274877906944.0 < 18446744073709551616.0, "radix 26 bound");

bool variant_fits(const PowmVariant &v, uint32_t precision) {
  const uint32_t nd = (uint32_t)(v.lanes * v.digits);
  // radix-2^28 groups of more than 127 digits run carry passes (Mont::kCarryPasses)
  const bool passes = v.radix == 28 && v.lanes > 1;
  return (nd <= max_digits(v.radix) || passes) && variant_rbits(v) >= precision + 2;
}

size_t powm_table_words(const PowmVariant &v, uint32_t count, uint32_t window_bits) {
  return (size_t)count * ((size_t)1 << window_bits) * (size_t)(v.lanes * v.digits);
}

bool powm_variant_batchable(int variant) { return batched_variant(variant); }

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
  if (!kVariants[variant].kernel) return launch_powm_latency(v.lanes, v.digits, args, stream);
  const uint32_t per_block = 64u / (uint32_t)v.lanes;
  void (*kernel)(PowmArgs) = kVariants[variant].kernel;
  if (kVariants[variant].qf_kernel && powm_quotient_free(v.lanes, v.digits, v.radix, args.mod.len))
    kernel = kVariants[variant].qf_kernel;
  hipLaunchKernelGGL(kernel, dim3(div_round_up(args.count, per_block)), dim3(64), 0, stream, args);
  return hipGetLastError();
}

hipError_t launch_r2_canonical(uint32_t *r2, uint32_t r2_stride, const uint32_t *v, uint32_t v_stride,
                               const uint32_t *m, uint32_t m_len, uint32_t m_stride, uint32_t count,
                               hipStream_t stream) {
  hipLaunchKernelGGL(r2_canonical_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, r2, r2_stride, v,
                     v_stride, m, m_len, m_stride, count);
  return hipGetLastError();
}

hipError_t launch_montgomery_r2(int variant, uint32_t *r2, uint32_t r2_stride, const uint32_t *m, uint32_t m_len,
                                uint32_t m_stride, uint32_t count, uint32_t *ws, hipStream_t stream) {
  if (variant < 0 || variant >= kNumVariants || count == 0) return hipErrorInvalidValue;
  const PowmVariant &v = kVariants[variant].v;
  const uint32_t rbits = variant_active_rbits(v, m_len);
  if (m_len <= kMaxRegDivisorLimbs) return launch_pow2_mod(r2, m_len, r2_stride, m, m_len, m_stride, count, 2 * rbits,
                                                           nullptr, stream);
  void (*kernel)(R2Args) = nullptr;
  for (const R2Entry &e : kR2Variants)
    if (e.v.lanes == v.lanes && e.v.digits == v.digits && e.v.radix == v.radix) kernel = e.kernel;
  if (!kernel || !ws) return hipErrorInvalidValue;
  const uint32_t j = (uint32_t)__builtin_ctz(rbits), t = rbits >> j;
  // Y = 2^(rbits+t) mod m into r2, then j squarings into the workspace (whose
  // division use is over), then the canonical value back into r2
  hipError_t e = launch_pow2_mod(r2, m_len, r2_stride, m, m_len, m_stride, count, rbits + t, ws, stream);
  if (e != hipSuccess) return e;
  const uint32_t ws_stride = round_up(count, 256);
  R2Args a;
  a.mod = Operand{m, m_len, m_stride, count, nullptr, 0};
  a.y = r2;
  a.y_stride = r2_stride;
  a.out = ws;
  a.out_stride = ws_stride;
  a.out_len = m_len + 1;
  a.count = count;
  a.squarings = j;
  hipLaunchKernelGGL(kernel, dim3(div_round_up(count, 64u / (uint32_t)v.lanes)), dim3(64), 0, stream, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_r2_canonical(r2, r2_stride, ws, ws_stride, m, m_len, m_stride, count, stream);
}

// The recombination layout for n-bit moduli and h-bit residues: the
// smallest whose digits hold n (R >= 4n is not needed here, only n < 2^(D*ND)
// with room for 2n) and whose lazy sums bear 3*steps products; steps = the
// digits of R >= 2^(h+1).  Returns false when none fits.
bool crt_combine_plan(uint32_t n_bits, uint32_t h_bits, int *which, uint32_t *steps) {
  for (int w = 0; w < (int)(sizeof(kCrtVariants) / sizeof(kCrtVariants[0])); w++) {
    const PowmVariant &v = kCrtVariants[w].v;
    const uint32_t nd = (uint32_t)(v.lanes * v.digits);
    const uint32_t ns = (h_bits + 1u + (uint32_t)v.radix - 1u) / (uint32_t)v.radix;
    const double bound = 3.0 * ns * std::ldexp(1.0, 2 * v.radix) + std::ldexp(1.0, 64 - v.radix);
    if ((uint32_t)v.radix * nd >= n_bits + 1 && ns <= nd && bound < std::ldexp(1.0, 64)) {
      *which = w;
      *steps = ns;
      return true;
    }
  }
  return false;
}

uint32_t crt_combine_radix(int which) { return (uint32_t)kCrtVariants[which].v.radix; }

hipError_t launch_crt_combine(int which, const CrtArgs &args, hipStream_t stream) {
  if (which < 0 || which >= (int)(sizeof(kCrtVariants) / sizeof(kCrtVariants[0])) || args.count == 0)
    return hipErrorInvalidValue;
  const uint32_t per_block = 64u / (uint32_t)kCrtVariants[which].v.lanes;
  hipLaunchKernelGGL(kCrtVariants[which].kernel, dim3(div_round_up(args.count, per_block)), dim3(64), 0, stream,
                     args);
  return hipGetLastError();
}

}  // namespace xmp
#endif  // XMP_POWM_LATENCY_UNIT
