// k_div.hip -- per-instance long division: q = floor(a/b), r = a mod b.
//
// Reference: src/div.cu:35-517 (dispatch), digitmp/div.h:108-423 (normalise,
// reciprocal of the top digit, quotient-digit estimate and refinement,
// multiply-subtract, add-back).  This is a different realisation of the same
// schoolbook method (Knuth vol.2, 4.3.1 algorithm D) on 32-bit limbs:
//
//  * the normalised divisor (NB limbs) and the NB+1-limb remainder window
//    live in registers; the dividend is streamed from HBM one limb per
//    quotient digit, so nothing but the operands and results touches memory
//    (the reference keeps its working set in a global scratch buffer,
//    div.cu:149-171);
//  * each quotient digit comes from a 2-by-1 division by the top divisor limb
//    using a precomputed reciprocal (Moller & Granlund, "Improved division by
//    invariant integers", IEEE TC 2011, algorithm 4) refined with the second
//    divisor limb, so at most one add-back can follow (rare, data-dependent);
//  * a divisor shorter than NB limbs is normalised by a whole-limb shift that
//    is folded into the addressing of the divisor load and of the remainder
//    store, so no register array is ever indexed at run time.
//
// Division by zero writes q = 0 and r = a (truncated), so a = q*b + r still
// holds; the reference leaves the outputs unwritten (digitmp/div.h:120-122).
//
// The same engine computes R^2 mod m for the Montgomery set-up, with the
// dividend 2^E generated on the fly.
#include "devmp.hpp"
#include "kernels.hpp"
#include "powm.hpp"

namespace xmp {

namespace {

// dividend sources ----------------------------------------------------------
// Limb k of the dividend, 0 outside [0, len): one unconditional load at a
// clamped, in-range limb and a select, with no branch around the load, so
// the loads the division loops issue ahead of their use stay in flight
// across loop iterations (a load under a branch made the compiler wait for
// every load at the loop's back edge).
// A limb fetched ahead of its use: `v & m` is the limb (m = 0 outside the
// dividend), applied only when the limb is used, so the load stays in flight.
struct RawLimb {
  uint32_t v, m;
  __device__ __forceinline__ uint32_t get() const { return v & m; }
};

struct OperandDividend {
  const uint32_t *data;
  uint32_t stride, inst, len;
  __device__ __forceinline__ RawLimb raw(int32_t k) const {
    const int32_t kc = k < 0 ? 0 : ((uint32_t)k < len ? k : (int32_t)len - 1);
    return RawLimb{data[(size_t)kc * stride + inst], (k >= 0 && (uint32_t)k < len) ? ~0u : 0u};
  }
  __device__ __forceinline__ uint32_t operator()(int32_t k) const { return raw(k).get(); }
  __device__ __forceinline__ uint32_t length() const { return len; }
};

struct PowerOfTwoDividend {   // 2^e
  uint32_t e;
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return (k >= 0 && (uint32_t)k == (e >> 5)) ? (1u << (e & 31)) : 0u;
  }
  __device__ __forceinline__ RawLimb raw(int32_t k) const { return RawLimb{(*this)(k), ~0u}; }
  __device__ __forceinline__ uint32_t length() const { return (e >> 5) + 1; }
};

// limb g of an interleaved operand, 0 outside [0, len): an unconditional
// load at a clamped, in-range limb and a select (see OperandDividend)
__device__ __forceinline__ uint32_t limb_or_zero(const uint32_t *data, uint32_t stride, uint32_t inst, uint32_t len,
                                                 int32_t g) {
  const int32_t gc = g < 0 ? 0 : ((uint32_t)g < len ? g : (int32_t)len - 1);
  const uint32_t v = data[(size_t)gc * stride + inst];
  return (g >= 0 && (uint32_t)g < len) ? v : 0u;
}

__device__ __forceinline__ uint32_t shl_limb(uint32_t hi, uint32_t lo, uint32_t s) {
  // low 32 bits of ((hi:lo) << s) >> 32, i.e. (hi << s) | (lo >> (32-s)); s in [0,31]
  return s ? funnel(hi, lo, 32u - s) : hi;
}

// floor((2^64-1)/d) - 2^32 for d with its top bit set
__device__ __forceinline__ uint32_t reciprocal(uint32_t d) {
  return (uint32_t)(~0ull / (uint64_t)d - (1ull << 32));
}

// (u1:u0) / d for u1 < d, d normalised; returns quotient, remainder in r
__device__ __forceinline__ uint32_t div_2by1(uint32_t u1, uint32_t u0, uint32_t d, uint32_t v, uint32_t &r) {
  uint64_t q = (uint64_t)v * u1 + (((uint64_t)u1 << 32) | u0);
  uint32_t q1 = (uint32_t)(q >> 32) + 1u;
  uint32_t q0 = (uint32_t)q;
  uint32_t rr = u0 - q1 * d;
  if (rr > q0) {
    q1--;
    rr += d;
  }
  if (rr >= d) {
    q1++;
    rr -= d;
  }
  r = rr;
  return q1;
}

struct DivOut {
  uint32_t *q;   // nullable
  uint32_t q_len, q_stride, q_inst;
  uint32_t *r;   // nullable
  uint32_t r_len, r_stride, r_inst;
};

// One quotient digit j of Knuth's algorithm D: estimate q from the window's
// top limbs (2-by-1 division by the top divisor limb with its reciprocal,
// refined twice with the second limb), subtract q*V from the window Win with
// two hardware borrow chains (low and high product words), add V back in the
// rare case the estimate was one too large, and write the difference shifted
// up one limb into Wout with the next dividend limb below it.
template <int NB>
__device__ __forceinline__ void digit_step(const uint32_t (&Win)[NB + 1], uint32_t (&Wout)[NB + 1], uint32_t next,
                                           const uint32_t (&v)[NB], uint32_t dinv, const DivOut &o, int32_t j) {
  uint32_t qhat, rhat;
  bool rhat_ok = true;
  if (Win[NB] >= v[NB - 1]) {
    qhat = 0xFFFFFFFFu;
    uint64_t rr = (uint64_t)Win[NB - 1] + v[NB - 1];
    rhat = (uint32_t)rr;
    rhat_ok = (rr >> 32) == 0;
  } else {
    qhat = div_2by1(Win[NB], Win[NB - 1], v[NB - 1], dinv, rhat);
  }
  if (NB > 1) {
#pragma unroll
    for (int it = 0; it < 2; it++) {
      if (rhat_ok && (uint64_t)qhat * v[NB - 2] > (((uint64_t)rhat << 32) | Win[NB > 1 ? NB - 2 : 0])) {
        qhat--;
        uint64_t rr = (uint64_t)rhat + v[NB - 1];
        rhat = (uint32_t)rr;
        rhat_ok = (rr >> 32) == 0;
      }
    }
  }
  // Win - qhat * V: limb k loses lo(qhat*v[k]) and hi(qhat*v[k-1])
  uint32_t b1 = 0, b2 = 0, ph = 0;
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const uint64_t p = (uint64_t)qhat * v[k];
    const uint32_t x = __builtin_subc(Win[k], (uint32_t)p, b1, &b1);
    Wout[k + 1] = __builtin_subc(x, ph, b2, &b2);
    ph = (uint32_t)(p >> 32);
  }
  {
    const uint32_t x = __builtin_subc(Win[NB], ph, b1, &b1);
    (void)__builtin_subc(x, 0u, b2, &b2);
  }
  if (b1 | b2) {   // estimate one too large: add V back (the carry out cancels the top limb)
    qhat--;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < NB; k++) Wout[k + 1] = __builtin_addc(Wout[k + 1], v[k], c, &c);
  }
  Wout[0] = next;
  if (o.q && (uint32_t)j < o.q_len) o.q[(size_t)j * o.q_stride + o.q_inst] = qhat;
}

// Divide the streamed dividend by the NB-limb divisor loaded from (bdata,
// bstride, binst, blen).  Returns false on division by zero.
template <int NB, class Dividend>
__device__ __forceinline__ void divide_reg(const Dividend &A, const uint32_t *bdata, uint32_t bstride, uint32_t binst,
                                           uint32_t blen, const DivOut &o) {
  uint32_t v[NB];
  int top = -1;
#pragma unroll
  for (int k = 0; k < NB; k++) {
    v[k] = limb_or_zero(bdata, bstride, binst, blen, k);
    if (v[k]) top = k;
  }
  const uint32_t la = A.length();
  if (top < 0) {   // division by zero
    if (o.q)
      for (uint32_t k = 0; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
    if (o.r)
      for (uint32_t k = 0; k < o.r_len; k++) o.r[(size_t)k * o.r_stride + o.r_inst] = A((int32_t)k);
    return;
  }
  // normalise: whole-limb shift sl (reload with offset) and bit shift s
  const int sl = NB - 1 - top;
  uint32_t vtop = 0;
#pragma unroll
  for (int k = 0; k < NB; k++)
    if (k == top) vtop = v[k];
  const uint32_t s = (uint32_t)__builtin_clz(vtop);
  if (sl) {
#pragma unroll
    for (int k = 0; k < NB; k++) {
      v[k] = limb_or_zero(bdata, bstride, binst, blen, k - sl);
    }
  }
#pragma unroll
  for (int k = NB - 1; k >= 0; k--) v[k] = shl_limb(v[k], k > 0 ? v[k - 1] : 0u, s);
  const uint32_t dinv = reciprocal(v[NB - 1]);

  // normalised dividend U = A << (32*sl + s): U limb k is the funnel of A
  // limbs k-sl and k-sl-1
  const int32_t lu = (int32_t)la + sl + 1;   // limbs of U
  const int32_t nq = lu - NB;                // quotient limbs computed
  if (nq <= 0) {
    // |a| < 2^(32*NB) * 2^-shift: quotient 0, remainder a
    if (o.q)
      for (uint32_t k = 0; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
    if (o.r)
      for (uint32_t k = 0; k < o.r_len; k++) o.r[(size_t)k * o.r_stride + o.r_inst] = A((int32_t)k);
    return;
  }
  // window = U[nq-1 .. nq-1+NB]; two windows in turn: each digit's
  // multiply-subtract writes the next (shifted) window, so no limb is moved
  uint32_t Wa[NB + 1], Wb[NB + 1];
#pragma unroll
  for (int k = 0; k <= NB; k++) Wb[k] = A(nq - 2 - sl + k);   // raw limbs, every load issued before any use
  {
    const uint32_t raw_top = A(nq - 1 - sl + NB);
#pragma unroll
    for (int k = 0; k <= NB; k++) Wa[k] = shl_limb(k < NB ? Wb[k + 1] : raw_top, Wb[k], s);
  }
  // The dividend limbs that enter the window: digit j takes U(j-1), the
  // funnel of raw limbs j-1-sl and j-2-sl.  Raw limbs are loaded six ahead
  // and only shifted when used, so their loads stay in flight for two
  // iterations of the loop.
  RawLimb r0 = A.raw(nq - 2 - sl), r1 = A.raw(nq - 3 - sl), r2 = A.raw(nq - 4 - sl), r3 = A.raw(nq - 5 - sl),
          r4 = A.raw(nq - 6 - sl), r5 = A.raw(nq - 7 - sl);
  int32_t j = nq - 1;
  for (; j >= 1; j -= 2) {
    digit_step<NB>(Wa, Wb, shl_limb(r0.get(), r1.get(), s), v, dinv, o, j);
    digit_step<NB>(Wb, Wa, shl_limb(r1.get(), r2.get(), s), v, dinv, o, j - 1);
    r0 = r2;
    r1 = r3;
    r2 = r4;
    r3 = r5;
    r4 = A.raw(j - 7 - sl);
    r5 = A.raw(j - 8 - sl);
  }
  if (j == 0) {
    digit_step<NB>(Wa, Wb, 0u, v, dinv, o, 0);
#pragma unroll
    for (int k = 0; k <= NB; k++) Wa[k] = Wb[k];
  }
  // the remainder is Wa[1 .. NB] (the last step shifted it up one limb)
  uint32_t W[NB + 1];
#pragma unroll
  for (int k = 0; k < NB; k++) W[k] = Wa[k + 1];
  W[NB] = 0;
  if (o.q)
    for (uint32_t k = (uint32_t)nq; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
  if (o.r) {
    // remainder = W[0..NB-1] >> (32*sl + s)
#pragma unroll
    for (int k = 0; k < NB; k++) {
      uint32_t hi = (k + 1 < NB) ? W[k + 1] : 0u;
      uint32_t val = s ? funnel(hi, W[k], s) : W[k];
      int32_t dst = k - sl;
      if (dst >= 0 && (uint32_t)dst < o.r_len) o.r[(size_t)dst * o.r_stride + o.r_inst] = val;
    }
    for (uint32_t k = (uint32_t)(NB - sl > 0 ? NB - sl : 0); k < o.r_len; k++)
      o.r[(size_t)k * o.r_stride + o.r_inst] = 0u;
  }
}

// ---- two lanes per division ----------------------------------------------
// Knuth D with the NB = 2H limb divisor and the remainder window split over
// a pair of lanes (lane h of the pair holds limbs h*H .. h*H+H-1, lane 1 also
// the window's top limb NB).  A batch too small to give every SIMD two waves
// at one lane per division (65,536 divisions fill only one wave per SIMD of
// the 1,024) gets twice the waves this way, and each lane issues about half
// the instructions of a quotient digit.  Per digit:
//  - lane 1 holds the window's top limbs and the divisor's top limbs, so it
//    estimates q (as the one-lane kernel does) and both lanes take its value;
//  - each lane subtracts q*V over its own limbs with two borrow chains; lane 1
//    starts its high-word chain from hi(q * v[H-1]) of lane 0 (kept since the
//    set-up), and then subtracts lane 0's outgoing borrows (0..2) with a
//    ripple that only continues past its first limb in the rare case that
//    limb underflows (a uniform branch on a wave ballot);
//  - the rare add-back adds V on both lanes and ripples lane 0's carry into
//    lane 1 the same way;
//  - the window moves up one limb: lane 1's first limb takes lane 0's last.
// Moves inside the pair are DPP quad_perm moves: [1,1,3,3] = 0xF5 gives both
// lanes lane 1's value, [0,0,2,2] = 0xA0 lane 0's.  A move reads 0 from a
// lane that is switched off, so every move runs on both lanes of the pair
// (never under a condition on h) and its result is selected afterwards.
__device__ __forceinline__ uint32_t pair_hi(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xF5, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t pair_lo(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xA0, 0xF, 0xF, true);
}

template <int H>
struct PairDiv {
  static constexpr int NB = 2 * H;
  uint32_t v[H];         // this lane's normalised divisor limbs
  uint32_t vt, vt2, dinv;   // the divisor's two top limbs and the top limb's reciprocal (meaningful on lane 1)
  uint32_t vbelow;       // lane 1: lane 0's top limb v[H-1]; lane 0: 0
  uint32_t h;            // 0 or 1

  // subtract c (< 2^32) from the limbs x[1..H] and then `top`: returns
  // whether the subtraction borrowed out of the top.  The ripple past x[1]
  // is rare, so it is a uniform branch on the wave's borrows.
  __device__ __forceinline__ static bool sub_ripple(uint32_t (&x)[H + 1], uint32_t c, uint32_t &top) {
    uint32_t br;
    x[1] = __builtin_subc(x[1], c, 0u, &br);
    if (__builtin_amdgcn_ballot_w64(br != 0) == 0) return false;
#pragma unroll
    for (int k = 2; k <= H; k++) x[k] = __builtin_subc(x[k], 0u, br, &br);
    top = __builtin_subc(top, 0u, br, &br);
    return br != 0;
  }

  // One quotient digit: Win (this lane's window limbs, lane 1's Win[H] the
  // window top) minus q*V into Wout[1..H], Wout[0] = the next limb (lane 0:
  // the dividend's next limb, lane 1: lane 0's Wout[H]).  Returns q.
  __device__ __forceinline__ uint32_t step(const uint32_t (&Win)[H + 1], uint32_t (&Wout)[H + 1],
                                           uint32_t next) const {
    // estimate from the window's three top limbs (lane 1's), as the one-lane kernel
    uint32_t qhat, rhat;
    bool rhat_ok = true;
    if (Win[H] >= vt) {
      qhat = 0xFFFFFFFFu;
      const uint64_t rr = (uint64_t)Win[H - 1] + vt;
      rhat = (uint32_t)rr;
      rhat_ok = (rr >> 32) == 0;
    } else {
      qhat = div_2by1(Win[H], Win[H - 1], vt, dinv, rhat);
    }
#pragma unroll
    for (int it = 0; it < 2; it++) {
      if (rhat_ok && (uint64_t)qhat * vt2 > (((uint64_t)rhat << 32) | Win[H - 2])) {
        qhat--;
        const uint64_t rr = (uint64_t)rhat + vt;
        rhat = (uint32_t)rr;
        rhat_ok = (rr >> 32) == 0;
      }
    }
    uint32_t q = pair_hi(qhat);
    // Win - q*V on this lane's limbs: limb k loses lo(q*v[k]) and hi(q*v[k-1])
    uint32_t b1 = 0, b2 = 0, ph = (uint32_t)(((uint64_t)q * vbelow) >> 32);
#pragma unroll
    for (int k = 0; k < H; k++) {
      const uint64_t p = (uint64_t)q * v[k];
      const uint32_t x = __builtin_subc(Win[k], (uint32_t)p, b1, &b1);
      Wout[k + 1] = __builtin_subc(x, ph, b2, &b2);
      ph = (uint32_t)(p >> 32);
    }
    const uint32_t out_borrow = b1 + b2;   // lane 0's borrows into lane 1's first limb
    uint32_t top;
    {
      const uint32_t x = __builtin_subc(Win[H], ph, b1, &b1);
      top = __builtin_subc(x, 0u, b2, &b2);
    }
    bool neg = (b1 | b2) != 0;
    const uint32_t lo_borrow = pair_lo(out_borrow);   // every move unconditional: see pair_hi
    const uint32_t c = h ? lo_borrow : 0u;
    if (sub_ripple(Wout, c, top)) neg = true;
    neg = pair_hi(neg ? 1u : 0u) != 0;   // lane 1's sign decides for the pair
    if (__builtin_amdgcn_ballot_w64(neg) != 0) {
      // the estimate was one too large (rare): add V back.  Both lanes of a
      // pair take the same branch (neg is lane 1's), so the pair's move reads
      // an active lane.
      if (neg) {
        q--;
        uint32_t cy = 0;
#pragma unroll
        for (int k = 0; k < H; k++) Wout[k + 1] = __builtin_addc(Wout[k + 1], v[k], cy, &cy);
        const uint32_t lo_carry = pair_lo(cy);
        uint32_t cc = h ? lo_carry : 0u;   // lane 0's carry into lane 1's limbs
#pragma unroll
        for (int k = 0; k < H; k++) Wout[k + 1] = __builtin_addc(Wout[k + 1], 0u, cc, &cc);
      }
    }
    const uint32_t from0 = pair_lo(Wout[H]);
    Wout[0] = h ? from0 : next;
    return q;
  }
};

template <int H, class Dividend>
__device__ __forceinline__ void divide_pair(const Dividend &A, const uint32_t *bdata, uint32_t bstride, uint32_t binst,
                                            uint32_t blen, const DivOut &o, uint32_t h) {
  constexpr int NB = 2 * H;
  PairDiv<H> d;
  d.h = h;
  // the pair's top non-zero divisor limb
  int top = -1;
#pragma unroll
  for (int k = 0; k < H; k++) {
    const uint32_t g = h * H + (uint32_t)k;
    d.v[k] = limb_or_zero(bdata, bstride, binst, blen, (int32_t)g);
    if (d.v[k]) top = (int)g;
  }
  top = max(top, (int)pair_hi((uint32_t)top));   // lane 1's top when it has one, else lane 0's
  top = max(top, (int)pair_lo((uint32_t)top));
  const uint32_t la = A.length();
  if (top < 0 || (int32_t)la + (NB - 1 - top) + 1 - NB <= 0) {
    // division by zero (q = 0, r = a) or a dividend shorter than the
    // normalised divisor (q = 0, r = a): lane 0 writes everything
    if (h == 0) {
      if (o.q)
        for (uint32_t k = 0; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
      if (o.r)
        for (uint32_t k = 0; k < o.r_len; k++) o.r[(size_t)k * o.r_stride + o.r_inst] = A((int32_t)k);
    }
    return;
  }
  uint32_t vtop = 0;
#pragma unroll
  for (int k = 0; k < H; k++)
    if ((int)(h * H + (uint32_t)k) == top) vtop = d.v[k];
  vtop |= pair_hi(vtop) | pair_lo(vtop);
  const uint32_t s = (uint32_t)__builtin_clz(vtop);
  const int sl = NB - 1 - top;
  // normalised divisor: limb g is (b << (32*sl + s)) limb g, read from memory
  {
    // raw limbs h*H-sl-1 .. h*H-sl+H-1 first (every load issued before any use)
    uint32_t raw[H + 1];
#pragma unroll
    for (int k = 0; k <= H; k++) raw[k] = limb_or_zero(bdata, bstride, binst, blen, (int32_t)(h * H) - sl - 1 + k);
#pragma unroll
    for (int k = 0; k < H; k++) d.v[k] = shl_limb(raw[k + 1], raw[k], s);
  }
  d.vt = pair_hi(d.v[H - 1]);
  d.vt2 = pair_hi(d.v[H - 2]);
  d.dinv = reciprocal(d.vt);
  const uint32_t lo_top = pair_lo(d.v[H - 1]);
  d.vbelow = h ? lo_top : 0u;

  const int32_t nq = (int32_t)la + sl + 1 - NB;   // quotient limbs computed (>= 1)
  uint32_t Wa[H + 1], Wb[H + 1];
  // window = U[nq-1 .. nq-1+NB]: lane h's limbs from h*H (lane 0's Wa[H] duplicates lane 1's Wa[0])
#pragma unroll
  for (int k = 0; k <= H; k++) Wb[k] = A(nq - 2 - sl + (int32_t)(h * H) + k);   // raw limbs first
  {
    const uint32_t raw_top = A(nq - 1 - sl + (int32_t)(h * H) + H);
#pragma unroll
    for (int k = 0; k <= H; k++) Wa[k] = shl_limb(k < H ? Wb[k + 1] : raw_top, Wb[k], s);
  }
  const bool store_q = o.q && h == 1;
  // the dividend limbs that enter the window (lane 0's), raw limbs loaded six
  // ahead and shifted when used (as in the one-lane kernel)
  RawLimb r0 = A.raw(nq - 2 - sl), r1 = A.raw(nq - 3 - sl), r2 = A.raw(nq - 4 - sl), r3 = A.raw(nq - 5 - sl),
          r4 = A.raw(nq - 6 - sl), r5 = A.raw(nq - 7 - sl);
  int32_t j = nq - 1;
  for (; j >= 1; j -= 2) {
    uint32_t q = d.step(Wa, Wb, shl_limb(r0.get(), r1.get(), s));
    if (store_q && (uint32_t)j < o.q_len) o.q[(size_t)j * o.q_stride + o.q_inst] = q;
    q = d.step(Wb, Wa, shl_limb(r1.get(), r2.get(), s));
    if (store_q && (uint32_t)(j - 1) < o.q_len) o.q[(size_t)(j - 1) * o.q_stride + o.q_inst] = q;
    r0 = r2;
    r1 = r3;
    r2 = r4;
    r3 = r5;
    r4 = A.raw(j - 7 - sl);
    r5 = A.raw(j - 8 - sl);
  }
  if (j == 0) {
    const uint32_t q = d.step(Wa, Wb, 0u);
    if (store_q && 0u < o.q_len) o.q[o.q_inst] = q;
#pragma unroll
    for (int k = 0; k <= H; k++) Wa[k] = Wb[k];
  }
  if (o.q && h == 0)
    for (uint32_t k = (uint32_t)nq; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
  if (o.r) {
    // remainder limbs h*H + k are Wa[k + 1] (the last step moved them up one
    // limb); unnormalise by s bits (lane 0's last limb takes lane 1's first)
    // and sl limbs
    const uint32_t hi_first = pair_hi(Wa[1]);
    const uint32_t next_hi = h ? 0u : hi_first;
#pragma unroll
    for (int k = 0; k < H; k++) {
      const uint32_t hi = k + 1 < H ? Wa[k + 2] : next_hi;
      const uint32_t val = s ? funnel(hi, Wa[k + 1], s) : Wa[k + 1];
      const int32_t dst = (int32_t)(h * H) + k - sl;
      if (dst >= 0 && (uint32_t)dst < o.r_len) o.r[(size_t)dst * o.r_stride + o.r_inst] = val;
    }
    if (h == 0)
      for (uint32_t k = (uint32_t)(NB - sl); k < o.r_len; k++) o.r[(size_t)k * o.r_stride + o.r_inst] = 0u;
  }
}

template <int H>
__global__ void __launch_bounds__(256) divmod_pair_kernel(OutOperand q, OutOperand r, Operand a, Operand b,
                                                          uint32_t count) {
  // pairs of lanes; a pair past `count` takes division count-1 and writes the same values
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t h = gt & 1u;
  const uint32_t pt = gt >> 1;
  const uint32_t t = pt < count ? pt : count - 1;
  const uint32_t ai = in_index(a, t), bi = in_index(b, t);
  DivOut o;
  o.q = q.data;
  o.q_len = q.len;
  o.q_stride = q.stride;
  o.q_inst = q.data ? out_index(q, t) : 0;
  o.r = r.data;
  o.r_len = r.len;
  o.r_stride = r.stride;
  o.r_inst = r.data ? out_index(r, t) : 0;
  OperandDividend A{a.data, a.stride, ai, a.len};
  divide_pair<H>(A, b.data, b.stride, bi, b.len, o, h);
}

template <int NB>
__global__ void __launch_bounds__(256) divmod_reg_kernel(OutOperand q, OutOperand r, Operand a, Operand b,
                                                         uint32_t count) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t ai = in_index(a, t), bi = in_index(b, t);
  DivOut o;
  o.q = q.data;
  o.q_len = q.len;
  o.q_stride = q.stride;
  o.q_inst = q.data ? out_index(q, t) : 0;
  o.r = r.data;
  o.r_len = r.len;
  o.r_stride = r.stride;
  o.r_inst = r.data ? out_index(r, t) : 0;
  OperandDividend A{a.data, a.stride, ai, a.len};
  divide_reg<NB>(A, b.data, b.stride, bi, b.len, o);
}

// R^2 mod m (R = 2^e/2) for every modulus, written interleaved to out
template <int NB>
__global__ void __launch_bounds__(256) pow2_mod_reg_kernel(uint32_t *out, uint32_t out_len, uint32_t out_stride,
                                                           const uint32_t *m, uint32_t m_len, uint32_t m_stride,
                                                           uint32_t count, uint32_t e) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  DivOut o{nullptr, 0, 0, 0, out, out_len, out_stride, t};
  PowerOfTwoDividend A{e};
  divide_reg<NB>(A, m, m_stride, t, m_len, o);
}

// ---- memory-resident fallback for divisors longer than the register kernels
// hold.  Per lane scratch: normalised divisor (nb limbs) and remainder window
// (nb+1 limbs), interleaved by lane so accesses stay coalesced.
template <class Dividend>
__device__ void divide_mem(const Dividend &A, const uint32_t *bdata, uint32_t bstride, uint32_t binst, uint32_t blen,
                           const DivOut &o, uint32_t *scr, uint32_t scr_stride) {
  const uint32_t NB = blen;
  uint32_t *v = scr;                                  // NB limbs
  uint32_t *W = scr + (size_t)NB * scr_stride;        // NB+1 limbs
#define V(k) v[(size_t)(k) * scr_stride]
#define WW(k) W[(size_t)(k) * scr_stride]
  int top = -1;
  for (uint32_t k = 0; k < NB; k++)
    if (bdata[(size_t)k * bstride + binst]) top = (int)k;
  const uint32_t la = A.length();
  if (top < 0) {
    if (o.q)
      for (uint32_t k = 0; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
    if (o.r)
      for (uint32_t k = 0; k < o.r_len; k++) o.r[(size_t)k * o.r_stride + o.r_inst] = A((int32_t)k);
    return;
  }
  const int sl = (int)NB - 1 - top;
  const uint32_t s = (uint32_t)__builtin_clz(bdata[(size_t)top * bstride + binst]);
  for (int k = (int)NB - 1; k >= 0; k--) {
    int src = k - sl;
    uint32_t hi = (src >= 0) ? bdata[(size_t)src * bstride + binst] : 0u;
    uint32_t lo = (src - 1 >= 0) ? bdata[(size_t)(src - 1) * bstride + binst] : 0u;
    V(k) = shl_limb(hi, lo, s);
  }
  const uint32_t vt = V(NB - 1), vt2 = NB > 1 ? V(NB - 2) : 0u;
  const uint32_t dinv = reciprocal(vt);
  auto U = [&](int32_t k) -> uint32_t {
    int32_t x = k - sl;
    return shl_limb(A(x), A(x - 1), s);
  };
  const int32_t lu = (int32_t)la + sl + 1;
  const int32_t nq = lu - (int32_t)NB;
  if (nq <= 0) {
    if (o.q)
      for (uint32_t k = 0; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
    if (o.r)
      for (uint32_t k = 0; k < o.r_len; k++) o.r[(size_t)k * o.r_stride + o.r_inst] = A((int32_t)k);
    return;
  }
  for (uint32_t k = 0; k <= NB; k++) WW(k) = U(nq - 1 + (int32_t)k);
  for (int32_t j = nq - 1; j >= 0; j--) {
    uint32_t wt = WW(NB), wt1 = WW(NB - 1), wt2 = NB > 1 ? WW(NB - 2) : 0u;
    uint32_t qhat, rhat;
    bool rhat_ok = true;
    if (wt >= vt) {
      qhat = 0xFFFFFFFFu;
      uint64_t rr = (uint64_t)wt1 + vt;
      rhat = (uint32_t)rr;
      rhat_ok = (rr >> 32) == 0;
    } else {
      qhat = div_2by1(wt, wt1, vt, dinv, rhat);
    }
    if (NB > 1) {
      for (int it = 0; it < 2; it++) {
        if (rhat_ok && (uint64_t)qhat * vt2 > (((uint64_t)rhat << 32) | wt2)) {
          qhat--;
          uint64_t rr = (uint64_t)rhat + vt;
          rhat = (uint32_t)rr;
          rhat_ok = (rr >> 32) == 0;
        }
      }
    }
    uint32_t carry = 0, borrow = 0;
    for (uint32_t k = 0; k < NB; k++) {
      uint64_t p = (uint64_t)qhat * V(k) + carry;
      carry = (uint32_t)(p >> 32);
      WW(k) = subb(WW(k), (uint32_t)p, borrow, borrow);
    }
    WW(NB) = subb(WW(NB), carry, borrow, borrow);
    if (borrow) {
      qhat--;
      uint32_t c = 0;
      for (uint32_t k = 0; k < NB; k++) WW(k) = addc(WW(k), V(k), c, c);
      WW(NB) = WW(NB) + c;
    }
    if (o.q && (uint32_t)j < o.q_len) o.q[(size_t)j * o.q_stride + o.q_inst] = qhat;
    if (j > 0) {
      for (uint32_t k = NB; k > 0; k--) WW(k) = WW(k - 1);
      WW(0) = U(j - 1);
    }
  }
  if (o.q)
    for (uint32_t k = (uint32_t)nq; k < o.q_len; k++) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
  if (o.r) {
    for (uint32_t k = 0; k < NB; k++) {
      uint32_t hi = (k + 1 < NB) ? WW(k + 1) : 0u;
      uint32_t lo = WW(k);
      uint32_t val = s ? funnel(hi, lo, s) : lo;
      int32_t dst = (int32_t)k - sl;
      if (dst >= 0 && (uint32_t)dst < o.r_len) o.r[(size_t)dst * o.r_stride + o.r_inst] = val;
    }
    for (uint32_t k = (uint32_t)((int)NB - sl > 0 ? (int)NB - sl : 0); k < o.r_len; k++)
      o.r[(size_t)k * o.r_stride + o.r_inst] = 0u;
  }
#undef V
#undef WW
}

__global__ void __launch_bounds__(256) divmod_mem_kernel(OutOperand q, OutOperand r, Operand a, Operand b,
                                                         uint32_t count, uint32_t *scratch, uint32_t scr_stride) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t ai = in_index(a, t), bi = in_index(b, t);
  DivOut o;
  o.q = q.data;
  o.q_len = q.len;
  o.q_stride = q.stride;
  o.q_inst = q.data ? out_index(q, t) : 0;
  o.r = r.data;
  o.r_len = r.len;
  o.r_stride = r.stride;
  o.r_inst = r.data ? out_index(r, t) : 0;
  OperandDividend A{a.data, a.stride, ai, a.len};
  divide_mem(A, b.data, b.stride, bi, b.len, o, scratch + t, scr_stride);
}

// ---- one wave per division: divisors beyond the register kernels ---------
// Knuth D as divide_mem (the normalised divisor v, the normalised dividend U
// divided in place, the window U[j .. j+NB] for quotient digit j), with v and
// U in LDS and each digit's multiply-subtract split over the wave's 64 lanes:
// lane l takes the C = ceil((NB+1)/64) window limbs from l*C, with the low and
// high product words on two borrow chains as in the pair kernel (the high
// word entering its first limb, hi(q*v[l*C-1]), it computes itself), and
// hands its outgoing borrows (0..2) to lane l+1, which subtracts them from
// its first limb.  A borrow crosses a whole chunk only when its limbs are all
// zero, so the rounds of hand-over end after one almost always (a wave
// ballot decides).  Every lane computes the same estimate from the window's
// top limbs (broadcast LDS reads).  With the borrows in, the window's top
// limb is 0, or all ones when the estimate was one too large (the difference
// is then in (-V, 0)): V is added back the same way.  One wave per division
// fills the GPU where the workspace kernel's one lane per division gave
// 16,384 divisions 256 waves.  The LDS operations of one wave complete in
// program order, so a wavefront fence between the steps keeps them ordered.
constexpr uint32_t kWaveDivMaxWords = 16384;   // v and U in 64 KiB of LDS

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// hand each lane's outgoing borrow (or carry, ADD) to the next lane, a round
// at a time, until none is left below the window's top limb (lane topl, whose
// outgoing borrow or carry leaves the window)
template <bool ADD>
__device__ __forceinline__ void wave_ripple(uint32_t *W, uint32_t k0, uint32_t k1, uint32_t out, uint32_t l,
                                            uint32_t topl) {
  if (l >= topl) out = 0;
  while (__builtin_amdgcn_ballot_w64(out != 0) != 0) {
    uint32_t cin = (uint32_t)__shfl_up((int)out, 1, 64);
    if (l == 0) cin = 0;
    for (uint32_t k = k0; k < k1 && cin != 0; k++) {
      if (ADD) {
        const uint64_t x = (uint64_t)W[k] + cin;
        W[k] = (uint32_t)x;
        cin = (uint32_t)(x >> 32);
      } else {
        const int64_t x = (int64_t)W[k] - (int64_t)cin;
        W[k] = (uint32_t)x;
        cin = x < 0 ? 1u : 0u;
      }
    }
    out = l >= topl ? 0u : cin;
  }
}

// LDS words divide_wave needs: v, then U (at most la + NB limbs)
__host__ __device__ inline uint32_t wave_div_words(uint32_t la, uint32_t nb) { return 2 * nb + la; }

template <class Dividend>
__device__ void divide_wave(const Dividend &A, const uint32_t *bdata, uint32_t bstride, uint32_t binst, uint32_t NB,
                            const DivOut &o, uint32_t *lds) {
  const uint32_t l = threadIdx.x & 63u;
  int top = -1;
  for (uint32_t k = l; k < NB; k += 64)
    if (bdata[(size_t)k * bstride + binst]) top = (int)k;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) top = max(top, __shfl_xor(top, d, 64));
  const uint32_t la = A.length();
  const int sl = top < 0 ? 0 : (int)NB - 1 - top;
  const int32_t lu = (int32_t)la + sl + 1;
  const int32_t nq = lu - (int32_t)NB;
  if (top < 0 || nq <= 0) {
    // division by zero, or a dividend shorter than the normalised divisor: q = 0, r = a
    if (o.q)
      for (uint32_t k = l; k < o.q_len; k += 64) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
    if (o.r)
      for (uint32_t k = l; k < o.r_len; k += 64) o.r[(size_t)k * o.r_stride + o.r_inst] = A((int32_t)k);
    return;
  }
  const uint32_t s = (uint32_t)__builtin_clz(bdata[(size_t)top * bstride + binst]);
  uint32_t *v = lds;        // NB limbs
  uint32_t *U = lds + NB;   // lu <= la + NB limbs
  for (uint32_t k = l; k < NB; k += 64) {
    const int src = (int)k - sl;
    const uint32_t hi = src >= 0 ? bdata[(size_t)src * bstride + binst] : 0u;
    const uint32_t lo = src >= 1 ? bdata[(size_t)(src - 1) * bstride + binst] : 0u;
    v[k] = shl_limb(hi, lo, s);
  }
  for (int32_t k = (int32_t)l; k < lu; k += 64) U[k] = shl_limb(A(k - sl), A(k - sl - 1), s);
  wave_fence();
  const uint32_t vt = v[NB - 1], vt2 = v[NB - 2];
  const uint32_t dinv = reciprocal(vt);
  const uint32_t C = (NB + 1 + 63) / 64;
  const uint32_t k0 = min(l * C, NB + 1), k1 = min(k0 + C, NB + 1);
  const uint32_t topl = NB / C;   // the lane holding the window's top limb
  for (int32_t j = nq - 1; j >= 0; j--) {
    uint32_t *W = U + j;
    const uint32_t wt = W[NB], wt1 = W[NB - 1], wt2 = W[NB - 2];
    uint32_t qhat, rhat;
    bool rhat_ok = true;
    if (wt >= vt) {
      qhat = 0xFFFFFFFFu;
      const uint64_t rr = (uint64_t)wt1 + vt;
      rhat = (uint32_t)rr;
      rhat_ok = (rr >> 32) == 0;
    } else {
      qhat = div_2by1(wt, wt1, vt, dinv, rhat);
    }
#pragma unroll
    for (int it = 0; it < 2; it++) {
      if (rhat_ok && (uint64_t)qhat * vt2 > (((uint64_t)rhat << 32) | wt2)) {
        qhat--;
        const uint64_t rr = (uint64_t)rhat + vt;
        rhat = (uint32_t)rr;
        rhat_ok = (rr >> 32) == 0;
      }
    }
    // W - qhat*V over this lane's limbs: limb k loses lo(q*v[k]) and hi(q*v[k-1])
    uint32_t b1 = 0, b2 = 0;
    uint32_t ph = (k0 >= 1 && k0 - 1 < NB) ? (uint32_t)(((uint64_t)qhat * v[k0 - 1]) >> 32) : 0u;
    for (uint32_t k = k0; k < k1; k++) {
      const uint64_t p = k < NB ? (uint64_t)qhat * v[k] : 0ull;
      const uint32_t x = subb(W[k], (uint32_t)p, b1, b1);
      W[k] = subb(x, ph, b2, b2);
      ph = (uint32_t)(p >> 32);
    }
    wave_ripple<false>(W, k0, k1, b1 + b2, l, topl);
    wave_fence();
    if (W[NB] != 0u) {
      // the estimate was one too large (rare; the same on every lane): add V back
      qhat--;
      uint32_t c = 0;
      for (uint32_t k = k0; k < k1; k++) W[k] = addc(W[k], k < NB ? v[k] : 0u, c, c);
      wave_ripple<true>(W, k0, k1, c, l, topl);
      wave_fence();
    }
    if (o.q && l == 0 && (uint32_t)j < o.q_len) o.q[(size_t)j * o.q_stride + o.q_inst] = qhat;
  }
  if (o.q)
    for (uint32_t k = (uint32_t)nq + l; k < o.q_len; k += 64) o.q[(size_t)k * o.q_stride + o.q_inst] = 0u;
  if (o.r) {
    // remainder limb k: U[k+sl .. k+sl+1] >> s, within U's NB low limbs
    for (uint32_t k = l; k < o.r_len; k += 64) {
      const uint32_t kk = k + (uint32_t)sl;
      const uint32_t lo = kk < NB ? U[kk] : 0u, hi = kk + 1 < NB ? U[kk + 1] : 0u;
      o.r[(size_t)k * o.r_stride + o.r_inst] = s ? funnel(hi, lo, s) : lo;
    }
  }
}

// one 64-lane workgroup per division, v and U in dynamic LDS
__global__ void __launch_bounds__(64) divmod_wave_kernel(OutOperand q, OutOperand r, Operand a, Operand b) {
  extern __shared__ uint32_t lds_words[];
  const uint32_t t = blockIdx.x;
  DivOut o;
  o.q = q.data;
  o.q_len = q.len;
  o.q_stride = q.stride;
  o.q_inst = q.data ? out_index(q, t) : 0;
  o.r = r.data;
  o.r_len = r.len;
  o.r_stride = r.stride;
  o.r_inst = r.data ? out_index(r, t) : 0;
  OperandDividend A{a.data, a.stride, in_index(a, t), a.len};
  divide_wave(A, b.data, b.stride, in_index(b, t), b.len, o, lds_words);
}

__global__ void __launch_bounds__(64) pow2_mod_wave_kernel(uint32_t *out, uint32_t out_len, uint32_t out_stride,
                                                           const uint32_t *m, uint32_t m_len, uint32_t m_stride,
                                                           uint32_t e) {
  extern __shared__ uint32_t lds_words[];
  const uint32_t t = blockIdx.x;
  DivOut o{nullptr, 0, 0, 0, out, out_len, out_stride, t};
  PowerOfTwoDividend A{e};
  divide_wave(A, m, m_stride, t, m_len, o, lds_words);
}

__global__ void __launch_bounds__(256) pow2_mod_mem_kernel(uint32_t *out, uint32_t out_len, uint32_t out_stride,
                                                           const uint32_t *m, uint32_t m_len, uint32_t m_stride,
                                                           uint32_t count, uint32_t e, uint32_t *scratch,
                                                           uint32_t scr_stride) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  DivOut o{nullptr, 0, 0, 0, out, out_len, out_stride, t};
  PowerOfTwoDividend A{e};
  divide_mem(A, m, m_stride, t, m_len, o, scratch + t, scr_stride);
}

template <int NB>
hipError_t launch_divmod_reg(OutOperand q, OutOperand r, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL(divmod_reg_kernel<NB>, dim3(div_round_up(count, 256)), dim3(256), 0, stream, q, r, a, b, count);
  return hipGetLastError();
}

template <int NB>
hipError_t launch_pow2_reg(uint32_t *out, uint32_t out_len, uint32_t out_stride, const uint32_t *m, uint32_t m_len,
                           uint32_t m_stride, uint32_t count, uint32_t e, hipStream_t stream) {
  hipLaunchKernelGGL(pow2_mod_reg_kernel<NB>, dim3(div_round_up(count, 256)), dim3(256), 0, stream, out, out_len,
                     out_stride, m, m_len, m_stride, count, e);
  return hipGetLastError();
}

}  // namespace

size_t divmod_scratch_words(uint32_t b_len, uint32_t count) {
  if (b_len <= kMaxRegDivisorLimbs) return 0;
  return (size_t)(2 * b_len + 1) * round_up(count, 256);
}

template <int H>
hipError_t launch_divmod_pair(OutOperand q, OutOperand r, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL(divmod_pair_kernel<H>, dim3(div_round_up(2 * count, 256)), dim3(256), 0, stream, q, r, a, b,
                     count);
  return hipGetLastError();
}

hipError_t launch_divmod_ws(OutOperand q, OutOperand r, Operand a, Operand b, uint32_t count, uint32_t *scratch,
                            hipStream_t stream) {
  if (count == 0) return hipSuccess;
  const uint32_t nb = b.len;
  // Below 131,072 divisions one lane each gives fewer than two waves per
  // SIMD (1,024 SIMDs x 2 x 64 lanes): two lanes per division instead
  if (count < 131072 && nb > 16 && nb <= 128) {
    if (nb <= 32) return launch_divmod_pair<16>(q, r, a, b, count, stream);
    if (nb <= 64) return launch_divmod_pair<32>(q, r, a, b, count, stream);
    if (nb <= 96) return launch_divmod_pair<48>(q, r, a, b, count, stream);
    return launch_divmod_pair<64>(q, r, a, b, count, stream);
  }
  if (nb <= 1) return launch_divmod_reg<1>(q, r, a, b, count, stream);
  if (nb <= 2) return launch_divmod_reg<2>(q, r, a, b, count, stream);
  if (nb <= 4) return launch_divmod_reg<4>(q, r, a, b, count, stream);
  if (nb <= 8) return launch_divmod_reg<8>(q, r, a, b, count, stream);
  if (nb <= 16) return launch_divmod_reg<16>(q, r, a, b, count, stream);
  if (nb <= 32) return launch_divmod_reg<32>(q, r, a, b, count, stream);
  if (nb <= 48) return launch_divmod_reg<48>(q, r, a, b, count, stream);
  if (nb <= 64) return launch_divmod_reg<64>(q, r, a, b, count, stream);
  if (nb <= 96) return launch_divmod_reg<96>(q, r, a, b, count, stream);
  if (nb <= 128) return launch_divmod_reg<128>(q, r, a, b, count, stream);
  const uint32_t words = wave_div_words(a.len, nb);
  if (words <= kWaveDivMaxWords) {
    hipLaunchKernelGGL(divmod_wave_kernel, dim3(count), dim3(64), words * sizeof(uint32_t), stream, q, r, a, b);
    return hipGetLastError();
  }
  if (!scratch) return hipErrorInvalidValue;
  hipLaunchKernelGGL(divmod_mem_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, q, r, a, b, count,
                     scratch, round_up(count, 256));
  return hipGetLastError();
}

hipError_t launch_pow2_mod(uint32_t *out, uint32_t out_len, uint32_t out_stride, const uint32_t *m, uint32_t m_len,
                           uint32_t m_stride, uint32_t count, uint32_t e, uint32_t *scratch, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (m_len <= 8) return launch_pow2_reg<8>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 16) return launch_pow2_reg<16>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 32) return launch_pow2_reg<32>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 64) return launch_pow2_reg<64>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 96) return launch_pow2_reg<96>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 128) return launch_pow2_reg<128>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  const uint32_t words = wave_div_words((e >> 5) + 1, m_len);
  if (words <= kWaveDivMaxWords) {
    hipLaunchKernelGGL(pow2_mod_wave_kernel, dim3(count), dim3(64), words * sizeof(uint32_t), stream, out, out_len,
                       out_stride, m, m_len, m_stride, e);
    return hipGetLastError();
  }
  if (!scratch) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pow2_mod_mem_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, out, out_len,
                     out_stride, m, m_len, m_stride, count, e, scratch, round_up(count, 256));
  return hipGetLastError();
}

}  // namespace xmp
