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
struct OperandDividend {
  const uint32_t *data;
  uint32_t stride, inst, len;
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return (k >= 0 && (uint32_t)k < len) ? data[(size_t)k * stride + inst] : 0u;
  }
  __device__ __forceinline__ uint32_t length() const { return len; }
};

struct PowerOfTwoDividend {   // 2^e
  uint32_t e;
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return (k >= 0 && (uint32_t)k == (e >> 5)) ? (1u << (e & 31)) : 0u;
  }
  __device__ __forceinline__ uint32_t length() const { return (e >> 5) + 1; }
};

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
    v[k] = (uint32_t)k < blen ? bdata[(size_t)k * bstride + binst] : 0u;
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
      int src = k - sl;
      v[k] = (src >= 0 && (uint32_t)src < blen) ? bdata[(size_t)src * bstride + binst] : 0u;
    }
  }
#pragma unroll
  for (int k = NB - 1; k >= 0; k--) v[k] = shl_limb(v[k], k > 0 ? v[k - 1] : 0u, s);
  const uint32_t dinv = reciprocal(v[NB - 1]);

  // normalised dividend U = A << (32*sl + s); U limb k:
  auto U = [&](int32_t k) -> uint32_t {
    int32_t x = k - sl;
    return shl_limb(A(x), A(x - 1), s);
  };
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
  for (int k = 0; k <= NB; k++) Wa[k] = U(nq - 1 + k);
  // the dividend limbs that enter the window, read four digits ahead
  uint32_t n0 = U(nq - 2), n1 = U(nq - 3), n2 = U(nq - 4), n3 = U(nq - 5);
  int32_t j = nq - 1;
  for (; j >= 1; j -= 2) {
    digit_step<NB>(Wa, Wb, n0, v, dinv, o, j);
    digit_step<NB>(Wb, Wa, n1, v, dinv, o, j - 1);
    // digits j-2 and j-3 take U(j-3) and U(j-4); U(j-5), U(j-6) are read for the pair after
    n0 = n2;
    n1 = n3;
    n2 = U(j - 5);
    n3 = U(j - 6);
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

hipError_t launch_divmod_ws(OutOperand q, OutOperand r, Operand a, Operand b, uint32_t count, uint32_t *scratch,
                            hipStream_t stream) {
  const uint32_t nb = b.len;
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
  if (!scratch) return hipErrorInvalidValue;
  hipLaunchKernelGGL(divmod_mem_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, q, r, a, b, count,
                     scratch, round_up(count, 256));
  return hipGetLastError();
}

hipError_t launch_pow2_mod(uint32_t *out, uint32_t out_len, uint32_t out_stride, const uint32_t *m, uint32_t m_len,
                           uint32_t m_stride, uint32_t count, uint32_t e, uint32_t *scratch, hipStream_t stream) {
  if (m_len <= 8) return launch_pow2_reg<8>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 16) return launch_pow2_reg<16>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 32) return launch_pow2_reg<32>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 64) return launch_pow2_reg<64>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 96) return launch_pow2_reg<96>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (m_len <= 128) return launch_pow2_reg<128>(out, out_len, out_stride, m, m_len, m_stride, count, e, stream);
  if (!scratch) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pow2_mod_mem_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, out, out_len,
                     out_stride, m, m_len, m_stride, count, e, scratch, round_up(count, 256));
  return hipGetLastError();
}

}  // namespace xmp
