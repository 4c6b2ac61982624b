// powm.hpp -- Montgomery exponentiation plans and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "internal.hpp"

namespace xmp {

constexpr uint32_t kMaxRegDivisorLimbs = 128;

// division workspace (words) for divisors beyond the register kernels
size_t divmod_scratch_words(uint32_t b_len, uint32_t count);
hipError_t launch_divmod_ws(OutOperand q, OutOperand r, Operand a, Operand b, uint32_t count, uint32_t *scratch,
                            hipStream_t stream);
// out[:, t] = 2^e mod m[:, t] for t < count (interleaved, m_stride/out_stride)
hipError_t launch_pow2_mod(uint32_t *out, uint32_t out_len, uint32_t out_stride, const uint32_t *m, uint32_t m_len,
                           uint32_t m_stride, uint32_t count, uint32_t e, uint32_t *scratch, hipStream_t stream);

// One Montgomery layout: an instance spans `lanes` lanes of a wave64, each
// holding `digits` radix-2^radix digits (lanes*digits*radix >= precision+2).
struct PowmVariant {
  int lanes;    // L: 1, 2, 4, 8
  int digits;   // K: digits per lane
  int radix;    // D: 28 or 27
};

// Everything a powm launch needs (device pointers, interleaved layouts).
struct PowmArgs {
  OutOperand out;       // 32-bit limbs, out precision == mod precision
  Operand base, exp, mod;
  const uint32_t *r2;   // R^2 mod m per modulus (mod.len limbs, interleaved by modulus index)
  uint32_t r2_stride;
  uint32_t *table;      // window table workspace
  uint32_t start;       // first instance of this launch
  uint32_t count;       // instances in this launch
  uint32_t window_bits;
};

// variants compiled into the library
int powm_variant_count();
PowmVariant powm_variant(int i);
// index of the compiled variant (L, K, D) or -1
int powm_find_variant(int lanes, int digits, int radix);
// largest R exponent (bits) a variant gives: every lane's digits
inline uint32_t variant_rbits(const PowmVariant &v) { return (uint32_t)(v.lanes * v.digits * v.radix); }
// Lanes of a group that hold digits of a modulus of m_len limbs: the fewest
// whose K*D-bit share reaches 32*m_len + 2 bits (R >= 4m), at most L.  The
// Montgomery steps run over these lanes' digits only (the digits above them
// are zero), so R = 2^(D*K*lanes): a 1024-bit modulus on 16 lanes of 3
// radix-2^28 digits takes 13 lanes' 39 steps per product instead of 48.
__host__ __device__ inline uint32_t powm_active_lanes(uint32_t L, uint32_t K, uint32_t D, uint32_t m_len) {
  const uint32_t nd = (32u * m_len + 2u + D - 1u) / D;
  const uint32_t nl = (nd + K - 1u) / K;
  return nl < 1u ? 1u : (nl > L ? L : nl);
}
// Whether those lanes hold D bits more than R >= 4m needs, so that the
// Montgomery reduction can run by m~ = m * (-m^-1 mod 2^D) (< 2^(32 m_len +
// D)) at the same R: m~ is -1 mod 2^D and each step's quotient digit is the
// low digit itself (the quotient-free form; products stay congruent mod m
// and the last one reduces by m).
__host__ __device__ inline bool powm_quotient_free(uint32_t L, uint32_t K, uint32_t D, uint32_t m_len) {
  return K * D * powm_active_lanes(L, K, D, m_len) >= 32u * m_len + 2u + D;
}
// the R exponent (bits) a variant runs with for moduli of m_len limbs
inline uint32_t variant_active_rbits(const PowmVariant &v, uint32_t m_len) {
  return (uint32_t)(v.digits * v.radix) * powm_active_lanes(v.lanes, v.digits, v.radix, m_len);
}
// table workspace in 32-bit words for `count` instances
size_t powm_table_words(const PowmVariant &v, uint32_t count, uint32_t window_bits);
// launch the exponentiation kernel for one variant
hipError_t launch_powm_variant(int variant, const PowmArgs &args, hipStream_t stream);
// the few-digit latency layouts (4x10, 8x5, 8x10, 16x5, 16x3 at radix 2^28), built in
// k_powm_latency.hip; launch_powm_variant forwards to it
hipError_t launch_powm_latency(int lanes, int digits, const PowmArgs &args, hipStream_t stream);
// R^2 mod m (R = 2^rbits of the variant) for m[:, t], t < count, canonical, into
// r2 (m_len limbs, interleaved with r2_stride).  Moduli longer than
// kMaxRegDivisorLimbs need `ws` of divmod_scratch_words(m_len, count) words.
hipError_t launch_montgomery_r2(int variant, uint32_t *r2, uint32_t r2_stride, const uint32_t *m, uint32_t m_len,
                                uint32_t m_stride, uint32_t count, uint32_t *ws, hipStream_t stream);
// r2 = v - m if v >= m else v (v < 2m, m_len + 1 limbs), one lane per modulus
hipError_t launch_r2_canonical(uint32_t *r2, uint32_t r2_stride, const uint32_t *v, uint32_t v_stride,
                               const uint32_t *m, uint32_t m_len, uint32_t m_stride, uint32_t count,
                               hipStream_t stream);

// Moduli wider than every lane layout: one workgroup per instance.  A tier
// is its digit count in units of 256: 8, 16, 32, 64 keep radix-2^24 digits
// in registers (up to 393,214 bits), 4C (C >= 20) keeps C radix-2^20 digits
// per thread of 1024 in global memory (up to 83,886,078 bits).
// the tier for a precision, or 0 above the widest
uint32_t powm_wide_digits_per_thread(uint32_t precision);
uint32_t powm_wide_rbits(uint32_t tier);
// the table workspace (for the memory tier, each instance's whole block)
size_t powm_wide_table_words(uint32_t tier, uint32_t count, uint32_t window_bits);
// the memory tier's blocks for R^2 set-up over `count` moduli (0 for register tiers)
size_t powm_wide_r2_scratch_words(uint32_t tier, uint32_t count);
hipError_t launch_powm_wide(uint32_t tier, const PowmArgs &args, hipStream_t stream);
// R^2 mod m for the wide layout (needs `ws` of divmod_scratch_words(m_len, count)
// words, and for the memory tier `tier_ws` of powm_wide_r2_scratch_words words)
hipError_t launch_wide_r2(uint32_t tier, uint32_t *r2, uint32_t r2_stride, const uint32_t *m, uint32_t m_len,
                          uint32_t m_stride, uint32_t count, uint32_t *ws, uint32_t *tier_ws, hipStream_t stream);
// variants usable for a modulus precision (in bits): those with R >= 2^(precision+2)
bool variant_fits(const PowmVariant &v, uint32_t precision);
// whether the persistent mixed-size kernel runs this variant
bool powm_variant_batchable(int variant);
// LDS staging bytes one wave of the variant needs
size_t powm_variant_lds_bytes(int variant);

// A powm call's workspace for callers that cannot allocate: a fixed device
// region used as it is (never grown or freed by the call).
struct FixedWorkspace {
  uint8_t *base;
  size_t bytes;
};
// xmpIntegersPowmAsync's validation, planning and launches on the handle's
// stream.  ws == nullptr: the handle's grow-only scratch (allocating as
// needed, within the policy's scratch limit); otherwise the fixed region, and
// the call allocates nothing (xmpErrorIncreaseScratchLimit when even its
// smallest plan does not fit).
xmpError_t powm_run(xmpHandle_t handle, xmpIntegers_t out, xmpIntegers_t a, xmpIntegers_t exp, xmpIntegers_t mod,
                    uint32_t count, const FixedWorkspace *ws);

// CRT recombination out = (mp*cp + mq*cq) mod n, as REDC(mp*cp2 + mq*cq2)
// with cp2 = cp*R mod n, cq2 = cq*R mod n and R = 2^(radix*steps).
struct CrtArgs {
  OutOperand out;        // n-bit results
  Operand mp, mq;        // residues below 2^h_bits
  Operand cp2, cq2, n;   // n-bit, gathered by key
  uint32_t steps;
  uint32_t count;
};
// the recombination layout and step count for n-bit moduli and h-bit residues
bool crt_combine_plan(uint32_t n_bits, uint32_t h_bits, int *which, uint32_t *steps);
uint32_t crt_combine_radix(int which);
hipError_t launch_crt_combine(int which, const CrtArgs &args, hipStream_t stream);

// Persistent mixed-size launch: tiles (class, tile) in queue order.
struct BatchArgs {
  const PowmArgs *classes;        // per class (device array)
  const int32_t *class_variant;   // per class variant index (device array)
  const uint32_t *tile_class;     // per queue entry
  const uint32_t *tile_index;     // per queue entry: tile within the class
  uint32_t ntiles;
  uint32_t *counter;              // queue head, reset by the launch
};
// queue reset + persistent kernel on `stream` (graph-capturable: no memset)
hipError_t launch_powm_batch(const BatchArgs &ba, uint32_t grid, size_t lds_bytes, hipStream_t stream);

}  // namespace xmp
