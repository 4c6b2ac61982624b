// kernels.hpp -- host-callable launchers for the HIP kernels.
//
// Every launcher takes plain device pointers / operand descriptors and the
// stream, launches on that stream and returns the launch status.  Operands
// use the interleaved layout (limb j of instance i at data[j*stride + i]).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace xmp {

// ---- transfer (k_xfer.hip) ----
// compact words (count x words of `size` bytes, order/endian/nails, optional
// gather indices) -> interleaved limbs.  Restates xmpTransform
// (reference src/xmp.cu:461-489) fused with the compact->strided transpose
// (src/xmp.cu:817-828).
hipError_t launch_import(uint32_t *dst, uint32_t dst_stride, uint32_t nlimbs, const uint8_t *src, uint32_t count,
                         uint32_t words, int32_t order, int32_t endian, uint32_t nails, uint32_t size,
                         const uint32_t *in_idx, uint32_t in_idx_count, hipStream_t stream);
// interleaved -> compact words with optional scatter indices (out row =
// idx[i % idx_count]); reference src/xmp.cu:586-682 + S2C (xmp.cu:832-843).
hipError_t launch_export(uint8_t *dst, const uint32_t *src, uint32_t src_stride, uint32_t nlimbs, uint32_t count,
                         uint32_t words, int32_t order, int32_t endian, uint32_t nails, uint32_t size,
                         const uint32_t *out_idx, uint32_t out_idx_count, hipStream_t stream);
// out[o][j] = in[i][j] (zero-extended) with o = out_idx[k % c0], i = in_idx[k % c1] % in_count;
// reference xmpReorder_compact_kernel (src/xmp.cu:693-716), interleaved here.
hipError_t launch_set(uint32_t *out, uint32_t out_stride, uint32_t out_limbs, const uint32_t *in, uint32_t in_stride,
                      uint32_t in_limbs, uint32_t in_count, uint32_t count, const uint32_t *out_idx,
                      uint32_t out_idx_count, const uint32_t *in_idx, uint32_t in_idx_count, hipStream_t stream);
hipError_t launch_copy_words(uint32_t *dst, const uint32_t *src, size_t words, hipStream_t stream);

// ---- elementwise arithmetic (k_arith.hip) ----
hipError_t launch_add(OutOperand s, Operand a, Operand b, uint32_t count, hipStream_t stream);
hipError_t launch_sub(OutOperand d, Operand a, Operand b, uint32_t count, hipStream_t stream);
hipError_t launch_mul(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream);
// p = a^2 (truncated to p.len limbs): the symmetric kernels where a fits them
hipError_t launch_sqr(OutOperand p, Operand a, uint32_t count, hipStream_t stream);
// Products whose longer operand has kKaratsubaMinLimbs limbs or more split by
// Karatsuba into products the row kernels hold; `ws` holds
// mul_karatsuba_words(a.len, b.len, count) words (0 below the threshold).
// square: b is a (the sub-products are squares too).
constexpr uint32_t kKaratsubaMinLimbs = 513;
size_t mul_karatsuba_words(uint32_t a_len, uint32_t b_len, uint32_t count);
hipError_t launch_mul_ws(OutOperand p, Operand a, Operand b, uint32_t count, bool square, uint32_t *ws,
                         hipStream_t stream);

// ---- division (k_div.hip): launch_divmod_ws and launch_pow2_mod in powm.hpp ----

// ---- bitwise (k_logic.hip), reference src/logic.cu ----
enum : int { kBitIor = 0, kBitAnd = 1, kBitXor = 2, kBitNot = 3 };
// c = a op b (zero-extended to c.len; not ignores b and fills with ones)
hipError_t launch_bitwise(int op, OutOperand c, Operand a, Operand b, uint32_t count, hipStream_t stream);
// c = a shifted by shift[i % shift_count] bits (positive left, negative right); shift on the device
hipError_t launch_shift(OutOperand c, Operand a, const int32_t *shift, uint32_t shift_count, uint32_t count,
                        hipStream_t stream);
hipError_t launch_popc(uint32_t *out, Operand a, uint32_t count, hipStream_t stream);
hipError_t launch_cmp(int32_t *out, Operand a, Operand b, uint32_t count, hipStream_t stream);

// ---- Montgomery powm (k_powm.hip) ----
struct PowmPlan;  // powm.hpp

}  // namespace xmp
