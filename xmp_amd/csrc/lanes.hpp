// lanes.hpp -- an integer spread over a group of L lanes of a wave64: the
// cross-lane moves (DPP) and the carry resolution of lazy 64-bit digit sums
// that the Montgomery and multiply kernels share.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "devmp.hpp"

namespace xmp {

// Cross-lane moves inside an instance's group of L lanes (groups are aligned:
// L <= 16 stays inside a 16-lane DPP row), done with DPP moves on the VALU
// instead of LDS permutes:
//   bcast0    -- the group's first lane's value to every lane of the group
//                (bcast0_and: the same, masked, as the Montgomery step uses it)
//   from_next -- lane l receives lane l+1's value, the group's last lane 0
//                (next_digit: the Montgomery step's masked digit hand-over)
//   from_prev -- lane l receives lane l-1's value, the group's first lane 0
// DPP controls: quad_perm [0,0,2,2] = 0xA0, [0,0,0,0] = 0x00, row_shl:1 =
// 0x101, row_shr:1 = 0x111, row_shr:4 = 0x114, row_newbcast:0 = 0x150 (lane 0
// of each row to the whole row), row_bcast:15 = 0x142 (lane 15 of each row to
// the next row), wave_shl:1 = 0x130 and wave_shr:1 = 0x138 (whole-wave shifts,
// for groups that span rows).  A 64-lane group reads its first lane with
// readfirstlane, which leaves the quotient digit in a scalar register.
// Moves whose source lane is always valid (quad_perm, row_newbcast) are
// still marked bound_ctrl: the compiler folds such a move into the VALU
// instruction that uses it (as a DPP operand) only when it is.
template <int L>
struct Group {
  static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16 || L == 32 || L == 64, "group size");
  static constexpr int SHL1 = L <= 16 ? 0x101 : 0x130;
  static constexpr int SHR1 = L <= 16 ? 0x111 : 0x138;
  int l;   // lane within the instance group
  __device__ __forceinline__ explicit Group(int lane) : l(lane & (L - 1)) {}
  __device__ __forceinline__ uint32_t bcast0(uint32_t x) const {
    if constexpr (L == 1) {
      return x;
    } else if constexpr (L == 2) {
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xA0, 0xF, 0xF, true);
    } else if constexpr (L == 4) {
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x00, 0xF, 0xF, true);
    } else if constexpr (L == 8) {
      uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x00, 0xF, 0xF, false);
      uint32_t z = (uint32_t)__builtin_amdgcn_mov_dpp((int)y, 0x114, 0xF, 0xF, true);
      return (l & 4) ? z : y;
    } else if constexpr (L == 16) {
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x150, 0xF, 0xF, true);
    } else if constexpr (L == 32) {
      // every row takes its lane 0; rows 1 and 3 then take lane 15 of rows 0 and 2
      uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x150, 0xF, 0xF, false);
      return (uint32_t)__builtin_amdgcn_update_dpp((int)y, (int)y, 0x142, 0xA, 0xF, false);
    } else {
      return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
    }
  }
  __device__ __forceinline__ uint32_t from_next(uint32_t x) const {
    if constexpr (L == 1) {
      return 0;
    } else {
      uint32_t v = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, SHL1, 0xF, 0xF, true);
      if constexpr (L == 16 || L == 64) return v;   // the row's / wave's last lane already reads 0
      return l == L - 1 ? 0u : v;
    }
  }
  // The Montgomery step's digit hand-over: lane l receives lane l+1's
  // x & mask, the group's last lane 0, for an x whose masked value is zero
  // on the group's first lane (the retired digit).  The move may then read
  // the first lane of a group, its own (a rotation inside the group: L = 2,
  // quad_perm [1,0,3,2] = 0xB1; L = 4, quad_perm [1,2,3,0] = 0x39) or the next
  // one (L = 8 in a row, L = 32 across the wave), so no group needs a select,
  // and the move folds into the AND as its DPP operand.
  __device__ __forceinline__ uint32_t next_digit(uint32_t x, uint32_t mask) const {
    if constexpr (L == 1) {
      return 0;
    } else {
      constexpr int ctrl = L == 2 ? 0xB1 : L == 4 ? 0x39 : SHL1;
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, ctrl, 0xF, 0xF, true) & mask;
    }
  }
  // bcast0(x) & mask.  The mask is applied as the first move's DPP operand, so
  // for L = 8 and 32 the broadcast is two instructions, the second a move
  // that writes only the lanes still to be filled (by bank or row mask),
  // instead of two moves, a select and the AND: for L = 8, quad_perm
  // [0,0,0,0] gives each quad its lane 0, then banks 1 and 3 of each row
  // (lanes 4-7, 12-15) take row_shr:4 of that (lanes 0 and 8); for L = 32,
  // row_newbcast:0 gives each row its lane 0, then rows 1 and 3 take
  // row_bcast:15 (lane 15 of rows 0 and 2).
  __device__ __forceinline__ uint32_t bcast0_and(uint32_t x, uint32_t mask) const {
    if constexpr (L == 8) {
      const uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x00, 0xF, 0xF, true) & mask;
      return (uint32_t)__builtin_amdgcn_update_dpp((int)y, (int)y, 0x114, 0xF, 0xA, false);
    } else if constexpr (L == 32) {
      const uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x150, 0xF, 0xF, true) & mask;
      return (uint32_t)__builtin_amdgcn_update_dpp((int)y, (int)y, 0x142, 0xA, 0xF, false);
    } else {
      return bcast0(x) & mask;
    }
  }
  __device__ __forceinline__ uint64_t from_prev(uint64_t x) const {
    if constexpr (L == 1) {
      return 0;
    } else {
      uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)x, SHR1, 0xF, 0xF, true);
      uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(x >> 32), SHR1, 0xF, 0xF, true);
      uint64_t v = ((uint64_t)hi << 32) | lo;
      if constexpr (L == 16 || L == 64) return v;   // the row's / wave's first lane already reads 0
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

// digit k of an integer given by 32-bit limbs (interleaved) as radix-2^D,
// one guarded load per limb: the form for prologues whose register budget is
// tight (the Montgomery kernels), where loads in flight would cost spills.
template <int D>
__device__ __forceinline__ uint32_t limb_digit_guarded(const uint32_t *data, uint32_t stride, uint32_t inst,
                                                       uint32_t len, uint32_t digit) {
  const uint32_t bit = (uint32_t)D * digit;
  const uint32_t w = bit >> 5, s = bit & 31;
  const uint32_t lo = w < len ? data[(size_t)w * stride + inst] : 0u;
  const uint32_t hi = w + 1 < len ? data[(size_t)(w + 1) * stride + inst] : 0u;
  return funnel(hi, lo, s) & ((1u << D) - 1);
}

// digit k of an integer given by len >= 1 32-bit limbs (interleaved) as
// radix-2^D.  The loads are unconditional (at a clamped, in-range limb) and
// the zero extension is a select, so a run of calls issues all of its loads
// before the first wait instead of one branch and one memory latency each.
template <int D>
__device__ __forceinline__ uint32_t limb_digit(const uint32_t *data, uint32_t stride, uint32_t inst, uint32_t len,
                                               uint32_t digit) {
  const uint32_t bit = (uint32_t)D * digit;
  const uint32_t w = bit >> 5, s = bit & 31;
  const uint32_t w0 = w < len ? w : len - 1, w1 = w + 1 < len ? w + 1 : len - 1;
  uint32_t lo = data[(size_t)w0 * stride + inst];
  uint32_t hi = data[(size_t)w1 * stride + inst];
  lo = w < len ? lo : 0u;
  hi = w + 1 < len ? hi : 0u;
  return funnel(hi, lo, s) & ((1u << D) - 1);
}

// Digits first .. first+K-1 of an integer given by len >= 1 32-bit limbs
// (interleaved) as radix-2^D.  Each limb the lane needs is loaded once (all
// loads unconditional at clamped, in-range limbs, zero-extended by a select);
// the lane's bit offset (D*first) mod 32 only decides, per digit, whether the
// digit starts one limb further and the funnel-shift amount.
template <int K, int D>
__device__ __forceinline__ void load_lane_digits(uint32_t (&d)[K], const uint32_t *data, uint32_t stride,
                                                 uint32_t inst, uint32_t len, uint32_t first) {
  constexpr int NW = (D * K + 31) / 32 + 2;
  const uint32_t bit0 = (uint32_t)D * first;
  const uint32_t w0 = bit0 >> 5, off = bit0 & 31;
  uint32_t lw[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) {
    const uint32_t w = w0 + (uint32_t)j;
    const uint32_t v = data[(size_t)(w < len ? w : len - 1) * stride + inst];
    lw[j] = w < len ? v : 0u;
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int jb = (D * k) >> 5;
    const uint32_t s = (uint32_t)((D * k) & 31) + off;   // 0 .. 62
    const bool up = s >= 32;
    const uint32_t lo = up ? lw[jb + 1] : lw[jb];
    const uint32_t hi = up ? lw[jb + 2] : lw[jb + 1];
    d[k] = funnel(hi, lo, s & 31) & ((1u << D) - 1);
  }
}

// Resolve the lazy sums into D-bit digits.  Each lane first carries through
// its own K digits; the lane carries (< 2^(64-D)) then ripple up the group
// one lane per round.  After the first round a lane passes a carry on only
// if its digits were all ones, so the rounds stop as soon as no lane of the
// wave holds a carry (almost always after one round).
template <int L, int K, int D>
__device__ __forceinline__ void normalize_group(uint32_t (&r)[K], const uint64_t (&t)[K], const Group<L> &g) {
  constexpr uint32_t MASK = (1u << D) - 1;
  uint64_t c = 0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    uint64_t x = t[k] + c;
    r[k] = (uint32_t)x & MASK;
    c = x >> D;
  }
  if constexpr (L > 1) {
    for (int round = 1; round < L; round++) {
      c = g.from_prev(c);
      if (round > 1 && __builtin_amdgcn_ballot_w64(c != 0) == 0) break;
      // c < 2^(64-D): past the first digit the carry fits in 32 bits, and
      // past the second it is below 2 and almost never set (a digit would
      // have to be within the carry of 2^D), so the rest of the ripple is a
      // uniform branch on the wave's carries.  Measured A/B, the branch made
      // the 16x3 and 10- to 38-digit layouts' exponentiations 1.5-2 % faster
      // and the 5-digit ones up to 4.6 % slower, which keep the full ripple.
      uint64_t x = (uint64_t)r[0] + c;
      r[0] = (uint32_t)x & MASK;
      uint32_t c32 = (uint32_t)(x >> D);
      if constexpr (K == 3 || K >= 10) {
        const uint32_t y = r[1] + c32;
        r[1] = y & MASK;
        c32 = y >> D;
        if (__builtin_amdgcn_ballot_w64(c32 != 0) != 0) {
#pragma unroll
          for (int k = 2; k < K; k++) {
            const uint32_t z = r[k] + c32;
            r[k] = z & MASK;
            c32 = z >> D;
          }
        }
      } else {
#pragma unroll
        for (int k = 1; k < K; k++) {
          uint32_t y = r[k] + c32;
          r[k] = y & MASK;
          c32 = y >> D;
        }
      }
      c = c32;
    }
  }
}

}  // namespace xmp
