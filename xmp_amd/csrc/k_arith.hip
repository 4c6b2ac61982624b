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
// 2^P(p).  The longer operand is split into radix-2^28 (2^27 above 4096 bits)
// digits held in the registers of one lane or a group of lanes, and the
// product is formed row by row in 64-bit "lazy" accumulators: each partial
// product is one v_mad_u64_u32 and carries are resolved once per retired
// digit, because up to 255 products of 56 bits fit in 64 bits.
#include <algorithm>
#include <utility>

#include "devmp.hpp"
#include "kernels.hpp"
#include "lanes.hpp"

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

// Row scanning (operand scanning) over radix-2^D digits: the multiplicand's
// L*K digits and the accumulator window are spread over a group of L lanes
// (lane l holds digits l*K .. l*K+K-1; L = 1 up to 1024 bits), so the
// registers per lane stay near 4K and three waves fit per SIMD.  The
// multiplier's digits are staged in LDS first and read four at a time (an LDS
// broadcast within the group).  Each row multiplies the window by one
// multiplier digit and retires the lowest digit, which lane 0 packs into
// output limbs; the window then moves down one digit with the same
// split-carry DPP shift as the Montgomery product.  Rows stop once the
// retired digits cover the output (truncated products skip the rest); else
// the remaining window is normalised and each lane writes the limbs that
// start in its digits.
// Bound: every accumulator receives at most ND_b products < 2^(2D) plus a
// carry < 2^(64-D) and a D-bit digit (ND_b <= 255 at D = 28, <= 1023 at D = 27).
template <int L, int K, int D>
struct MulRow {
  static constexpr uint32_t MASK = (1u << D) - 1;
  const Group<L> &g;
  uint64_t obuf = 0;   // product bits not yet stored (meaningful on lane 0)
  uint32_t onb = 0, limb = 0;
  __device__ __forceinline__ explicit MulRow(const Group<L> &grp) : g(grp) {}
  // one row: window += a * d, then retire (and return) the lowest digit
  __device__ __forceinline__ uint32_t retire(uint64_t (&t)[K], const uint32_t (&da)[K], uint32_t d) const {
    uint64_t f = t[0] + (uint64_t)da[0] * d;
    const uint32_t lo = (uint32_t)f & MASK;
    const uint32_t up = g.from_next(lo);
#pragma unroll
    for (int j = 1; j < K; j++) t[j - 1] = t[j] + (uint64_t)da[j] * d;
    t[K - 1] = up;
    t[0] += f >> D;
    return lo;
  }
  // For a square, row i = li K + KI (the digit a_i) takes column j = l K + J
  // (lane l's digit J) once if J == KI, and twice or not at all otherwise,
  // as a tournament on the within-lane indices decides: of a pair i != j,
  // exactly one of row i and row j takes 2 a_i a_j (when J != KI), or both
  // take it once (J == KI).  Both rows come before the row that retires
  // position i + j, so every position holds the plain product's total when
  // it is retired: the same digits, carries and accumulator bound.
  static constexpr bool beats(int x, int y) {
    const int dd = (y - x + K) % K;
    return 2 * dd < K || (2 * dd == K && x < K / 2);
  }
  // up to 10 digits per lane keep their doubles in registers; else a row
  // doubles its multiplier digit once
  static constexpr bool kOwnDoubled = K <= 10;
  template <int KI>
  __device__ __forceinline__ uint32_t retire_sq(uint64_t (&t)[K], const uint32_t (&da)[K], const uint32_t (&a2)[K],
                                                uint32_t d) const {
    const uint32_t d2 = kOwnDoubled ? 0u : d << 1;
    auto twice = [&](int j) { return kOwnDoubled ? (uint64_t)a2[j] * d : (uint64_t)da[j] * d2; };
    uint64_t f = t[0];
    if constexpr (KI == 0) f += (uint64_t)da[0] * d;
    else if constexpr (beats(KI, 0)) f += twice(0);
    const uint32_t lo = (uint32_t)f & MASK;
    const uint32_t up = g.from_next(lo);
    sq_cols<KI>(t, da, a2, d, d2, std::make_integer_sequence<int, K - 1>());
    t[K - 1] = up;
    t[0] += f >> D;
    return lo;
  }
  template <int KI, int... J>
  __device__ __forceinline__ void sq_cols(uint64_t (&t)[K], const uint32_t (&da)[K], const uint32_t (&a2)[K],
                                          uint32_t d, uint32_t d2, std::integer_sequence<int, J...>) const {
    (sq_col<KI, J + 1>(t, da, a2, d, d2), ...);
  }
  template <int KI, int J>
  __device__ __forceinline__ void sq_col(uint64_t (&t)[K], const uint32_t (&da)[K], const uint32_t (&a2)[K],
                                         uint32_t d, uint32_t d2) const {
    uint64_t x = t[J];
    if constexpr (J == KI) x += (uint64_t)da[J] * d;
    else if constexpr (beats(KI, J)) x += kOwnDoubled ? (uint64_t)a2[J] * d : (uint64_t)da[J] * d2;
    t[J - 1] = x;
  }
  // pack a retired digit into the bit buffer, storing each completed limb
  __device__ __forceinline__ void pack(uint32_t lo, const OutOperand &p, uint32_t pi, bool store) {
    obuf |= (uint64_t)lo << onb;
    onb += D;
    if (onb >= 32) {
      if (store && limb < p.len) p.data[(size_t)limb * p.stride + pi] = (uint32_t)obuf;
      limb++;
      obuf >>= 32;
      onb -= 32;
    }
  }
  template <int KI>
  __device__ __forceinline__ void row_sq(uint64_t (&t)[K], const uint32_t (&da)[K], const uint32_t (&a2)[K],
                                         uint32_t d, const OutOperand &p, uint32_t pi, bool store) {
    pack(retire_sq<KI>(t, da, a2, d), p, pi, store);
  }
  template <int... KI>
  __device__ __forceinline__ void rows_sq(uint64_t (&t)[K], const uint32_t (&da)[K], const uint32_t (&a2)[K],
                                          const uint32_t *st, uint32_t r0, uint32_t rows, const OutOperand &p,
                                          uint32_t pi, bool store, std::integer_sequence<int, KI...>) {
    // the block's digits first, unconditionally (r0 + KI < ceil(rows/K) K
    // <= L K, within the stage), so no row waits on its own LDS load
    uint32_t dd[K];
    ((dd[KI] = st[r0 + KI]), ...);
    ((r0 + KI < rows ? row_sq<KI>(t, da, a2, dd[KI], p, pi, store) : void()), ...);
  }
  // one row with the retired digit packed into the bit buffer
  __device__ __forceinline__ void row(uint64_t (&t)[K], const uint32_t (&da)[K], uint32_t d, const OutOperand &p,
                                      uint32_t pi, bool store) {
    const uint32_t lo = retire(t, da, d);
    obuf |= (uint64_t)lo << onb;
    onb += D;
    if (onb >= 32) {
      if (store && limb < p.len) p.data[(size_t)limb * p.stride + pi] = (uint32_t)obuf;
      limb++;
      obuf >>= 32;
      onb -= 32;
    }
  }
};

template <int L, int K, int D, bool SQR = false>
__global__ void __launch_bounds__(64, 2) mul_lanes_kernel(OutOperand p, Operand a, Operand b, uint32_t count) {
  constexpr int ND = L * K;
  constexpr int G = 64 / L;
  constexpr int NDP = (ND + 3) / 4 * 4;   // the multiplier's digits; 16-byte rows
  __shared__ __attribute__((aligned(16))) uint32_t stage[G * NDP];
  const int lane = threadIdx.x;
  const int grp = lane / L;
  const Group<L> g(lane);
  const int l = g.l;
  const uint32_t inst = blockIdx.x * G + grp;
  const bool active = inst < count;
  const uint32_t i = active ? inst : count - 1;   // inactive groups compute a valid instance and store nothing
  const uint32_t ai = in_index(a, i), bi = in_index(b, i);
  const uint32_t pi = active ? out_index(p, i) : 0u;
  const bool store = active && l == 0;
  uint32_t *st = stage + grp * NDP;

  uint32_t da[K];
  uint64_t t[K];
  {
    uint32_t db[K];   // the multiplier's digits (b.len <= a.len), staged for the rows
    load_lane_digits<K, D>(db, b.data, b.stride, bi, b.len, (uint32_t)(l * K));
#pragma unroll
    for (int k = 0; k < K; k++) st[l * K + k] = db[k];
  }
  load_lane_digits<K, D>(da, a.data, a.stride, ai, a.len, (uint32_t)(l * K));
#pragma unroll
  for (int k = 0; k < K; k++) t[k] = 0;
  __syncthreads();
  const uint32_t ndb = (32u * b.len + D - 1) / D;
  const uint32_t nout = (32u * p.len + D - 1) / D;   // digits that reach the output
  const uint32_t rows = ndb < nout ? ndb : nout;
  MulRow<L, K, D> m(g);
  if constexpr (SQR) {
    // b is a: rows li K .. li K + K - 1 with the tournament of their
    // within-lane indices, lane by lane (the loop over lanes is kept)
    uint32_t a2[K];
#pragma unroll
    for (int k = 0; k < K; k++) a2[k] = MulRow<L, K, D>::kOwnDoubled ? da[k] << 1 : 0u;
#pragma unroll 1
    for (uint32_t r0 = 0; r0 < rows; r0 += K)
      m.rows_sq(t, da, a2, st, r0, rows, p, pi, store, std::make_integer_sequence<int, K>());
  } else {
    uint32_t r = 0;
    for (; r + 4 <= rows; r += 4) {
      const uint4 bb = *reinterpret_cast<const uint4 *>(st + r);
      m.row(t, da, bb.x, p, pi, store);
      m.row(t, da, bb.y, p, pi, store);
      m.row(t, da, bb.z, p, pi, store);
      m.row(t, da, bb.w, p, pi, store);
    }
    for (; r < rows; r++) m.row(t, da, st[r], p, pi, store);
  }
  if (rows < ndb) {
    // truncated: the retired digits already cover every output bit
    return;
  }
  // The rest of the product: the window's ND digits, from bit D*rows upward.
  // Lane l's digits start at bit S_l = D*(rows + l*K); it writes the limbs
  // whose first bit lies in [S_l, S_l+1) (lane 0 from limb `limb`, after its
  // pending bits), reading the next lane's two lowest digits for the last one.
  uint32_t rd[K];
  normalize_group<L, K, D>(rd, t, g);
  const uint32_t n0 = g.from_next(rd[0]), n1 = g.from_next(rd[1]);
  const uint32_t s_here = (uint32_t)D * (rows + (uint32_t)(l * K));
  const uint32_t s_next = s_here + (uint32_t)(D * K);
  const uint32_t wend = (l == L - 1) ? p.len : min(p.len, (s_next + 31) / 32);
  uint64_t buf;
  uint32_t nb, w;
  auto emit = [&]() {
    if (nb >= 32) {
      if (active && w < wend) p.data[(size_t)w * p.stride + pi] = (uint32_t)buf;
      w++;
      buf >>= 32;
      nb -= 32;
    }
  };
  if (l == 0) {
    // after the pending bits of limb `limb`
    buf = m.obuf;
    nb = m.onb;
    w = m.limb;
    buf |= (uint64_t)rd[0] << nb;
    nb += D;
    emit();
    buf |= (uint64_t)rd[1] << nb;
    nb += D;
    emit();
  } else {
    // from the first limb boundary at or above S_l
    w = (s_here + 31) / 32;
    const uint32_t skip = 32 * w - s_here;
    buf = (((uint64_t)rd[1] << D) | rd[0]) >> skip;
    nb = 2 * D - skip;
    emit();
  }
#pragma unroll
  for (int k = 2; k < K + 2; k++) {
    const uint32_t x = k < K ? rd[k < K ? k : 0] : (k == K ? n0 : n1);
    buf |= (uint64_t)x << nb;
    nb += D;
    emit();
  }
  if (l == L - 1 && active) {
    // past the product: the partial limb, then zeros
    if (nb > 0 && w < wend) p.data[(size_t)w * p.stride + pi] = (uint32_t)buf;
    for (w += (nb > 0 ? 1u : 0u); w < p.len; w++) p.data[(size_t)w * p.stride + pi] = 0u;
  }
}

template <int L, int K, int D, bool SQR = false>
hipError_t launch_mul_lanes(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  constexpr uint32_t G = 64 / L;
  hipLaunchKernelGGL((mul_lanes_kernel<L, K, D, SQR>), dim3(div_round_up(count, G)), dim3(64), 0, stream, p, a, b,
                     count);
  return hipGetLastError();
}

// One limb of a row of the interleaved layout at a lane's byte offset, as a
// raw buffer access: the row's resource is built in scalar registers and the
// lane's offset is one loop-invariant VGPR, so walking the rows of an
// operand costs no vector address arithmetic.  Offsets stay below 4 GiB
// (instance slots below 2^30, checked by the launcher).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const uint32_t *row) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(row), 0, -1, 0x00020000);
}
__device__ __forceinline__ uint32_t ld_row(const uint32_t *row, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(row_rsrc(row), (int)off, 0, 0);
}
__device__ __forceinline__ void st_row(uint32_t *row, uint32_t off, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, row_rsrc(row), (int)off, 0, 0);
}

// Product scanning with one lane per product (up to 74 radix-2^28 digits per
// operand).  Both operands' digits sit in the lane's registers and the
// product's columns are formed in order: column c is the lazy 64-bit sum of
// every a_i b_j with i + j = c plus the previous column's carry (74 products
// of < 2^56 and a carry < 2^36 stay below 2^63), its low D bits are the
// product's digit c, and each output limb is packed and stored as soon as
// its second digit is final.  Every register index is a compile-time
// constant (the columns are unrolled), so there is no cross-lane traffic, no
// window shift and no per-row packing branch: the instruction stream is the
// multiply-adds plus two instructions per column and two per output limb.
// A square (SQR) adds each product of two different digits once, against
// the doubled digit (a_i (2 a_j) for i < j, < 2^57 each), and the squares
// a_k^2 on the diagonal: about half the multiply-adds.
template <int NA, int NB, int D, bool SQR>
struct ColMul {
  static constexpr uint32_t MASK = (1u << D) - 1;
  static constexpr int NC = NA + NB - 1;                  // product columns
  static constexpr int NOUT = (D * (NC + 1) + 31) / 32;   // limbs the product's NC+1 digits reach
  static_assert(NA <= 74 && NB <= 74, "column sums must stay below 2^63");
  static_assert(!SQR || NA == NB, "a square has one operand");

  const uint32_t (&a)[NA];
  const uint32_t (&b)[NB];   // the multiplier's digits; for a square the doubled digits 2 a_j
  uint32_t *out;             // row of the next limb to store (uniform; limbs are stored in order)
  size_t stride;
  uint32_t inst;             // the product's slot within a row
  uint32_t len;
  uint32_t dig[NC + 2];   // the product's digits; dig[NC + 1] = 0 pads the last partial limb
  uint64_t acc;           // the running column sum

  // the product a_i b_j of column i + j (for a square only i < j)
  template <int I, int J>
  __device__ __forceinline__ void term(uint64_t &acc) const {
    if constexpr (!SQR || I < J) acc += (uint64_t)a[I] * b[J];
  }

  // column C: the products a_i b_{C-i} (for a square, then a_{C/2}^2)
  template <int C, int... I>
  __device__ __forceinline__ uint64_t col(uint64_t acc, std::integer_sequence<int, I...>) const {
    constexpr int LO = C - NB + 1 > 0 ? C - NB + 1 : 0;
    (term<LO + I, C - LO - I>(acc), ...);
    if constexpr (SQR && C % 2 == 0) acc += (uint64_t)a[C / 2] * a[C / 2];
    return acc;
  }

  // the limb that becomes complete with digit K (limb w needs digits
  // floor(32w/D) and the next one; D = 28 < 32 so at most one per digit)
  template <int K>
  __device__ __forceinline__ bool emit() {
    constexpr int W = (D * K + 31) / 32 - 1;   // last limb whose second digit is K ...
    constexpr int K0 = W >= 0 ? (32 * W) / D : 0;
    if constexpr (W < 0 || K0 + 1 != K) {
      return false;
    } else {
      constexpr int S = 32 * W - D * K0;   // 0 .. D-4
      static_assert(S + 32 <= 2 * D, "a limb spans at most two digits");
      if ((uint32_t)W >= len) return true;
      static_assert(K0 + 1 <= NC + 1, "digit index");
      const uint32_t v = (dig[K0] >> S) | (dig[K0 + 1] << (D - S));
      st_row(out, inst * 4u, v);
      out += stride;
      return (uint32_t)W + 1 >= len;   // every limb written: stop
    }
  }

  template <int C>
  __device__ __forceinline__ bool step() {
    constexpr int LO = C - NB + 1 > 0 ? C - NB + 1 : 0;
    constexpr int HI = C < NA - 1 ? C : NA - 1;
    acc = col<C>(acc, std::make_integer_sequence<int, HI - LO + 1>());
    dig[C] = (uint32_t)acc & MASK;
    acc >>= D;
    return emit<C>();
  }

  // columns OFF, OFF+1, ...; true once every output limb is written (the
  // columns after that are skipped)
  template <int OFF, int... C>
  __device__ __forceinline__ bool cols(std::integer_sequence<int, C...>) {
    return (step<OFF + C>() || ...);
  }

  // the top digit (< 2^D: the product has NC+1 digits) and the last limbs
  __device__ __forceinline__ bool finish() {
    dig[NC] = (uint32_t)acc;
    dig[NC + 1] = 0;
    return emit<NC>() || emit<NC + 1>();
  }

  // every column; true if the output ended inside the product
  __device__ __forceinline__ bool run() { return cols<0>(std::make_integer_sequence<int, NC>()) || finish(); }
};

// 32-bit limbs 0 .. NL-1 of an integer of len limbs at a lane's slot (limbs
// at or above len read as zero).  The row pointer is uniform (scalar) and
// the slot a 32-bit lane offset, so a load needs no vector address
// arithmetic; the branches are uniform, so every load is in flight at once.
template <int NL>
__device__ __forceinline__ void load_limbs(uint32_t (&w)[NL], const uint32_t *data, uint32_t stride, uint32_t inst,
                                           uint32_t len) {
  const uint32_t *row = data;
  const uint32_t off = inst * 4u;
#pragma unroll
  for (int j = 0; j < NL; j++) {
    w[j] = 0;
    if ((uint32_t)j < len) w[j] = ld_row(row, off);   // uniform branch
    row += stride;
  }
}

// digits 0 .. N-1 (radix 2^D) of the limbs w (the lane's digits start at
// bit 0, so every shift is a constant)
template <int N, int D, int NL>
__device__ __forceinline__ void limbs_to_digits(uint32_t (&d)[N], const uint32_t (&w)[NL]) {
  static_assert(NL == (D * N + 31) / 32, "limb count");
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int bit = D * k, j = bit >> 5, s = bit & 31;
    const uint32_t hi = j + 1 < NL ? w[j + 1] : 0u;
    d[k] = (s + D <= 32 ? (w[j] >> s) : __builtin_amdgcn_alignbit(hi, w[j], s)) & ((1u << D) - 1);
  }
}

// digits 0 .. N-1 (radix 2^D) of an integer of len 32-bit limbs, lane-private;
// limbs at or above len read as zero
template <int N, int D>
__device__ __forceinline__ void load_digits(uint32_t (&d)[N], const uint32_t *data, uint32_t stride, uint32_t inst,
                                            uint32_t len) {
  constexpr int NL = (D * N + 31) / 32;
  uint32_t w[NL];
  load_limbs<NL>(w, data, stride, inst, len);
  limbs_to_digits<N, D, NL>(d, w);
}

// Workgroups of blockDim.x lanes (64 or 256).  The hardware deals
// consecutive blocks to the eight XCDs in turn; with `xcd_order` block b
// takes tile b / 8 of XCD (b mod 8)'s share, the shares being contiguous
// runs of tiles in XCD order (the first nb mod 8 XCDs one tile more), so
// each XCD works through one contiguous range of instances (and a 256-lane
// block's four waves write 1 KB runs of each limb row).
template <int NA, int NB, int D, bool SQR>
__global__ void __launch_bounds__(256, 3) mul_cols_kernel(OutOperand p, Operand a, Operand b, uint32_t count,
                                                          bool xcd_order) {
  uint32_t tile = blockIdx.x;
  if (xcd_order) {
    const uint32_t nb = gridDim.x, x = tile & 7u, per = nb >> 3, rem = nb & 7u;
    tile = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + (tile >> 3);
  }
  // waves wholly past `count` have nothing to do (there is no barrier)
  if (tile * blockDim.x + (threadIdx.x & ~63u) >= count) return;
  // Lanes past `count` (in the last wave only) take instance count-1, which
  // is in the same wave: they compute and store exactly its limbs, so their
  // stores need no mask (the duplicate writes carry identical values).
  const uint32_t t = tile * blockDim.x + threadIdx.x;
  const uint32_t i = t < count ? t : count - 1;
  const uint32_t ai = in_index(a, i);
  const uint32_t pi = out_index(p, i);
  // instance slots are below 2^30 (the launcher checks the strides), so a
  // slot is a 32-bit byte offset from a row
  __builtin_assume(ai < (1u << 30));
  __builtin_assume(pi < (1u << 30));
  uint32_t da[NA], db[NB];
  load_digits<NA, D>(da, a.data, a.stride, ai, a.len);
  if constexpr (SQR) {
#pragma unroll
    for (int k = 0; k < NB; k++) db[k] = da[k] << 1;
  } else {
    const uint32_t bi = in_index(b, i);
    __builtin_assume(bi < (1u << 30));
    load_digits<NB, D>(db, b.data, b.stride, bi, b.len);
  }
  using CM = ColMul<NA, NB, D, SQR>;
  CM m{da, db, p.data, p.stride, pi, p.len, {}, 0};
  if (m.run()) return;
  // past the product: zero limbs
  for (uint32_t w = (uint32_t)CM::NOUT; w < p.len; w++) p.data[(size_t)w * p.stride + pi] = 0u;
}

template <int NA, int NB, bool SQR>
hipError_t launch_mul_cols(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  // instance slots must stay below 2^30 (every slot is below its stride)
  if (a.stride >= (1u << 30) || b.stride >= (1u << 30) || p.stride >= (1u << 30)) return hipErrorInvalidValue;
  // Batches of 2^18 and more: 256-lane workgroups in XCD order, measured A/B
  // on two MI355X boxes for 1,048,576 instances: 1024-bit squares 0.089-0.090
  // against 0.100-0.101 ms, 1024-bit products 0.120-0.122 against
  // 0.128-0.129, 2048-bit squares 0.207-0.212 against 0.214-0.217.  The
  // 2048 x 2048-bit product (three waves per SIMD) came out 1 % faster on
  // one box and 2 % slower on the other, so it keeps 64-lane workgroups, as
  // do smaller batches, which spread over more CUs that way (8,192 1024-bit
  // squares: 0.0066 against 0.0073 ms).
  const bool big = count >= (1u << 18) && (SQR || NA <= 37);
  const uint32_t lanes = big ? 256u : 64u;
  hipLaunchKernelGGL((mul_cols_kernel<NA, NB, 28, SQR>), dim3(div_round_up(count, lanes)), dim3(lanes), 0, stream, p,
                     a, b, count, big);
  return hipGetLastError();
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

namespace {

// register-resident row kernels, the longer operand (a) as the multiplicand:
// one lane per instance up to 1024 bits, then lane groups; b.len <= a.len
// keeps the multiplier within the accumulator bound.  SQR: b is a, and on
// the one-lane and 32-lane layouts each row takes the symmetric half of its
// products.  Measured A/B on one MI355X: 2^20 512-bit squares 0.035 against
// 0.038 ms, 1,024 16384-bit ones (32 x 19) 0.076 against 0.081, and the
// 65536-bit Karatsuba squares whose half products they are 0.87 against
// 0.91; the 2-, 4-, 8- and 16-lane groups keep full rows (4 x 37: 0.093
// against 0.087 ms for 65,536 4096-bit squares; 8 x 38 and 16 x 38 even).
// False when a is longer than every group holds.
template <bool SQR>
bool launch_rows(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream, hipError_t *e) {
  if (a.len <= 4) *e = launch_mul_lanes<1, 5, 28, SQR>(p, a, b, count, stream);          // 140 bits
  else if (a.len <= 8) *e = launch_mul_lanes<1, 10, 28, SQR>(p, a, b, count, stream);    // 280 bits
  else if (a.len <= 16) *e = launch_mul_lanes<1, 19, 28, SQR>(p, a, b, count, stream);   // 532 bits
  else if (a.len <= 32) *e = launch_mul_lanes<1, 37, 28>(p, a, b, count, stream);        // 1036 bits
  else if (a.len <= 48) *e = launch_mul_lanes<2, 28, 28>(p, a, b, count, stream);        // 56 digits >= 1536 bits
  else if (a.len <= 64) *e = launch_mul_lanes<2, 37, 28>(p, a, b, count, stream);        // 74 digits >= 2048 bits
  else if (a.len <= 128) *e = launch_mul_lanes<4, 37, 28>(p, a, b, count, stream);       // 148 digits >= 4096 bits
  // Above 4096 bits a batch of fewer than 4,096 products takes 32-lane groups
  // (count / 2 waves where 8 or 16 lanes leave SIMDs without a wave).
  // Measured on one MI355X: 1,024 x 16384-bit products 0.082 against 0.100
  // ms, 2,048 0.091 against 0.110, 1,024 x 8192-bit 0.038 against 0.048; at
  // 4,096 and above the extra lanes' per-row moves cost more than the waves
  // gain (16384 bits x 8,191: 0.23 against 0.20 ms; 16 x 19 digits at 8192
  // bits was slower than 8 x 38 at every batch from 8,191 to 16,383).
  else if (a.len <= 256)   // 304 digits >= 8192 bits
    *e = count < 4096 ? launch_mul_lanes<32, 10, 27, SQR>(p, a, b, count, stream)
                      : launch_mul_lanes<8, 38, 27>(p, a, b, count, stream);
  else if (a.len <= 512)   // 608 digits >= 16384 bits
    *e = count < 4096 ? launch_mul_lanes<32, 19, 27, SQR>(p, a, b, count, stream)
                      : launch_mul_lanes<16, 38, 27>(p, a, b, count, stream);
  else
    return false;
  return true;
}

}  // namespace

hipError_t launch_mul(OutOperand p, Operand a, Operand b, uint32_t count, hipStream_t stream) {
  // one lane per product with both operands in registers (product scanning)
  // for 513-1024-bit by 513-1024-bit and 1537-2048-bit by 513-2048-bit
  if (a.len > 16 && a.len <= 32 && b.len > 16) return launch_mul_cols<37, 37, false>(p, a, b, count, stream);
  if (a.len > 48 && a.len <= 64 && b.len > 16) {
    if (b.len <= 32) return launch_mul_cols<74, 37, false>(p, a, b, count, stream);
    return launch_mul_cols<74, 74, false>(p, a, b, count, stream);
  }
  hipError_t e;
  if (launch_rows<false>(p, a, b, count, stream, &e)) return e;
  hipLaunchKernelGGL(mul_general_kernel, dim3(div_round_up(count, 256)), dim3(256), 0, stream, p, a, b, count);
  return hipGetLastError();
}

hipError_t launch_sqr(OutOperand p, Operand a, uint32_t count, hipStream_t stream) {
  // symmetric product scanning: each product of two different digits once
  if (a.len > 16 && a.len <= 32) return launch_mul_cols<37, 37, true>(p, a, a, count, stream);
  if (a.len > 48 && a.len <= 64) return launch_mul_cols<74, 74, true>(p, a, a, count, stream);
  // symmetric rows on the one-lane and 32-lane layouts, full rows elsewhere
  hipError_t e;
  if (launch_rows<true>(p, a, a, count, stream, &e)) return e;
  return launch_mul(p, a, a, count, stream);
}

// ---- Karatsuba above the register kernels --------------------------------
// The reference has a register-level Karatsuba (regmp/kar.h:24-207) that no
// instantiated variant enables (powm.cu:67-70).  Here Karatsuba splits the
// operands of a whole batch beyond the lane-group row kernels (more than 512
// limbs, where the fallback is the schoolbook with its row accumulator in HBM):
// with a = a1 B + a0, b = b1 B + b0 and B = 2^(32h), h = ceil(len(a)/2),
//   a b = z2 B^2 + (z1 - z0 - z2) B + z0,  z0 = a0 b0, z2 = a1 b1,
//   z1 = (a0 + a1)(b0 + b1).
// The sums have h limbs and a carry bit each, sa = sa' + ca B and
// sb = sb' + cb B, so z1 = sa' sb' + (ca sb' + cb sa') B + ca cb B^2: the
// product stays at h limbs (within the row kernels up to 1,024-limb
// operands) and the carry terms join the others.  Three half-size products
// of the whole batch (recursively, until they fit the row kernels), two sums
// and a recombination that adds the shifted terms with a signed carry (the
// sums run on it too; each is a chunk pass and a carry pass).  When b has
// no more than h limbs, a alone is split: a b = (a1 b) B + a0 b.  The terms,
// and the chunks' carry-outs, live in a workspace, instance i in column i of
// rows of stride round_up(count, 64).

namespace {

struct KarTerm {
  const uint32_t *data;     // term limb k of instance i: data[k * stride + column(i)]
  const uint32_t *factor;   // null, or a row whose limb i (0 or 1) scales the term
  uint32_t len, shift;      // limbs, and the term's offset in limbs
  int32_t sign;             // +1 or -1
  uint32_t stride;
  // column(i): i for the workspace's terms; for an input operand's half,
  // the operand's own gather (idx[i % idx_count] % count, or i % count)
  const uint32_t *idx;
  uint32_t idx_count, count;
};
constexpr int kKarTerms = 8;
struct KarTerms {
  KarTerm t[kKarTerms];
  uint32_t n;
};

__device__ __forceinline__ uint32_t kar_column(const KarTerm &t, uint32_t i) {
  if (t.idx) return t.idx[i % t.idx_count] % t.count;
  return t.count ? i % t.count : i;
}

// p = sum of the shifted terms, truncated to p.len limbs, in two passes.
// The first sums chunks of kKarChunk limbs: a 64-lane wave per 64 instances
// (every access a coalesced row) and chunk, so that a small batch's many
// limbs still fill the GPU.  Whether a term covers a limb of the chunk is
// the same for the whole wave, so every load is issued unconditionally (at
// a clamped, in-bounds limb, masked to zero where the term does not cover
// it) and a lane has all of its chunk's loads in flight at once; it then
// sums the limbs with a signed carry (the arithmetic shift, which clang
// defines for signed values, floors it) and leaves the chunk's carry-out in
// `carries`.  The second, one lane per instance, walks the chunks from the
// bottom: the carry entering a chunk is added at its first limb (it crosses
// the chunk only if all of the chunk's limbs are ones, or zeros for a
// borrow) and what leaves it joins that chunk's own carry-out.  The sum is
// known to fit `limbs` limbs (the product, or a half sum): limbs from there
// up are written as zeros without summing the terms (ca cb B^2 can reach one
// limb past; the carries below cancel it) and the carry past limb
// limbs - 1 is dropped, which leaves the sum mod 2^(32 limbs), that is the
// sum itself.
constexpr uint32_t kKarChunk = 16;
// Two waves per SIMD: the eight-term form holds its 128 loads in 256 VGPRs
// (a few values spill); measured on one MI355X, 0.48 ms for 4,096 products of
// 32,768 x 32,768 bits against 0.50 at the compiler's one wave, and 1.13 ms
// against 1.15 for 1,024 of 65,536 bits.
template <int NT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) kar_chunk_kernel(OutOperand p, KarTerms c, uint32_t count, int32_t *carries,
                                                       uint32_t cstride, uint32_t limbs, uint32_t groups) {
  // block b: instances 64 (b mod groups) on, chunk b / groups
  const uint32_t chunk = blockIdx.x / groups;
  const uint32_t i = (blockIdx.x - chunk * groups) * 64u + threadIdx.x;
  if (i >= count) return;
  const uint32_t k0 = chunk * kKarChunk, k1 = min(k0 + kKarChunk, p.len);
  uint32_t *out = p.data + out_index(p, i);
  if (k0 >= limbs) {
    for (uint32_t k = k0; k < k1; k++) out[(size_t)k * p.stride] = 0u;
    return;
  }
  int64_t acc[kKarChunk];
#pragma unroll
  for (uint32_t kk = 0; kk < kKarChunk; kk++) acc[kk] = 0;
#pragma unroll
  for (int j = 0; j < NT; j++) {
    const KarTerm &t = c.t[j];
    // the factor is a carry limb, 0 or 1
    const uint32_t keep = (t.factor && t.factor[i] == 0u) ? 0u : 0xFFFFFFFFu;
    const uint32_t *d = t.data + kar_column(t, i);
#pragma unroll
    for (uint32_t kk = 0; kk < kKarChunk; kk++) {
      const uint32_t k = k0 + kk;
      const bool in = k >= t.shift && k - t.shift < t.len;
      const uint32_t x = d[(size_t)(in ? k - t.shift : 0u) * t.stride] & (in ? keep : 0u);
      acc[kk] += t.sign > 0 ? (int64_t)x : -(int64_t)x;
    }
  }
  const uint32_t ks = min(k1, limbs);
  int64_t carry = 0;
#pragma unroll
  for (uint32_t kk = 0; kk < kKarChunk; kk++) {
    if (k0 + kk < ks) {
      const int64_t v = acc[kk] + carry;
      out[(size_t)(k0 + kk) * p.stride] = (uint32_t)v;
      carry = v >> 32;
    } else if (k0 + kk < k1) {
      out[(size_t)(k0 + kk) * p.stride] = 0u;
    }
  }
  carries[(size_t)chunk * cstride + i] = (int32_t)carry;
}

// the chunks' carry-outs and the first limbs they enter are fetched kKarBatch
// chunks at a time; the ripple beyond a chunk's first limb, rare, goes
// through memory
constexpr uint32_t kKarBatch = 16;
__global__ void __launch_bounds__(64) kar_carry_kernel(OutOperand p, uint32_t count, const int32_t *carries,
                                                       uint32_t cstride, uint32_t limbs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t *out = p.data + out_index(p, i);
  const uint32_t top = min(p.len, limbs);
  const uint32_t nchunks = (top + kKarChunk - 1) / kKarChunk;
  int64_t cin = 0;
  // carry-out of chunk ch enters chunk ch + 1, for ch + 1 < nchunks
  for (uint32_t base = 0; base + 1 < nchunks; base += kKarBatch) {
    int32_t co[kKarBatch];
    uint32_t first[kKarBatch];
#pragma unroll
    for (uint32_t j = 0; j < kKarBatch; j++) {
      const uint32_t ch = min(base + j, nchunks - 2);   // in bounds; unused past the end
      co[j] = carries[(size_t)ch * cstride + i];
      first[j] = out[(size_t)(ch + 1) * kKarChunk * p.stride];
    }
#pragma unroll
    for (uint32_t j = 0; j < kKarBatch; j++) {
      const uint32_t ch = base + j;
      if (ch + 1 < nchunks) {
        cin += co[j];
        const uint32_t k = (ch + 1) * kKarChunk;
        const int64_t x = (int64_t)first[j] + cin;
        out[(size_t)k * p.stride] = (uint32_t)x;
        cin = x >> 32;
        const uint32_t k1 = min(k + kKarChunk, top);
        for (uint32_t kk = k + 1; kk < k1 && cin != 0; kk++) {
          const int64_t y = (int64_t)out[(size_t)kk * p.stride] + cin;
          out[(size_t)kk * p.stride] = (uint32_t)y;
          cin = y >> 32;
        }
      }
    }
  }
}

// workspace limbs per instance column for a product of a_len x b_len limbs
// (a_len >= b_len): this level's terms, then the largest child's
size_t kar_limbs(uint32_t a_len, uint32_t b_len) {
  if (a_len < kKaratsubaMinLimbs) return 0;
  const uint32_t h = (a_len + 1) / 2;
  const size_t rows = (a_len + b_len + kKarChunk - 1) / kKarChunk;   // chunk carries
  if (b_len <= h) {
    const size_t own = rows + (size_t)(h + b_len) + (a_len - h + b_len);
    return own + std::max(kar_limbs(h, b_len), kar_limbs(std::max(a_len - h, b_len), std::min(a_len - h, b_len)));
  }
  const size_t own = rows + 2 * (size_t)(h + 1) + 2 * (size_t)h + 2 * (size_t)h + (a_len - h) + (b_len - h);
  return own + std::max(kar_limbs(h, h), kar_limbs(a_len - h, b_len - h));
}

}  // namespace

size_t mul_karatsuba_words(uint32_t a_len, uint32_t b_len, uint32_t count) {
  return kar_limbs(std::max(a_len, b_len), std::min(a_len, b_len)) * round_up(count, kStrideAlign);
}

hipError_t launch_mul_ws(OutOperand p, Operand a, Operand b, uint32_t count, bool square, uint32_t *ws,
                         hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (a.len < b.len) std::swap(a, b);
  if (a.len < kKaratsubaMinLimbs) return square ? launch_sqr(p, a, count, stream) : launch_mul(p, a, b, count, stream);
  const uint32_t st = round_up(count, kStrideAlign);
  const uint32_t h = (a.len + 1) / 2;
  auto take = [&](uint32_t limbs) {
    uint32_t *d = ws;
    ws += (size_t)limbs * st;
    return d;
  };
  auto in = [&](const uint32_t *d, uint32_t len) { return Operand{d, len, st, count, nullptr, 0}; };
  auto out = [&](uint32_t *d, uint32_t len) { return OutOperand{d, len, st, nullptr}; };
  Operand a0 = a, a1 = a;
  a0.len = h;
  a1.data = a.data + (size_t)h * a.stride;
  a1.len = a.len - h;
  // a workspace term, and an input operand's half with its own gather
  auto term = [&](const uint32_t *d, const uint32_t *factor, uint32_t len, uint32_t shift, int32_t sign) {
    return KarTerm{d, factor, len, shift, sign, st, nullptr, 0, 0};
  };
  auto half = [&](const Operand &o) { return KarTerm{o.data, nullptr, o.len, 0, 1, o.stride, o.idx, o.idx_count, o.count}; };
  // the chunks' carry-outs, one row per chunk of the sum's `limbs` (the
  // product, or a half sum), in this level's workspace
  int32_t *carries = reinterpret_cast<int32_t *>(take(div_round_up(a.len + b.len, kKarChunk)));
  auto combine = [&](OutOperand dst, const KarTerms &c, uint32_t limbs) {
    const uint32_t groups = div_round_up(count, 64);
    const dim3 grid(groups * div_round_up(dst.len, kKarChunk));
    if (c.n == 2)
      hipLaunchKernelGGL(kar_chunk_kernel<2>, grid, dim3(64), 0, stream, dst, c, count, carries, st, limbs, groups);
    else
      hipLaunchKernelGGL(kar_chunk_kernel<kKarTerms>, grid, dim3(64), 0, stream, dst, c, count, carries, st, limbs,
                         groups);
    hipLaunchKernelGGL(kar_carry_kernel, dim3(groups), dim3(64), 0, stream, dst, count, carries, st, limbs);
    return hipGetLastError();
  };
  KarTerms c{};
  hipError_t e;
  if (b.len <= h) {
    uint32_t *z0 = take(h + b.len), *z2 = take(a1.len + b.len);
    if ((e = launch_mul_ws(out(z0, h + b.len), a0, b, count, false, ws, stream)) != hipSuccess) return e;
    if ((e = launch_mul_ws(out(z2, a1.len + b.len), a1, b, count, false, ws, stream)) != hipSuccess) return e;
    c.t[0] = term(z0, nullptr, h + b.len, 0, 1);
    c.t[1] = term(z2, nullptr, a1.len + b.len, h, 1);
    c.n = 2;
  } else {
    Operand b0 = b, b1 = b;
    b0.len = h;
    b1.data = b.data + (size_t)h * b.stride;
    b1.len = b.len - h;
    // the sums with their carry limb (row h), then the three products
    uint32_t *sa = take(h + 1), *sbuf = take(h + 1);
    uint32_t *sb = square ? sa : sbuf;
    uint32_t *z0 = take(2 * h), *z1 = take(2 * h), *z2 = take(a1.len + b1.len);
    const uint32_t *ca = sa + (size_t)h * st, *cb = sb + (size_t)h * st;
    // the sums on the combine kernel too: its chunks fill the GPU where the
    // add kernel's one instance per lane would leave a small batch latency-bound
    KarTerms s{};
    s.t[0] = half(a0);
    s.t[1] = half(a1);
    s.n = 2;
    if ((e = combine(out(sa, h + 1), s, h + 1)) != hipSuccess) return e;
    if (!square) {
      s.t[0] = half(b0);
      s.t[1] = half(b1);
      if ((e = combine(out(sb, h + 1), s, h + 1)) != hipSuccess) return e;
    }
    if ((e = launch_mul_ws(out(z0, 2 * h), a0, b0, count, square, ws, stream)) != hipSuccess) return e;
    if ((e = launch_mul_ws(out(z2, a1.len + b1.len), a1, b1, count, square, ws, stream)) != hipSuccess) return e;
    if ((e = launch_mul_ws(out(z1, 2 * h), in(sa, h), in(sb, h), count, square, ws, stream)) != hipSuccess) return e;
    c.t[0] = term(z0, nullptr, 2 * h, 0, 1);
    c.t[1] = term(z1, nullptr, 2 * h, h, 1);
    c.t[2] = term(sb, ca, h, 2 * h, 1);    // ca sb' B
    c.t[3] = term(sa, cb, h, 2 * h, 1);    // cb sa' B
    c.t[4] = term(ca, cb, 1, 3 * h, 1);    // ca cb B^2
    c.t[5] = term(z0, nullptr, 2 * h, h, -1);
    c.t[6] = term(z2, nullptr, a1.len + b1.len, h, -1);
    c.t[7] = term(z2, nullptr, a1.len + b1.len, 2 * h, 1);
    c.n = 8;
  }
  return combine(p, c, a.len + b.len);
}

}  // namespace xmp
