/*
 * oracle.c -- CPU restatement of XMP's batched big-integer arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the smoke
 * entry point and the CPU-baseline leg of the benchmark may load it; the
 * product library never does.  It is deliberately simple scalar C.
 *
 * Conventions (reference src/include/xmp_internal.h:27): limbs are 32-bit,
 * least significant limb first.  An integer of precision P bits is held in
 * n = P/32 limbs.  All functions take explicit limb counts for every operand
 * so the zero-extension and truncation rules of the reference are explicit.
 *
 * Each function cites the reference file:line whose behaviour it restates.
 * The one deliberate difference from the reference is the final canonical
 * reduction of powm (results are always < m, as GMP's mpz_powm gives); the
 * reference's Montgomery tail can leave m itself for inputs congruent to 0
 * (src/include/warpmp/mulmod_*.h tail, src/include/powm/warp_distributed.h:212-223).
 *
 * Pinning (tests/test_oracle.py): every function here is checked against
 * GMP (the reference test suite's own oracle, unit_tests/test.cu:1214,
 * 1616, 1780, 2317-2322, 2784-2813), against Python's int, against the
 * committed fixtures in tests/golden/, and against closed forms.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t limb_t;
typedef uint64_t dlimb_t;

#define XO_API __attribute__((visibility("default")))

static inline limb_t get_limb(const limb_t *x, int xn, int i) { return i < xn ? x[i] : 0; }

/* s = (a + b) mod 2^(32*sn).  Inputs are zero-extended to sn limbs and the
 * carry lands in limb max(an,bn) when sn is larger, limbs above are zero.
 * Restates src/include/implementations.h:86-100 and digitmp/add.h:24-41. */
XO_API void xo_add(limb_t *s, int sn, const limb_t *a, int an, const limb_t *b, int bn) {
  dlimb_t carry = 0;
  for (int i = 0; i < sn; i++) {
    dlimb_t t = (dlimb_t)get_limb(a, an, i) + get_limb(b, bn, i) + carry;
    s[i] = (limb_t)t;
    carry = t >> 32;
  }
}

/* d = (a - b) mod 2^(32*dn): the borrow propagates through all dn limbs, so a
 * negative difference wraps (two's complement).  Restates
 * src/include/implementations.h:195-206 and digitmp/sub.h:24-41; the
 * reference test builds the same value as mpz_sub + 2^P (unit_tests/test.cu:2317-2322). */
XO_API void xo_sub(limb_t *d, int dn, const limb_t *a, int an, const limb_t *b, int bn) {
  dlimb_t borrow = 0;
  for (int i = 0; i < dn; i++) {
    dlimb_t t = (dlimb_t)get_limb(a, an, i) - get_limb(b, bn, i) - borrow;
    d[i] = (limb_t)t;
    borrow = (t >> 32) & 1;
  }
}

/* p = (a * b) mod 2^(32*pn): schoolbook product truncated to the output's
 * precision.  Restates digitmp/mul.h:24-55 (output truncated to p.digits()). */
XO_API void xo_mul(limb_t *p, int pn, const limb_t *a, int an, const limb_t *b, int bn) {
  limb_t *t = (limb_t *)calloc((size_t)(an + bn + 1), sizeof(limb_t));
  for (int i = 0; i < an; i++) {
    dlimb_t carry = 0;
    for (int j = 0; j < bn; j++) {
      dlimb_t x = (dlimb_t)a[i] * b[j] + t[i + j] + carry;
      t[i + j] = (limb_t)x;
      carry = x >> 32;
    }
    t[i + bn] = (limb_t)carry;
  }
  for (int i = 0; i < pn; i++) p[i] = i < an + bn ? t[i] : 0;
  free(t);
}

static int nlz32(limb_t x) {
  int n = 0;
  if (x == 0) return 32;
  while (!(x & 0x80000000u)) { x <<= 1; n++; }
  return n;
}

static int top_nonzero(const limb_t *x, int xn) {
  int i = xn - 1;
  while (i >= 0 && x[i] == 0) i--;
  return i + 1; /* number of significant limbs */
}

/* Long division q = floor(a/b), r = a mod b (Knuth vol.2 4.3.1 algorithm D on
 * 32-bit limbs with a normalised divisor).  Restates the normalise /
 * estimate / multiply-subtract / add-back structure of digitmp/div.h:108-210
 * (div) and :319-423 (rem).  q is truncated to qn limbs and r to rn limbs.
 * Division by zero returns -1 and writes q = 0, r = a (truncated): the
 * reference leaves the output unwritten (digitmp/div.h:120-122); this build
 * defines the value so that a = q*b + r still holds. */
XO_API int xo_divmod(limb_t *q, int qn, limb_t *r, int rn,
                     const limb_t *a, int an, const limb_t *b, int bn) {
  int m = top_nonzero(b, bn);
  int na = top_nonzero(a, an);
  if (m == 0) {
    if (q) for (int i = 0; i < qn; i++) q[i] = 0;
    if (r) for (int i = 0; i < rn; i++) r[i] = get_limb(a, an, i);
    return -1;
  }
  limb_t *qq = (limb_t *)calloc((size_t)(na + 1), sizeof(limb_t));
  limb_t *rr = (limb_t *)calloc((size_t)(m + 1), sizeof(limb_t));
  if (na < m) {
    for (int i = 0; i < na; i++) rr[i] = a[i];
  } else if (m == 1) {
    dlimb_t rem = 0;
    for (int i = na - 1; i >= 0; i--) {
      dlimb_t cur = (rem << 32) | a[i];
      qq[i] = (limb_t)(cur / b[0]);
      rem = cur % b[0];
    }
    rr[0] = (limb_t)rem;
  } else {
    int s = nlz32(b[m - 1]);
    limb_t *vn = (limb_t *)calloc((size_t)m, sizeof(limb_t));
    limb_t *un = (limb_t *)calloc((size_t)(na + 1), sizeof(limb_t));
    for (int i = m - 1; i > 0; i--) vn[i] = s ? (b[i] << s) | (b[i - 1] >> (32 - s)) : b[i];
    vn[0] = b[0] << s;
    un[na] = s ? a[na - 1] >> (32 - s) : 0;
    for (int i = na - 1; i > 0; i--) un[i] = s ? (a[i] << s) | (a[i - 1] >> (32 - s)) : a[i];
    un[0] = a[0] << s;
    for (int j = na - m; j >= 0; j--) {
      dlimb_t num = ((dlimb_t)un[j + m] << 32) | un[j + m - 1];
      dlimb_t qhat = num / vn[m - 1];
      dlimb_t rhat = num % vn[m - 1];
      while (qhat >= ((dlimb_t)1 << 32) ||
             qhat * vn[m - 2] > ((rhat << 32) | un[j + m - 2])) {
        qhat--;
        rhat += vn[m - 1];
        if (rhat >= ((dlimb_t)1 << 32)) break;
      }
      /* multiply and subtract */
      int64_t borrow = 0;
      dlimb_t carry = 0;
      for (int i = 0; i < m; i++) {
        dlimb_t p = qhat * vn[i] + carry;
        carry = p >> 32;
        int64_t t = (int64_t)un[i + j] - (int64_t)(limb_t)p - borrow;
        un[i + j] = (limb_t)t;
        borrow = t < 0 ? 1 : 0;
      }
      int64_t t = (int64_t)un[j + m] - (int64_t)carry - borrow;
      un[j + m] = (limb_t)t;
      if (t < 0) { /* add back */
        qhat--;
        dlimb_t c = 0;
        for (int i = 0; i < m; i++) {
          dlimb_t x = (dlimb_t)un[i + j] + vn[i] + c;
          un[i + j] = (limb_t)x;
          c = x >> 32;
        }
        un[j + m] += (limb_t)c;
      }
      qq[j] = (limb_t)qhat;
    }
    for (int i = 0; i < m; i++) rr[i] = s ? (un[i] >> s) | (un[i + 1] << (32 - s)) : un[i];
    free(vn);
    free(un);
  }
  if (q) for (int i = 0; i < qn; i++) q[i] = i <= na ? qq[i] : 0;
  if (r) for (int i = 0; i < rn; i++) r[i] = i < m ? rr[i] : 0;
  free(qq);
  free(rr);
  return 0;
}

/* np0 = -m0^{-1} mod 2^32 by Newton iteration; restates src/include/regmp/np0.h:24-33. */
XO_API limb_t xo_np0(limb_t m0) {
  limb_t inv = m0;            /* correct to 3 bits for odd m0 */
  for (int i = 0; i < 4; i++) inv = inv * (2u - m0 * inv);
  return (limb_t)(0u - inv);
}

/* Window width chosen from the exponent precision; restates
 * src/include/powm_operators.h:15-31. */
XO_API int xo_window_bits(uint32_t exp_bits) {
  if (exp_bits < 400) return 4;
  if (exp_bits < 800) return 5;
  if (exp_bits < 1600) return 6;
  if (exp_bits <= 4096) return 7;
  return 8;
}

/* Montgomery product r = a*b*2^(-32n) mod m (CIOS), almost reduced: inputs
 * and output are < 2^(32n) and the output is only congruent mod m.
 * Restates the arithmetic of src/include/warpmp/mulmod_imad.h:25-183 (a
 * distributed CIOS there; scalar here) including the conditional -m when the
 * accumulator overflows 2^(32n) (mulmod_imad.h tail). */
static void mont_mul(limb_t *r, const limb_t *a, const limb_t *b, const limb_t *m, limb_t np0, int n, limb_t *t) {
  memset(t, 0, sizeof(limb_t) * (size_t)(n + 2));
  for (int i = 0; i < n; i++) {
    dlimb_t c = 0;
    for (int j = 0; j < n; j++) {
      dlimb_t x = (dlimb_t)a[j] * b[i] + t[j] + c;
      t[j] = (limb_t)x;
      c = x >> 32;
    }
    dlimb_t x = (dlimb_t)t[n] + c;
    t[n] = (limb_t)x;
    t[n + 1] = (limb_t)(x >> 32);
    limb_t qd = t[0] * np0;
    x = (dlimb_t)qd * m[0] + t[0];
    c = x >> 32;
    for (int j = 1; j < n; j++) {
      x = (dlimb_t)qd * m[j] + t[j] + c;
      t[j - 1] = (limb_t)x;
      c = x >> 32;
    }
    x = (dlimb_t)t[n] + c;
    t[n - 1] = (limb_t)x;
    t[n] = t[n + 1] + (limb_t)(x >> 32);
    t[n + 1] = 0;
  }
  if (t[n]) { /* t >= 2^(32n): subtract m once, result < 2^(32n) */
    dlimb_t borrow = 0;
    for (int j = 0; j < n; j++) {
      dlimb_t x = (dlimb_t)t[j] - m[j] - borrow;
      t[j] = (limb_t)x;
      borrow = (x >> 32) & 1;
    }
  }
  memcpy(r, t, sizeof(limb_t) * (size_t)n);
}

static int geq(const limb_t *a, const limb_t *b, int n) {
  for (int i = n - 1; i >= 0; i--) {
    if (a[i] != b[i]) return a[i] > b[i];
  }
  return 1;
}

/* out = base^exp mod m for one instance.
 *   n limbs for base, out and m; en limbs for the exponent (the exponent's
 *   precision may differ, samples/crt_rsa/rsa.cc:141).
 * Fixed-window, left-to-right exponentiation over ALL 32*en exponent bits
 * (leading zeros included) with a full 2^w table whose slot 0 is the
 * Montgomery one; restates fwe<> in src/include/powm/powm.h:39-127:
 *   slot0 = R mod m                 (warp_distributed.h:66-79, computeRModuloN)
 *   slot1 = base*R mod m            (AR kernel, powm_implementations.h:3-87)
 *   slot2 = slot1^2, odd i = slot(i-2)*slot2, even 2j = slot(j)^2  (powm.h:76-107)
 *   then per window: w squarings and one multiply by slot[bits] (powm.h:59-123)
 *   and finally a multiply by one to leave Montgomery form.
 * Even moduli are outside the reference's contract (docs/xmp.docx para 192) and
 * return -1 with out = 0.  m == 1 gives 0 (as GMP). */
XO_API int xo_powm(limb_t *out, const limb_t *base, const limb_t *exp, int en, const limb_t *m, int n) {
  if (!(m[0] & 1)) {
    for (int i = 0; i < n; i++) out[i] = 0;
    return -1;
  }
  int bits = 32 * en;
  int w = xo_window_bits((uint32_t)bits);
  int tsize = 1 << w;
  limb_t np0 = xo_np0(m[0]);
  limb_t *table = (limb_t *)calloc((size_t)tsize * (size_t)n, sizeof(limb_t));
  limb_t *t = (limb_t *)calloc((size_t)(n + 2), sizeof(limb_t));
  limb_t *cur = (limb_t *)calloc((size_t)n, sizeof(limb_t));
  limb_t *num = (limb_t *)calloc((size_t)(2 * n + 1), sizeof(limb_t));
  limb_t *one = (limb_t *)calloc((size_t)n, sizeof(limb_t));

  /* slot0 = 2^(32n) mod m, slot1 = base * 2^(32n) mod m, both by long division */
  num[n] = 1;
  xo_divmod(NULL, 0, table, n, num, n + 1, m, n);
  memset(num, 0, sizeof(limb_t) * (size_t)(2 * n + 1));
  for (int i = 0; i < n; i++) num[n + i] = base[i];
  xo_divmod(NULL, 0, table + n, n, num, 2 * n, m, n);

  if (tsize > 2) mont_mul(table + 2 * n, table + n, table + n, m, np0, n, t);
  for (int i = 3; i < tsize; i += 2) mont_mul(table + (size_t)i * n, table + (size_t)(i - 2) * n, table + 2 * (size_t)n, m, np0, n, t);
  for (int i = 4; i < tsize; i += 2) mont_mul(table + (size_t)i * n, table + (size_t)(i / 2) * n, table + (size_t)(i / 2) * n, m, np0, n, t);

  /* top window holds bits % w bits (or w when w divides bits) */
  int offset = (bits % w == 0) ? w : bits % w;
  int pos = bits - offset;
#define EXPBITS(p, len) ((uint32_t)(((((dlimb_t)get_limb(exp, en, ((p) >> 5) + 1) << 32) | get_limb(exp, en, (p) >> 5)) >> ((p) & 31)) & ((1u << (len)) - 1)))
  memcpy(cur, table + (size_t)EXPBITS(pos, offset) * n, sizeof(limb_t) * (size_t)n);
  while (pos > 0) {
    pos -= w;
    for (int k = 0; k < w; k++) mont_mul(cur, cur, cur, m, np0, n, t);
    mont_mul(cur, cur, table + (size_t)EXPBITS(pos, w) * n, m, np0, n, t);
  }
#undef EXPBITS
  one[0] = 1;
  mont_mul(cur, cur, one, m, np0, n, t);
  if (geq(cur, m, n)) { /* canonical result < m (GMP semantics) */
    dlimb_t borrow = 0;
    for (int j = 0; j < n; j++) {
      dlimb_t x = (dlimb_t)cur[j] - m[j] - borrow;
      cur[j] = (limb_t)x;
      borrow = (x >> 32) & 1;
    }
  }
  memcpy(out, cur, sizeof(limb_t) * (size_t)n);
  free(table); free(t); free(cur); free(num); free(one);
  return 0;
}

/* Word repack used by import/export; restates xmpTransform
 * (src/xmp.cu:461-489): output word j of integer i is input word j (order -1,
 * least significant first) or words-1-j (order +1) of integer in_idx, byte
 * swapped when endian is +1 (big endian) and masked to the low size*8-nails
 * bits.  in_idx = in_indices[i % in_count] when given, out_idx likewise. */
XO_API void xo_transform(void *output, const void *input, uint32_t count, uint32_t words, int32_t order,
                         int32_t endian, uint32_t nails, uint32_t size,
                         const uint32_t *out_indices, uint32_t out_count,
                         const uint32_t *in_indices, uint32_t in_count) {
  for (uint32_t i = 0; i < count; i++) {
    uint32_t in_idx = in_indices ? in_indices[i % in_count] : i;
    uint32_t out_idx = out_indices ? out_indices[i % out_count] : i;
    for (uint32_t j = 0; j < words; j++) {
      uint32_t off = (order == -1) ? j : words - j - 1;
      const uint8_t *src = (const uint8_t *)input + ((size_t)in_idx * words + off) * size;
      uint8_t *dst = (uint8_t *)output + ((size_t)out_idx * words + j) * size;
      uint64_t w = 0;
      for (uint32_t k = 0; k < size; k++) {
        uint32_t byte = (endian == -1) ? k : size - 1 - k;
        w |= (uint64_t)src[byte] << (8 * k);
      }
      uint64_t mask = (size == 8 ? ~0ull : ((1ull << (8 * size)) - 1)) >> nails;
      w &= mask;
      for (uint32_t k = 0; k < size; k++) dst[k] = (uint8_t)(w >> (8 * k));
    }
  }
}

/* ---- bitwise operations (reference src/logic.cu, kernels in
 * src/include/implementations.h:766-1067).  Inputs are zero-extended to the
 * output length; outputs longer than both inputs are zero-filled (ior, and,
 * xor) or one-filled (not). ---- */

/* c = a | b (implementations.h:850-894). */
XO_API void xo_ior(limb_t *c, int cn, const limb_t *a, int an, const limb_t *b, int bn) {
  for (int i = 0; i < cn; i++) c[i] = get_limb(a, an, i) | get_limb(b, bn, i);
}

/* c = a & b (implementations.h:896-936). */
XO_API void xo_and(limb_t *c, int cn, const limb_t *a, int an, const limb_t *b, int bn) {
  for (int i = 0; i < cn; i++) c[i] = get_limb(a, an, i) & get_limb(b, bn, i);
}

/* c = a ^ b (implementations.h:938-982). */
XO_API void xo_xor(limb_t *c, int cn, const limb_t *a, int an, const limb_t *b, int bn) {
  for (int i = 0; i < cn; i++) c[i] = get_limb(a, an, i) ^ get_limb(b, bn, i);
}

/* c = ~a over cn limbs, limbs above a's length all ones (implementations.h:984-1018,
 * unit_tests/test.cu:1480-1488). */
XO_API void xo_not(limb_t *c, int cn, const limb_t *a, int an) {
  for (int i = 0; i < cn; i++) c[i] = i < an ? ~a[i] : ~(limb_t)0;
}

/* c = a * 2^shift mod 2^(32 cn) for shift >= 0, floor(a / 2^-shift) for
 * shift < 0 (implementations.h:1020-1067; the reference test builds the
 * same value with mpz_mul_2exp / mpz_tdiv_q_2exp, unit_tests/test.cu:1366-1372).
 * Output limb i holds bits [32 i - shift, 32 i - shift + 32) of a. */
XO_API void xo_shf(limb_t *c, int cn, const limb_t *a, int an, int32_t shift) {
  for (int i = 0; i < cn; i++) {
    limb_t v = 0;
    for (int k = 0; k < 32; k++) {
      int64_t bit = 32 * (int64_t)i + k - (int64_t)shift;
      if (bit >= 0 && bit < 32 * (int64_t)an && ((a[bit >> 5] >> (bit & 31)) & 1u)) v |= (limb_t)1 << k;
    }
    c[i] = v;
  }
}

/* number of one bits (implementations.h:821-848). */
XO_API uint32_t xo_popc(const limb_t *a, int an) {
  uint32_t n = 0;
  for (int i = 0; i < an; i++)
    for (limb_t x = a[i]; x; x &= x - 1) n++;
  return n;
}

/* -1, 0, +1 as a <, ==, > b, with the shorter operand zero-extended
 * (implementations.h:766-819; the reference test compares sgn of mpz_cmp,
 * unit_tests/test.cu:2091-2098). */
XO_API int32_t xo_cmp(const limb_t *a, int an, const limb_t *b, int bn) {
  int n = an > bn ? an : bn;
  for (int i = n - 1; i >= 0; i--) {
    limb_t x = get_limb(a, an, i), y = get_limb(b, bn, i);
    if (x != y) return x > y ? 1 : -1;
  }
  return 0;
}

/* ---- batch drivers over compact arrays ([count][limbs]) with the reference's
 * broadcast rule: input instance i reads element i % array_count
 * (src/include/implementations.h:58,68 and docs/xmp.docx para 15-17). ---- */

XO_API void xo_powm_batch(limb_t *out, uint32_t count, const limb_t *base, uint32_t base_count,
                          const limb_t *exp, uint32_t exp_count, int en,
                          const limb_t *mod, uint32_t mod_count, int n) {
  for (uint32_t i = 0; i < count; i++)
    xo_powm(out + (size_t)i * n, base + (size_t)(i % base_count) * n, exp + (size_t)(i % exp_count) * en, en,
            mod + (size_t)(i % mod_count) * n, n);
}
