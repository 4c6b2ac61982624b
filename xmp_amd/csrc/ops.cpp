// ops.cpp -- math entry points: add, sub, mul, sqr, div, mod, divmod, powm.
//
// Validation order and error codes follow the reference (add_sub.cu:35-142,
// mul.cu:26-303, div.cu:35-517, powm.cu:250-307).  In-place operation (the
// output aliases an input) goes through the handle's scratch buffer as in the
// reference; the reference's in-place DivMod stores its scratch pointers in
// the index fields (div.cu:429-435) -- here it uses scratch outputs.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"
#include "powm.hpp"
#include "tuner.hpp"

using namespace xmp;

namespace {

inline size_t out_bytes(xmpIntegers_t x) { return sizeof(uint32_t) * (size_t)x->nlimbs * x->stride; }

// Common prologue of the two-input one-output ops.
xmpError_t check_binary(xmpHandle_t h, xmpIntegers_t o, xmpIntegers_t a, xmpIntegers_t b, uint32_t count) {
  if (o->device != h->device || a->device != h->device || b->device != h->device) return xmpErrorInvalidDevice;
  return xmpErrorSuccess;
}

xmpError_t finish(xmpHandle_t h, bool inplace, xmpIntegers_t o, uint32_t *buffer) {
  if (inplace &&
      hipMemcpyAsync(o->limbs, buffer, out_bytes(o), hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
    return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t sync(xmpHandle_t h, xmpError_t e) {
  if (e != xmpErrorSuccess) return e;
  XMP_SET_DEVICE(h);
  (void)hipStreamSynchronize(h->stream);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

enum class Bin { Add, Sub, Mul };

xmpError_t binary_async(Bin op, xmpHandle_t h, xmpIntegers_t o, xmpIntegers_t a, xmpIntegers_t b, uint32_t count) {
  xmpExecutionPolicy_t p = h->policy;
  xmpError_t e = check_binary(h, o, a, b, count);
  if (e != xmpErrorSuccess) return e;
  if (count == 0) return xmpErrorSuccess;
  XMP_SET_DEVICE(h);
  if (o->count < count) return xmpErrorInvalidCount;
  if (p->indices[0] && p->indices_count[0] < count) return xmpErrorInvalidCount;
  if (o->precision < std::max(a->precision, b->precision)) return xmpErrorInvalidPrecision;
  XMP_CHECK_HIP();
  bool inplace = (o == a || o == b);
  uint32_t *buffer = o->limbs;
  uint32_t *ws = nullptr;
  // products beyond the row kernels run Karatsuba in a workspace after the
  // in-place buffer (without one, within the scratch limit, the schoolbook)
  const size_t kw = op == Bin::Mul ? mul_karatsuba_words(a->nlimbs, b->nlimbs, count) : 0;
  const size_t obytes = inplace ? (out_bytes(o) + 255) / 256 * 256 : 0;
  if (kw && set_necessary_scratch(h, obytes + kw * 4) == xmpErrorSuccess) {
    ws = (uint32_t *)((uint8_t *)h->scratch + obytes);
  } else if (inplace) {
    e = set_necessary_scratch(h, out_bytes(o));
    if (e != xmpErrorSuccess) return e;
  }
  if (inplace) {
    buffer = (uint32_t *)h->scratch;
    // indexed outputs leave unselected instances untouched: start from o
    if (p->indices[0] &&
        hipMemcpyAsync(buffer, o->limbs, out_bytes(o), hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
      return xmpErrorCuda;
  }
  OutOperand out = make_out(o, buffer, p, 0);
  Operand A = make_in(a, p, 1), B = make_in(b, p, 2);
  hipError_t st = hipSuccess;
  switch (op) {
    case Bin::Add: st = launch_add(out, A, B, count, h->stream); break;
    case Bin::Sub: st = launch_sub(out, A, B, count, h->stream); break;
    case Bin::Mul:
      if (A.len < B.len) std::swap(A, B);   // reference mul.cu:168-175 orders by precision
      st = ws ? launch_mul_ws(out, A, B, count, false, ws, h->stream) : launch_mul(out, A, B, count, h->stream);
      break;
  }
  if (st != hipSuccess) return xmpErrorCuda;
  return finish(h, inplace, o, buffer);
}

}  // namespace

extern "C" {

xmpError_t XMPAPI xmpIntegersAddAsync(xmpHandle_t handle, xmpIntegers_t s, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersAddAsync");
  return binary_async(Bin::Add, handle, s, a, b, count);
}
xmpError_t XMPAPI xmpIntegersAdd(xmpHandle_t handle, xmpIntegers_t s, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersAddAsync(handle, s, a, b, count));
}
xmpError_t XMPAPI xmpIntegersSubAsync(xmpHandle_t handle, xmpIntegers_t d, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersSubAsync");
  return binary_async(Bin::Sub, handle, d, a, b, count);
}
xmpError_t XMPAPI xmpIntegersSub(xmpHandle_t handle, xmpIntegers_t d, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersSubAsync(handle, d, a, b, count));
}
xmpError_t XMPAPI xmpIntegersSqrAsync(xmpHandle_t handle, xmpIntegers_t p, const xmpIntegers_t a, uint32_t count) {
  XMP_TRACE("xmpIntegersSqrAsync");
  xmpExecutionPolicy_t pol = handle->policy;
  if (p->device != handle->device || a->device != handle->device) return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  XMP_SET_DEVICE(handle);
  if (p->count < count) return xmpErrorInvalidCount;
  if (pol->indices[0] && pol->indices_count[0] < count) return xmpErrorInvalidCount;
  if (p->precision < a->precision) return xmpErrorInvalidPrecision;
  bool inplace = (p == a);
  uint32_t *buffer = p->limbs;
  uint32_t *ws = nullptr;
  // as in binary_async: Karatsuba's workspace after the in-place buffer
  const size_t kw = mul_karatsuba_words(a->nlimbs, a->nlimbs, count);
  const size_t obytes = inplace ? (out_bytes(p) + 255) / 256 * 256 : 0;
  if (kw && set_necessary_scratch(handle, obytes + kw * 4) == xmpErrorSuccess) {
    ws = (uint32_t *)((uint8_t *)handle->scratch + obytes);
  } else if (inplace) {
    xmpError_t e = set_necessary_scratch(handle, out_bytes(p));
    if (e != xmpErrorSuccess) return e;
  }
  if (inplace) {
    buffer = (uint32_t *)handle->scratch;
    if (pol->indices[0] &&
        hipMemcpyAsync(buffer, p->limbs, out_bytes(p), hipMemcpyDeviceToDevice, handle->stream) != hipSuccess)
      return xmpErrorCuda;
  }
  Operand A = make_in(a, pol, 1);
  const OutOperand P = make_out(p, buffer, pol, 0);
  const hipError_t st = ws ? launch_mul_ws(P, A, A, count, true, ws, handle->stream)
                           : launch_sqr(P, A, count, handle->stream);
  if (st != hipSuccess) return xmpErrorCuda;
  return finish(handle, inplace, p, buffer);
}
xmpError_t XMPAPI xmpIntegersMulAsync(xmpHandle_t handle, xmpIntegers_t p, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersMulAsync");
  xmpExecutionPolicy_t pol = handle->policy;
  // reference mul.cu:147-149: a == b without input indices is a square
  if (a == b && !pol->indices[1] && !pol->indices[2]) return xmpIntegersSqrAsync(handle, p, a, count);
  return binary_async(Bin::Mul, handle, p, a, b, count);
}
xmpError_t XMPAPI xmpIntegersMul(xmpHandle_t handle, xmpIntegers_t p, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersMulAsync(handle, p, a, b, count));
}

// ---- division ----------------------------------------------------------
static xmpError_t divmod_async(xmpHandle_t h, xmpIntegers_t q, xmpIntegers_t m, xmpIntegers_t a, xmpIntegers_t b,
                               uint32_t count, bool divmod) {
  xmpExecutionPolicy_t p = h->policy;
  if ((q && q->device != h->device) || (m && m->device != h->device) || a->device != h->device ||
      b->device != h->device)
    return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  XMP_SET_DEVICE(h);
  if ((q && q->count < count) || (m && m->count < count)) return xmpErrorInvalidCount;
  if (p->indices[0] && p->indices_count[0] < count) return xmpErrorInvalidCount;
  if (divmod && p->indices[1] && p->indices_count[1] < count) return xmpErrorInvalidCount;
  if (q && q->precision < a->precision) return xmpErrorInvalidPrecision;
  if (m && m->precision < b->precision) return xmpErrorInvalidPrecision;
  // index slots: Div/Mod use 0 = out, 1 = a, 2 = b; DivMod 0 = q, 1 = r, 2 = a, 3 = b (div.cu:407-412)
  const int sa = divmod ? 2 : 1, sb = divmod ? 3 : 2, sq = 0, sr = divmod ? 1 : 0;
  bool qin = q && (q == a || q == b), min_ = m && (m == a || m == b);
  size_t qbytes = qin ? out_bytes(q) : 0, mbytes = min_ ? out_bytes(m) : 0;
  size_t ws = divmod_scratch_words(b->nlimbs, count) * sizeof(uint32_t);
  xmpError_t e = set_necessary_scratch(h, qbytes + mbytes + ws);
  if (e != xmpErrorSuccess) return e;
  uint8_t *scr = (uint8_t *)h->scratch;
  uint32_t *qbuf = q ? (qin ? (uint32_t *)scr : q->limbs) : nullptr;
  uint32_t *mbuf = m ? (min_ ? (uint32_t *)(scr + qbytes) : m->limbs) : nullptr;
  uint32_t *wsp = ws ? (uint32_t *)(scr + qbytes + mbytes) : nullptr;
  if (qin && p->indices[sq] && hipMemcpyAsync(qbuf, q->limbs, qbytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
    return xmpErrorCuda;
  if (min_ && p->indices[sr] &&
      hipMemcpyAsync(mbuf, m->limbs, mbytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
    return xmpErrorCuda;
  OutOperand Q = q ? make_out(q, qbuf, p, sq) : OutOperand{nullptr, 0, 0, nullptr};
  OutOperand R = m ? make_out(m, mbuf, p, sr) : OutOperand{nullptr, 0, 0, nullptr};
  if (launch_divmod_ws(Q, R, make_in(a, p, sa), make_in(b, p, sb), count, wsp, h->stream) != hipSuccess)
    return xmpErrorCuda;
  if (qin && hipMemcpyAsync(q->limbs, qbuf, qbytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
    return xmpErrorCuda;
  if (min_ && hipMemcpyAsync(m->limbs, mbuf, mbytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
    return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersDivAsync(xmpHandle_t handle, xmpIntegers_t q, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersDivAsync");
  return divmod_async(handle, q, nullptr, a, b, count, false);
}
xmpError_t XMPAPI xmpIntegersDiv(xmpHandle_t handle, xmpIntegers_t q, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersDivAsync(handle, q, a, b, count));
}
xmpError_t XMPAPI xmpIntegersModAsync(xmpHandle_t handle, xmpIntegers_t m, const xmpIntegers_t a,
                                      const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersModAsync");
  return divmod_async(handle, nullptr, m, a, b, count, false);
}
xmpError_t XMPAPI xmpIntegersMod(xmpHandle_t handle, xmpIntegers_t m, const xmpIntegers_t a, const xmpIntegers_t b,
                                 uint32_t count) {
  return sync(handle, xmpIntegersModAsync(handle, m, a, b, count));
}
xmpError_t XMPAPI xmpIntegersDivModAsync(xmpHandle_t handle, xmpIntegers_t q, xmpIntegers_t m,
                                         const xmpIntegers_t a, const xmpIntegers_t b, uint32_t count) {
  XMP_TRACE("xmpIntegersDivModAsync");
  if (q == nullptr || m == nullptr) return xmpErrorInvalidParameter;
  return divmod_async(handle, q, m, a, b, count, true);
}
xmpError_t XMPAPI xmpIntegersDivMod(xmpHandle_t handle, xmpIntegers_t q, xmpIntegers_t m, const xmpIntegers_t a,
                                    const xmpIntegers_t b, uint32_t count) {
  return sync(handle, xmpIntegersDivModAsync(handle, q, m, a, b, count));
}

// ---- powm --------------------------------------------------------------
}  // extern "C"

namespace xmp {

// Plan and launch one powm call (reference powm.cu:250-307 and the scratch
// planning of powm_operators.h:308-328).  The workspace holds R^2 mod m for
// the moduli the call can touch, the division workspace of moduli above the
// register division, the window table of one chunk of instances and, for an
// in-place call with a fixed workspace, the output.  ws == nullptr: the
// handle's grow-only scratch, limited by the policy's scratch limit (chunks
// above 65,536 instances halve first, then the window narrows); otherwise the
// fixed region, which is never grown (the window narrows to 3 bits first,
// then chunks halve down to 64 instances, then the window to 2 bits).
xmpError_t powm_run(xmpHandle_t handle, xmpIntegers_t out, xmpIntegers_t a, xmpIntegers_t exp, xmpIntegers_t mod,
                    uint32_t count, const FixedWorkspace *ws) {
  xmpExecutionPolicy_t p = handle->policy;
  const int dev = handle->device;
  if (out->device != dev || a->device != dev || exp->device != dev || mod->device != dev)
    return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  if (out->count < count) return xmpErrorInvalidCount;
  if (p->indices[0] && p->indices_count[0] < count) return xmpErrorInvalidCount;
  if (a->precision != out->precision || mod->precision != out->precision) return xmpErrorInvalidPrecision;
  XMP_SET_DEVICE(handle);
  PowmChoice c;
  xmpError_t e = choose_powm(handle, out->precision, exp->precision, count, p->algorithm, &c);
  // moduli wider than every lane layout take the workgroup layout (under the
  // default and digit-serial policies, as the reference's DigitMP fallback)
  uint32_t wide = 0;
  if (e == xmpErrorUnsupported && (p->algorithm == xmpAlgorithmDefault || p->algorithm == xmpAlgorithmDigitMP)) {
    wide = powm_wide_digits_per_thread(out->precision);
    if (wide) {
      c.variant = -1;
      c.window_bits = best_window_bits(exp->precision);
      e = xmpErrorSuccess;
    }
  }
  if (e != xmpErrorSuccess) return e;
  const PowmVariant v = wide ? PowmVariant{0, 0, 0} : powm_variant(c.variant);
  auto table_words = [&](uint32_t n, uint32_t w) {
    return wide ? powm_wide_table_words(wide, n, w) : powm_table_words(v, round_up(n, kStrideAlign), w);
  };

  const bool inplace = (out == a || out == exp || out == mod);
  const size_t obytes = inplace ? out_bytes(out) : 0;
  // R^2 mod m for the moduli this call can reach: instance i < count reads
  // modulus i % mod->count without a modulus index array, any of them with one
  const uint32_t nmod = p->indices[3] ? mod->count : std::min(count, mod->count);
  const uint32_t mstride = round_up(nmod, kStrideAlign);
  const size_t r2_words = (size_t)mod->nlimbs * mstride;
  const size_t div_words = mod->nlimbs > kMaxRegDivisorLimbs ? divmod_scratch_words(mod->nlimbs, nmod) : 0;
  // the memory tier's R^2 squarings need a block per modulus of their own
  const size_t r2blk_words = wide ? powm_wide_r2_scratch_words(wide, nmod) : 0;
  const size_t obuf_words = ws ? round_up((uint32_t)(obytes / sizeof(uint32_t)), 64) : 0;
  const size_t limit = ws ? ws->bytes : (p->scratch_size_limit ? p->scratch_size_limit : handle->memorySize / 4 * 3);
  const size_t fixed = (r2_words + div_words + r2blk_words + obuf_words) * sizeof(uint32_t);
  uint32_t wbits = c.window_bits;
  uint32_t chunk = count;
  while (fixed + table_words(chunk, wbits) * sizeof(uint32_t) > limit) {
    // (a workgroup-layout instance can need megabytes: its chunks may shrink further)
    if (!ws && chunk > (wide ? kStrideAlign : 65536u)) {
      chunk = round_up(chunk / 2, kStrideAlign);
    } else if (ws && wbits > 3) {
      wbits--;
    } else if (ws && chunk > kStrideAlign) {
      chunk = round_up(chunk / 2, kStrideAlign);
    } else if (wbits > 2) {
      wbits--;
    } else {
      return xmpErrorIncreaseScratchLimit;
    }
  }
  const size_t need = fixed + table_words(chunk, wbits) * sizeof(uint32_t);
  uint32_t *base;
  uint32_t *obuf = out->limbs;
  if (ws) {
    base = (uint32_t *)ws->base;
    if (inplace) obuf = base + r2_words + div_words + r2blk_words;
  } else {
    e = set_necessary_scratch(handle, need);
    if (e != xmpErrorSuccess) return e;
    base = (uint32_t *)handle->scratch;
    if (inplace) {
      e = set_necessary_tmp_out(handle, obytes);
      if (e != xmpErrorSuccess) return e;
      obuf = (uint32_t *)handle->tmpOut;
    }
  }
  // indexed outputs leave unselected instances untouched: start from out
  if (inplace && p->indices[0] &&
      hipMemcpyAsync(obuf, out->limbs, obytes, hipMemcpyDeviceToDevice, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  uint32_t *r2 = base;
  uint32_t *divws = r2 + r2_words;
  uint32_t *r2blk = divws + div_words;
  uint32_t *table = r2blk + r2blk_words + obuf_words;
  if ((wide ? launch_wide_r2(wide, r2, mstride, mod->limbs, mod->nlimbs, mod->stride, nmod,
                             div_words ? divws : nullptr, r2blk_words ? r2blk : nullptr, handle->stream)
            : launch_montgomery_r2(c.variant, r2, mstride, mod->limbs, mod->nlimbs, mod->stride, nmod,
                                   div_words ? divws : nullptr, handle->stream)) != hipSuccess)
    return xmpErrorCuda;
  PowmArgs args;
  args.out = make_out(out, obuf, p, 0);
  args.base = make_in(a, p, 1);
  args.exp = make_in(exp, p, 2);
  args.mod = make_in(mod, p, 3);
  args.r2 = r2;
  args.r2_stride = mstride;
  args.table = table;
  args.window_bits = wbits;
  for (uint32_t start = 0; start < count; start += chunk) {
    args.start = start;
    args.count = std::min(chunk, count - start);
    if ((wide ? launch_powm_wide(wide, args, handle->stream) : launch_powm_variant(c.variant, args, handle->stream)) !=
        hipSuccess)
      return xmpErrorCuda;
  }
  if (inplace && hipMemcpyAsync(out->limbs, obuf, obytes, hipMemcpyDeviceToDevice, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

}  // namespace xmp

extern "C" {

xmpError_t XMPAPI xmpIntegersPowmAsync(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t a,
                                       const xmpIntegers_t exp, const xmpIntegers_t mod, uint32_t count) {
  XMP_TRACE("xmpIntegersPowmAsync");
  return powm_run(handle, out, a, exp, mod, count, nullptr);
}

xmpError_t XMPAPI xmpIntegersPowm(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t base,
                                  const xmpIntegers_t exp, const xmpIntegers_t mod, uint32_t count) {
  return sync(handle, xmpIntegersPowmAsync(handle, out, base, exp, mod, count));
}

}  // extern "C"
