// ext.cpp -- Montgomery contexts and the autotuner (include/xmp_ext.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xmp_ext.h"
#include "internal.hpp"
#include "kernels.hpp"
#include "powm.hpp"
#include "tuner.hpp"

namespace xmp {
int tuner_load_text(const char *text);
}

using namespace xmp;

struct _xmpMontgomery_t {
  xmpIntegers_t mod = nullptr;
  uint32_t *r2 = nullptr;
  uint32_t r2_stride = 0;
  int variant = -1;
  uint32_t window_bits = 0;
  uint32_t exp_precision = 0;
  uint32_t *table = nullptr;
  size_t table_words = 0;
};

namespace xmp {

// R^2 mod m for every instance of `mod` into a new device buffer
xmpError_t montgomery_prepare(xmpHandle_t h, xmpIntegers_t mod, int variant, uint32_t **r2_out,
                              uint32_t *stride_out) {
  const uint32_t stride = round_up(mod->count, kStrideAlign);
  uint32_t *r2 = (uint32_t *)h->da(sizeof(uint32_t) * (size_t)mod->nlimbs * stride);
  if (!r2) return xmpErrorInvalidCudaMalloc;
  uint32_t *ws = nullptr;
  if (mod->nlimbs > kMaxRegDivisorLimbs) {
    size_t words = (size_t)(2 * mod->nlimbs + 1) * round_up(mod->count, 256);
    xmpError_t e = set_necessary_scratch(h, words * sizeof(uint32_t));
    if (e != xmpErrorSuccess) {
      h->df(r2);
      return e;
    }
    ws = (uint32_t *)h->scratch;
  }
  if (launch_montgomery_r2(variant, r2, stride, mod->limbs, mod->nlimbs, mod->stride, mod->count, ws, h->stream) !=
      hipSuccess) {
    h->df(r2);
    return xmpErrorCuda;
  }
  *r2_out = r2;
  *stride_out = stride;
  return xmpErrorSuccess;
}

}  // namespace xmp

extern "C" {

xmpError_t XMPAPI xmpMontgomeryCreate(xmpHandle_t handle, xmpMontgomery_t *ctx, xmpIntegers_t mod,
                                      uint32_t exp_precision, uint32_t count_hint) {
  XMP_TRACE("xmpMontgomeryCreate");
  XMP_CHECK_NE(ctx, nullptr);
  XMP_CHECK_NE(mod, nullptr);
  if (mod->device != handle->device) return xmpErrorInvalidDevice;
  if (exp_precision == 0 || exp_precision % 32) return xmpErrorInvalidPrecision;
  XMP_SET_DEVICE(handle);
  PowmChoice c;
  xmpError_t e = choose_powm(handle, mod->precision, exp_precision, std::max(count_hint, 1u),
                             handle->policy->algorithm, &c);
  if (e != xmpErrorSuccess) return e;
  void *mem = handle->ha(sizeof(_xmpMontgomery_t));
  if (!mem) return xmpErrorInvalidMalloc;
  xmpMontgomery_t m = new (mem) _xmpMontgomery_t();
  m->mod = mod;
  m->variant = c.variant;
  m->window_bits = c.window_bits;
  m->exp_precision = exp_precision;
  e = montgomery_prepare(handle, mod, c.variant, &m->r2, &m->r2_stride);
  if (e != xmpErrorSuccess) {
    m->~_xmpMontgomery_t();
    handle->hf(m);
    return e;
  }
  (void)hipStreamSynchronize(handle->stream);
  *ctx = m;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpMontgomeryDestroy(xmpHandle_t handle, xmpMontgomery_t ctx) {
  XMP_CHECK_NE(ctx, nullptr);
  XMP_SET_DEVICE(handle);
  if (ctx->r2) handle->df(ctx->r2);
  if (ctx->table) handle->df(ctx->table);
  ctx->~_xmpMontgomery_t();
  handle->hf(ctx);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpMontgomeryGetLayout(xmpHandle_t handle, xmpMontgomery_t ctx, uint32_t *lanes, uint32_t *digits,
                                         uint32_t *radix, uint32_t *window_bits) {
  (void)handle;
  XMP_CHECK_NE(ctx, nullptr);
  PowmVariant v = powm_variant(ctx->variant);
  if (lanes) *lanes = (uint32_t)v.lanes;
  if (digits) *digits = (uint32_t)v.digits;
  if (radix) *radix = (uint32_t)v.radix;
  if (window_bits) *window_bits = ctx->window_bits;
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpMontgomeryGetR2(xmpHandle_t handle, xmpMontgomery_t ctx, xmpIntegers_t out) {
  XMP_CHECK_NE(ctx, nullptr);
  XMP_CHECK_NE(out, nullptr);
  if (out->device != handle->device) return xmpErrorInvalidDevice;
  const xmpIntegers_t m = ctx->mod;
  if (out->precision != m->precision) return xmpErrorInvalidPrecision;
  if (out->count < m->count) return xmpErrorInvalidCount;
  XMP_SET_DEVICE(handle);
  if (hipMemcpy2DAsync(out->limbs, sizeof(uint32_t) * out->stride, ctx->r2, sizeof(uint32_t) * ctx->r2_stride,
                       sizeof(uint32_t) * m->count, m->nlimbs, hipMemcpyDeviceToDevice, handle->stream) != hipSuccess ||
      hipStreamSynchronize(handle->stream) != hipSuccess)
    return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpIntegersPowmMontgomeryAsync(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t base,
                                                 const xmpIntegers_t exp, xmpMontgomery_t ctx, uint32_t count) {
  XMP_TRACE("xmpIntegersPowmMontgomeryAsync");
  XMP_CHECK_NE(ctx, nullptr);
  xmpExecutionPolicy_t p = handle->policy;
  const int dev = handle->device;
  if (out->device != dev || base->device != dev || exp->device != dev) return xmpErrorInvalidDevice;
  if (count == 0) return xmpErrorSuccess;
  if (out->count < count) return xmpErrorInvalidCount;
  if (p->indices[0] && p->indices_count[0] < count) return xmpErrorInvalidCount;
  if (base->precision != out->precision || ctx->mod->precision != out->precision) return xmpErrorInvalidPrecision;
  if (exp->precision != ctx->exp_precision) return xmpErrorInvalidPrecision;
  XMP_SET_DEVICE(handle);
  PowmVariant v = powm_variant(ctx->variant);
  size_t need = powm_table_words(v, round_up(count, kStrideAlign), ctx->window_bits);
  if (need > ctx->table_words) {
    if (ctx->table) handle->df(ctx->table);
    ctx->table = (uint32_t *)handle->da(need * sizeof(uint32_t));
    ctx->table_words = ctx->table ? need : 0;
    if (!ctx->table) return xmpErrorInvalidCudaMalloc;
  }
  bool inplace = (out == base || out == exp || out == ctx->mod);
  size_t obytes = sizeof(uint32_t) * (size_t)out->nlimbs * out->stride;
  uint32_t *obuf = out->limbs;
  if (inplace) {
    xmpError_t e = set_necessary_tmp_out(handle, obytes);
    if (e != xmpErrorSuccess) return e;
    obuf = (uint32_t *)handle->tmpOut;
    if (p->indices[0] && hipMemcpyAsync(obuf, out->limbs, obytes, hipMemcpyDeviceToDevice, handle->stream) != hipSuccess)
      return xmpErrorCuda;
  }
  PowmArgs args;
  args.out = make_out(out, obuf, p, 0);
  args.base = make_in(base, p, 1);
  args.exp = make_in(exp, p, 2);
  args.mod = make_in(ctx->mod, p, 3);
  args.r2 = ctx->r2;
  args.r2_stride = ctx->r2_stride;
  args.table = ctx->table;
  args.window_bits = ctx->window_bits;
  args.start = 0;
  args.count = count;
  if (launch_powm_variant(ctx->variant, args, handle->stream) != hipSuccess) return xmpErrorCuda;
  if (inplace && hipMemcpyAsync(out->limbs, obuf, obytes, hipMemcpyDeviceToDevice, handle->stream) != hipSuccess)
    return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

// ---------------------------------------------------------------- tuner
// Replaces the reference's offline tool (src/tune/tune.cu:70-275): instead of
// sweeping blocks/SM per algorithm on a broadcast modulus, it times every
// compiled lane layout and nearby window widths on random per-instance
// moduli at the requested batch size and records throughput.
xmpError_t XMPAPI xmpTunerRun(xmpHandle_t handle, uint32_t precision, uint32_t exp_precision, uint32_t count,
                              char *json, size_t json_bytes, uint32_t *measured) {
  XMP_CHECK_NE(json, nullptr);
  if (precision == 0 || precision % 32 || exp_precision == 0 || exp_precision % 32 || count == 0)
    return xmpErrorInvalidParameter;
  XMP_SET_DEVICE(handle);
  const uint32_t n32 = precision / 32, e32 = exp_precision / 32;
  // deterministic pseudo-random inputs (splitmix64), odd moduli with the top bit set
  std::vector<uint32_t> hm((size_t)count * n32), hb((size_t)count * n32), he((size_t)count * e32);
  uint64_t s = 0x9E3779B97F4A7C15ull ^ ((uint64_t)precision << 32) ^ count;
  auto next = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
  };
  for (auto &x : hm) x = next();
  for (auto &x : hb) x = next();
  for (auto &x : he) x = next();
  for (uint32_t i = 0; i < count; i++) {
    hm[(size_t)i * n32] |= 1u;
    hm[(size_t)i * n32 + n32 - 1] |= 0x80000000u;
    hb[(size_t)i * n32 + n32 - 1] &= 0x7FFFFFFFu;
  }
  xmpIntegers_t M = nullptr, B = nullptr, E = nullptr, O = nullptr;
  xmpError_t e;
  auto cleanup = [&]() {
    if (M) xmpIntegersDestroy(handle, M);
    if (B) xmpIntegersDestroy(handle, B);
    if (E) xmpIntegersDestroy(handle, E);
    if (O) xmpIntegersDestroy(handle, O);
  };
  if ((e = xmpIntegersCreate(handle, &M, precision, count)) != xmpErrorSuccess ||
      (e = xmpIntegersCreate(handle, &B, precision, count)) != xmpErrorSuccess ||
      (e = xmpIntegersCreate(handle, &E, exp_precision, count)) != xmpErrorSuccess ||
      (e = xmpIntegersCreate(handle, &O, precision, count)) != xmpErrorSuccess ||
      (e = xmpIntegersImport(handle, M, n32, -1, 4, 0, 0, hm.data(), count)) != xmpErrorSuccess ||
      (e = xmpIntegersImport(handle, B, n32, -1, 4, 0, 0, hb.data(), count)) != xmpErrorSuccess ||
      (e = xmpIntegersImport(handle, E, e32, -1, 4, 0, 0, he.data(), count)) != xmpErrorSuccess) {
    cleanup();
    return e;
  }
  // gold standard for the self-check: the library's own call (as the
  // reference tuner checks every timed run, src/tune/tune.cu:93-97,131-140)
  std::vector<uint32_t> gold((size_t)count * n32), test((size_t)count * n32);
  uint32_t words = n32;
  if ((e = xmpIntegersPowm(handle, O, B, E, M, count)) != xmpErrorSuccess ||
      (e = xmpIntegersExport(handle, gold.data(), &words, -1, 4, 0, 0, O, count)) != xmpErrorSuccess) {
    cleanup();
    return e;
  }
  hipEvent_t ev0, ev1;
  (void)hipEventCreate(&ev0);
  (void)hipEventCreate(&ev1);
  std::string out;
  uint32_t nmeasured = 0, mismatched = 0;
  double best_rate = 0;
  std::string best_entry;
  const uint32_t w0 = best_window_bits(exp_precision);
  for (int vi = 0; vi < powm_variant_count(); vi++) {
    PowmVariant v = powm_variant(vi);
    if (!variant_fits(v, precision)) continue;
    // a layout much wider than the size leaves most of its lanes without
    // digits (its steps stop at the lanes that hold them, but the idle lanes
    // still take wave slots); up to three times the width can still win on
    // small batches, where the extra lanes fill otherwise idle SIMDs
    if (variant_rbits(v) > 3 * precision + 128) continue;
    uint32_t *r2 = nullptr, r2s = 0;
    if ((e = montgomery_prepare(handle, M, vi, &r2, &r2s)) != xmpErrorSuccess) break;
    // the window widths around the cost-optimal one share one table (sized
    // for the widest) and are timed in turns: a warm-up round, then four
    // rounds of one launch per width, so that clock drift over the sweep
    // falls on every width alike; each width keeps its median
    const uint32_t wlo = w0 > 2 ? w0 - 1 : 1, whi = std::min(w0 + 1, 8u);
    const uint32_t nw = whi - wlo + 1;
    uint32_t *table = (uint32_t *)handle->da(powm_table_words(v, round_up(count, kStrideAlign), whi) *
                                             sizeof(uint32_t));
    if (!table) {
      handle->df(r2);
      continue;
    }
    PowmArgs args;
    args.out = OutOperand{O->limbs, O->nlimbs, O->stride, nullptr};
    args.base = Operand{B->limbs, B->nlimbs, B->stride, B->count, nullptr, 0};
    args.exp = Operand{E->limbs, E->nlimbs, E->stride, E->count, nullptr, 0};
    args.mod = Operand{M->limbs, M->nlimbs, M->stride, M->count, nullptr, 0};
    args.r2 = r2;
    args.r2_stride = r2s;
    args.table = table;
    args.start = 0;
    args.count = count;
    constexpr int kRounds = 4;
    std::vector<float> times((size_t)nw * kRounds, 0.f);
    std::vector<bool> same(nw, true);
    for (int round = 0; round <= kRounds && e == xmpErrorSuccess; round++) {
      for (uint32_t k = 0; k < nw; k++) {
        args.window_bits = wlo + k;
        (void)hipEventRecord(ev0, handle->stream);
        (void)launch_powm_variant(vi, args, handle->stream);
        (void)hipEventRecord(ev1, handle->stream);
        (void)hipEventSynchronize(ev1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev0, ev1);
        if (round > 0) times[(size_t)k * kRounds + round - 1] = ms;
        if (round == kRounds) {
          // self-check: each width's results against the gold standard
          words = n32;
          if ((e = xmpIntegersExport(handle, test.data(), &words, -1, 4, 0, 0, O, count)) != xmpErrorSuccess) break;
          same[k] = test == gold;
        }
      }
    }
    handle->df(table);
    for (uint32_t k = 0; k < nw && e == xmpErrorSuccess; k++) {
      if (!same[k]) {
        mismatched++;
        continue;
      }
      float *tk = times.data() + (size_t)k * kRounds;
      std::sort(tk, tk + kRounds);
      const float med_ms = 0.5f * (tk[kRounds / 2 - 1] + tk[kRounds / 2]);
      nmeasured++;
      const uint32_t w = wlo + k;
      double rate = count / (med_ms * 1e-3);
      char buf[512];
      std::snprintf(buf, sizeof buf,
                    "{\"precision\": %u, \"exp_bits\": %u, \"count\": %u, \"lanes\": %d, \"digits\": %d, "
                    "\"radix\": %d, \"window\": %u, \"ms\": %.4f, \"powm_per_s\": %.1f}",
                    precision, exp_precision, count, v.lanes, v.digits, v.radix, w, med_ms, rate);
      if (!out.empty()) out += ",\n";
      out += buf;
      if (rate > best_rate) {
        best_rate = rate;
        best_entry = buf;
      }
    }
    handle->df(r2);
    if (e != xmpErrorSuccess) break;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  cleanup();
  if (e != xmpErrorSuccess) return e;
  if (out.size() + 1 > json_bytes) return xmpErrorInvalidParameter;
  std::memcpy(json, out.c_str(), out.size() + 1);
  if (measured) *measured = nmeasured;
  XMP_CHECK_HIP();
  // a layout whose results differ from the library's own is left out and fails the run
  if (mismatched) return xmpErrorUnspecified;
  return best_entry.empty() ? xmpErrorUnsupported : xmpErrorSuccess;
}

xmpError_t XMPAPI xmpTunerLoad(const char *json) {
  XMP_CHECK_NE(json, nullptr);
  return tuner_load_text(json) >= 0 ? xmpErrorSuccess : xmpErrorInvalidParameter;
}

}  // extern "C"
