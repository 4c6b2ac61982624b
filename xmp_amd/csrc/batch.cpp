// batch.cpp -- mixed-size batches in one persistent, graph-captured launch.
//
// No reference counterpart: XMP computes one precision per call and a user
// with several sizes issues one xmpIntegersPowm per size class
// (src/powm.cu:250-307), each with its own set-up launches and its own
// partially idle last wave.  A batch binds the per-class arrays once,
// keeps R^2 mod m for every class resident, and runs every class's tiles from
// one device queue ordered by decreasing modelled cost (longest tiles first),
// so short tiles fill the SIMDs that long ones leave idle.  The queue reset
// and the persistent kernel are captured in a HIP graph on the first launch
// and replayed afterwards.  A one-class batch captures that class's own
// kernel instead: with nothing to interleave, the queue only costs time.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "../../include/xmp_ext.h"
#include "internal.hpp"
#include "kernels.hpp"
#include "powm.hpp"
#include "tuner.hpp"

namespace xmp {
xmpError_t montgomery_prepare(xmpHandle_t h, xmpIntegers_t mod, int variant, uint32_t **r2_out,
                              uint32_t *stride_out);
}

using namespace xmp;

struct _xmpBatch_t {
  uint32_t classes = 0;
  std::vector<xmpIntegers_t> out, base, exp, mod;
  std::vector<uint32_t> counts;
  std::vector<int> variant;
  std::vector<uint32_t> wbits;
  std::vector<uint32_t *> r2, table;
  std::vector<uint32_t> r2_stride;
  PowmArgs *d_classes = nullptr;
  PowmArgs single;   // a one-class batch's arguments (that class's kernel runs directly)
  int32_t *d_variant = nullptr;
  uint32_t *d_tile_class = nullptr, *d_tile_index = nullptr, *d_counter = nullptr;
  uint32_t ntiles = 0, grid = 0;
  size_t lds_bytes = 0;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void release(xmpHandle_t h, xmpBatch_t b) {
  for (auto p : b->r2)
    if (p) h->df(p);
  for (auto p : b->table)
    if (p) h->df(p);
  if (b->d_classes) h->df(b->d_classes);
  if (b->d_variant) h->df(b->d_variant);
  if (b->d_tile_class) h->df(b->d_tile_class);
  if (b->d_tile_index) h->df(b->d_tile_index);
  if (b->d_counter) h->df(b->d_counter);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
}

// modelled cost of one tile (one wave) of a class: products x per-product work
double tile_cost(const PowmVariant &v, uint32_t exp_bits, uint32_t wbits) {
  const double nd = (double)(v.lanes * v.digits);
  const double products = (double)(1u << wbits) + exp_bits + std::ceil((double)exp_bits / wbits) + 2;
  return products * nd * (2.0 * v.digits + (v.lanes > 1 ? 14.0 : 5.0));
}

}  // namespace

extern "C" {

xmpError_t XMPAPI xmpBatchCreate(xmpHandle_t handle, xmpBatch_t *batch, uint32_t classes, xmpIntegers_t *out,
                                 xmpIntegers_t *base, xmpIntegers_t *exp, xmpIntegers_t *mod,
                                 const uint32_t *counts) {
  XMP_TRACE("xmpBatchCreate");
  XMP_CHECK_NE(batch, nullptr);
  if (classes == 0 || !out || !base || !exp || !mod || !counts) return xmpErrorInvalidParameter;
  const int dev = handle->device;
  for (uint32_t c = 0; c < classes; c++) {
    if (!out[c] || !base[c] || !exp[c] || !mod[c]) return xmpErrorInvalidParameter;
    if (out[c]->device != dev || base[c]->device != dev || exp[c]->device != dev || mod[c]->device != dev)
      return xmpErrorInvalidDevice;
    if (counts[c] == 0 || out[c]->count < counts[c]) return xmpErrorInvalidCount;
    if (base[c]->precision != out[c]->precision || mod[c]->precision != out[c]->precision)
      return xmpErrorInvalidPrecision;
  }
  XMP_SET_DEVICE(handle);
  void *mem = handle->ha(sizeof(_xmpBatch_t));
  if (!mem) return xmpErrorInvalidMalloc;
  xmpBatch_t b = new (mem) _xmpBatch_t();
  b->classes = classes;
  b->out.assign(out, out + classes);
  b->base.assign(base, base + classes);
  b->exp.assign(exp, exp + classes);
  b->mod.assign(mod, mod + classes);
  b->counts.assign(counts, counts + classes);
  b->r2.assign(classes, nullptr);
  b->table.assign(classes, nullptr);
  b->r2_stride.assign(classes, 0);
  xmpError_t e = xmpErrorSuccess;
  std::vector<PowmArgs> args(classes);
  struct Tile {
    uint32_t cls, idx;
    double cost;
  };
  std::vector<Tile> tiles;
  for (uint32_t c = 0; c < classes && e == xmpErrorSuccess; c++) {
    PowmChoice ch;
    e = choose_powm(handle, out[c]->precision, exp[c]->precision, counts[c], xmpAlgorithmDefault, &ch, true);
    if (e != xmpErrorSuccess) break;
    PowmVariant v = powm_variant(ch.variant);
    b->variant.push_back(ch.variant);
    b->wbits.push_back(ch.window_bits);
    b->lds_bytes = std::max(b->lds_bytes, powm_variant_lds_bytes(ch.variant));
    size_t tw = powm_table_words(v, round_up(counts[c], kStrideAlign), ch.window_bits);
    b->table[c] = (uint32_t *)handle->da(tw * sizeof(uint32_t));
    if (!b->table[c]) {
      e = xmpErrorInvalidCudaMalloc;
      break;
    }
    e = montgomery_prepare(handle, mod[c], ch.variant, &b->r2[c], &b->r2_stride[c]);
    if (e != xmpErrorSuccess) break;
    PowmArgs &a = args[c];
    a.out = OutOperand{out[c]->limbs, out[c]->nlimbs, out[c]->stride, nullptr};
    a.base = Operand{base[c]->limbs, base[c]->nlimbs, base[c]->stride, base[c]->count, nullptr, 0};
    a.exp = Operand{exp[c]->limbs, exp[c]->nlimbs, exp[c]->stride, exp[c]->count, nullptr, 0};
    a.mod = Operand{mod[c]->limbs, mod[c]->nlimbs, mod[c]->stride, mod[c]->count, nullptr, 0};
    a.r2 = b->r2[c];
    a.r2_stride = b->r2_stride[c];
    a.table = b->table[c];
    a.start = 0;
    a.count = counts[c];
    a.window_bits = ch.window_bits;
    const uint32_t per_tile = 64u / (uint32_t)v.lanes;
    const uint32_t nt = div_round_up(counts[c], per_tile);
    const double cost = tile_cost(v, exp[c]->precision, ch.window_bits);
    for (uint32_t t = 0; t < nt; t++) tiles.push_back({c, t, cost});
  }
  if (e == xmpErrorSuccess && classes == 1) b->single = args[0];
  if (e == xmpErrorSuccess) {
    // longest tiles first (LPT order); stable so each class's tiles stay in order
    std::stable_sort(tiles.begin(), tiles.end(), [](const Tile &x, const Tile &y) { return x.cost > y.cost; });
    b->ntiles = (uint32_t)tiles.size();
    std::vector<uint32_t> tc(b->ntiles), ti(b->ntiles);
    for (uint32_t i = 0; i < b->ntiles; i++) {
      tc[i] = tiles[i].cls;
      ti[i] = tiles[i].idx;
    }
    b->d_classes = (PowmArgs *)handle->da(sizeof(PowmArgs) * classes);
    b->d_variant = (int32_t *)handle->da(sizeof(int32_t) * classes);
    b->d_tile_class = (uint32_t *)handle->da(sizeof(uint32_t) * b->ntiles);
    b->d_tile_index = (uint32_t *)handle->da(sizeof(uint32_t) * b->ntiles);
    b->d_counter = (uint32_t *)handle->da(256);
    if (!b->d_classes || !b->d_variant || !b->d_tile_class || !b->d_tile_index || !b->d_counter) {
      e = xmpErrorInvalidCudaMalloc;
    } else if (hipMemcpy(b->d_classes, args.data(), sizeof(PowmArgs) * classes, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(b->d_variant, b->variant.data(), sizeof(int32_t) * classes, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(b->d_tile_class, tc.data(), sizeof(uint32_t) * b->ntiles, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(b->d_tile_index, ti.data(), sizeof(uint32_t) * b->ntiles, hipMemcpyHostToDevice) != hipSuccess) {
      e = xmpErrorCuda;
    }
    // persistent grid: two waves per SIMD, four SIMDs per CU (one wave per workgroup)
    b->grid = std::min<uint32_t>(b->ntiles, (handle->cuCount ? handle->cuCount : 256) * 8);
  }
  if (e == xmpErrorSuccess && hipStreamSynchronize(handle->stream) != hipSuccess) e = xmpErrorCuda;
  if (e != xmpErrorSuccess) {
    release(handle, b);
    b->~_xmpBatch_t();
    handle->hf(b);
    *batch = nullptr;
    return e;
  }
  *batch = b;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpBatchDestroy(xmpHandle_t handle, xmpBatch_t batch) {
  XMP_CHECK_NE(batch, nullptr);
  XMP_SET_DEVICE(handle);
  (void)hipDeviceSynchronize();
  release(handle, batch);
  batch->~_xmpBatch_t();
  handle->hf(batch);
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpBatchPrepare(xmpHandle_t handle, xmpBatch_t batch) {
  XMP_TRACE("xmpBatchPrepare");
  XMP_CHECK_NE(batch, nullptr);
  XMP_SET_DEVICE(handle);
  for (uint32_t c = 0; c < batch->classes; c++) {
    xmpIntegers_t m = batch->mod[c];
    uint32_t *ws = nullptr;
    if (m->nlimbs > kMaxRegDivisorLimbs) {
      xmpError_t e = set_necessary_scratch(handle, (size_t)(2 * m->nlimbs + 1) * round_up(m->count, 256) * 4);
      if (e != xmpErrorSuccess) return e;
      ws = (uint32_t *)handle->scratch;
    }
    if (launch_montgomery_r2(batch->variant[c], batch->r2[c], batch->r2_stride[c], m->limbs, m->nlimbs, m->stride,
                             m->count, ws, handle->stream) != hipSuccess)
      return xmpErrorCuda;
  }
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpBatchLaunchAsync(xmpHandle_t handle, xmpBatch_t batch) {
  XMP_TRACE("xmpBatchLaunchAsync");
  XMP_CHECK_NE(batch, nullptr);
  XMP_SET_DEVICE(handle);
  BatchArgs ba{batch->d_classes, batch->d_variant, batch->d_tile_class, batch->d_tile_index, batch->ntiles,
               batch->d_counter};
  if (!batch->exec) {
    // capture on a private stream (the legacy default stream cannot be captured)
    if (!batch->capture_stream && hipStreamCreate(&batch->capture_stream) != hipSuccess) return xmpErrorCuda;
    if (hipStreamBeginCapture(batch->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return xmpErrorCuda;
    // one class has no tiles of other classes to interleave: its own kernel
    // runs without the queue (and without the persistent kernel's calls)
    hipError_t st = batch->classes == 1
                        ? launch_powm_variant(batch->variant[0], batch->single, batch->capture_stream)
                        : launch_powm_batch(ba, batch->grid, batch->lds_bytes, batch->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t st2 = hipStreamEndCapture(batch->capture_stream, &g);
    if (st != hipSuccess || st2 != hipSuccess || !g) return xmpErrorCuda;
    batch->graph = g;
    if (hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0) != hipSuccess) return xmpErrorCuda;
  }
  if (hipGraphLaunch(batch->exec, handle->stream) != hipSuccess) return xmpErrorCuda;
  XMP_CHECK_HIP();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpBatchLaunchPlainAsync(xmpHandle_t handle, xmpBatch_t batch) {
  XMP_TRACE("xmpBatchLaunchPlainAsync");
  XMP_CHECK_NE(batch, nullptr);
  xmpExecutionPolicy_t saved = handle->policy;
  handle->policy = const_cast<xmpExecutionPolicy_t>(&kDefaultPolicy);
  xmpError_t e = xmpErrorSuccess;
  for (uint32_t c = 0; c < batch->classes && e == xmpErrorSuccess; c++)
    e = xmpIntegersPowmAsync(handle, batch->out[c], batch->base[c], batch->exp[c], batch->mod[c], batch->counts[c]);
  handle->policy = saved;
  return e;
}

}  // extern "C"
