// serve.cpp -- continuous-batching request loop over resident keys.
//
// No reference counterpart.  The nearest reference mechanisms are the Async
// API on a user stream (src/include/xmp.h:125-138), policy index arrays that
// gather a per-instance modulus from a resident array
// (powm_implementations.h:417-418) and the CRT sequence of
// samples/crt_rsa/rsa.cc:205-219 (c mod p, powm, c mod q, powm, mul, mul,
// add, mod).
//
// Keys stay resident on the device with their Montgomery constants.  A
// scheduler thread takes whatever requests have arrived, admits as many as
// the HBM budget allows into a free batch slot, and launches that batch on
// the slot's own pair of streams (one per request kind); slots run
// concurrently, so requests arriving while
// others are on the GPU are admitted as soon as a slot and memory free up and
// their kernels overlap with the ones already running.  Requests are
//   kind 0: out = base^exp mod n_k                         (2048-bit)
//   kind 1: out = CRT(c^dp_k mod p_k, c^dq_k mod q_k) mod n_k (two 1024-bit
//           exponentiations recombined on the device as (mp*cp + mq*cq) mod n).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
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

namespace {

using Clock = std::chrono::steady_clock;

// per-slot index staging: keyA[cap] keyPQ[2cap] cidx[2cap] lo[cap] hi[cap] keyB[cap]
constexpr uint32_t kIdxWords = 8;

struct Request {
  uint32_t kind, key;
  uint64_t tag;
  std::vector<uint32_t> base, exp;   // n32 words each (exp unused for kind 1)
  Clock::time_point arrival;
};

// a finished batch's requests in slot order: tags, latencies and results
// (n32 words each); Poll consumes them from `next` on
struct DoneChunk {
  std::vector<uint64_t> tags;
  std::vector<double> latency_us;
  std::vector<uint32_t> words;
  size_t next = 0;
};

// device array in the interleaved layout
struct DevArray {
  uint32_t *data = nullptr;
  uint32_t limbs = 0, stride = 0, count = 0;
};

}  // namespace

struct _xmpServer_t {
  xmpHandle_t h = nullptr;
  xmpServerConfig_t cfg{};
  uint32_t bits = 0, n32 = 0, h32 = 0, nkeys = 0;
  // resident keys
  xmpIntegers_t N = nullptr, PQ = nullptr, DPQ = nullptr, CPQ = nullptr;
  // the kind-1 recombination as one Montgomery reduction (crt >= 0: its
  // layout; cp*R mod n and cq*R mod n resident, R = 2^(radix*crt_steps)),
  // else the multiply, add and division chain
  xmpIntegers_t CP2 = nullptr, CQ2 = nullptr;
  int crt = -1;
  uint32_t crt_steps = 0;
  // Montgomery layouts per request kind, chosen per batch by its size: the
  // tuner's choice at a few batch sizes up to the slot capacity, each with
  // its own R^2 mod m (R depends on the layout)
  struct Layout {
    uint32_t count;   // batch size the choice was made for
    int variant;
    uint32_t window;
    uint32_t *r2;     // shared between entries with the same variant
    uint32_t r2_stride;
  };
  std::vector<Layout> lay_n, lay_pq;
  size_t bytes_per_request = 0, budget = 0;
  uint32_t cap = 0, max_inflight = 0;

  struct Slot {
    hipStream_t stream = nullptr;     // the kind-0 (single powm) exponentiation
    hipStream_t stream_b = nullptr;   // the kind-1 (CRT) exponentiation, concurrent with kind 0
    // the short kernels (index staging, imports, the CRT chain's divisions,
    // multiply and add, exports): the same two streams, or streams confined
    // to the reserved CUs so that they never wait behind exponentiations
    hipStream_t short_a = nullptr, short_b = nullptr;
    hipEvent_t done = nullptr, fork = nullptr, join = nullptr;
    hipEvent_t in_a = nullptr, pow_a = nullptr, in_b = nullptr, pow_b = nullptr;
    // pinned staging
    uint32_t *h_in = nullptr;     // [A base | A exp | B c] compact, cap*n32 each
    uint32_t *h_idx = nullptr;    // index arrays
    uint32_t *h_out = nullptr;    // [A out | B out] compact
    // device addresses of the three pinned buffers: the slot's kernels read the
    // requests and write the results there directly, with no copy-engine transfer
    uint32_t *g_in = nullptr, *g_idx = nullptr, *g_out = nullptr;
    uint32_t *d_idx = nullptr;
    DevArray a_base, a_exp, a_out, b_c, b_cm, b_pow, b_prod, b_sum, b_out;
    uint32_t *table_a = nullptr, *table_b = nullptr;
    uint32_t *div_ws = nullptr;   // division workspace for divisors above the register kernels (else null)
    std::vector<Request> reqs;
    uint32_t na = 0, nb = 0;
  };
  std::vector<Slot> slots;

  // Two host threads: the scheduler admits arrived requests into free slots
  // and launches them; the completer retires finished slots and publishes
  // their results.  Everything below is guarded by mu.
  std::mutex mu;
  std::condition_variable cv_sched, cv_drain;
  std::deque<Request> pending;
  std::deque<uint32_t> free_slots;      // slots the scheduler may fill
  std::vector<uint32_t> launched;       // slots on the device, for the completer
  std::deque<DoneChunk> done;           // completed batches in completion order
  uint64_t submitted = 0, completed = 0, batches = 0;
  uint32_t inflight = 0, peak_inflight = 0, peak_batch = 0;
  bool stopping = false, scheduler_exited = false;
  xmpError_t worker_error = xmpErrorSuccess;
  std::thread scheduler, completer;
  // failure-handling test hook: with XMP_SERVER_FAIL_BATCH=k set at creation,
  // the k-th batch launch reports an error after enqueuing its kind-0 work
  uint32_t fail_batch = 0, launches = 0;
};

namespace {

hipError_t alloc_array(xmpHandle_t h, DevArray &a, uint32_t limbs, uint32_t count) {
  a.limbs = limbs;
  a.count = count;
  a.stride = round_up(count, kStrideAlign);
  a.data = (uint32_t *)h->da(sizeof(uint32_t) * (size_t)limbs * a.stride);
  return a.data ? hipSuccess : hipErrorOutOfMemory;
}

Operand op(const DevArray &a, const uint32_t *idx = nullptr, uint32_t idx_count = 0) {
  return Operand{a.data, a.limbs, a.stride, a.count, idx, idx_count};
}
Operand op(xmpIntegers_t x, const uint32_t *idx = nullptr, uint32_t idx_count = 0) {
  return Operand{x->limbs, x->nlimbs, x->stride, x->count, idx, idx_count};
}
OutOperand outop(const DevArray &a) { return OutOperand{a.data, a.limbs, a.stride, nullptr}; }

// The slots' pinned buffers are mapped and coherent (not cached on the
// device): the host rewrites them between batches, and the kernels read the
// requests from them and write the results into them directly.
constexpr unsigned kPinned = hipHostMallocMapped | hipHostMallocCoherent;

// the layout prepared for the batch size nearest (log scale) to `count`
const _xmpServer_t::Layout &pick_layout(const std::vector<_xmpServer_t::Layout> &set, uint32_t count) {
  size_t best = 0;
  double bd = 1e30;
  for (size_t i = 0; i < set.size(); i++) {
    double d = std::fabs(std::log((double)set[i].count) - std::log((double)(count ? count : 1)));
    if (d < bd) {
      bd = d;
      best = i;
    }
  }
  return set[best];
}

// `later` waits for what `earlier` has enqueued so far (nothing when they are one stream)
hipError_t order(hipStream_t earlier, hipStream_t later, hipEvent_t ev) {
  if (earlier == later) return hipSuccess;
  hipError_t e = hipEventRecord(ev, earlier);
  return e != hipSuccess ? e : hipStreamWaitEvent(later, ev, 0);
}

// Launch one batch on its slot's streams: the index staging, then the two
// request kinds concurrently, each importing its requests from the pinned
// input buffer and exporting its results into the pinned output buffer, then
// the completion event once both have finished.  The exponentiations run on
// the slot's two main streams and everything else on its short streams.
hipError_t launch_batch(xmpServer_t s, _xmpServer_t::Slot &sl, uint32_t scale) {
  XMP_TRACE("xmpServer batch");
  const uint32_t n32 = s->n32;
  const uint32_t na = sl.na, nb = sl.nb, cap = s->cap;
  hipStream_t sa = sl.short_a, sb = sl.short_b;
  // index layout in h_idx: keyA[cap] | keyPQ[2cap] | cidx[2cap] | lo[cap] | hi[cap] | keyB[cap]
  uint32_t *keyA = sl.h_idx, *keyPQ = keyA + cap, *cidx = keyPQ + 2 * cap, *lo = cidx + 2 * cap, *hi = lo + cap,
           *keyB = hi + cap;
  for (uint32_t i = 0; i < na; i++) keyA[i] = sl.reqs[i].key;
  for (uint32_t j = 0; j < nb; j++) {
    const Request &r = sl.reqs[na + j];
    keyPQ[j] = r.key;
    keyPQ[nb + j] = s->nkeys + r.key;
    cidx[j] = j;
    cidx[nb + j] = j;
    lo[j] = j;
    hi[j] = nb + j;
    keyB[j] = r.key;
  }
  uint32_t *d_keyA = sl.d_idx, *d_keyPQ = d_keyA + cap, *d_cidx = d_keyPQ + 2 * cap, *d_lo = d_cidx + 2 * cap,
           *d_hi = d_lo + cap, *d_keyB = d_hi + cap;
  hipError_t e;
  // the index arrays are re-read by the kernels, so they are staged in device
  // memory (a one-limb import is a word copy)
  if ((e = launch_import(sl.d_idx, kIdxWords * cap, 1, (const uint8_t *)sl.g_idx, kIdxWords * cap, 1, -1, -1, 0, 4,
                         nullptr, 0, sa)) != hipSuccess)
    return e;
  if ((e = hipEventRecord(sl.fork, sa)) || (e = hipStreamWaitEvent(sb, sl.fork, 0))) return e;
  const bool inject_failure = s->fail_batch != 0 && ++s->launches == s->fail_batch;
  const size_t words = (size_t)cap * n32;
  // the imports read the rows this batch uses straight from the pinned buffer:
  // bases and exponents of the na powm requests, c of the nb CRT requests
  const uint8_t *in_base = (const uint8_t *)sl.g_in, *in_exp = (const uint8_t *)(sl.g_in + words),
                *in_c = (const uint8_t *)(sl.g_in + 2 * words);
  if (na) {
    if ((e = launch_import(sl.a_base.data, sl.a_base.stride, n32, in_base, na, n32, -1, -1, 0, 4, nullptr, 0, sa)) ||
        (e = launch_import(sl.a_exp.data, sl.a_exp.stride, n32, in_exp, na, n32, -1, -1, 0, 4, nullptr, 0, sa)) ||
        (e = order(sa, sl.stream, sl.in_a)))
      return e;
    PowmArgs a;
    a.out = outop(sl.a_out);
    a.base = op(sl.a_base);
    a.exp = op(sl.a_exp);
    a.mod = op(s->N, d_keyA, na);
    const _xmpServer_t::Layout &ly = pick_layout(s->lay_n, na * scale);
    a.r2 = ly.r2;
    a.r2_stride = ly.r2_stride;
    a.table = sl.table_a;
    a.start = 0;
    a.count = na;
    a.window_bits = ly.window;
    if ((e = launch_powm_variant(ly.variant, a, sl.stream)) != hipSuccess || (e = order(sl.stream, sa, sl.pow_a)))
      return e;
    if ((e = launch_export((uint8_t *)sl.g_out, sl.a_out.data, sl.a_out.stride, n32, na, n32, -1, -1, 0, 4, nullptr,
                           0, sa)) != hipSuccess)
      return e;
  }
  if (inject_failure) return hipErrorLaunchFailure;
  if (nb) {
    if ((e = launch_import(sl.b_c.data, sl.b_c.stride, n32, in_c, nb, n32, -1, -1, 0, 4, nullptr, 0, sb))) return e;
    // c mod p_k and c mod q_k in one launch over 2*nb instances
    OutOperand none{nullptr, 0, 0, nullptr};
    if ((e = launch_divmod_ws(none, outop(sl.b_cm), op(sl.b_c, d_cidx, 2 * nb), op(s->PQ, d_keyPQ, 2 * nb), 2 * nb,
                              sl.div_ws, sb)) != hipSuccess ||
        (e = order(sb, sl.stream_b, sl.in_b)))
      return e;
    // mp = cm^dp mod p, mq = cm^dq mod q in one launch
    PowmArgs a;
    a.out = outop(sl.b_pow);
    a.base = op(sl.b_cm);
    a.exp = op(s->DPQ, d_keyPQ, 2 * nb);
    a.mod = op(s->PQ, d_keyPQ, 2 * nb);
    const _xmpServer_t::Layout &ly = pick_layout(s->lay_pq, 2 * nb * scale);
    a.r2 = ly.r2;
    a.r2_stride = ly.r2_stride;
    a.table = sl.table_b;
    a.start = 0;
    a.count = 2 * nb;
    a.window_bits = ly.window;
    if ((e = launch_powm_variant(ly.variant, a, sl.stream_b)) != hipSuccess ||
        (e = order(sl.stream_b, sb, sl.pow_b)))
      return e;
    if (s->crt >= 0) {
      // r = (mp*cp + mq*cq) mod n in one launch
      CrtArgs ca;
      ca.out = outop(sl.b_out);
      ca.mp = op(sl.b_pow, d_lo, nb);
      ca.mq = op(sl.b_pow, d_hi, nb);
      ca.cp2 = op(s->CP2, d_keyB, nb);
      ca.cq2 = op(s->CQ2, d_keyB, nb);
      ca.n = op(s->N, d_keyB, nb);
      ca.steps = s->crt_steps;
      ca.count = nb;
      if ((e = launch_crt_combine(s->crt, ca, sb)) != hipSuccess) return e;
    } else {
      // sp = mp*cp, sq = mq*cq; r = (sp + sq) mod n
      if ((e = launch_mul(outop(sl.b_prod), op(s->CPQ, d_keyPQ, 2 * nb), op(sl.b_pow), 2 * nb, sb)) != hipSuccess)
        return e;
      if ((e = launch_add(outop(sl.b_sum), op(sl.b_prod, d_lo, nb), op(sl.b_prod, d_hi, nb), nb, sb)) != hipSuccess)
        return e;
      if ((e = launch_divmod_ws(none, outop(sl.b_out), op(sl.b_sum), op(s->N, d_keyB, nb), nb, sl.div_ws, sb)) !=
          hipSuccess)
        return e;
    }
    if ((e = launch_export((uint8_t *)(sl.g_out + words), sl.b_out.data, sl.b_out.stride, n32, nb, n32, -1, -1, 0, 4,
                           nullptr, 0, sb)) != hipSuccess)
      return e;
  }
  if ((e = hipEventRecord(sl.join, sb)) || (e = hipStreamWaitEvent(sa, sl.join, 0))) return e;
  return hipEventRecord(sl.done, sa);
}

// Wait until nothing of a slot's batch is still queued on any of its streams.
void sync_slot(_xmpServer_t::Slot &sl) {
  for (hipStream_t st : {sl.stream, sl.stream_b, sl.short_a, sl.short_b})
    if (st) (void)hipStreamSynchronize(st);
  (void)hipGetLastError();
}

// A batch whose launch or execution failed (the slot's streams are idle):
// count its requests as completed without publishing results, make the error
// sticky (Poll, Drain and RunPoisson return it; no further batch is
// admitted) and free the slot.  Called with mu held.
void fail_slot_locked(xmpServer_t s, uint32_t slot) {
  _xmpServer_t::Slot &sl = s->slots[slot];
  s->worker_error = xmpErrorCuda;
  s->completed += sl.reqs.size();
  s->inflight -= (uint32_t)sl.reqs.size();
  sl.reqs.clear();
  s->free_slots.push_back(slot);
}

// Scheduler thread: whenever requests are pending, a slot is free and the
// budget allows, move up to a slot's capacity of them into it and launch.
void scheduler_loop(xmpServer_t s) {
  (void)hipSetDevice(s->h->device);
  roctxNameOsThread("xmpServer scheduler");
  const uint32_t n32 = s->n32;
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    auto ready = [&] {
      return !s->pending.empty() && !s->free_slots.empty() && s->inflight < s->max_inflight;
    };
    s->cv_sched.wait(lk, [&] {
      return s->worker_error != xmpErrorSuccess || ready() || (s->stopping && s->pending.empty());
    });
    if (s->worker_error != xmpErrorSuccess || !ready()) break;   // an error, or stopping with nothing pending
    const uint32_t slot = s->free_slots.front();
    s->free_slots.pop_front();
    _xmpServer_t::Slot &sl = s->slots[slot];
    const uint32_t take = std::min<uint32_t>({s->cap, s->max_inflight - s->inflight, (uint32_t)s->pending.size()});
    // kind-0 requests first, then kind 1 (slot layout); order within a kind is arrival order
    std::vector<Request> b;
    sl.reqs.clear();
    for (uint32_t i = 0; i < take; i++) {
      Request r = std::move(s->pending.front());
      s->pending.pop_front();
      (r.kind == 0 ? sl.reqs : b).push_back(std::move(r));
    }
    sl.na = (uint32_t)sl.reqs.size();
    sl.nb = (uint32_t)b.size();
    for (auto &r : b) sl.reqs.push_back(std::move(r));
    s->inflight += take;
    // the layout is chosen as for a batch `scale` times this one: the slots in
    // flight share the device (or the configured factor)
    const uint32_t scale = s->cfg.layout_scale ? s->cfg.layout_scale : (uint32_t)s->launched.size() + 1;
    s->peak_inflight = std::max(s->peak_inflight, s->inflight);
    s->peak_batch = std::max(s->peak_batch, take);
    s->batches++;
    lk.unlock();
    const size_t words = (size_t)s->cap * n32;
    for (uint32_t i = 0; i < sl.na; i++) {
      std::memcpy(sl.h_in + (size_t)i * n32, sl.reqs[i].base.data(), n32 * 4);
      std::memcpy(sl.h_in + words + (size_t)i * n32, sl.reqs[i].exp.data(), n32 * 4);
    }
    for (uint32_t j = 0; j < sl.nb; j++)
      std::memcpy(sl.h_in + 2 * words + (size_t)j * n32, sl.reqs[sl.na + j].base.data(), n32 * 4);
    const bool ok = launch_batch(s, sl, scale) == hipSuccess;
    // part of a failed batch may be enqueued: wait for all of it before the
    // slot's buffers can be reused
    if (!ok) sync_slot(sl);
    lk.lock();
    if (ok) {
      s->launched.push_back(slot);
    } else {
      fail_slot_locked(s, slot);
      s->cv_drain.notify_all();
    }
  }
  s->scheduler_exited = true;
}

// Completer thread: polls the launched slots' completion events and
// publishes every finished slot's results (in whatever order slots finish).
void completer_loop(xmpServer_t s) {
  (void)hipSetDevice(s->h->device);
  roctxNameOsThread("xmpServer completer");
  const uint32_t n32 = s->n32;
  std::vector<uint32_t> slots;
  while (true) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->launched.empty() && s->scheduler_exited) break;
      slots = s->launched;
    }
    bool progressed = false;
    for (uint32_t slot : slots) {
      _xmpServer_t::Slot &sl = s->slots[slot];
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      progressed = true;
      const auto now = Clock::now();
      if (q != hipSuccess) {
        // the batch failed on the device: its results are not published
        sync_slot(sl);
        std::lock_guard<std::mutex> lk(s->mu);
        s->launched.erase(std::find(s->launched.begin(), s->launched.end(), slot));
        fail_slot_locked(s, slot);
      } else {
        // the results are copied out before the lock is taken, so that
        // submitters and the scheduler never wait for the copy
        DoneChunk c;
        const size_t words = (size_t)s->cap * n32, nreq = sl.reqs.size();
        c.words.resize(nreq * n32);
        std::memcpy(c.words.data(), sl.h_out, sizeof(uint32_t) * (size_t)sl.na * n32);
        std::memcpy(c.words.data() + (size_t)sl.na * n32, sl.h_out + words, sizeof(uint32_t) * (size_t)sl.nb * n32);
        c.tags.resize(nreq);
        c.latency_us.resize(nreq);
        for (size_t i = 0; i < nreq; i++) {
          c.tags[i] = sl.reqs[i].tag;
          c.latency_us[i] = std::chrono::duration<double, std::micro>(now - sl.reqs[i].arrival).count();
        }
        sl.reqs.clear();   // the slot is not free yet: no other thread touches it
        std::lock_guard<std::mutex> lk(s->mu);
        s->launched.erase(std::find(s->launched.begin(), s->launched.end(), slot));
        s->done.push_back(std::move(c));
        s->completed += nreq;
        s->inflight -= (uint32_t)nreq;
        s->free_slots.push_back(slot);
      }
      s->cv_sched.notify_one();
      s->cv_drain.notify_all();
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

void free_server(xmpServer_t s) {
  xmpHandle_t h = s->h;
  for (auto &sl : s->slots) {
    sync_slot(sl);
    for (DevArray *a : {&sl.a_base, &sl.a_exp, &sl.a_out, &sl.b_c, &sl.b_cm, &sl.b_pow, &sl.b_prod, &sl.b_sum,
                        &sl.b_out})
      if (a->data) h->df(a->data);
    if (sl.table_a) h->df(sl.table_a);
    if (sl.table_b) h->df(sl.table_b);
    if (sl.div_ws) h->df(sl.div_ws);
    if (sl.d_idx) h->df(sl.d_idx);
    if (sl.h_in) (void)hipHostFree(sl.h_in);
    if (sl.h_idx) (void)hipHostFree(sl.h_idx);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
    for (hipEvent_t ev : {sl.done, sl.fork, sl.join, sl.in_a, sl.pow_a, sl.in_b, sl.pow_b})
      if (ev) (void)hipEventDestroy(ev);
    if (sl.short_a && sl.short_a != sl.stream) (void)hipStreamDestroy(sl.short_a);
    if (sl.short_b && sl.short_b != sl.stream_b) (void)hipStreamDestroy(sl.short_b);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.stream_b) (void)hipStreamDestroy(sl.stream_b);
  }
  for (xmpIntegers_t x : {s->N, s->PQ, s->DPQ, s->CPQ, s->CP2, s->CQ2})
    if (x) xmpIntegersDestroy(h, x);
  for (auto *set : {&s->lay_n, &s->lay_pq}) {
    std::vector<uint32_t *> freed;
    for (auto &ly : *set)
      if (ly.r2 && std::find(freed.begin(), freed.end(), ly.r2) == freed.end()) {
        h->df(ly.r2);
        freed.push_back(ly.r2);
      }
    set->clear();
  }
}

}  // namespace

extern "C" {

xmpError_t XMPAPI xmpServerCreate(xmpHandle_t handle, xmpServer_t *server, const xmpServerConfig_t *config,
                                  uint32_t bits, uint32_t nkeys, const uint32_t *n, const uint32_t *p,
                                  const uint32_t *q, const uint32_t *dp, const uint32_t *dq, const uint32_t *cp,
                                  const uint32_t *cq) {
  XMP_CHECK_NE(server, nullptr);
  XMP_CHECK_NE(config, nullptr);
  if (!n || !p || !q || !dp || !dq || !cp || !cq) return xmpErrorInvalidParameter;
  if (bits == 0 || bits % 64 != 0 || nkeys == 0) return xmpErrorInvalidParameter;
  if (config->streams == 0 || config->max_batch == 0) return xmpErrorInvalidParameter;
  XMP_SET_DEVICE(handle);
  xmpServer_t s = new _xmpServer_t();
  s->h = handle;
  s->cfg = *config;
  s->bits = bits;
  s->n32 = bits / 32;
  s->h32 = bits / 64;
  s->nkeys = nkeys;
  if (const char *f = std::getenv("XMP_SERVER_FAIL_BATCH")) s->fail_batch = (uint32_t)std::strtoul(f, nullptr, 10);
  xmpError_t e = xmpErrorSuccess;
  auto fail = [&](xmpError_t err) {
    free_server(s);
    delete s;
    return err;
  };
  // resident keys: n (bits), [p;q] and [dp;dq] (bits/2), [cp;cq] (bits)
  std::vector<uint32_t> pq(2 * (size_t)nkeys * s->h32), dpq(2 * (size_t)nkeys * s->h32),
      cpq(2 * (size_t)nkeys * s->n32);
  std::memcpy(pq.data(), p, sizeof(uint32_t) * nkeys * s->h32);
  std::memcpy(pq.data() + (size_t)nkeys * s->h32, q, sizeof(uint32_t) * nkeys * s->h32);
  std::memcpy(dpq.data(), dp, sizeof(uint32_t) * nkeys * s->h32);
  std::memcpy(dpq.data() + (size_t)nkeys * s->h32, dq, sizeof(uint32_t) * nkeys * s->h32);
  std::memcpy(cpq.data(), cp, sizeof(uint32_t) * nkeys * s->n32);
  std::memcpy(cpq.data() + (size_t)nkeys * s->n32, cq, sizeof(uint32_t) * nkeys * s->n32);
  if ((e = xmpIntegersCreate(handle, &s->N, bits, nkeys)) || (e = xmpIntegersCreate(handle, &s->PQ, bits / 2, 2 * nkeys)) ||
      (e = xmpIntegersCreate(handle, &s->DPQ, bits / 2, 2 * nkeys)) ||
      (e = xmpIntegersCreate(handle, &s->CPQ, bits, 2 * nkeys)) ||
      (e = xmpIntegersImport(handle, s->N, s->n32, -1, 4, 0, 0, const_cast<uint32_t *>(n), nkeys)) ||
      (e = xmpIntegersImport(handle, s->PQ, s->h32, -1, 4, 0, 0, pq.data(), 2 * nkeys)) ||
      (e = xmpIntegersImport(handle, s->DPQ, s->h32, -1, 4, 0, 0, dpq.data(), 2 * nkeys)) ||
      (e = xmpIntegersImport(handle, s->CPQ, s->n32, -1, 4, 0, 0, cpq.data(), 2 * nkeys)))
    return fail(e);
  // The kind-1 recombination as one Montgomery reduction when a layout holds
  // n and every n is odd: cp*R mod n and cq*R mod n (a shift by the R
  // exponent, then one division by n) stay resident.
  {
    bool odd = true;
    for (uint32_t k = 0; k < nkeys; k++) odd = odd && (n[(size_t)k * s->n32] & 1u);
    int which = -1;
    uint32_t steps = 0;
    if (odd && crt_combine_plan(bits, bits / 2, &which, &steps)) {
      const int32_t sh = (int32_t)(crt_combine_radix(which) * steps);
      const uint32_t tbits = bits + ((uint32_t)sh + 31u) / 32u * 32u;
      xmpIntegers_t T = nullptr;
      int32_t *d_sh = (int32_t *)handle->da(sizeof(int32_t));
      if (!d_sh) return fail(xmpErrorInvalidCudaMalloc);
      auto done = [&](xmpError_t err) {
        if (T) xmpIntegersDestroy(handle, T);
        handle->df(d_sh);
        return err;
      };
      if ((e = xmpIntegersCreate(handle, &s->CP2, bits, nkeys)) || (e = xmpIntegersCreate(handle, &s->CQ2, bits, nkeys)) ||
          (e = xmpIntegersCreate(handle, &T, tbits, 2 * nkeys)))
        return fail(done(e));
      OutOperand none{nullptr, 0, 0, nullptr};
      const Operand t_p{T->limbs, T->nlimbs, T->stride, nkeys, nullptr, 0};
      const Operand t_q{T->limbs + nkeys, T->nlimbs, T->stride, nkeys, nullptr, 0};
      if (hipMemcpy(d_sh, &sh, sizeof sh, hipMemcpyHostToDevice) != hipSuccess ||
          launch_shift(OutOperand{T->limbs, T->nlimbs, T->stride, nullptr}, op(s->CPQ), d_sh, 1, 2 * nkeys,
                       handle->stream) != hipSuccess ||
          launch_divmod_ws(none, OutOperand{s->CP2->limbs, s->CP2->nlimbs, s->CP2->stride, nullptr}, t_p, op(s->N),
                           nkeys, nullptr, handle->stream) != hipSuccess ||
          launch_divmod_ws(none, OutOperand{s->CQ2->limbs, s->CQ2->nlimbs, s->CQ2->stride, nullptr}, t_q, op(s->N),
                           nkeys, nullptr, handle->stream) != hipSuccess ||
          hipStreamSynchronize(handle->stream) != hipSuccess)
        return fail(done(xmpErrorCuda));
      done(xmpErrorSuccess);
      s->crt = which;
      s->crt_steps = steps;
    }
  }
  // Layouts for batch sizes 64, 128, 256, ... up to the capacity; with prepare,
  // R^2 mod m for each distinct layout; returns the table words one slot
  // needs per kind and the device bytes one in-flight request holds.
  size_t tab_a = 0, tab_b = 0;
  auto plan = [&](uint32_t cap, bool prepare, size_t *per_request) -> xmpError_t {
    size_t req_a = 0, req_b = 0;
    tab_a = tab_b = 0;
    for (int kind = 0; kind < 2; kind++) {
      auto &set = kind == 0 ? s->lay_n : s->lay_pq;
      const uint32_t pbits = kind == 0 ? bits : bits / 2, mult = kind == 0 ? 1 : 2;
      xmpIntegers_t mods = kind == 0 ? s->N : s->PQ;
      std::vector<uint32_t> counts;
      // up to the largest batch size a layout is chosen for: the capacity
      // times the layout scale (at most the number of slots)
      const uint32_t top = cap * (config->layout_scale ? config->layout_scale : config->streams);
      for (uint32_t c = 64; c < top; c *= 2) counts.push_back(c * mult);
      counts.push_back(top * mult);
      for (uint32_t c : counts) {
        PowmChoice ch;
        xmpError_t err = choose_powm(handle, pbits, pbits, c, xmpAlgorithmDefault, &ch);
        if (err != xmpErrorSuccess) return err;
        const PowmVariant v = powm_variant(ch.variant);
        size_t &tab = kind == 0 ? tab_a : tab_b;
        size_t &req = kind == 0 ? req_a : req_b;
        tab = std::max(tab, powm_table_words(v, round_up(cap * mult, kStrideAlign), ch.window_bits));
        req = std::max(req, powm_table_words(v, mult, ch.window_bits) * 4);
        if (!prepare) continue;
        _xmpServer_t::Layout ly{c, ch.variant, ch.window_bits, nullptr, 0};
        for (auto &o : set)
          if (o.variant == ch.variant) {
            ly.r2 = o.r2;
            ly.r2_stride = o.r2_stride;
          }
        if (!ly.r2 && (err = montgomery_prepare(handle, mods, ch.variant, &ly.r2, &ly.r2_stride))) return err;
        set.push_back(ly);
      }
    }
    // staging, intermediates and the table share of one request (kind-1 worst case)
    *per_request = 4 * ((size_t)s->n32 * 8 + (size_t)s->h32 * 4 + 2 * ((size_t)s->n32 + s->h32) + s->n32 + 2 +
                        kIdxWords) +
                   std::max(req_a, req_b);
    return xmpErrorSuccess;
  };
  // The HBM budget: the configured bytes, else half of the device memory free
  // now.  Every slot's buffers are allocated here for its capacity, so the
  // capacity shrinks until all slots fit the budget, and the scheduler never
  // has more requests in flight than the budget holds.
  size_t budget = config->hbm_budget;
  if (!budget) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(xmpErrorCuda);
    budget = free_b / 2;
  }
  size_t bpr = 0;
  if ((e = plan(config->max_batch, false, &bpr))) return fail(e);
  const size_t fit = budget / (bpr * config->streams);
  s->cap = (uint32_t)std::max<size_t>(1, std::min<size_t>(config->max_batch, fit));
  if ((e = plan(s->cap, true, &s->bytes_per_request))) return fail(e);
  s->budget = budget;
  s->max_inflight = (uint32_t)std::max<size_t>(1, std::min<size_t>(budget / s->bytes_per_request,
                                                                     (size_t)s->cap * config->streams));
  const uint32_t cap = s->cap;
  // CU masks: the exponentiations avoid the reserved CUs, the short kernels use only them
  std::vector<uint32_t> mask_main, mask_short;
  if (config->reserved_cus) {
    const uint32_t cus = handle->cuCount;
    if (config->reserved_cus >= cus) return fail(xmpErrorInvalidParameter);
    mask_main.assign((cus + 31) / 32, 0);
    mask_short.assign((cus + 31) / 32, 0);
    for (uint32_t c = 0; c < cus; c++) (c < config->reserved_cus ? mask_short : mask_main)[c / 32] |= 1u << (c % 32);
  }
  auto make_stream = [&](hipStream_t *st, const std::vector<uint32_t> &mask) {
    return mask.empty() ? hipStreamCreate(st) : hipExtStreamCreateWithCUMask(st, (uint32_t)mask.size(), mask.data());
  };
  s->slots.resize(config->streams);
  for (uint32_t k = 0; k < config->streams; k++) {
    auto &sl = s->slots[k];
    const size_t words = (size_t)cap * s->n32;
    if (make_stream(&sl.stream, mask_main) != hipSuccess || make_stream(&sl.stream_b, mask_main) != hipSuccess)
      return fail(xmpErrorCuda);
    if (config->reserved_cus) {
      if (make_stream(&sl.short_a, mask_short) != hipSuccess || make_stream(&sl.short_b, mask_short) != hipSuccess)
        return fail(xmpErrorCuda);
    } else {
      sl.short_a = sl.stream;
      sl.short_b = sl.stream_b;
    }
    for (hipEvent_t *ev : {&sl.done, &sl.fork, &sl.join, &sl.in_a, &sl.pow_a, &sl.in_b, &sl.pow_b})
      if (hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess) return fail(xmpErrorCuda);
    if (hipHostMalloc((void **)&sl.h_in, sizeof(uint32_t) * 3 * words, kPinned) != hipSuccess ||
        hipHostMalloc((void **)&sl.h_idx, sizeof(uint32_t) * kIdxWords * cap, kPinned) != hipSuccess ||
        hipHostMalloc((void **)&sl.h_out, sizeof(uint32_t) * 2 * words, kPinned) != hipSuccess ||
        hipHostGetDevicePointer((void **)&sl.g_in, sl.h_in, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&sl.g_idx, sl.h_idx, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&sl.g_out, sl.h_out, 0) != hipSuccess)
      return fail(xmpErrorCuda);
    sl.d_idx = (uint32_t *)handle->da(sizeof(uint32_t) * kIdxWords * cap);
    sl.table_a = (uint32_t *)handle->da(sizeof(uint32_t) * tab_a);
    sl.table_b = (uint32_t *)handle->da(sizeof(uint32_t) * tab_b);
    if (!sl.d_idx || !sl.table_a || !sl.table_b) return fail(xmpErrorInvalidCudaMalloc);
    // the CRT chain's two divisions (c mod p/q over 2*cap instances, the sum
    // mod n over cap) share one workspace, needed above 128-limb divisors
    const size_t div_words = std::max(divmod_scratch_words(s->h32, 2 * cap), divmod_scratch_words(s->n32, cap));
    if (div_words && !(sl.div_ws = (uint32_t *)handle->da(sizeof(uint32_t) * div_words)))
      return fail(xmpErrorInvalidCudaMalloc);
    if (alloc_array(handle, sl.a_base, s->n32, cap) || alloc_array(handle, sl.a_exp, s->n32, cap) ||
        alloc_array(handle, sl.a_out, s->n32, cap) || alloc_array(handle, sl.b_c, s->n32, cap) ||
        alloc_array(handle, sl.b_cm, s->h32, 2 * cap) || alloc_array(handle, sl.b_pow, s->h32, 2 * cap) ||
        alloc_array(handle, sl.b_out, s->n32, cap))
      return fail(xmpErrorInvalidCudaMalloc);
    // the products and their sum only for the multiply, add and division chain
    if (s->crt < 0 && (alloc_array(handle, sl.b_prod, s->n32 + s->h32, 2 * cap) ||
                       alloc_array(handle, sl.b_sum, s->n32 + s->h32 + 1, cap)))
      return fail(xmpErrorInvalidCudaMalloc);
    s->free_slots.push_back(k);
  }
  if (hipDeviceSynchronize() != hipSuccess) return fail(xmpErrorCuda);
  s->scheduler = std::thread(scheduler_loop, s);
  s->completer = std::thread(completer_loop, s);
  *server = s;
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpServerSubmit(xmpServer_t server, uint32_t kind, uint32_t key, const uint32_t *base,
                                  const uint32_t *exp, uint64_t tag) {
  XMP_CHECK_NE(server, nullptr);
  if (kind > 1 || key >= server->nkeys || !base || (kind == 0 && !exp)) return xmpErrorInvalidParameter;
  Request r;
  r.kind = kind;
  r.key = key;
  r.tag = tag;
  r.base.assign(base, base + server->n32);
  if (kind == 0) r.exp.assign(exp, exp + server->n32);
  r.arrival = Clock::now();
  {
    std::lock_guard<std::mutex> lk(server->mu);
    server->pending.push_back(std::move(r));
    server->submitted++;
  }
  server->cv_sched.notify_one();
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpServerPoll(xmpServer_t server, uint32_t max, uint64_t *tags, uint32_t *results,
                                double *latency_us, uint32_t *n_out) {
  XMP_CHECK_NE(server, nullptr);
  XMP_CHECK_NE(n_out, nullptr);
  const uint32_t n32 = server->n32;
  std::lock_guard<std::mutex> lk(server->mu);
  uint32_t k = 0;
  while (k < max && !server->done.empty()) {
    DoneChunk &c = server->done.front();
    const size_t n = std::min<size_t>(max - k, c.tags.size() - c.next);
    if (tags) std::memcpy(tags + k, c.tags.data() + c.next, sizeof(uint64_t) * n);
    if (results)
      std::memcpy(results + (size_t)k * n32, c.words.data() + c.next * n32, sizeof(uint32_t) * n * n32);
    if (latency_us) std::memcpy(latency_us + k, c.latency_us.data() + c.next, sizeof(double) * n);
    c.next += n;
    k += (uint32_t)n;
    if (c.next == c.tags.size()) server->done.pop_front();
  }
  *n_out = k;
  return server->worker_error;
}

xmpError_t XMPAPI xmpServerDrain(xmpServer_t server) {
  XMP_CHECK_NE(server, nullptr);
  std::unique_lock<std::mutex> lk(server->mu);
  server->cv_drain.wait(lk, [&] { return server->completed == server->submitted || server->worker_error; });
  return server->worker_error;
}

xmpError_t XMPAPI xmpServerGetStats(xmpServer_t server, xmpServerStats_t *stats) {
  XMP_CHECK_NE(server, nullptr);
  XMP_CHECK_NE(stats, nullptr);
  std::lock_guard<std::mutex> lk(server->mu);
  stats->batches = server->batches;
  stats->slot_capacity = server->cap;
  stats->max_inflight = server->max_inflight;
  stats->peak_inflight = server->peak_inflight;
  stats->peak_batch = server->peak_batch;
  stats->bytes_per_request = server->bytes_per_request;
  stats->hbm_budget = server->budget;
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpServerDestroy(xmpServer_t server) {
  XMP_CHECK_NE(server, nullptr);
  {
    std::lock_guard<std::mutex> lk(server->mu);
    server->stopping = true;
  }
  server->cv_sched.notify_all();
  if (server->scheduler.joinable()) server->scheduler.join();
  if (server->completer.joinable()) server->completer.join();
  XMP_SET_DEVICE(server->h);
  free_server(server);
  delete server;
  return xmpErrorSuccess;
}

xmpError_t XMPAPI xmpServerRunPoisson(xmpServer_t server, double rate_per_s, uint32_t requests, uint64_t seed,
                                      const uint32_t *kinds, const uint32_t *keys, const uint32_t *bases,
                                      const uint32_t *exps, uint32_t *results, double *latency_us, double *wall_s) {
  XMP_CHECK_NE(server, nullptr);
  if (rate_per_s <= 0 || requests == 0 || !kinds || !keys || !bases || !exps) return xmpErrorInvalidParameter;
  const uint32_t n32 = server->n32;
  // the requests and their arrival offsets are prepared before the clock starts
  std::vector<Request> reqs(requests);
  std::vector<double> due(requests);
  {
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_per_s);
    double t = 0;
    for (uint32_t i = 0; i < requests; i++) {
      if (kinds[i] > 1 || keys[i] >= server->nkeys) return xmpErrorInvalidParameter;
      t += gap(rng);
      due[i] = t;
      Request &r = reqs[i];
      r.kind = kinds[i];
      r.key = keys[i];
      r.tag = i;
      r.base.assign(bases + (size_t)i * n32, bases + (size_t)(i + 1) * n32);
      if (r.kind == 0) r.exp.assign(exps + (size_t)i * n32, exps + (size_t)(i + 1) * n32);
    }
  }
  // every request is submitted at its arrival time; the ones already due
  // when the generator gets to them are submitted together, under one lock
  const auto t0 = Clock::now();
  uint32_t i = 0;
  while (i < requests) {
    const auto first = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(due[i]));
    while (Clock::now() < first) {
    }
    const auto now = Clock::now();
    const double elapsed = std::chrono::duration<double>(now - t0).count();
    uint32_t end = i + 1;
    while (end < requests && due[end] <= elapsed) end++;
    {
      std::lock_guard<std::mutex> lk(server->mu);
      for (uint32_t k = i; k < end; k++) {
        reqs[k].arrival = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(due[k]));
        server->pending.push_back(std::move(reqs[k]));
      }
      server->submitted += end - i;
    }
    server->cv_sched.notify_one();
    i = end;
  }
  xmpError_t e = xmpServerDrain(server);
  const auto t1 = Clock::now();
  if (wall_s) *wall_s = std::chrono::duration<double>(t1 - t0).count();
  std::vector<uint64_t> tags(requests);
  std::vector<uint32_t> res((size_t)requests * n32);
  std::vector<double> lat(requests);
  uint32_t got = 0;
  while (got < requests) {
    uint32_t k = 0;
    xmpError_t pe = xmpServerPoll(server, requests - got, tags.data() + got, res.data() + (size_t)got * n32,
                                  lat.data() + got, &k);
    if (pe != xmpErrorSuccess) return pe;
    if (k == 0) break;
    got += k;
  }
  for (uint32_t k = 0; k < got; k++) {
    const uint64_t tg = tags[k];
    if (tg >= requests) continue;
    if (results) std::memcpy(results + tg * n32, res.data() + (size_t)k * n32, sizeof(uint32_t) * n32);
    if (latency_us) latency_us[tg] = lat[k];
  }
  return e != xmpErrorSuccess ? e : (got == requests ? xmpErrorSuccess : xmpErrorUnspecified);
}

}  // extern "C"
