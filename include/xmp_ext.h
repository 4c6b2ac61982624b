/*
 * xmp_ext.h -- extensions of the MI355X big-integer library beyond the
 * reference API (include/xmp.h).  None of these has a counterpart in the
 * reference; each is listed with the reference mechanism it builds on.
 *
 *  - Montgomery contexts: moduli kept resident on the device together with
 *    R^2 mod m and the chosen layout, so repeated exponentiations skip the
 *    per-call set-up (the reference recomputes a*R mod m by long division on
 *    every call, src/include/powm_implementations.h:379-443).
 *  - Mixed-size batches: one persistent, graph-captured launch over several
 *    precision classes (the reference issues one xmpIntegersPowm per class,
 *    src/powm.cu:250-307).
 *  - Tuner: measured choice of lane layout and window per precision
 *    (replaces src/tune/tune.cu and the compiled tables of src/powm.cu:117-236).
 *  - Serving: a continuous-batching request loop over resident moduli.
 */
#pragma once

#include "xmp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Montgomery contexts ---------------- */
typedef struct _xmpMontgomery_t *xmpMontgomery_t;

/* Precompute R^2 mod m for the `mod` array (all of its instances) and pick
 * the Montgomery layout for exponents of exp_precision bits and about
 * count_hint instances per call.  `mod` must outlive the context and must not
 * be modified while the context is used.  Moduli must be odd. */
xmpError_t XMPAPI xmpMontgomeryCreate(xmpHandle_t handle, xmpMontgomery_t *ctx, xmpIntegers_t mod,
                                      uint32_t exp_precision, uint32_t count_hint);
xmpError_t XMPAPI xmpMontgomeryDestroy(xmpHandle_t handle, xmpMontgomery_t ctx);
/* layout chosen for the context: lanes per instance, digits per lane, digit bits, window bits */
xmpError_t XMPAPI xmpMontgomeryGetLayout(xmpHandle_t handle, xmpMontgomery_t ctx, uint32_t *lanes,
                                         uint32_t *digits, uint32_t *radix, uint32_t *window_bits);

/* copy the context's resident R^2 mod m (R = 2^(lanes*digits*radix) of its
 * layout) into `out` (same precision as the moduli, count >= theirs); waits
 * for the copy.  For inspection and tests. */
xmpError_t XMPAPI xmpMontgomeryGetR2(xmpHandle_t handle, xmpMontgomery_t ctx, xmpIntegers_t out);

/* out = base^exp mod m using a context: one kernel launch on the handle's
 * stream.  Index semantics as xmpIntegersPowm (policy slots 0..3; slot 3
 * selects the modulus inside the context's array). */
xmpError_t XMPAPI xmpIntegersPowmMontgomeryAsync(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t base,
                                                 const xmpIntegers_t exp, xmpMontgomery_t ctx, uint32_t count);

/* ---------------- mixed-size batches ---------------- */
typedef struct _xmpBatch_t *xmpBatch_t;

/* A batch of `classes` precision classes; class c computes
 * out[c][i] = base[c][i]^exp[c][i] mod mod[c][i] for i < counts[c].
 * The arrays are bound at creation (they must outlive the batch); their
 * contents may change between launches.  The launch is one persistent kernel
 * whose workgroups pull tiles from a device queue ordered by cost, captured
 * in a HIP graph on first launch. */
xmpError_t XMPAPI xmpBatchCreate(xmpHandle_t handle, xmpBatch_t *batch, uint32_t classes, xmpIntegers_t *out,
                                 xmpIntegers_t *base, xmpIntegers_t *exp, xmpIntegers_t *mod,
                                 const uint32_t *counts);
xmpError_t XMPAPI xmpBatchDestroy(xmpHandle_t handle, xmpBatch_t batch);
/* recompute the per-modulus constants (after the moduli changed) */
xmpError_t XMPAPI xmpBatchPrepare(xmpHandle_t handle, xmpBatch_t batch);
/* one persistent launch (graph replay) on the handle's stream */
xmpError_t XMPAPI xmpBatchLaunchAsync(xmpHandle_t handle, xmpBatch_t batch);
/* the same work as separate plain launches, one per class (reference behaviour) */
xmpError_t XMPAPI xmpBatchLaunchPlainAsync(xmpHandle_t handle, xmpBatch_t batch);

/* ---------------- serving ---------------- */
typedef struct _xmpServer_t *xmpServer_t;

typedef struct {
  uint32_t streams;       /* batch slots in flight at once (each on two streams: the single powm
                             requests and the CRT requests of a batch run concurrently; set
                             GPU_MAX_HW_QUEUES >= 2 * streams for all of them to overlap) */
  uint32_t max_batch;     /* requests per batch slot (lowered to fit the budget) */
  size_t hbm_budget;      /* device bytes for the slots' request state and the requests in flight;
                             0 = half of the device memory free at creation */
  uint32_t reserved_cus;  /* CUs kept for the short kernels (imports, the CRT chain's divisions,
                             multiply and add, exports) on two more streams per slot, which the
                             exponentiations do not use; 0 = none (set GPU_MAX_HW_QUEUES >=
                             4 * streams when used) */
  uint32_t layout_scale;  /* a batch's Montgomery layout is the tuner's choice for a batch this many
                             times larger; 0 = the number of slots on the device when it launches
                             (they share it), 1 = the batch alone */
} xmpServerConfig_t;

typedef struct {
  uint64_t batches;          /* batches launched */
  uint32_t slot_capacity;    /* requests per batch after the budget */
  uint32_t max_inflight;     /* requests the budget lets be in flight at once */
  uint32_t peak_inflight;    /* most requests that were in flight at once */
  uint32_t peak_batch;       /* largest batch launched */
  size_t bytes_per_request;  /* device bytes one request in flight holds */
  size_t hbm_budget;         /* the budget in force */
} xmpServerStats_t;

/* Keys (host arrays, `nkeys` each, 32-bit little-endian limbs, compact): n
 * (bits), p and q (bits/2, odd, coprime), dp and dq (bits/2), cp and cq
 * (bits, with cp = 1 mod p, 0 mod q and cq = 0 mod p, 1 mod q).  They stay
 * resident on the device with their Montgomery constants.  A scheduler
 * thread batches arriving requests into free slots within the HBM budget and
 * a completer thread publishes each finished batch's results. */
xmpError_t XMPAPI xmpServerCreate(xmpHandle_t handle, xmpServer_t *server, const xmpServerConfig_t *config,
                                  uint32_t bits, uint32_t nkeys, const uint32_t *n, const uint32_t *p,
                                  const uint32_t *q, const uint32_t *dp, const uint32_t *dq, const uint32_t *cp,
                                  const uint32_t *cq);
xmpError_t XMPAPI xmpServerDestroy(xmpServer_t server);
/* kind 0: base^exp mod n[key]; kind 1: (c^dp mod p)*cp + (c^dq mod q)*cq mod n
 * for c = base (exp ignored).  base/exp: bits/32 words. */
xmpError_t XMPAPI xmpServerSubmit(xmpServer_t server, uint32_t kind, uint32_t key, const uint32_t *base,
                                  const uint32_t *exp, uint64_t tag);
/* up to `max` completed requests: tags, results (bits/32 words each), latency in microseconds */
xmpError_t XMPAPI xmpServerPoll(xmpServer_t server, uint32_t max, uint64_t *tags, uint32_t *results,
                                double *latency_us, uint32_t *n_out);
/* wait until every submitted request has completed */
xmpError_t XMPAPI xmpServerDrain(xmpServer_t server);
/* admission and budget counters since creation */
xmpError_t XMPAPI xmpServerGetStats(xmpServer_t server, xmpServerStats_t *stats);
/* load generator: `requests` requests arriving as a Poisson process of the
 * given rate (seeded); request i is (kinds[i], keys[i], bases[i], exps[i]).
 * Fills results and per-request latency (indexed by request) and the wall time. */
xmpError_t XMPAPI xmpServerRunPoisson(xmpServer_t server, double rate_per_s, uint32_t requests, uint64_t seed,
                                      const uint32_t *kinds, const uint32_t *keys, const uint32_t *bases,
                                      const uint32_t *exps, uint32_t *results, double *latency_us, double *wall_s);

/* ---------------- tuner ---------------- */
/* Measure every compiled layout that fits (and is at most about three times
 * the precision's width), with the window widths next to the modelled best, for
 * powm at `precision` bits with `count` instances on this device;
 * write one JSON object per measurement, comma separated, to `json`
 * (capacity `json_bytes`).  Returns the number of measurements in *measured. */
xmpError_t XMPAPI xmpTunerRun(xmpHandle_t handle, uint32_t precision, uint32_t exp_precision, uint32_t count,
                              char *json, size_t json_bytes, uint32_t *measured);
/* Load a tuning table (JSON text as written by xmpTunerRun or the tune tool). */
xmpError_t XMPAPI xmpTunerLoad(const char *json);

/* ---------------- allocation-free exponentiation ---------------- */
/* No counterpart in the reference (whose xmpIntegersPowm grows the handle's
 * scratch with cudaMalloc, src/xmp.cu:105-129).  For callers that cannot call
 * hipMalloc: code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation.  Same arguments, index slots, results and errors as
 * xmpIntegersPowmAsync, but the call allocates nothing: its workspace (R^2
 * mod m, the window table, an in-place output) is a 256 MiB device array
 * inside the library, one per device, held from the first call on that
 * device until the process ends.  A call whose smallest plan (window down to
 * 2 bits, chunks down to 64 instances) does not fit 256 MiB returns
 * xmpErrorIncreaseScratchLimit.  Calls on one device share that array:
 * calls that overlap on one device race (order them on one stream). */
xmpError_t XMPAPI xmpIntegersPowmNoAllocAsync(xmpHandle_t handle, xmpIntegers_t out, const xmpIntegers_t base,
                                              const xmpIntegers_t exp, const xmpIntegers_t mod, uint32_t count);

/* ---------------- diagnostics ---------------- */
/* Issue-rate probe: independent v_mad_u64_u32 streams on every CU with
 * `waves_per_simd` waves per SIMD (1..8); returns multiply-adds per clock per
 * CU (64 = one per lane per clock) and the probe's kernel time. */
xmpError_t XMPAPI xmpProbeMadRate(xmpHandle_t handle, uint32_t waves_per_simd, double *mads_per_clk_per_cu,
                                  double *seconds);

#ifdef __cplusplus
}
#endif
