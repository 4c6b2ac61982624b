// tuner.hpp -- choice of Montgomery layout and window for a powm launch.
//
// Replaces the reference's compiled-in latency tables and their lookup
// (src/powm.cu:110-236, src/tune/tune_*.h).  The runtime consults a table
// measured on the device by the project autotuner (xmpTunerRun in
// xmp_ext.h) and falls back to an analytic cost model when no measurement
// covers the request.
#pragma once

#include <cstdint>

#include "internal.hpp"
#include "powm.hpp"

namespace xmp {

struct PowmChoice {
  int variant;
  uint32_t window_bits;
};

// window width minimising table build + one multiply per window for an
// exponent of `exp_bits` bits
uint32_t best_window_bits(uint32_t exp_bits);

// modelled time (arbitrary units) of `count` instances with variant v
double model_cost(const PowmVariant &v, uint32_t precision, uint32_t exp_bits, uint32_t count, uint32_t window_bits,
                  uint32_t cu_count);

// batch_only: restrict to layouts the persistent mixed-size kernel runs
xmpError_t choose_powm(xmpHandle_t h, uint32_t precision, uint32_t exp_bits, uint32_t count, xmpAlgorithm_t alg,
                       PowmChoice *out, bool batch_only = false);

}  // namespace xmp
