// k_powm_latency.hip -- the Montgomery engine's few-digit latency layouts
// (4x10, 8x5 and 16x3 for 1024-bit moduli, 8x10 and 16x5 for 2048-bit), built from
// the templates of k_powm.hip in a translation unit of their own.  k_powm.hip
// is compiled with LLVM's ILP-oriented scheduler, which speeds up the
// many-digit throughput layouts; these layouts, whose steps are short and
// latency-bound, run 1-3 % faster with the default scheduler, so this unit
// keeps it.  Only launch_powm_latency is defined here.
#define XMP_POWM_LATENCY_UNIT 1
#include "k_powm.hip"
