/* The public headers compile as C (C11): a C caller sees the same API. */
#include "xmp.h"
#include "xmp_ext.h"

int xmp_header_c_check(xmpHandle_t h, xmpIntegers_t x) {
  uint32_t precision = 0;
  xmpExecutionPolicyValue_t v;
  v.size = 0;
  (void)v;
  return (int)xmpIntegersGetPrecision(h, x, &precision);
}
