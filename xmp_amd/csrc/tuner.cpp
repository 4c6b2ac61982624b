// tuner.cpp -- runtime choice of powm layout: measured table, else model.
#include "tuner.hpp"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace xmp {

namespace {

struct TunedEntry {
  uint32_t precision, exp_bits, count;
  int lanes, digits, radix;
  uint32_t window;
  double rate;
};

std::mutex g_mu;
std::vector<TunedEntry> g_table;
bool g_loaded = false;

bool find_number(const std::string &obj, const char *key, double *out) {
  std::string k = std::string("\"") + key + "\"";
  size_t p = obj.find(k);
  if (p == std::string::npos) return false;
  p = obj.find(':', p);
  if (p == std::string::npos) return false;
  *out = std::strtod(obj.c_str() + p + 1, nullptr);
  return true;
}

void parse_table(const std::string &text, std::vector<TunedEntry> *dst) {
  size_t p = text.find("\"entries\"");
  if (p == std::string::npos) return;
  while (true) {
    size_t a = text.find('{', p);
    if (a == std::string::npos) break;
    size_t b = text.find('}', a);
    if (b == std::string::npos) break;
    std::string obj = text.substr(a, b - a + 1);
    double pr, eb, cnt, l, k, d, w, rate = 0;
    if (find_number(obj, "precision", &pr) && find_number(obj, "exp_bits", &eb) && find_number(obj, "count", &cnt) &&
        find_number(obj, "lanes", &l) && find_number(obj, "digits", &k) && find_number(obj, "radix", &d) &&
        find_number(obj, "window", &w)) {
      find_number(obj, "powm_per_s", &rate);
      dst->push_back({(uint32_t)pr, (uint32_t)eb, (uint32_t)cnt, (int)l, (int)k, (int)d, (uint32_t)w, rate});
    }
    p = b + 1;
  }
}

// the measured table shipped next to the library (xmp_amd/tune_gfx950.json),
// overridden by $XMP_TUNE_FILE
std::string default_table_path() {
  Dl_info info;
  if (dladdr((void *)&default_table_path, &info) == 0 || !info.dli_fname) return std::string();
  std::string so(info.dli_fname);
  size_t slash = so.rfind('/');
  return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tune_gfx950.json";
}

void ensure_loaded() {
  if (g_loaded) return;
  g_loaded = true;
  const char *env = std::getenv("XMP_TUNE_FILE");
  std::string path = env ? std::string(env) : default_table_path();
  if (path.empty()) return;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return;
  std::string text;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
  std::fclose(f);
  parse_table(text, &g_table);
}

}  // namespace

// Load (or replace) the measured table from JSON text.
int tuner_load_text(const char *text) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_loaded = true;
  std::vector<TunedEntry> t;
  parse_table(text ? std::string(text) : std::string(), &t);
  g_table.swap(t);
  return (int)g_table.size();
}

uint32_t best_window_bits(uint32_t exp_bits) {
  uint32_t best = 1;
  double best_cost = 1e30;
  for (uint32_t w = 1; w <= 8; w++) {
    // table build (2^w products) + one multiply per window; squarings are fixed
    double cost = (double)(1u << w) + std::ceil((double)exp_bits / w);
    if (cost < best_cost) {
      best_cost = cost;
      best = w;
    }
  }
  return best;
}

double model_cost(const PowmVariant &v, uint32_t precision, uint32_t exp_bits, uint32_t count, uint32_t window_bits,
                  uint32_t cu_count) {
  (void)precision;
  const double nd = (double)(v.lanes * v.digits);
  // per CIOS step: 2K multiply-adds plus the quotient/shift work (cross-lane when L > 1)
  const double step = 2.0 * v.digits + (v.lanes > 1 ? 14.0 : 5.0);
  const double mulmod = nd * step + 4.0 * v.digits * v.lanes;   // + normalisation
  const double products = (double)(1u << window_bits) + exp_bits + std::ceil((double)exp_bits / window_bits) + 2;
  const double per_wave = products * mulmod;
  // waves in flight: two per SIMD (register budget), four SIMDs per CU
  const double slots = 2.0 * 4.0 * cu_count;
  const double waves = std::ceil((double)count * v.lanes / 64.0);
  // a wave alone on its SIMD issues the 64-bit multiply at half rate
  const double rounds = std::ceil(waves / slots);
  const double fill = std::min(1.0, waves / (slots / 2.0));
  // more than 40 digits per lane overflow the two-wave register budget: the
  // 2x55 layout spills over a kilobyte per lane and measured 4-8x slower than
  // 4x28 at the same 3072-bit sizes
  const double spill = v.digits > 40 ? 4.0 : 1.0;
  return rounds * per_wave * (fill < 1.0 ? 2.0 : 1.0) * spill;
}

xmpError_t choose_powm(xmpHandle_t h, uint32_t precision, uint32_t exp_bits, uint32_t count, xmpAlgorithm_t alg,
                       PowmChoice *out, bool batch_only) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  auto allowed = [&](const PowmVariant &v) {
    if (!variant_fits(v, precision)) return false;
    if (batch_only && !powm_variant_batchable(powm_find_variant(v.lanes, v.digits, v.radix))) return false;
    if (alg == xmpAlgorithmRegMP) return v.lanes == 1;
    if (alg == xmpAlgorithmDistributedMP) return v.lanes > 1;
    return true;
  };
  // Measured entries for this precision, at the exponent size measured
  // nearest the request's, taken at the smallest measured batch size that is
  // at least `count` (the largest one for batches beyond every measured
  // size); the fastest allowed layout there wins.  A layout's time grows in
  // steps as its waves outgrow the resident wave slots, so a batch between
  // two measured sizes can cost up to the larger size's time: pricing it at
  // that bound picks the layout whose worst case is best.  (Interpolating
  // log-log between the two sizes sent 12,936 8192-bit instances to 16x19,
  // 772 ms, where 8x37 takes about 742.)
  {
    uint32_t eb = 0;
    double ebd = 1e30;
    for (const TunedEntry &e : g_table) {
      if (e.precision != precision || !e.exp_bits) continue;
      double d = std::fabs(std::log((double)e.exp_bits) - std::log((double)std::max(exp_bits, 1u)));
      if (d < ebd) {
        ebd = d;
        eb = e.exp_bits;
      }
    }
    uint32_t lo = 0, hi = 0;   // largest measured size below count, smallest at or above it
    for (const TunedEntry &e : g_table) {
      if (e.precision != precision || e.exp_bits != eb || !eb) continue;
      if (e.count < count) lo = std::max(lo, e.count);
      if (e.count >= count) hi = hi ? std::min(hi, e.count) : e.count;
    }
    const uint32_t size = hi ? hi : lo;
    const TunedEntry *pick = nullptr;
    for (const TunedEntry &a : g_table) {
      if (!size || a.precision != precision || a.exp_bits != eb || a.count != size || a.rate <= 0) continue;
      int vi = powm_find_variant(a.lanes, a.digits, a.radix);
      if (vi < 0 || !allowed(powm_variant(vi))) continue;
      if (!pick || a.rate > pick->rate) pick = &a;
    }
    if (pick) {
      out->variant = powm_find_variant(pick->lanes, pick->digits, pick->radix);
      out->window_bits = eb == exp_bits ? pick->window : best_window_bits(exp_bits);
      return xmpErrorSuccess;
    }
  }
  // entries without a rate: the batch size nearest the request (log scale,
  // same exponent size preferred), then the fastest allowed layout there
  const TunedEntry *best = nullptr;
  double best_d = 1e30, best_rate = -1;
  for (const TunedEntry &e : g_table) {
    if (e.precision != precision) continue;
    int vi = powm_find_variant(e.lanes, e.digits, e.radix);
    if (vi < 0 || !allowed(powm_variant(vi))) continue;
    double d = std::fabs(std::log((double)e.count + 1) - std::log((double)count + 1)) +
               (e.exp_bits == exp_bits ? 0.0 : 10.0);
    if (d < best_d - 1e-9 || (std::fabs(d - best_d) <= 1e-9 && e.rate > best_rate)) {
      best_d = d;
      best_rate = e.rate;
      best = &e;
    }
  }
  if (best) {
    out->variant = powm_find_variant(best->lanes, best->digits, best->radix);
    out->window_bits = best->exp_bits == exp_bits ? best->window : best_window_bits(exp_bits);
    return xmpErrorSuccess;
  }
  const uint32_t w = best_window_bits(exp_bits);
  int bi = -1;
  double bc = 1e300;
  for (int i = 0; i < powm_variant_count(); i++) {
    PowmVariant v = powm_variant(i);
    if (!allowed(v)) continue;
    double c = model_cost(v, precision, exp_bits, count, w, h->cuCount ? h->cuCount : 256);
    if (c < bc) {
      bc = c;
      bi = i;
    }
  }
  if (bi < 0) return xmpErrorUnsupported;
  out->variant = bi;
  out->window_bits = w;
  return xmpErrorSuccess;
}

}  // namespace xmp
