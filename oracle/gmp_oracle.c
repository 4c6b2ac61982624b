/*
 * gmp_oracle.c -- GMP mpz as the reference test suite uses it.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline, never the product).
 *
 * XMP's unit tests compare every device result with GMP mpz on the host
 * (unit_tests/test.cu:1214 mpz_powm, :1616 and :1780 mpz_mul/mpz_add,
 * :2317-2322 mpz_sub plus 2^P wrap, :2598-2813 mpz_div/mpz_mod/mpz_divmod)
 * and its perf harness times mpz_powm_sec with N host threads
 * (perf_tests/main.cu:153-229).  The reference pins gmp-6.0.0a
 * (unit_tests/Makefile:2-3); this image has the GMP 6.2.1 runtime
 * (libgmp.so.10) but no gmp.h, so the library is opened with dlopen and
 * the handful of entry points used here are declared locally.  All results
 * are exact integer functions, so the version does not change them.
 */
#include <dlfcn.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define GO_API __attribute__((visibility("default")))

/* Layout of GMP's __mpz_struct (public ABI, stable since GMP 4). */
typedef struct {
  int alloc;
  int size;
  void *d;
} mpz_struct;
typedef mpz_struct mpz_t[1];

static void (*p_init)(mpz_struct *);
static void (*p_clear)(mpz_struct *);
static void (*p_import)(mpz_struct *, size_t, int, size_t, int, size_t, const void *);
static void *(*p_export)(void *, size_t *, int, size_t, int, size_t, const mpz_struct *);
static void (*p_powm)(mpz_struct *, const mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_powm_sec)(mpz_struct *, const mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_add)(mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_sub)(mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_mul)(mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_fdiv_qr)(mpz_struct *, mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_mod)(mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_mul_2exp)(mpz_struct *, const mpz_struct *, unsigned long);
static void (*p_fdiv_r_2exp)(mpz_struct *, const mpz_struct *, unsigned long);
static void (*p_tdiv_q_2exp)(mpz_struct *, const mpz_struct *, unsigned long);
static void (*p_ior)(mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_and)(mpz_struct *, const mpz_struct *, const mpz_struct *);
static void (*p_xor)(mpz_struct *, const mpz_struct *, const mpz_struct *);
static unsigned long (*p_popcount)(const mpz_struct *);
static int (*p_cmp)(const mpz_struct *, const mpz_struct *);
static const char *gmp_version_str;

static int loaded = 0;

GO_API int go_load(void) {
  if (loaded) return loaded > 0 ? 0 : -1;
  void *h = dlopen("libgmp.so.10", RTLD_NOW | RTLD_LOCAL);
  if (!h) { loaded = -1; return -1; }
#define SYM(var, name) do { *(void **)(&var) = dlsym(h, name); if (!var) { loaded = -1; return -1; } } while (0)
  SYM(p_init, "__gmpz_init");
  SYM(p_clear, "__gmpz_clear");
  SYM(p_import, "__gmpz_import");
  SYM(p_export, "__gmpz_export");
  SYM(p_powm, "__gmpz_powm");
  SYM(p_powm_sec, "__gmpz_powm_sec");
  SYM(p_add, "__gmpz_add");
  SYM(p_sub, "__gmpz_sub");
  SYM(p_mul, "__gmpz_mul");
  SYM(p_fdiv_qr, "__gmpz_fdiv_qr");
  SYM(p_mod, "__gmpz_mod");
  SYM(p_mul_2exp, "__gmpz_mul_2exp");
  SYM(p_fdiv_r_2exp, "__gmpz_fdiv_r_2exp");
  SYM(p_tdiv_q_2exp, "__gmpz_tdiv_q_2exp");
  SYM(p_ior, "__gmpz_ior");
  SYM(p_and, "__gmpz_and");
  SYM(p_xor, "__gmpz_xor");
  SYM(p_popcount, "__gmpz_popcount");
  SYM(p_cmp, "__gmpz_cmp");
#undef SYM
  const char **ver = (const char **)dlsym(h, "__gmp_version");
  gmp_version_str = ver ? *ver : "unknown";
  loaded = 1;
  return 0;
}

GO_API const char *go_version(void) { return loaded > 0 ? gmp_version_str : "not loaded"; }

static void imp(mpz_struct *z, const uint32_t *x, int n) { p_import(z, (size_t)n, -1, 4, 0, 0, x); }

/* export |z| mod 2^(32n) into n limbs (truncation = the reference's low-limb compare, test.cu:1223-1228) */
static void exp_trunc(uint32_t *x, int n, const mpz_struct *z) {
  memset(x, 0, sizeof(uint32_t) * (size_t)n);
  size_t cnt = 0;
  uint32_t *tmp = (uint32_t *)p_export(NULL, &cnt, -1, 4, 0, 0, z);
  for (size_t i = 0; i < cnt && i < (size_t)n; i++) x[i] = tmp[i];
  free(tmp);
}

enum { GO_ADD = 0, GO_SUB = 1, GO_MUL = 2, GO_DIV = 3, GO_MOD = 4, GO_POWM = 5, GO_POWM_SEC = 6,
       GO_IOR = 7, GO_AND = 8, GO_XOR = 9, GO_SHF = 10, GO_POPC = 11, GO_CMP = 12 };

typedef struct {
  int op;
  uint32_t lo, hi;           /* instance range */
  uint32_t *out; int on;     /* compact outputs */
  uint32_t *out2; int on2;   /* remainder for divmod */
  const uint32_t *a; int an; uint32_t ac;
  const uint32_t *b; int bn; uint32_t bc;
  const uint32_t *c; int cn; uint32_t cc;
} job_t;

static void *worker(void *arg) {
  job_t *j = (job_t *)arg;
  mpz_t za, zb, zc, zr, zs;
  p_init(za); p_init(zb); p_init(zc); p_init(zr); p_init(zs);
  for (uint32_t i = j->lo; i < j->hi; i++) {
    imp(za, j->a + (size_t)(i % j->ac) * j->an, j->an);
    if (j->b && j->op != GO_SHF) imp(zb, j->b + (size_t)(i % j->bc) * j->bn, j->bn);
    if (j->c) imp(zc, j->c + (size_t)(i % j->cc) * j->cn, j->cn);
    switch (j->op) {
      case GO_ADD: p_add(zr, za, zb); break;
      case GO_SUB:
        /* d = a - b mod 2^(32*on) (test.cu:2317-2322 adds 2^P when negative) */
        p_sub(zr, za, zb);
        p_fdiv_r_2exp(zr, zr, 32ul * (unsigned long)j->on);
        break;
      case GO_MUL: p_mul(zr, za, zb); break;
      case GO_DIV: case GO_MOD: p_fdiv_qr(zr, zs, za, zb); break;
      case GO_POWM: p_powm(zr, za, zb, zc); break;       /* a^b mod c */
      case GO_POWM_SEC: p_powm_sec(zr, za, zb, zc); break;
      /* bitwise ops as the reference tests compute them (unit_tests/test.cu:2516-2595
       * mpz_and/ior/xor, :1366-1372 mpz_mul_2exp / mpz_tdiv_q_2exp, :2091 mpz_cmp,
       * :2191 mpz_popcount); shift and the popc / cmp results are one int32 limb */
      case GO_IOR: p_ior(zr, za, zb); break;
      case GO_AND: p_and(zr, za, zb); break;
      case GO_XOR: p_xor(zr, za, zb); break;
      case GO_SHF: {
        int32_t sh = (int32_t)j->b[i % j->bc];
        if (sh >= 0) p_mul_2exp(zr, za, (unsigned long)sh);
        else p_tdiv_q_2exp(zr, za, (unsigned long)(-(int64_t)sh));
        break;
      }
      case GO_POPC: j->out[(size_t)i * j->on] = (uint32_t)p_popcount(za); continue;
      case GO_CMP: {
        int r = p_cmp(za, zb);
        j->out[(size_t)i * j->on] = (uint32_t)(int32_t)(r > 0 ? 1 : r < 0 ? -1 : 0);
        continue;
      }
    }
    if (j->out) exp_trunc(j->out + (size_t)i * j->on, j->on, j->op == GO_MOD ? zs : zr);
    if (j->out2) exp_trunc(j->out2 + (size_t)i * j->on2, j->on2, zs);
  }
  p_clear(za); p_clear(zb); p_clear(zc); p_clear(zr); p_clear(zs);
  return NULL;
}

/* Run one GMP op over count instances split statically over nthreads
 * pthreads (the perf harness's static OpenMP split, perf_tests/main.cu:200-210).
 * Returns wall seconds for the compute (inputs already in memory), or -1.
 * Divisors must be nonzero (GMP traps on division by zero) and powm moduli odd
 * for GO_POWM_SEC. */
GO_API double go_batch(int op, uint32_t count, int nthreads,
                       uint32_t *out, int on, uint32_t *out2, int on2,
                       const uint32_t *a, int an, uint32_t ac,
                       const uint32_t *b, int bn, uint32_t bc,
                       const uint32_t *c, int cn, uint32_t cc) {
  if (go_load() != 0) return -1.0;
  if (nthreads < 1) nthreads = 1;
  if ((uint32_t)nthreads > count && count > 0) nthreads = (int)count;
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  job_t *jobs = (job_t *)calloc((size_t)nthreads, sizeof(job_t));
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < nthreads; t++) {
    job_t *j = &jobs[t];
    j->op = op;
    j->lo = (uint32_t)((uint64_t)count * t / nthreads);
    j->hi = (uint32_t)((uint64_t)count * (t + 1) / nthreads);
    j->out = out; j->on = on; j->out2 = out2; j->on2 = on2;
    j->a = a; j->an = an; j->ac = ac ? ac : 1;
    j->b = b; j->bn = bn; j->bc = bc ? bc : 1;
    j->c = c; j->cn = cn; j->cc = cc ? cc : 1;
    pthread_create(&th[t], NULL, worker, j);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  free(jobs);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
