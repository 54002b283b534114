/* host_asan.c -- the host C layer (sba_host.c, sba_div.c) under
 * AddressSanitizer, on the CPU only: the CPU multiply tiers against a
 * schoolbook of this file at shapes across the Karatsuba and strip-mining
 * edges, and the error contract's CPU paths with GPU failures injected
 * before any device is touched (simd_mpn_mul and the division finish on
 * the CPU tier; nt_* / p30_* / p50_* return 0). Exits 0 when every check
 * passed; ASan aborts on any invalid access or leak it sees. */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "simd_bigint_amd.h"

typedef unsigned __int128 u128;
static uint64_t g_x = 42;
static uint64_t xr(void) {
  g_x ^= g_x << 13;
  g_x ^= g_x >> 7;
  g_x ^= g_x << 17;
  return g_x;
}
static void school(uint64_t *r, const uint64_t *a, size_t an, const uint64_t *b, size_t bn) {
  memset(r, 0, (an + bn) * 8);
  for (size_t j = 0; j < bn; ++j) {
    uint64_t c = 0;
    for (size_t i = 0; i < an; ++i) {
      u128 t = (u128)a[i] * b[j] + r[i + j] + c;
      r[i + j] = (uint64_t)t;
      c = (uint64_t)(t >> 64);
    }
    r[j + an] = c;
  }
}
static int g_fail;
static void check(const char *what, size_t an, size_t bn, int ok) {
  if (!ok) {
    printf("FAILED %s %zu x %zu\n", what, an, bn);
    g_fail = 1;
  }
}

/* the library's parallel host copy (exported for tests; the staged
 * transfers of the host entries use it) */
void sba_par_copy(void *dst, const void *src, size_t bytes);
static void check_copy(size_t bytes) {
  unsigned char *s = malloc(bytes + 1), *d = malloc(bytes + 1);
  for (size_t i = 0; i < bytes; ++i) s[i] = (unsigned char)xr();
  d[bytes] = 0xa5;
  sba_par_copy(d, s, bytes);
  check("par_copy", bytes, 0, !memcmp(d, s, bytes) && d[bytes] == 0xa5);
  free(s);
  free(d);
}
static void *copy_thread(void *arg) {
  (void)arg;
  for (int k = 0; k < 4; ++k) {
    const size_t bytes = ((size_t)(2 + k) << 20) + 9;
    unsigned char *s = malloc(bytes), *d = malloc(bytes);
    for (size_t i = 0; i < bytes; ++i) s[i] = (unsigned char)(i * 131 + k);
    sba_par_copy(d, s, bytes);
    if (memcmp(d, s, bytes)) check("par_copy from threads", bytes, 0, 0);
    free(s);
    free(d);
  }
  return NULL;
}

int main(void) {
  static const size_t shapes[][2] = {{1, 1},    {31, 31},   {32, 32},   {33, 7},   {64, 63},
                                     {100, 99}, {257, 31},  {1000, 32}, {777, 333}, {2048, 2047},
                                     {3000, 1500}, {4097, 33}, {5000, 4999}};
  const size_t ns = sizeof shapes / sizeof shapes[0];
  sba_set_gpu_threshold((uint64_t)1 << 62);          /* CPU tiers only */
  for (size_t k = 0; k < ns; ++k) {
    const size_t an = shapes[k][0], bn = shapes[k][1];
    uint64_t *a = malloc(an * 8), *b = malloc(bn * 8), *r = malloc((an + bn) * 8),
             *w = malloc((an + bn) * 8);
    for (int pat = 0; pat < 2; ++pat) {
      for (size_t i = 0; i < an; ++i) a[i] = pat ? ~0ull : xr();
      for (size_t i = 0; i < bn; ++i) b[i] = pat ? ~0ull : xr();
      school(w, a, an, b, bn);
      sba_cpu_mul(r, a, an, b, bn);
      check("sba_cpu_mul", an, bn, !memcmp(r, w, (an + bn) * 8));
      sba_cpu_mul(r, b, bn, a, an);
      check("sba_cpu_mul swapped", bn, an, !memcmp(r, w, (an + bn) * 8));
      simd_mpn_mul(r, a, an, b, bn);
      check("simd_mpn_mul (CPU)", an, bn, !memcmp(r, w, (an + bn) * 8));
    }
    free(a); free(b); free(r); free(w);
  }
  /* GPU failures injected: the CPU tier finishes the void entries */
  sba_set_gpu_threshold(1024);
  {
    const size_t an = 3000, bn = 2500;
    uint64_t *a = malloc(an * 8), *b = malloc(bn * 8), *r = malloc((an + bn) * 8),
             *w = malloc((an + bn) * 8);
    for (size_t i = 0; i < an; ++i) a[i] = xr();
    for (size_t i = 0; i < bn; ++i) b[i] = xr();
    school(w, a, an, b, bn);
    const uint64_t f0 = sba_gpu_fallbacks();
    sba_test_inject_failure(SBA_FAIL_ENTRY, 1);
    simd_mpn_mul(r, a, an, b, bn);
    check("simd_mpn_mul fallback", an, bn, !memcmp(r, w, (an + bn) * 8) &&
                                           sba_gpu_fallbacks() == f0 + 1);
    sba_test_inject_failure(SBA_FAIL_ENTRY, 1);
    check("nt_mul on failure", an, bn, nt_mul(r, a, an, b, bn) == 0);
    sba_test_inject_failure(SBA_FAIL_ENTRY, 1);
    check("p50_sqr_r on failure", an, an, p50_sqr_r(r, a, (ptrdiff_t)bn, NULL) == 0);
    check("p30_mul out of band", an, bn, p30_mul(r, a, 0, b, bn) == 0);
    /* division N = A B + (B - 1) by A through the host Newton branch */
    uint64_t *N = malloc((an + bn) * 8), *q = malloc((bn + 1) * 8), *rr = malloc(an * 8);
    memcpy(N, w, (an + bn) * 8);
    uint64_t c = 0;
    for (size_t i = 0; i < an + bn; ++i) {
      u128 t = (u128)N[i] + (i < bn ? b[i] : 0) + c - (i == 0 ? 1 : 0);
      N[i] = (uint64_t)t;
      c = (uint64_t)(t >> 64);
    }
    sba_test_inject_failure(SBA_FAIL_ENTRY, 1000);
    uint64_t qn = 0, rn = 0;
    divrem_u64(q, rr, N, an + bn, a, an, &qn, &rn);
    sba_test_inject_failure(SBA_FAIL_ENTRY, 0);
    int ok = qn == bn && !memcmp(q, b, bn * 8) && rr[0] == b[0] - 1 &&
             !memcmp(rr + 1, b + 1, (bn - 1) * 8);
    for (size_t i = bn; i < an; ++i) ok &= rr[i] == 0;
    check("divrem_u64 fallback", an + bn, an, ok);
    free(N); free(q); free(rr);
    free(a); free(b); free(r); free(w);
  }
  /* the staged copies' thread pool: one caller, then three at once */
  {
    static const size_t sizes[] = {0, 1, 4096, (size_t)1 << 19, ((size_t)1 << 20) + 3,
                                   ((size_t)8 << 20) + 5, ((size_t)24 << 20) + 7};
    for (size_t k = 0; k < sizeof sizes / sizeof sizes[0]; ++k) check_copy(sizes[k]);
    pthread_t th[3];
    for (int t = 0; t < 3; ++t) pthread_create(&th[t], NULL, copy_thread, NULL);
    for (int t = 0; t < 3; ++t) pthread_join(th[t], NULL);
  }
  printf(g_fail ? "host asan test: FAILED\n" : "host asan test: all checks passed\n");
  return g_fail;
}
