/* drop_in_caller.c -- a reference-style C caller of the drop-in library:
 * it includes only include/simd_bigint_amd.h (in place of the reference's
 * ntmul.h / p30.h / p50.h / u64cvt.h / dc_div.h), links
 * libsimd_bigint_amd.so, and calls every reference-named entry in band:
 * nt_mul(_r), nt_sqr(_r), p30_*, p50_*, simd_mpn_mul, divrem_u64(_dc).
 * Each product is checked against a schoolbook multiply in this file; the
 * division by Q * D + R == N and R < D. Inputs: xorshift64 (13, 7, 17)
 * seeded 42, as the reference's tests fill them (tests/nt_test.c:16-30).
 * Prints one line per entry and exits 0 only if every check passed. */
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

static int g_fail = 0;
static void check(const char *what, int ok) {
  printf("%-28s %s\n", what, ok ? "ok" : "FAILED");
  if (!ok) g_fail = 1;
}

int main(void) {
  const size_t an = 3000, bn = 2100, sn = 2500;
  uint64_t *a = malloc(an * 8), *b = malloc(bn * 8), *r = malloc((an + bn) * 8),
           *w = malloc((an + bn) * 8), *s = malloc(2 * sn * 8), *sw = malloc(2 * sn * 8);
  for (size_t i = 0; i < an; ++i) a[i] = xr();
  for (size_t i = 0; i < bn; ++i) b[i] = xr();
  school(w, a, an, b, bn);
  school(sw, a, sn, a, sn);
  int (*muls[])(uint64_t *, const uint64_t *, ptrdiff_t, const uint64_t *, ptrdiff_t) = {
      nt_mul, p30_mul, p50_mul};
  const char *mul_names[] = {"nt_mul", "p30_mul", "p50_mul"};
  for (int k = 0; k < 3; ++k) {
    memset(r, 0, (an + bn) * 8);
    check(mul_names[k], muls[k](r, a, an, b, bn) == 1 && !memcmp(r, w, (an + bn) * 8));
  }
  int (*muls_r[])(uint64_t *, const uint64_t *, ptrdiff_t, const uint64_t *, ptrdiff_t,
                  struct scratch *) = {nt_mul_r, p30_mul_r, p50_mul_r};
  const char *mulr_names[] = {"nt_mul_r", "p30_mul_r", "p50_mul_r"};
  for (int k = 0; k < 3; ++k) {
    memset(r, 0, (an + bn) * 8);
    check(mulr_names[k], muls_r[k](r, b, bn, a, an, NULL) == 1 && !memcmp(r, w, (an + bn) * 8));
  }
  int (*sqrs[])(uint64_t *, const uint64_t *, ptrdiff_t) = {nt_sqr, p30_sqr, p50_sqr};
  const char *sqr_names[] = {"nt_sqr", "p30_sqr", "p50_sqr"};
  for (int k = 0; k < 3; ++k) {
    memset(s, 0, 2 * sn * 8);
    check(sqr_names[k], sqrs[k](s, a, sn) == 1 && !memcmp(s, sw, 2 * sn * 8));
  }
  int (*sqrs_r[])(uint64_t *, const uint64_t *, ptrdiff_t, struct scratch *) = {
      nt_sqr_r, p30_sqr_r, p50_sqr_r};
  const char *sqrr_names[] = {"nt_sqr_r", "p30_sqr_r", "p50_sqr_r"};
  for (int k = 0; k < 3; ++k) {
    memset(s, 0, 2 * sn * 8);
    check(sqrr_names[k], sqrs_r[k](s, a, sn, NULL) == 1 && !memcmp(s, sw, 2 * sn * 8));
  }
  /* out of band: 0 (include/ntmul.h:19, include/p30.h:1851-1858) */
  check("nt_mul an = 0 -> 0", nt_mul(r, a, 0, b, bn) == 0);
  /* simd_mpn_mul: either order, exactly an + bn limbs written */
  uint64_t *g = malloc((an + bn + 1) * 8);
  g[an + bn] = 0xABABABABABABABABull;
  simd_mpn_mul(g, b, bn, a, an);
  check("simd_mpn_mul", !memcmp(g, w, (an + bn) * 8) && g[an + bn] == 0xABABABABABABABABull);
  /* division: N = (A * B) + B - 1, D = A; Q = B, R = B - 1 where B < A */
  {
    uint64_t *N = malloc((an + bn) * 8), *q = malloc((bn + 1) * 8), *rr = malloc(an * 8);
    memcpy(N, w, (an + bn) * 8);
    /* N += B - 1 (B < A since an > bn): remainder B - 1 */
    uint64_t c = 0;
    for (size_t i = 0; i < an + bn; ++i) {
      u128 t = (u128)N[i] + (i < bn ? b[i] : 0) + c - (i == 0 ? 1 : 0);
      N[i] = (uint64_t)t;
      c = (uint64_t)(t >> 64);
    }
    void (*divs[])(uint64_t *, uint64_t *, const uint64_t *, uint64_t, const uint64_t *,
                   uint64_t, uint64_t *, uint64_t *) = {divrem_u64, divrem_u64_dc};
    const char *div_names[] = {"divrem_u64", "divrem_u64_dc"};
    for (int k = 0; k < 2; ++k) {
      uint64_t qn = 0, rn = 0;
      memset(q, 0, (bn + 1) * 8);
      memset(rr, 0, an * 8);
      divs[k](q, rr, N, an + bn, a, an, &qn, &rn);
      int ok = qn == bn && !memcmp(q, b, bn * 8) && q[bn] == 0 && rn == bn &&
               rr[0] == b[0] - 1 && !memcmp(rr + 1, b + 1, (bn - 1) * 8);
      for (size_t i = bn; i < an; ++i) ok &= rr[i] == 0;
      check(div_names[k], ok);
    }
    free(N);
    free(q);
    free(rr);
  }
  printf("fallbacks to the CPU tier: %llu\n", (unsigned long long)sba_gpu_fallbacks());
  if (sba_gpu_fallbacks() != 0) g_fail = 1;
  free(a); free(b); free(r); free(w); free(s); free(sw); free(g);
  printf(g_fail ? "drop-in caller: FAILED\n" : "drop-in caller: all entries ok\n");
  return g_fail;
}
