/* sba_div.c -- division on the GPU multiply (Newton reciprocal).
 *
 * Contract of the reference's divrem_u64 / divrem_u64_dc
 * (include/divrem.h:212-280, include/dc_div.h:150-208): Q = floor(N/D)
 * written as nn64-dn64+1 limbs, R = N mod D as dn64 limbs, trimmed lengths
 * (>= 1) returned; when bitlen(N) < bitlen(D) only qp[0] = 0 is written and
 * the remainder length is dn64 (divrem.h:222-227). The reference divides by
 * block Knuth-D / divide-and-conquer over its Toom multiply; here the
 * quotient comes from one Newton reciprocal of precision qn+1 limbs
 * (approximate reciprocal of Brent & Zimmermann, Modern Computer
 * Arithmetic, Algorithm 3.5), one quotient multiply and an exact remainder
 * check -- every large multiply goes through the threshold dispatch, i.e. the
 * GPU NTT above the crossover.
 */
#include "sba_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "simd_bigint_amd: out of host memory in division\n");
    abort();
  }
  return p;
}

static size_t trim(const u64 *a, size_t n) {
  while (n > 0 && a[n - 1] == 0) --n;
  return n;
}

static int cmp_n(const u64 *a, const u64 *b, size_t n) {
  for (size_t i = n; i-- > 0;)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

/* r = a - b over n limbs, returns borrow */
static u64 sub_n(u64 *r, const u64 *a, const u64 *b, size_t n) {
  u64 bw = 0;
  for (size_t i = 0; i < n; ++i) {
    u64 x = a[i], y = b[i];
    r[i] = x - y - bw;
    bw = (x < y) | ((x == y) & bw);
  }
  return bw;
}
/* a[0..n) -= 1, returns borrow */
static u64 dec_n(u64 *a, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (a[i]-- != 0) return 0;
  return 1;
}
static u64 inc_n(u64 *a, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (++a[i] != 0) return 0;
  return 1;
}
/* r[0..n) += a[0..m) (m <= n), returns carry */
static u64 add_into(u64 *r, size_t n, const u64 *a, size_t m) {
  u64 c = 0;
  size_t i = 0;
  for (; i < m; ++i) {
    u128 s = (u128)r[i] + a[i] + c;
    r[i] = (u64)s;
    c = (u64)(s >> 64);
  }
  for (; c && i < n; ++i) c = ++r[i] == 0;
  return c;
}

/* ------------------------------------------------------------------ */
/* schoolbook division (Knuth, TAOCP vol. 2, 4.3.1 algorithm D).       */
/* u: un limbs (un >= vn), v: vn limbs with v[vn-1] top bit set.       */
/* q[0..un-vn+1) = floor(u / v); u is overwritten by the remainder.   */
/* u must have one extra zero limb at u[un].                          */
/* ------------------------------------------------------------------ */
static void div_school_norm(u64 *q, u64 *u, size_t un, const u64 *v, size_t vn) {
  if (vn == 1) {
    u128 r = 0;
    for (size_t i = un + 1; i-- > 0;) {
      r = (r << 64) | u[i];
      u64 qi = (u64)(r / v[0]);
      r %= v[0];
      u[i] = 0;
      if (i <= un - 1) q[i] = qi;
    }
    u[0] = (u64)r;
    return;
  }
  const u64 vt = v[vn - 1], vs = v[vn - 2];
  for (size_t j = un - vn + 1; j-- > 0;) {
    u128 num = ((u128)u[j + vn] << 64) | u[j + vn - 1];
    u128 qh = num / vt, rh = num % vt;
    while ((qh >> 64) || (u128)(u64)qh * vs > ((rh << 64) | u[j + vn - 2])) {
      --qh;
      rh += vt;
      if (rh >> 64) break;
    }
    u64 carry = 0, borrow = 0;
    for (size_t i = 0; i < vn; ++i) {
      u128 p = (u128)(u64)qh * v[i] + carry;
      carry = (u64)(p >> 64);
      u64 pl = (u64)p, x = u[i + j];
      u64 d = x - pl;
      u64 b1 = x < pl;
      u64 d2 = d - borrow;
      u64 b2 = d < borrow;
      u[i + j] = d2;
      borrow = b1 + b2;
    }
    u64 top = u[j + vn];
    u64 d = top - carry;
    u64 b1 = top < carry;
    u64 d2 = d - borrow;
    u64 b2 = d < borrow;
    u[j + vn] = d2;
    if (b1 + b2) {
      --qh;
      add_into(u + j, vn + 1, v, vn);
      u[j + vn] = 0;
    }
    q[j] = (u64)qh;
  }
}

/* ------------------------------------------------------------------ */
/* approximate reciprocal (MCA Algorithm 3.5): A has n limbs, top bit  */
/* set. X[0..n] (n+1 limbs) with A*X < B^2n <= A*(X+2), B = 2^64.      */
/* ------------------------------------------------------------------ */
#define RECIP_BASE 16

static void approx_recip(u64 *X, const u64 *A, size_t n) {
  if (n <= RECIP_BASE) {
    /* X = floor((B^2n - 1) / A) */
    u64 *u = (u64 *)xmalloc((2 * n + 1) * 8);
    for (size_t i = 0; i < 2 * n; ++i) u[i] = ~(u64)0;
    u[2 * n] = 0;
    memset(X, 0, (n + 1) * 8);
    div_school_norm(X, u, 2 * n, A, n);
    free(u);
    return;
  }
  const size_t l = (n - 1) / 2, h = n - l;
  u64 *Xh = (u64 *)xmalloc((h + 1) * 8);
  approx_recip(Xh, A + l, h);
  /* T = A * Xh : n + h + 1 limbs */
  const size_t tn = n + h + 1;
  u64 *T = (u64 *)xmalloc(tn * 8);
  sba_mul(T, A, n, Xh, h + 1);
  while (T[n + h] != 0) {      /* T >= B^(n+h) */
    dec_n(Xh, h + 1);
    u64 bw = sub_n(T, T, A, n);
    for (size_t i = n; bw && i < tn; ++i) bw = T[i]-- == 0;
  }
  /* T = B^(n+h) - T over n+h limbs (T in (0, B^(n+h))) */
  {
    u64 bw = 0;
    for (size_t i = 0; i < n + h; ++i) {
      u64 x = T[i];
      T[i] = 0 - x - bw;
      bw = (x != 0) | bw;
    }
  }
  /* Tm = floor(T / B^l): 2h limbs; U = Tm * Xh: 3h+1 limbs */
  u64 *U = (u64 *)xmalloc((3 * h + 1) * 8);
  sba_mul(U, T + l, 2 * h, Xh, h + 1);
  /* X = Xh * B^l + floor(U / B^(2h-l)) */
  memset(X, 0, (n + 1) * 8);
  memcpy(X + l, Xh, (h + 1) * 8);
  const size_t sh = 2 * h - l;
  add_into(X, n + 1, U + sh, 3 * h + 1 - sh < n + 1 ? 3 * h + 1 - sh : n + 1);
  free(U);
  free(T);
  free(Xh);
}

void sba_approx_recip_host(u64 *X, const u64 *A, size_t n) { approx_recip(X, A, n); }

/* r = a << s (0 <= s < 64), n limbs in, writes n limbs + returns the spill */
static u64 lshift(u64 *r, const u64 *a, size_t n, unsigned s) {
  if (s == 0) {
    memmove(r, a, n * 8);
    return 0;
  }
  u64 out = a[n - 1] >> (64 - s);
  for (size_t i = n - 1; i > 0; --i) r[i] = (a[i] << s) | (a[i - 1] >> (64 - s));
  r[0] = a[0] << s;
  return out;
}
static void rshift(u64 *r, const u64 *a, size_t n, unsigned s) {
  if (s == 0) {
    memmove(r, a, n * 8);
    return;
  }
  for (size_t i = 0; i + 1 < n; ++i) r[i] = (a[i] >> s) | (a[i + 1] << (64 - s));
  r[n - 1] = a[n - 1] >> s;
}

static u64 bitlen(const u64 *a, size_t n) {
  n = trim(a, n);
  if (!n) return 0;
  return (u64)n * 64 - (u64)__builtin_clzll(a[n - 1]);
}

#define SCHOOL_MAX 32   /* schoolbook when min(qn, dn) is at most this */
#define DIV_DEV_MIN 8192 /* Newton steps on device buffers from this quotient size (limbs) */

/* the Newton branch runs on device buffers from DIV_DEV_MIN quotient limbs,
 * when its products are above the GPU threshold and in the GPU's band */
static int div_on_device(size_t qn, size_t dn) {
  const size_t m = qn + 1;
  return m >= DIV_DEV_MIN && 2 * m + 1 >= sba_get_gpu_threshold() &&
         sba_gpu_in_band((int64_t)m, (int64_t)m + 1) && sba_gpu_in_band((int64_t)qn, (int64_t)dn);
}

static void divrem_core(u64 *qp, u64 *rp, const u64 *np, u64 nn64, const u64 *dp,
                        u64 dn64, u64 *qn_out, u64 *rn_out) {
  const u64 Dbits = bitlen(dp, dn64), Nbits = bitlen(np, nn64);
  if (Dbits == 0) {
    fprintf(stderr, "simd_bigint_amd: division by zero\n");
    abort();
  }
  if (Nbits < Dbits) { /* divrem.h:222-227 */
    for (u64 i = 0; i < dn64; ++i) rp[i] = i < nn64 ? np[i] : 0;
    qp[0] = 0;
    *qn_out = 1;
    *rn_out = dn64;
    return;
  }
  const size_t dn = trim(dp, dn64), nt = trim(np, nn64);
  const unsigned s = (unsigned)__builtin_clzll(dp[dn - 1]);
  const size_t qn = nt + 1 - dn; /* Q < B^qn */
  const u64 qm = nn64 - dn64 + 1;
  if (qn > SCHOOL_MAX && dn > SCHOOL_MAX && div_on_device(qn, dn)) {
    /* the whole Newton branch on device buffers: operands up, Q and R down */
    const int rc = sba_div_dev(qp, (size_t)qm, rp, np, nt, dp, dn, s);
    if (rc == SBA_BAND) {
      fprintf(stderr, "simd_bigint_amd: quotient estimate out of bounds\n");
      abort();
    }
    if (rc == SBA_OK) {
      for (u64 i = qn; i < qm; ++i) qp[i] = 0;
      u64 qt = trim(qp, qm);
      *qn_out = qt ? qt : 1;
      for (u64 i = dn; i < dn64; ++i) rp[i] = 0;
      u64 rt = trim(rp, dn64);
      *rn_out = rt ? rt : 1;
      return;
    }
    /* a HIP failure (reported): the host Newton branch below finishes the
     * division, its multiplies through simd_mpn_mul */
    sba_note_fallback("division on device buffers", rc);
  }
  u64 *Dn = (u64 *)xmalloc(dn * 8);
  lshift(Dn, dp, dn, s);
  u64 *Nn = (u64 *)xmalloc((nt + 2) * 8); /* N << s: nt+1 limbs, +1 for the school path */
  Nn[nt] = lshift(Nn, np, nt, s);
  Nn[nt + 1] = 0;
  u64 *Q = (u64 *)xmalloc((qn + 1) * 8);
  memset(Q, 0, (qn + 1) * 8);
  u64 *R; /* remainder << s, dn limbs */

  if (qn <= SCHOOL_MAX || dn <= SCHOOL_MAX) {
    div_school_norm(Q, Nn, nt + 1, Dn, dn);
    R = Nn;
  } else {
    const size_t m = qn + 1;
    u64 *Dm = (u64 *)xmalloc(m * 8);
    u64 *X = (u64 *)xmalloc((m + 1) * 8);
    int exact_pow = 0;
    if (dn >= m) {
      memcpy(Dm, Dn + (dn - m), m * 8);
      if (inc_n(Dm, m)) exact_pow = 1; /* Dm + 1 == B^m */
    } else {
      memset(Dm, 0, (m - dn) * 8);
      memcpy(Dm + (m - dn), Dn, dn * 8);
    }
    if (exact_pow) {
      memset(X, 0, (m + 1) * 8);
      X[m] = 1; /* B^2m / B^m */
    } else {
      approx_recip(X, Dm, m);
    }
    /* Qest = floor(Nt * X / B^(qn+2)), Nt = N' >> 64(dn-1): qn+1 limbs */
    const size_t ntl = qn + 1;
    u64 *P = (u64 *)xmalloc((ntl + m + 1) * 8);
    sba_mul(P, Nn + (dn - 1), ntl, X, m + 1);
    memcpy(Q, P + (qn + 2), (qn + 1) * 8); /* P has 2qn+3 limbs */
    free(P);
    free(X);
    free(Dm);
    if (Q[qn] != 0) {
      fprintf(stderr, "simd_bigint_amd: quotient estimate overflow\n");
      abort();
    }
    /* R = N' - Qest * D' (>= 0), then fix Qest upward */
    u64 *QD = (u64 *)xmalloc((qn + dn) * 8);
    size_t qt = trim(Q, qn);
    if (qt) sba_mul(QD, Q, qt, Dn, dn);
    memset(QD + (qt ? qt + dn : 0), 0, (qn + dn - (qt ? qt + dn : 0)) * 8);
    R = (u64 *)xmalloc((qn + dn) * 8);
    if (sub_n(R, Nn, QD, qn + dn)) {
      fprintf(stderr, "simd_bigint_amd: quotient estimate too large\n");
      abort();
    }
    free(QD);
    int fix = 0;
    for (;;) {
      /* R >= D' ? (R fits dn+1 limbs here) */
      int ge = 0;
      if (trim(R + dn, qn) != 0) ge = 1;
      else ge = cmp_n(R, Dn, dn) >= 0;
      if (!ge) break;
      u64 bw = sub_n(R, R, Dn, dn);
      for (size_t i = dn; bw && i < qn + dn; ++i) bw = R[i]-- == 0;
      inc_n(Q, qn + 1);
      if (++fix > 64) {
        fprintf(stderr, "simd_bigint_amd: quotient correction did not converge\n");
        abort();
      }
    }
  }
  /* outputs: Q as nn64-dn64+1 limbs, R >> s as dn64 limbs */
  for (u64 i = 0; i < qm; ++i) qp[i] = i < qn ? Q[i] : 0;
  u64 qt = trim(qp, qm);
  *qn_out = qt ? qt : 1;
  u64 *Rs = (u64 *)xmalloc(dn * 8);
  rshift(Rs, R, dn, s);
  for (u64 i = 0; i < dn64; ++i) rp[i] = i < dn ? Rs[i] : 0;
  u64 rt = trim(rp, dn64);
  *rn_out = rt ? rt : 1;
  free(Rs);
  if (R != Nn) free(R);
  free(Q);
  free(Nn);
  free(Dn);
}

void divrem_u64(uint64_t *qp, uint64_t *rp, const uint64_t *np, uint64_t nn64,
                const uint64_t *dp, uint64_t dn64, uint64_t *qn64_out, uint64_t *rn64_out) {
  divrem_core(qp, rp, np, nn64, dp, dn64, qn64_out, rn64_out);
}

void divrem_u64_dc(uint64_t *qp, uint64_t *rp, const uint64_t *np, uint64_t nn64,
                   const uint64_t *dp, uint64_t dn64, uint64_t *qn64_out,
                   uint64_t *rn64_out) {
  divrem_core(qp, rp, np, nn64, dp, dn64, qn64_out, rn64_out);
}
