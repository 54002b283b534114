/* oracle.c -- CPU oracle for the large-operand multiply path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never the thing
 * measured or shipped. Reference citations are paths in the reference
 * repository.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint32_t u32;

/* ------------------------------------------------------------------ */
/* generator: bench/nt_sweep.c:15-16, tests/p30_test.c:14-15           */
/* ------------------------------------------------------------------ */
u64 orc_xorshift(u64 *s) {
  u64 x = *s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  *s = x;
  return x;
}

void orc_fill_xorshift(u64 *p, size_t n, u64 *s) {
  for (size_t i = 0; i < n; ++i) p[i] = orc_xorshift(s);
}

/* ------------------------------------------------------------------ */
/* schoolbook: tests/mpn_mul_u64_test.cpp:29-41                        */
/* ------------------------------------------------------------------ */
void orc_mul_school(u64 *rp, const u64 *ap, size_t an, const u64 *bp,
                    size_t bn) {
  memset(rp, 0, (an + bn) * sizeof(u64));
  for (size_t i = 0; i < an; ++i) {
    u64 c = 0;
    for (size_t j = 0; j < bn; ++j) {
      u128 t = (u128)ap[i] * bp[j] + rp[i + j] + c;
      rp[i + j] = (u64)t;
      c = (u64)(t >> 64);
    }
    rp[i + bn] = c;
  }
}

/* ------------------------------------------------------------------ */
/* mod-p scalar arithmetic (p < 2^30). Barrett by mu = floor(2^64/p):  */
/* for any x < 2^64, x - floor(x*mu/2^64)*p lies in [0, 2p).           */
/* ------------------------------------------------------------------ */
typedef struct {
  u64 p, mu;
} modp;

static inline u64 red64(const modp *m, u64 x) {
  u64 q = (u64)(((u128)x * m->mu) >> 64);
  u64 r = x - q * m->p;
  return r >= m->p ? r - m->p : r;
}
static inline u64 mulm(const modp *m, u64 a, u64 b) { return red64(m, a * b); }
static inline u64 addm(const modp *m, u64 a, u64 b) {
  u64 s = a + b;
  return s >= m->p ? s - m->p : s;
}
static inline u64 subm(const modp *m, u64 a, u64 b) {
  return a >= b ? a - b : a + m->p - b;
}
static u64 powm(const modp *m, u64 a, u64 e) {
  u64 r = 1;
  for (; e; e >>= 1, a = mulm(m, a, a))
    if (e & 1) r = mulm(m, r, a);
  return r;
}
static u64 invm(const modp *m, u64 a) { return powm(m, a, m->p - 2); }

/* smallest primitive root: include/p30.h:184-197 */
static u64 primitive_root(const modp *m) {
  u64 f[16];
  int nf = 0;
  u64 n = m->p - 1;
  for (u64 d = 2; d * d <= n; d += (d == 2 ? 1 : 2))
    if (n % d == 0) {
      f[nf++] = d;
      while (n % d == 0) n /= d;
    }
  if (n > 1) f[nf++] = n;
  for (u64 g = 2;; ++g) {
    int ok = 1;
    for (int i = 0; i < nf && ok; ++i)
      if (powm(m, g, (m->p - 1) / f[i]) == 1) ok = 0;
    if (ok) return g;
  }
}

/* ------------------------------------------------------------------ */
/* prime sets. p30's five (include/p30.h:70-71), ascending; the       */
/* six-prime set prepends 897581057 (= 107*2^23+1, also 2^22 | p-1)   */
/* so ascending order (load-bearing for Garner) is kept.              */
/* ------------------------------------------------------------------ */
#define ORC_W 16      /* lanes per vector, include/p30.h:69 */
#define ORC_MAX_LG 22 /* roots of order 2^22 exist for every prime */
static const u64 PR5[5] = {918552577u, 935329793u, 943718401u, 985661441u,
                           998244353u};
static const u64 PR6[6] = {897581057u, 918552577u, 935329793u,
                           943718401u, 985661441u, 998244353u};
/* floor(prod / (2^64-1)^2): include/p30.h:78 for five primes; the
 * six-prime value is computed the same way at first use. */
#define ORC_CAP5 2344425u

static u64 cap6(void) {
  /* floor(P / F), P = prod PR6 (< 2^179), F = (2^64-1)^2 = {1, 2^64-2}:
   * the largest q with q*F <= P, by bisection on 3-limb products. */
  u64 w[3] = {1, 0, 0};
  for (int i = 0; i < 6; ++i) {
    u128 c = 0;
    for (int k = 0; k < 3; ++k) {
      c += (u128)w[k] * PR6[i];
      w[k] = (u64)c;
      c >>= 64;
    }
  }
  const u64 F[2] = {1, ~(u64)1};
  u64 lo = 0, hi = (u64)1 << 62;
  while (lo < hi) {
    u64 mid = lo + (hi - lo + 1) / 2;
    u64 t[3];
    u128 c = (u128)mid * F[0];
    t[0] = (u64)c;
    c = (u128)mid * F[1] + (u64)(c >> 64);
    t[1] = (u64)c;
    t[2] = (u64)(c >> 64);
    int le;
    if (t[2] != w[2]) le = t[2] < w[2];
    else if (t[1] != w[1]) le = t[1] < w[1];
    else le = t[0] <= w[0];
    if (le) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

int orc_nt_nprimes(ptrdiff_t an, ptrdiff_t bn) {
  if (an <= 0 || bn <= 0) return 0;
  u64 mn = (u64)(an < bn ? an : bn);
  u64 zlen = (u64)an + (u64)bn - 1;
  u64 tv = (zlen + ORC_W - 1) / ORC_W;
  if (tv > ((u64)1 << ORC_MAX_LG)) return 0;
  if (mn <= ORC_CAP5) return 5;
  static u64 c6 = 0;
  if (!c6) c6 = cap6();
  return mn <= c6 ? 6 : 0;
}

/* ------------------------------------------------------------------ */
/* per-prime transform context: the prefix-stable bit-reversed root   */
/* table, include/p30.h:312-329 / tests/p30_tft_model_test.c:66-94:   */
/*   w[0] = 1; w[2^{t-1} + s] = zeta_t * w[s], zeta_t of order 2^t.   */
/* ------------------------------------------------------------------ */
typedef struct {
  modp m;
  u64 inv2;
  u32 *w, *winv;
} tctx;

static void ctx_build(tctx *c, u64 p, int lg) {
  c->m.p = p;
  c->m.mu = (u64)(((u128)1 << 64) / p);
  c->inv2 = (p + 1) >> 1;
  size_t M = (size_t)1 << lg;
  c->w = (u32 *)malloc(M * sizeof(u32));
  c->winv = (u32 *)malloc(M * sizeof(u32));
  u64 g = primitive_root(&c->m);
  c->w[0] = c->winv[0] = 1;
  for (int t = 1; t <= lg; ++t) {
    u64 z = powm(&c->m, g, (p - 1) >> t);
    u64 zi = invm(&c->m, z);
    size_t blk = (size_t)1 << (t - 1);
    for (size_t s = 0; s < blk; ++s) {
      c->w[blk + s] = (u32)mulm(&c->m, z, c->w[s]);
      c->winv[blk + s] = (u32)mulm(&c->m, zi, c->winv[s]);
    }
  }
}
static void ctx_free(tctx *c) {
  free(c->w);
  free(c->winv);
}

/* ------------------------------------------------------------------ */
/* truncated forward NTT on W-lane vectors, node (m, j), output trunc */
/* zo: tests/p30_tft_model_test.c:96-120 (FLINT convention, node      */
/* twiddle w[2j]; butterfly c = a + w b, d = a - w b).                */
/* ------------------------------------------------------------------ */
static void tft_fwd(const tctx *c, u32 *B, size_t m, size_t j, size_t zo) {
  const modp *md = &c->m;
  while (m > 1) {
    size_t h = m >> 1;
    u64 w = c->w[2 * j];
    if (zo <= h) {
      /* only the left child's outputs are wanted: fold the right half */
      for (size_t i = 0; i < h * ORC_W; ++i)
        B[i] = (u32)addm(md, B[i], mulm(md, w, B[i + h * ORC_W]));
      m = h;
      j = 2 * j;
      continue;
    }
    for (size_t i = 0; i < h * ORC_W; ++i) {
      u64 a = B[i], wb = mulm(md, w, B[i + h * ORC_W]);
      B[i] = (u32)addm(md, a, wb);
      B[i + h * ORC_W] = (u32)subm(md, a, wb);
    }
    tft_fwd(c, B, h, 2 * j, h);
    B += h * ORC_W;
    m = h;
    j = 2 * j + 1;
    zo -= h;
  }
}

/* ------------------------------------------------------------------ */
/* truncated inverse: tests/p30_tft_model_test.c:122-181 (general     */
/* form also include/p30_ref.h:556-611). Entry B[0..z) spectral,      */
/* B[z..m) = m*u tails (zero at the top call); exit B[0..z) = m*u.    */
/* ------------------------------------------------------------------ */
static void tft_inv(const tctx *c, u32 *B, size_t m, size_t j, size_t z) {
  const modp *md = &c->m;
  if (m == 1 || z == 0) return;
  size_t h = m >> 1;
  u64 w = c->w[2 * j];
  size_t HW = h * ORC_W;
  if (z < h) {
    for (size_t i = z * ORC_W; i < HW; ++i)
      B[i] = (u32)mulm(md, addm(md, B[i], mulm(md, w, B[i + HW])), c->inv2);
    tft_inv(c, B, h, 2 * j, z);
    for (size_t i = 0; i < z * ORC_W; ++i)
      B[i] = (u32)subm(md, addm(md, B[i], B[i]), mulm(md, w, B[i + HW]));
    return;
  }
  size_t zr = z - h;
  tft_inv(c, B, h, 2 * j, h);
  for (size_t i = zr * ORC_W; i < HW; ++i) {
    u64 t = subm(md, B[i], mulm(md, w, B[i + HW]));
    B[i] = (u32)addm(md, B[i], t);
    B[i + HW] = (u32)t;
  }
  tft_inv(c, B + HW, h, 2 * j + 1, zr);
  u64 wi = c->winv[2 * j];
  for (size_t i = 0; i < zr * ORC_W; ++i) {
    u64 s = addm(md, B[i], B[i + HW]);
    u64 d = subm(md, B[i], B[i + HW]);
    B[i] = (u32)s;
    B[i + HW] = (u32)mulm(md, d, wi);
  }
}

/* ------------------------------------------------------------------ */
/* pointwise: twisted cyclic convolution mod (x^W - ww) at spectral   */
/* position k, ww = w[k]: tests/p30_tft_model_test.c:183-198,         */
/* include/p30.h:541-598. Products of canonical residues are < 2^60,  */
/* so W = 16 of them sum below 2^64.                                   */
/* ------------------------------------------------------------------ */
static void conv_tw(const modp *md, u32 *fa, const u32 *fb, u64 ww) {
  u64 out[ORC_W];
  for (int l = 0; l < ORC_W; ++l) {
    u64 s1 = 0, s2 = 0;
    for (int i = 0; i <= l; ++i) s1 += (u64)fa[i] * fb[l - i];
    for (int i = l + 1; i < ORC_W; ++i) s2 += (u64)fa[i] * fb[l - i + ORC_W];
    out[l] = addm(md, red64(md, s1), mulm(md, ww, red64(md, s2)));
  }
  for (int l = 0; l < ORC_W; ++l) fa[l] = (u32)out[l];
}

/* ------------------------------------------------------------------ */
/* one prime's cyclic product of the zlen-coefficient result:         */
/* include/p30.h:1907-1932 (fwd a, fused fwd b + conv + inverse,      */
/* scale by M^-1).                                                     */
/* ------------------------------------------------------------------ */
static void prime_product(u32 *res, u64 p, int lg, const u64 *a, size_t an,
                          const u64 *b, size_t bn, int sq, size_t zlen) {
  size_t M = (size_t)1 << lg;
  size_t tv = (zlen + ORC_W - 1) / ORC_W;
  tctx c;
  ctx_build(&c, p, lg);
  u32 *va = (u32 *)calloc(M * ORC_W, sizeof(u32));
  u32 *vb = sq ? va : (u32 *)calloc(M * ORC_W, sizeof(u32));
  /* input residues: include/p30.h:1530-1574 computes a % p per limb */
  for (size_t i = 0; i < an; ++i) va[i] = (u32)(a[i] % p);
  if (!sq)
    for (size_t i = 0; i < bn; ++i) vb[i] = (u32)(b[i] % p);
  tft_fwd(&c, va, M, 0, tv);
  if (!sq) tft_fwd(&c, vb, M, 0, tv);
  if (sq) {
    u32 tmp[ORC_W];
    for (size_t k = 0; k < tv; ++k) {
      memcpy(tmp, va + k * ORC_W, sizeof tmp);
      conv_tw(&c.m, va + k * ORC_W, tmp, c.w[k]);
    }
  } else {
    for (size_t k = 0; k < tv; ++k)
      conv_tw(&c.m, va + k * ORC_W, vb + k * ORC_W, c.w[k]);
  }
  memset(va + tv * ORC_W, 0, (M - tv) * ORC_W * sizeof(u32));
  tft_inv(&c, va, M, 0, tv);
  u64 mi = invm(&c.m, M % p);
  for (size_t i = 0; i < zlen; ++i) res[i] = (u32)mulm(&c.m, va[i], mi);
  free(va);
  if (!sq) free(vb);
  ctx_free(&c);
}

/* ------------------------------------------------------------------ */
/* Garner (ascending primes) + Horner compose + accumulate at limb    */
/* offset i: tests/p30_tft_model_test.c:276-319, include/p30.h:1694-  */
/* 1839. Returns 0 if the result spills past an+bn limbs (p30.h:1935).*/
/* ------------------------------------------------------------------ */
static int garner_out(u64 *out, size_t outl, u32 *const *res, const u64 *pr,
                      int K, size_t zlen) {
  u64 ppm[6][6], invp[6];
  modp md[6];
  for (int i = 0; i < K; ++i) {
    md[i].p = pr[i];
    md[i].mu = (u64)(((u128)1 << 64) / pr[i]);
  }
  for (int i = 1; i < K; ++i) {
    u64 pp = 1;
    for (int k = 0; k < i; ++k) {
      ppm[i][k] = pp;
      pp = mulm(&md[i], pp, pr[k] % pr[i]);
    }
    invp[i] = invm(&md[i], pp);
  }
  memset(out, 0, outl * sizeof(u64));
  u64 spill = 0;
  for (size_t i = 0; i < zlen; ++i) {
    u64 v[6];
    v[0] = res[0][i];
    for (int q = 1; q < K; ++q) {
      const modp *m = &md[q];
      u64 t = v[0] % m->p;
      for (int k = 1; k < q; ++k) t = addm(m, t, mulm(m, v[k], ppm[q][k]));
      v[q] = mulm(m, subm(m, res[q][i] % m->p, t), invp[q]);
    }
    /* Horner: c = (..((v[K-1] p[K-2] + v[K-2]) p[K-3] + ..) p0 + v0,
     * kept as 3 u64 words (c < 2^180) */
    u64 c[3] = {v[K - 1], 0, 0};
    for (int q = K - 2; q >= 0; --q) {
      u128 acc = (u128)c[0] * pr[q] + v[q];
      c[0] = (u64)acc;
      acc = (u128)c[1] * pr[q] + (u64)(acc >> 64);
      c[1] = (u64)acc;
      acc = (u128)c[2] * pr[q] + (u64)(acc >> 64);
      c[2] = (u64)acc;
    }
    u128 s = 0;
    size_t j = i;
    for (int k = 0; k < 3; ++k, ++j) {
      s += c[k];
      if (j < outl) {
        s += out[j];
        out[j] = (u64)s;
        s >>= 64;
      } else {
        spill |= (u64)s;
        s >>= 64;
      }
    }
    for (; s; ++j) {
      if (j >= outl) {
        spill |= (u64)s;
        break;
      }
      s += out[j];
      out[j] = (u64)s;
      s >>= 64;
    }
  }
  return spill == 0;
}

static int nt_core(u64 *rp, const u64 *ap, ptrdiff_t an, const u64 *bp,
                   ptrdiff_t bn, int sq) {
  int K = orc_nt_nprimes(an, bn);
  if (!K) return 0;
  const u64 *pr = K == 5 ? PR5 : PR6;
  size_t zlen = (size_t)an + (size_t)bn - 1;
  size_t tv = (zlen + ORC_W - 1) / ORC_W;
  int lg = 0;
  while (((size_t)1 << lg) < tv) lg++;
  u32 *res[6];
  for (int q = 0; q < K; ++q) res[q] = (u32 *)malloc(zlen * sizeof(u32));
#pragma omp parallel for schedule(dynamic, 1)
  for (int q = 0; q < K; ++q)
    prime_product(res[q], pr[q], lg, ap, (size_t)an, bp, (size_t)bn, sq, zlen);
  int ok = garner_out(rp, (size_t)an + (size_t)bn, res, pr, K, zlen);
  for (int q = 0; q < K; ++q) free(res[q]);
  return ok;
}

int orc_nt_mul(u64 *rp, const u64 *ap, ptrdiff_t an, const u64 *bp,
               ptrdiff_t bn) {
  return nt_core(rp, ap, an, bp, bn, 0);
}
int orc_nt_sqr(u64 *rp, const u64 *ap, ptrdiff_t an) {
  return nt_core(rp, ap, an, ap, an, 1);
}

/* ------------------------------------------------------------------ */
/* brute-force residues and fingerprints                               */
/* ------------------------------------------------------------------ */
void orc_residues(u32 *out, const u64 *ap, size_t an, u32 p) {
  for (size_t i = 0; i < an; ++i) out[i] = (u32)(ap[i] % p);
}

uint64_t orc_fingerprint(const u64 *ap, size_t an, u64 q) {
  /* Horner from the top limb: r = (r * 2^64 + a_i) mod q */
  u128 base = ((u128)1 << 64) % q;
  u128 r = 0;
  for (size_t i = an; i-- > 0;) r = (r * base + ap[i] % q) % q;
  return (u64)r;
}

/* ------------------------------------------------------------------ */
/* Knuth algorithm D (TAOCP 4.3.1) on u64 limbs; contract of          */
/* include/divrem.h:212-221 (Q >= nn-dn+1 limbs, trimmed lengths >= 1)*/
/* ------------------------------------------------------------------ */
void orc_divrem(u64 *qp, u64 *rp, const u64 *np, size_t nn, const u64 *dp,
                size_t dn, size_t *qn_out, size_t *rn_out) {
  while (nn > 0 && np[nn - 1] == 0) --nn;
  /* compare N < D */
  int lt = 0;
  if (nn < dn) lt = 1;
  else if (nn == dn) {
    lt = 1;
    for (size_t i = dn; i-- > 0;) {
      if (np[i] != dp[i]) {
        lt = np[i] < dp[i];
        break;
      }
      if (i == 0) lt = 0;
    }
  }
  if (lt) {
    for (size_t i = 0; i < dn; ++i) rp[i] = i < nn ? np[i] : 0;
    qp[0] = 0;
    *qn_out = 1;
    size_t rn = dn;
    while (rn > 1 && rp[rn - 1] == 0) --rn;
    *rn_out = rn;
    return;
  }
  size_t qn = nn - dn + 1;
  if (dn == 1) {
    u128 r = 0;
    for (size_t i = nn; i-- > 0;) {
      r = (r << 64) | np[i];
      qp[i] = (u64)(r / dp[0]);
      r %= dp[0];
    }
    rp[0] = (u64)r;
  } else {
    int s = __builtin_clzll(dp[dn - 1]);
    u64 *v = (u64 *)malloc(dn * sizeof(u64));
    u64 *u = (u64 *)malloc((nn + 1) * sizeof(u64));
    for (size_t i = dn; i-- > 0;)
      v[i] = (dp[i] << s) | (s && i ? dp[i - 1] >> (64 - s) : 0);
    u[nn] = s ? np[nn - 1] >> (64 - s) : 0;
    for (size_t i = nn; i-- > 0;)
      u[i] = (np[i] << s) | (s && i ? np[i - 1] >> (64 - s) : 0);
    u64 vt = v[dn - 1], vs = v[dn - 2];
    for (size_t j = qn; j-- > 0;) {
      u128 num = ((u128)u[j + dn] << 64) | u[j + dn - 1];
      u128 qh = num / vt, rh = num % vt;
      while (qh >> 64 ||
             (u128)(u64)qh * vs > ((rh << 64) | u[j + dn - 2])) {
        qh--;
        rh += vt;
        if (rh >> 64) break;
      }
      /* multiply and subtract */
      u64 borrow = 0, carry = 0;
      for (size_t i = 0; i < dn; ++i) {
        u128 pr = (u128)(u64)qh * v[i] + carry;
        carry = (u64)(pr >> 64);
        u64 pl = (u64)pr;
        u64 t = u[i + j] - pl;
        u64 b1 = u[i + j] < pl;
        u64 t2 = t - borrow;
        u64 b2 = t < borrow;
        u[i + j] = t2;
        borrow = b1 + b2;
      }
      u64 top = u[j + dn];
      u64 t = top - carry;
      u64 b1 = top < carry;
      u64 t2 = t - borrow;
      u64 b2 = t < borrow;
      u[j + dn] = t2;
      if (b1 + b2) {
        /* add back */
        qh--;
        u64 c = 0;
        for (size_t i = 0; i < dn; ++i) {
          u128 sm = (u128)u[i + j] + v[i] + c;
          u[i + j] = (u64)sm;
          c = (u64)(sm >> 64);
        }
        u[j + dn] += c;
      }
      qp[j] = (u64)qh;
    }
    for (size_t i = 0; i < dn; ++i)
      rp[i] = (u[i] >> s) | (s ? u[i + 1] << (64 - s) : 0);
    free(v);
    free(u);
  }
  size_t q = qn;
  while (q > 1 && qp[q - 1] == 0) --q;
  *qn_out = q;
  size_t rn = dn;
  while (rn > 1 && rp[rn - 1] == 0) --rn;
  *rn_out = rn;
}
