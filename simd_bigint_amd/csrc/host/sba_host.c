/* sba_host.c -- host side (C) of the MI355X multiply: the reference's
 * mpn-style entries on host memory, the threshold dispatch, the CPU tiers
 * below the GPU crossover, and Newton division on the GPU multiply.
 *
 * Citations are files of the reference repository (iamstupid/simd_bigint).
 */
#include "sba_internal.h"

#include <hip/hip_runtime_api.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* device staging: grow-only, one set per process (guarded), the       */
/* device-side analogue of the reference's scratch arena rule "no      */
/* per-call allocation" (include/scratch.h:1-20).                      */
/* ------------------------------------------------------------------ */
static pthread_mutex_t g_stage_mu = PTHREAD_MUTEX_INITIALIZER;
static uint64_t *g_da, *g_db, *g_dr;
static size_t g_na, g_nb, g_nr;
static pthread_once_t g_dev_once = PTHREAD_ONCE_INIT;
static int g_dev_ok = 0;

/* Error contract. A HIP failure inside a host entry (allocation, copy,
 * launch) is reported on stderr and passed up: nt_* / p30_* / p50_* return
 * 0, simd_mpn_mul and the division finish that product on the CPU tier
 * (counted by sba_gpu_fallbacks). Only a process without any HIP device
 * aborts: the library never computes the GPU path on the CPU silently. */
static void die(const char *what, int rc) {
  fprintf(stderr, "simd_bigint_amd: %s failed (code %d); the library has no CPU "
                  "fallback for the GPU path\n", what, rc);
  abort();
}
static void report(const char *what, int rc) {
  fprintf(stderr, "simd_bigint_amd: %s failed (code %d)\n", what, rc);
}

static _Atomic uint64_t g_fallbacks;
uint64_t sba_gpu_fallbacks(void) { return atomic_load(&g_fallbacks); }
void sba_note_fallback(const char *what, int rc) {
  report(what, rc);
  fprintf(stderr, "simd_bigint_amd: that product is finished on the CPU tier\n");
  atomic_fetch_add(&g_fallbacks, 1);
}

/* failure injection for the tests of this contract */
static _Atomic int g_fail_entry, g_fail_alloc;
void sba_test_inject_failure(int where, int count) {
  if (where == SBA_FAIL_ENTRY) atomic_store(&g_fail_entry, count);
  else if (where == SBA_FAIL_ALLOC) atomic_store(&g_fail_alloc, count);
}
/* 1 when an armed injection point fires (and disarms one count) */
static int inject(_Atomic int *ctr) {
  int c = atomic_load(ctr);
  while (c > 0)
    if (atomic_compare_exchange_weak(ctr, &c, c - 1)) return 1;
  return 0;
}

static void device_probe(void) {
  int n = 0;
  g_dev_ok = hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

/* once per process, whichever thread gets here first */
static int device_ok(void) {
  pthread_once(&g_dev_once, device_probe);
  return g_dev_ok;
}

/* grow-only device buffer of >= n limbs; SBA_ERR_HIP leaves it empty (an
 * armed SBA_FAIL_ALLOC injection acts as a failed reallocation even when
 * the buffer is large enough) */
static int grow(uint64_t **p, size_t *cap, size_t n) {
  if (*cap >= n && atomic_load(&g_fail_alloc) <= 0) return 0;
  if (*p) (void)hipFree(*p);
  *p = NULL;
  *cap = 0;
  size_t want = n + n / 8 + 64;
  if (inject(&g_fail_alloc) || hipMalloc((void **)p, want * 8) != hipSuccess) {
    (void)hipGetLastError();             /* do not leave the error for a later check */
    *p = NULL;
    return SBA_ERR_HIP;
  }
  *cap = want;
  return 0;
}

/* pinned host staging (grow-only): the caller's pageable buffers are copied
 * into it by several threads, and the DMA runs from pinned memory at full
 * link rate */
static uint64_t *g_ha, *g_hb, *g_hr;
static size_t g_hna, g_hnb, g_hnr;

static int grow_pinned(uint64_t **p, size_t *cap, size_t n) {
  if (*cap >= n && atomic_load(&g_fail_alloc) <= 0) return 0;
  if (*p) (void)hipHostFree(*p);
  *p = NULL;
  *cap = 0;
  size_t want = n + n / 8 + 64;
  if (inject(&g_fail_alloc) ||
      hipHostMalloc((void **)p, want * 8, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *p = NULL;
    return SBA_ERR_HIP;
  }
  *cap = want;
  return 0;
}

#define COPY_THREADS 8
#define COPY_MIN_PER_THREAD ((size_t)1 << 19)     /* bytes */
struct copy_job { void *dst; const void *src; size_t n; };

/* Copy threads: COPY_THREADS - 1 workers started by the first parallel copy
 * and kept for the life of the process, so a staged copy costs a wake-up
 * per piece instead of thread creation. One parallel copy runs at a time
 * (g_pool_mu): it posts one slice per worker under a new generation number,
 * copies the last slice itself and waits until every worker has finished
 * that generation. A forked child (another pid) copies on one thread. */
static pthread_mutex_t g_pool_mu = PTHREAD_MUTEX_INITIALIZER;  /* one copy at a time */
static pthread_mutex_t g_pool_lk = PTHREAD_MUTEX_INITIALIZER;  /* guards the fields below */
static pthread_cond_t g_pool_go = PTHREAD_COND_INITIALIZER;
static pthread_cond_t g_pool_done = PTHREAD_COND_INITIALIZER;
static struct copy_job g_pool_job[COPY_THREADS];
static unsigned g_pool_gen;      /* generation of the posted slices */
static int g_pool_pending;       /* workers still on the current generation */
static int g_pool_n;             /* workers running */
static unsigned g_pool_gen0[COPY_THREADS];  /* each worker's generation at its start */
static atomic_int g_pool_pid;    /* the process that started them (0: none yet) */

static void *pool_worker(void *arg) {
  const int id = (int)(intptr_t)arg;
  pthread_mutex_lock(&g_pool_lk);
  unsigned seen = g_pool_gen0[id];
  for (;;) {
    while (g_pool_gen == seen) pthread_cond_wait(&g_pool_go, &g_pool_lk);
    seen = g_pool_gen;
    const struct copy_job j = g_pool_job[id];
    pthread_mutex_unlock(&g_pool_lk);
    if (j.n) memcpy(j.dst, j.src, j.n);
    pthread_mutex_lock(&g_pool_lk);
    if (--g_pool_pending == 0) pthread_cond_signal(&g_pool_done);
  }
  return NULL;
}

/* called with g_pool_mu held: the workers of this process, started once */
static int pool_workers(void) {
  const int me = (int)getpid();
  if (atomic_load(&g_pool_pid) == me) return g_pool_n;
  atomic_store(&g_pool_pid, me);
  g_pool_n = 0;
  pthread_attr_t at;
  if (pthread_attr_init(&at) != 0) return 0;
  (void)pthread_attr_setdetachstate(&at, PTHREAD_CREATE_DETACHED);
  for (int t = 0; t < COPY_THREADS - 1; ++t) {
    pthread_t th;
    const int id = g_pool_n;
    pthread_mutex_lock(&g_pool_lk);
    g_pool_gen0[id] = g_pool_gen;
    pthread_mutex_unlock(&g_pool_lk);
    if (pthread_create(&th, &at, pool_worker, (void *)(intptr_t)id) != 0) break;
    ++g_pool_n;
  }
  (void)pthread_attr_destroy(&at);
  return g_pool_n;
}

/* memcpy split over the calling thread and up to COPY_THREADS - 1 workers */
static void par_copy(void *dst, const void *src, size_t bytes) {
  size_t nt = bytes / COPY_MIN_PER_THREAD;
  if (nt > COPY_THREADS) nt = COPY_THREADS;
  const int pid = atomic_load(&g_pool_pid);
  if (nt < 2 || (pid && pid != (int)getpid())) {
    memcpy(dst, src, bytes);
    return;
  }
  pthread_mutex_lock(&g_pool_mu);
  const int nw = pool_workers();
  if (nt > (size_t)nw + 1) nt = (size_t)nw + 1;
  const size_t per = (bytes / nt + 63) & ~(size_t)63;
  size_t off = 0;
  pthread_mutex_lock(&g_pool_lk);
  for (int t = 0; t < nw; ++t) {      /* workers beyond nt - 1 get empty slices */
    const size_t n = (size_t)t + 1 < nt && off < bytes ? (bytes - off < per ? bytes - off : per) : 0;
    g_pool_job[t].dst = (char *)dst + off;
    g_pool_job[t].src = (const char *)src + off;
    g_pool_job[t].n = n;
    off += n;
  }
  g_pool_pending = nw;
  ++g_pool_gen;
  pthread_cond_broadcast(&g_pool_go);
  pthread_mutex_unlock(&g_pool_lk);
  if (off < bytes) memcpy((char *)dst + off, (const char *)src + off, bytes - off);
  pthread_mutex_lock(&g_pool_lk);
  while (g_pool_pending > 0) pthread_cond_wait(&g_pool_done, &g_pool_lk);
  pthread_mutex_unlock(&g_pool_lk);
  pthread_mutex_unlock(&g_pool_mu);
}

void sba_par_copy(void *dst, const void *src, size_t bytes) { par_copy(dst, src, bytes); }

/* Staged copies in pieces: the host copy of piece k+1 (caller memory <->
 * pinned staging, several threads) overlaps the DMA of piece k. */
#ifndef SBA_STAGE_PIECE_MB
#define SBA_STAGE_PIECE_MB 8
#endif
#define STAGE_PIECE ((size_t)SBA_STAGE_PIECE_MB << 20)   /* bytes */
#define STAGE_EVENTS 64
static hipEvent_t g_stage_ev[STAGE_EVENTS];

static size_t stage_piece(size_t bytes) {
  size_t piece = STAGE_PIECE;
  while (bytes / piece >= STAGE_EVENTS) piece *= 2;
  return piece;
}

/* dev <- host through the pinned buffer pin, on the default stream */
static int stage_up(void *dev, const void *host, void *pin, size_t bytes) {
  const size_t piece = stage_piece(bytes);
  for (size_t off = 0; off < bytes; off += piece) {
    const size_t n = bytes - off < piece ? bytes - off : piece;
    par_copy((char *)pin + off, (const char *)host + off, n);
    if (hipMemcpyAsync((char *)dev + off, (char *)pin + off, n, hipMemcpyHostToDevice, NULL) !=
        hipSuccess)
      return SBA_ERR_HIP;
  }
  return 0;
}

/* host <- dev through the pinned buffer pin: every piece's DMA is queued with
 * an event behind it, and each piece is copied out as soon as it lands */
static int stage_down(void *host, const void *dev, void *pin, size_t bytes) {
  const size_t piece = stage_piece(bytes);
  int k = 0;
  for (size_t off = 0; off < bytes; off += piece, ++k) {
    const size_t n = bytes - off < piece ? bytes - off : piece;
    if (!g_stage_ev[k] && hipEventCreateWithFlags(&g_stage_ev[k], hipEventDisableTiming) !=
                              hipSuccess) {
      g_stage_ev[k] = NULL;
      return SBA_ERR_HIP;
    }
    if (hipMemcpyAsync((char *)pin + off, (const char *)dev + off, n, hipMemcpyDeviceToHost,
                       NULL) != hipSuccess ||
        hipEventRecord(g_stage_ev[k], NULL) != hipSuccess)
      return SBA_ERR_HIP;
  }
  k = 0;
  for (size_t off = 0; off < bytes; off += piece, ++k) {
    const size_t n = bytes - off < piece ? bytes - off : piece;
    if (hipEventSynchronize(g_stage_ev[k]) != hipSuccess) return SBA_ERR_HIP;
    par_copy((char *)host + off, (const char *)pin + off, n);
  }
  return 0;
}

static int gpu_in_band(size_t an, size_t bn) {
  return sba_gpu_in_band((int64_t)an, (int64_t)bn);
}

/* H2D, multiply on the default stream, D2H. SBA_OK; SBA_BAND outside the
 * device band (nothing touched); < 0 on a HIP failure (reported on stderr;
 * rp unspecified). Aborts only when the process has no HIP device. */
static int gpu_mul_host(uint64_t *rp, const uint64_t *ap, size_t an,
                        const uint64_t *bp, size_t bn, int sqr) {
  if (!gpu_in_band(an, bn)) return SBA_BAND;   /* before touching any buffer */
  if (inject(&g_fail_entry)) {
    report("GPU multiply (injected failure)", SBA_ERR_HIP);
    return SBA_ERR_HIP;
  }
  if (!device_ok()) die("device lookup (no HIP device)", SBA_ERR_NODEV);
  pthread_mutex_lock(&g_stage_mu);
  int rc = 0;
  if (!rc) rc = grow(&g_da, &g_na, an);
  if (!rc && !sqr) rc = grow(&g_db, &g_nb, bn);
  if (!rc) rc = grow(&g_dr, &g_nr, an + bn);
  if (!rc) rc = grow_pinned(&g_ha, &g_hna, an);
  if (!rc && !sqr) rc = grow_pinned(&g_hb, &g_hnb, bn);
  if (!rc) rc = grow_pinned(&g_hr, &g_hnr, an + bn);
  if (rc) {
    report("device or pinned staging allocation", rc);
  } else {
    rc = stage_up(g_da, ap, g_ha, an * 8);
    if (!rc && !sqr) rc = stage_up(g_db, bp, g_hb, bn * 8);    /* overlaps a's last DMA */
    if (rc) report("hipMemcpyAsync H2D", rc);
  }
  if (!rc) {
    rc = sqr ? sba_gpu_sqr_dev(g_dr, g_da, (int64_t)an, NULL)
             : sba_gpu_mul_dev(g_dr, g_da, (int64_t)an, g_db, (int64_t)bn, NULL);
    if (rc == SBA_OK) {
      const int c = stage_down(rp, g_dr, g_hr, (an + bn) * 8);
      if (c) {
        report("hipMemcpyAsync D2H", c);
        rc = c;
      }
    } else if (rc < 0) {
      report("GPU multiply", rc);
    }
  }
  /* staging idle before the next caller reuses it */
  if (rc != SBA_OK && hipStreamSynchronize(NULL) != hipSuccess) (void)hipGetLastError();
  pthread_mutex_unlock(&g_stage_mu);
  return rc;
}

/* ------------------------------------------------------------------ */
/* NTT dispatcher entries: include/ntmul.h:47-79 (1 ok / 0 beyond band) */
/* ------------------------------------------------------------------ */
/* The reference's nt_* return 0 only beyond the p50 band
 * (include/ntmul.h:19,52-57; include/p50.h:1020-1046): the widest trunk
 * T in {88, 84, 80} whose CRT capacity floor(p0 p1 / 2^T) * floor(p2 p3 /
 * 2^T) (include/p50.h:161-169) admits ceil(64 min(an,bn) / T) trunks, and
 * ceil((ceil(64 an / T) + ceil(64 bn / T) - 1) / 8) <= 2^22 vectors. The
 * four 50-bit primes are include/p50.h:51-56; the capacities only depend on
 * them. The GPU band contains this band. */
static const uint64_t kP50Primes[4] = {1025844348715009ull, 1072023837081601ull,
                                       1086317488242689ull, 1108307720798209ull};
static int p50_band(ptrdiff_t an, ptrdiff_t bn) {
  if (an <= 0 || bn <= 0) return 0;
  const uint64_t mn = (uint64_t)(an < bn ? an : bn);
  const unsigned __int128 a = (unsigned __int128)kP50Primes[0] * kP50Primes[1];
  const unsigned __int128 b = (unsigned __int128)kP50Primes[2] * kP50Primes[3];
  static const int kT[3] = {80, 84, 88};
  int T = 0;
  for (int i = 2; i >= 0 && !T; --i) {
    const int t = kT[i];
    const unsigned __int128 cap = (a >> t) * (b >> t);
    const uint64_t capt = cap > (unsigned __int128)~0ull ? ~0ull : (uint64_t)cap;
    if ((mn * 64 + (uint64_t)t - 1) / (uint64_t)t <= capt) T = t;
  }
  if (!T) return 0;
  const uint64_t nat = ((uint64_t)an * 64 + T - 1) / T, nbt = ((uint64_t)bn * 64 + T - 1) / T;
  return (nat + nbt - 1 + 7) / 8 <= ((uint64_t)1 << 22);
}
int sba_ref_nt_band(int64_t an, int64_t bn) { return p50_band((ptrdiff_t)an, (ptrdiff_t)bn); }

int nt_mul_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
             ptrdiff_t bn, struct scratch *sc) {
  (void)sc;
  if (!p50_band(an, bn)) return 0;
  /* inside the reference band the device band always holds; a HIP failure
   * returns 0 (reported on stderr): the product is not written */
  return gpu_mul_host(rp, ap, (size_t)an, bp, (size_t)bn, 0) == SBA_OK;
}
int nt_sqr_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, struct scratch *sc) {
  (void)sc;
  if (!p50_band(an, an)) return 0;
  return gpu_mul_host(rp, ap, (size_t)an, ap, (size_t)an, 1) == SBA_OK;
}
int nt_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
           ptrdiff_t bn) {
  return nt_mul_r(rp, ap, an, bp, bn, NULL);
}
int nt_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an) {
  return nt_sqr_r(rp, ap, an, NULL);
}

/* p30's band: include/p30.h:1851-1858 */
#define P30_CAP 2344425u
static int p30_band(ptrdiff_t an, ptrdiff_t bn) {
  if (an <= 0 || bn <= 0) return 0;
  uint64_t mn = (uint64_t)(an < bn ? an : bn);
  uint64_t zlen = (uint64_t)an + (uint64_t)bn - 1;
  return mn <= P30_CAP && (zlen + 15) / 16 <= ((uint64_t)1 << 18);
}
int p30_mul_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
              ptrdiff_t bn, struct scratch *sc) {
  if (!p30_band(an, bn)) return 0;
  return nt_mul_r(rp, ap, an, bp, bn, sc);
}
int p30_sqr_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, struct scratch *sc) {
  if (!p30_band(an, an)) return 0;
  return nt_sqr_r(rp, ap, an, sc);
}
int p30_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
            ptrdiff_t bn) {
  return p30_mul_r(rp, ap, an, bp, bn, NULL);
}
int p30_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an) {
  return p30_sqr_r(rp, ap, an, NULL);
}
int p50_mul_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
              ptrdiff_t bn, struct scratch *sc) {
  return nt_mul_r(rp, ap, an, bp, bn, sc);
}
int p50_sqr_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, struct scratch *sc) {
  return nt_sqr_r(rp, ap, an, sc);
}
int p50_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
            ptrdiff_t bn) {
  return p50_mul_r(rp, ap, an, bp, bn, NULL);
}
int p50_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an) {
  return p50_sqr_r(rp, ap, an, NULL);
}

/* ------------------------------------------------------------------ */
/* CPU tiers below the GPU crossover (the library's own code):         */
/* u128 basecase and Karatsuba with strip-mining for unbalanced shapes  */
/* -- the roles of include/mul.h:233 (basecase), :326 (Karatsuba) and  */
/* :709 (strip-mining) in the reference's ladder.                      */
/* ------------------------------------------------------------------ */
#define SBA_KARA_MIN 32     /* basecase below this many limbs of the smaller operand */

static void cpu_basecase(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp,
                         size_t bn) {
  memset(rp, 0, (an + bn) * 8);
  for (size_t j = 0; j < bn; ++j) {
    uint64_t c = 0, b = bp[j];
    for (size_t i = 0; i < an; ++i) {
      u128 t = (u128)ap[i] * b + rp[i + j] + c;
      rp[i + j] = (uint64_t)t;
      c = (uint64_t)(t >> 64);
    }
    rp[j + an] = c;
  }
}

static uint64_t add_n(uint64_t *r, const uint64_t *a, const uint64_t *b, size_t n) {
  uint64_t c = 0;
  for (size_t i = 0; i < n; ++i) {
    u128 s = (u128)a[i] + b[i] + c;
    r[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  return c;
}
/* r = a(an) + b(bn), an >= bn; returns carry */
static uint64_t add_nm(uint64_t *r, const uint64_t *a, size_t an, const uint64_t *b,
                       size_t bn) {
  uint64_t c = add_n(r, a, b, bn);
  for (size_t i = bn; i < an; ++i) {
    u128 s = (u128)a[i] + c;
    r[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  return c;
}
/* r[0..n) += a[0..m), propagate carry through r; returns carry out */
static uint64_t add_into(uint64_t *r, size_t n, const uint64_t *a, size_t m) {
  uint64_t c = 0;
  size_t i = 0;
  for (; i < m; ++i) {
    u128 s = (u128)r[i] + a[i] + c;
    r[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  for (; c && i < n; ++i) {
    r[i] += 1;
    c = r[i] == 0;
  }
  return c;
}
static uint64_t sub_into(uint64_t *r, size_t n, const uint64_t *a, size_t m) {
  uint64_t bw = 0;
  size_t i = 0;
  for (; i < m; ++i) {
    uint64_t x = r[i], y = a[i];
    r[i] = x - y - bw;
    bw = (x < y) | ((x == y) & bw);
  }
  for (; bw && i < n; ++i) {
    bw = r[i] == 0;
    r[i] -= 1;
  }
  return bw;
}

static void cpu_mul_bal(uint64_t *rp, const uint64_t *ap, const uint64_t *bp, size_t n,
                        uint64_t *tmp);

/* balanced Karatsuba, both operands n limbs; tmp holds >= 4n + 64 limbs per level */
static void cpu_mul_bal(uint64_t *rp, const uint64_t *ap, const uint64_t *bp, size_t n,
                        uint64_t *tmp) {
  if (n < SBA_KARA_MIN) {
    cpu_basecase(rp, ap, n, bp, n);
    return;
  }
  size_t h = n / 2, l = n - h; /* low part l >= high part h */
  uint64_t *sa = tmp, *sb = tmp + l + 1, *mid = tmp + 2 * (l + 1);
  uint64_t *next = mid + 2 * (l + 1);
  sa[l] = add_nm(sa, ap, l, ap + l, h);
  sb[l] = add_nm(sb, bp, l, bp + l, h);
  cpu_mul_bal(rp, ap, bp, l, next);                 /* low  -> rp[0, 2l) */
  cpu_mul_bal(rp + 2 * l, ap + l, bp + l, h, next); /* high -> rp[2l, 2l+2h) */
  cpu_mul_bal(mid, sa, sb, l + 1, next);            /* (a0+a1)(b0+b1), 2l+2 limbs */
  sub_into(mid, 2 * l + 2, rp, 2 * l);
  sub_into(mid, 2 * l + 2, rp + 2 * l, 2 * h);
  add_into(rp + l, 2 * n - l, mid, 2 * l + 2 < 2 * n - l ? 2 * l + 2 : 2 * n - l);
}

static size_t kara_tmp(size_t n) {
  size_t t = 0;
  while (n >= SBA_KARA_MIN) {
    size_t l = n - n / 2;
    t += 4 * (l + 1);
    n = l + 1;
  }
  return t + 64;
}

/* scratch limbs cpu_mul_ws needs for an >= bn: the Karatsuba levels, one
 * 2bn-limb block product, and the strip-mined last block's own call */
static size_t cpu_need(size_t an, size_t bn) {
  if (bn < SBA_KARA_MIN) return 0;
  size_t t = kara_tmp(bn);
  if (an == bn) return t;
  t += 2 * bn + 8;
  const size_t w = an % bn;
  if (w) t += cpu_need(bn, w);
  return t;
}

/* an >= bn >= 1; tmp holds cpu_need(an, bn) limbs */
static void cpu_mul_ws(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp,
                       size_t bn, uint64_t *tmp) {
  if (bn < SBA_KARA_MIN) {
    cpu_basecase(rp, ap, an, bp, bn);
    return;
  }
  if (an == bn) {
    cpu_mul_bal(rp, ap, bp, bn, tmp);
    return;
  }
  /* strip-mine a in bn-limb blocks, overlap-adding each 2bn product */
  uint64_t *blk = tmp + kara_tmp(bn), *inner = blk + 2 * bn + 8;
  memset(rp, 0, (an + bn) * 8);
  size_t done = 0;
  while (done < an) {
    size_t w = an - done < bn ? an - done : bn;
    if (w == bn) cpu_mul_bal(blk, ap + done, bp, bn, tmp);
    else cpu_mul_ws(blk, bp, bn, ap + done, w, inner);
    add_into(rp + done, an + bn - done, blk, w + bn);
    done += w;
  }
}

/* per-thread grow-only scratch of the CPU tier (the role of the reference's
 * thread arena, include/scratch.h:72-77,339-356: no heap traffic once the
 * largest size has been seen). Host memory exhaustion aborts, as the
 * reference's arena does (include/scratch.h:66-70). */
static _Thread_local uint64_t *t_cpu_tmp;
static _Thread_local size_t t_cpu_cap;

/* an >= bn >= 1 */
static void cpu_mul(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp,
                    size_t bn) {
  const size_t need = cpu_need(an, bn);
  if (need > t_cpu_cap) {
    uint64_t *t = (uint64_t *)realloc(t_cpu_tmp, (need + need / 4) * 8);
    if (!t) die("host scratch allocation", 0);
    t_cpu_tmp = t;
    t_cpu_cap = need + need / 4;
  }
  cpu_mul_ws(rp, ap, an, bp, bn, t_cpu_tmp);
}

/* ------------------------------------------------------------------ */
/* simd_mpn_mul: include/u64cvt.h:94-132 semantics (swap to an >= bn,  */
/* smallest sizes on the basecase, threshold dispatch above).          */
/* ------------------------------------------------------------------ */
static _Atomic uint64_t g_gpu_min_total = SBA_GPU_MIN_TOTAL_DEFAULT;

void sba_set_gpu_threshold(uint64_t min_total_limbs) {
  atomic_store_explicit(&g_gpu_min_total, min_total_limbs, memory_order_relaxed);
}
uint64_t sba_get_gpu_threshold(void) {
  return atomic_load_explicit(&g_gpu_min_total, memory_order_relaxed);
}

/* ------------------------------------------------------------------ */
/* crossover tuner for the threshold above (the reference's tier       */
/* thresholds come from its own sweeps, include/u52-mparam.h:1-19):    */
/* balanced products on host buffers, this library's CPU multiply      */
/* against the GPU path (copies included), best of a few runs each.    */
/* ------------------------------------------------------------------ */
static const uint64_t kTuneTotals[] = {256,  384,   512,   768,   1024,  1536,  2048,  3072, 4096,
                                       6144, 8192, 12288, 16384, 24576, 32768, 49152, 65536};
#define TUNE_N ((int)(sizeof kTuneTotals / sizeof kTuneTotals[0]))

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

int sba_tune_sizes(uint64_t *totals, int max) {
  for (int i = 0; i < TUNE_N && i < max; ++i) totals[i] = kTuneTotals[i];
  return TUNE_N;
}

uint64_t sba_tune_gpu_threshold(int apply, double *cpu_ms, double *gpu_ms) {
  const uint64_t top = kTuneTotals[TUNE_N - 1];
  uint64_t *a = (uint64_t *)malloc(top * 8), *r = (uint64_t *)malloc(top * 8);
  if (!a || !r) die("host allocation in the tuner", 0);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  for (uint64_t i = 0; i < top; ++i) {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    a[i] = x;
  }
  int gpu_wins_from = TUNE_N;       /* first index from which the GPU always wins */
  for (int k = 0; k < TUNE_N; ++k) {
    const uint64_t T = kTuneTotals[k], an = T / 2, bn = T - an;
    const uint64_t *b = a + an;
    double bc = 1e30, bg = 1e30;
    int ok = gpu_mul_host(r, a, an, b, bn, 0) == SBA_OK;
    for (int rep = 0; rep < 5 && ok; ++rep) {
      double t0 = now_ms();
      cpu_mul(r, a, an, b, bn);
      double t1 = now_ms();
      ok = gpu_mul_host(r, a, an, b, bn, 0) == SBA_OK;
      double t2 = now_ms();
      if (t1 - t0 < bc) bc = t1 - t0;
      if (t2 - t1 < bg) bg = t2 - t1;
    }
    if (!ok) {                       /* a GPU failure (reported): no threshold */
      free(a);
      free(r);
      return 0;
    }
    if (cpu_ms) cpu_ms[k] = bc;
    if (gpu_ms) gpu_ms[k] = bg;
    if (bg < bc) {
      if (gpu_wins_from == TUNE_N) gpu_wins_from = k;
    } else {
      gpu_wins_from = TUNE_N;
    }
  }
  free(a);
  free(r);
  const uint64_t t = gpu_wins_from < TUNE_N ? kTuneTotals[gpu_wins_from] : 2 * top;
  if (apply) sba_set_gpu_threshold(t);
  return t;
}

void simd_mpn_mul(uint64_t *rp, const uint64_t *ap, uint64_t an, const uint64_t *bp,
                  uint64_t bn) {
  if (an < bn) {
    const uint64_t *t = ap; ap = bp; bp = t;
    uint64_t tn = an; an = bn; bn = tn;
  }
  if (bn == 0) {
    if (an) memset(rp, 0, an * 8);
    return;
  }
  if (an + bn >= sba_get_gpu_threshold()) {
    int rc = gpu_mul_host(rp, ap, an, bp, bn, ap == bp && an == bn);
    if (rc == SBA_OK) return;
    if (rc < 0) {
      /* a HIP failure (reported): the void entry still delivers the product */
      sba_note_fallback("simd_mpn_mul on the GPU", rc);
      cpu_mul(rp, ap, an, bp, bn);
      return;
    }
    /* beyond the GPU band (operands of > 43M limbs): split a in halves */
    size_t h = an / 2;
    uint64_t *t = (uint64_t *)malloc((an - h + bn) * 8);
    if (!t) die("host allocation", 0);
    simd_mpn_mul(rp, ap, h, bp, bn);
    memset(rp + h + bn, 0, (an - h) * 8);
    simd_mpn_mul(t, ap + h, an - h, bp, bn);
    add_into(rp + h, an + bn - h, t, an - h + bn);
    free(t);
    return;
  }
  cpu_mul(rp, ap, an, bp, bn);
}

/* internal multiply used by the division (either order, an,bn >= 1) */
void sba_mul(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp, size_t bn) {
  simd_mpn_mul(rp, ap, an, bp, bn);
}

/* ------------------------------------------------------------------ */
/* exported helpers for the tests of the host logic                    */
/* ------------------------------------------------------------------ */
void sba_cpu_mul(uint64_t *rp, const uint64_t *ap, uint64_t an, const uint64_t *bp,
                 uint64_t bn) {
  if (an < bn) {
    const uint64_t *t = ap; ap = bp; bp = t;
    uint64_t tn = an; an = bn; bn = tn;
  }
  if (bn == 0) {
    if (an) memset(rp, 0, an * 8);
    return;
  }
  cpu_mul(rp, ap, an, bp, bn);
}
