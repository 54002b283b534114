// div_dev.hip -- the large division's Newton branch kept in HBM.
//
// The host division (host/sba_div.c) computes the quotient from one
// approximate reciprocal (Brent & Zimmermann, Modern Computer Arithmetic,
// Algorithm 3.5) and one quotient multiply. Above a size threshold this file
// runs the whole branch on device buffers: every multiply is the GPU NTT on
// device pointers, and the linear steps between them are kernels -- the
// normalising shifts, an add/subtract with a carry-lookahead scan, the
// complement B^k - T (an add of ~T and 1), the comparison of the correction
// loop, the reciprocal's overflow fix-up. The operands go up and the
// quotient and remainder come back through pinned staging, small words
// through a pinned mailbox; only the reciprocal's recursion at or below
// DIV_HOST_RECIP limbs runs on the host.
#include "sba_internal.h"

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>

typedef uint64_t u64;

#define DNT 256
#define DLPT 8
#define DLPB (DNT * DLPT)          // limbs per block
#define DIV_HOST_RECIP 1024        // reciprocals of at most this many limbs: host

// ---------------------------------------------------------------------------
// add with carry: r = a + (sub ? ~b : b) + cin over n limbs (a or b may be
// null = zero; b has bn <= n limbs, zero-extended). Carries are 0/1, so a
// limb is a map {0,1} -> {0,1}: bit 0 = carry out for carry in 0, bit 1 =
// for carry in 1; maps compose associatively.
// ---------------------------------------------------------------------------
#define M2_ID 2u
__device__ __forceinline__ uint32_t m2_apply(uint32_t f, uint32_t c) { return (f >> c) & 1u; }
__device__ __forceinline__ uint32_t m2_then(uint32_t f, uint32_t g) {    // f, then g
  return m2_apply(g, f & 1u) | (m2_apply(g, (f >> 1) & 1u) << 1);
}

// block-wide exclusive scan of maps; *total = the block's composition
__device__ uint32_t m2_block_scan(uint32_t f, uint32_t *total) {
  __shared__ uint32_t wt[DNT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = f;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc = m2_then(o, inc);
  }
  uint32_t exc = __shfl_up(inc, 1, 64);
  if (lane == 0) exc = M2_ID;
  if (lane == 63) wt[wid] = inc;
  __syncthreads();
  uint32_t pre = M2_ID, tot = M2_ID;
  for (int w = 0; w < DNT / 64; ++w) {
    if (w < wid) pre = m2_then(pre, wt[w]);
    tot = m2_then(tot, wt[w]);
  }
  *total = tot;
  __syncthreads();
  return m2_then(pre, exc);
}

// pass 1: limb sums, one map byte per limb, one map per block
// (a or b may be the same buffer as s: each limb is read before it is written)
__global__ __launch_bounds__(DNT) void k_add_maps(const u64 *a, const u64 *b, size_t bn, size_t n,
                                                  int sub, u64 *s,
                                                  uint8_t *__restrict__ lm,
                                                  uint32_t *__restrict__ bm) {
  __shared__ uint8_t mp[DLPB];
  const size_t i0 = (size_t)blockIdx.x * DLPB;
#pragma unroll
  for (int k = 0; k < DLPT; ++k) {
    const uint32_t j = threadIdx.x + k * DNT;
    const size_t i = i0 + j;
    uint32_t f = M2_ID;
    if (i < n) {
      const u64 x = a ? a[i] : 0;
      u64 y = (b && i < bn) ? b[i] : 0;
      if (sub) y = ~y;
      const u64 t = x + y;
      const uint32_t g = t < x, p = t == ~(u64)0;
      f = g | ((g | p) << 1);
      s[i] = t;
    }
    mp[j] = (uint8_t)f;
  }
  __syncthreads();
  uint32_t f = M2_ID;
#pragma unroll
  for (int k = 0; k < DLPT; ++k) {
    const uint32_t m = mp[threadIdx.x * DLPT + k];
    lm[i0 + threadIdx.x * DLPT + k] = (uint8_t)m;
    f = m2_then(f, m);
  }
  uint32_t tot;
  m2_block_scan(f, &tot);
  if (threadIdx.x == 0) bm[blockIdx.x] = tot;
}

// pass 2: carry into the block (earlier block maps composed, applied to
// cin), carries into the limbs, r = s + c; *cout (if not null) = the carry
// out of the top limb
__global__ __launch_bounds__(DNT) void k_add_fix(const u64 *s, const uint8_t *__restrict__ lm,
                                                 const uint32_t *__restrict__ bm, size_t n,
                                                 uint32_t cin, u64 *r,
                                                 uint32_t *__restrict__ cout) {
  __shared__ uint8_t cj[DLPB];
  const size_t i0 = (size_t)blockIdx.x * DLPB;
  const uint32_t nprev = blockIdx.x;
  const uint32_t per = (nprev + DNT - 1) / DNT;
  uint32_t g = M2_ID;
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t idx = threadIdx.x * per + k;
    if (idx < nprev) g = m2_then(g, bm[idx]);
  }
  uint32_t gall;
  m2_block_scan(g, &gall);
  uint32_t f = M2_ID;
  uint8_t ms[DLPT];
#pragma unroll
  for (int k = 0; k < DLPT; ++k) {
    const size_t i = i0 + threadIdx.x * DLPT + k;
    ms[k] = i < n ? lm[i] : (uint8_t)M2_ID;
    f = m2_then(f, ms[k]);
  }
  uint32_t tot;
  const uint32_t pre = m2_block_scan(f, &tot);
  uint32_t c = m2_apply(pre, m2_apply(gall, cin));
#pragma unroll
  for (int k = 0; k < DLPT; ++k) {
    cj[threadIdx.x * DLPT + k] = (uint8_t)c;
    c = m2_apply(ms[k], c);
  }
  if (cout && blockIdx.x == gridDim.x - 1 && threadIdx.x == DNT - 1) *cout = c;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < DLPT; ++k) {
    const uint32_t j = threadIdx.x + k * DNT;
    const size_t i = i0 + j;
    if (i < n) r[i] = s[i] + cj[j];
  }
}

// *flag |= 1 when any of x[0..n) is nonzero
__global__ __launch_bounds__(DNT) void k_any_nonzero(const u64 *__restrict__ x, size_t n,
                                                     uint32_t *__restrict__ flag) {
  const size_t i = (size_t)blockIdx.x * DNT + threadIdx.x;
  if (i < n && x[i] != 0) atomicOr(flag, 1u);
}

// dst[0..n] = src[0..n) << s (0 <= s < 64); dst[n] = the bits shifted out
__global__ __launch_bounds__(DNT) void k_shl(u64 *__restrict__ dst, const u64 *__restrict__ src,
                                             size_t n, unsigned s) {
  const size_t i = (size_t)blockIdx.x * DNT + threadIdx.x;
  if (i > n) return;
  const u64 cur = i < n ? src[i] : 0, prev = i > 0 ? src[i - 1] : 0;
  dst[i] = s ? (cur << s) | (prev >> (64 - s)) : cur;
}

// dst[0..n) = src[0..n) >> s (0 <= s < 64)
__global__ __launch_bounds__(DNT) void k_shr(u64 *__restrict__ dst, const u64 *__restrict__ src,
                                             size_t n, unsigned s) {
  const size_t i = (size_t)blockIdx.x * DNT + threadIdx.x;
  if (i >= n) return;
  const u64 cur = src[i], next = i + 1 < n ? src[i + 1] : 0;
  dst[i] = s ? (cur >> s) | (next << (64 - s)) : cur;
}

// *top = max(*top, i + 1) over the limbs where a[i] != b[i] (b zero-extended
// from bn limbs): 0 when a and b are equal over n limbs; one atomic per block
__global__ __launch_bounds__(DNT) void k_top_diff(const u64 *__restrict__ a,
                                                  const u64 *__restrict__ b, size_t bn, size_t n,
                                                  unsigned long long *__restrict__ top) {
  __shared__ unsigned long long bmax;
  if (threadIdx.x == 0) bmax = 0;
  __syncthreads();
  const size_t i = (size_t)blockIdx.x * DNT + threadIdx.x;
  if (i < n && a[i] != (i < bn ? b[i] : 0)) atomicMax(&bmax, (unsigned long long)(i + 1));
  __syncthreads();
  if (threadIdx.x == 0 && bmax) atomicMax(top, bmax);
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
#define DCHK(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "simd_bigint_amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return SBA_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

// bump allocator over one grow-only device buffer (u64 units)
struct DivPool {
  u64 *base = nullptr;
  size_t cap = 0, top = 0;
  uint8_t *lm = nullptr;            // add scratch: limb maps, block maps, carry word
  uint32_t *bm = nullptr, *cw = nullptr;
  size_t lmcap = 0;
  unsigned long long *word = nullptr;   // k_top_diff's result
  u64 *mbox = nullptr;                  // pinned host words for small read-backs
  u64 *pin = nullptr;               // pinned host staging (grow-only)
  size_t pincap = 0;
  u64 *take(size_t n) {
    u64 *p = base + top;
    top += (n + 63) & ~(size_t)63;
    return top <= cap ? p : nullptr;
  }
};
static DivPool g_dp;
static std::mutex g_dp_mu;

// r = a + (sub ? ~b : b) + cin over n limbs; *cout_host (if not null) gets
// the carry out (synchronizes)
static int dev_add(u64 *r, const u64 *a, const u64 *b, size_t bn, size_t n, int sub,
                   uint32_t cin, uint32_t *cout_host, hipStream_t st) {
  if (n == 0) return SBA_OK;
  const size_t nb = (n + DLPB - 1) / DLPB;
  if (nb * DLPB > g_dp.lmcap) return SBA_ERR_WORKSPACE;
  hipLaunchKernelGGL(k_add_maps, dim3((unsigned)nb), dim3(DNT), 0, st, a, b, bn, n, sub, r,
                     g_dp.lm, g_dp.bm);
  hipLaunchKernelGGL(k_add_fix, dim3((unsigned)nb), dim3(DNT), 0, st, (const u64 *)r,
                     (const uint8_t *)g_dp.lm, (const uint32_t *)g_dp.bm, n, cin, r,
                     cout_host ? g_dp.cw : nullptr);
  DCHK(hipGetLastError());
  if (cout_host) {
    DCHK(hipMemcpyAsync(g_dp.mbox + 1, g_dp.cw, 4, hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    *cout_host = (uint32_t)(g_dp.mbox[1] & 1u);
  }
  return SBA_OK;
}

static int dev_mul(u64 *r, const u64 *a, size_t an, const u64 *b, size_t bn, hipStream_t st) {
  const int rc = sba_gpu_mul_dev(r, a, (int64_t)an, b, (int64_t)bn, st);
  return rc == SBA_OK ? SBA_OK : (rc == SBA_BAND ? SBA_ERR_WORKSPACE : rc);
}

// X[0..n] (device) = approximate reciprocal of A[0..n) (device, top bit set)
static int recip_dev(u64 *X, const u64 *A, size_t n, hipStream_t st) {
  if (n <= DIV_HOST_RECIP) {
    // through the pinned staging buffer (idle here: the uploads before this
    // point have completed once the stream is synchronized)
    u64 *ha = g_dp.pin, *hx = g_dp.pin + n;
    DCHK(hipStreamSynchronize(st));
    DCHK(hipMemcpyAsync(ha, A, n * 8, hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    sba_approx_recip_host(hx, ha, n);
    DCHK(hipMemcpyAsync(X, hx, (n + 1) * 8, hipMemcpyHostToDevice, st));
    DCHK(hipStreamSynchronize(st));
    return SBA_OK;
  }
  const size_t l = (n - 1) / 2, h = n - l;
  const size_t mark = g_dp.top;
  u64 *Xh = g_dp.take(h + 1);
  u64 *T = g_dp.take(n + h + 1);
  u64 *U = g_dp.take(3 * h + 1);
  if (!Xh || !T || !U) return SBA_ERR_WORKSPACE;
  int rc = recip_dev(Xh, A + l, h, st);
  if (rc != SBA_OK) return rc;
  // T = A * Xh (n + h + 1 limbs)
  if ((rc = dev_mul(T, A, n, Xh, h + 1, st)) != SBA_OK) return rc;
  // while T >= B^(n+h): Xh -= 1, T -= A, on the device. T < B^(n+h) + 2 B^n
  // (A_h Xh < B^2h, Xh < 2 B^h) and A >= B^n / 2, so at most 4 rounds.
  for (int round = 0;; ++round) {
    DCHK(hipMemcpyAsync(g_dp.mbox, T + n + h, 8, hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    if (g_dp.mbox[0] == 0) break;
    if (round == 4) return SBA_BAND;
    if ((rc = dev_add(Xh, Xh, nullptr, 0, h + 1, 1, 0, nullptr, st)) != SBA_OK) return rc;
    if ((rc = dev_add(T, T, A, n, n + h + 1, 1, 1, nullptr, st)) != SBA_OK) return rc;
  }
  // T = B^(n+h) - T over n + h limbs: ~T + 1
  if ((rc = dev_add(T, nullptr, T, n + h, n + h, 1, 1, nullptr, st)) != SBA_OK) return rc;
  // U = floor(T / B^l) * Xh: 3h + 1 limbs
  if ((rc = dev_mul(U, T + l, 2 * h, Xh, h + 1, st)) != SBA_OK) return rc;
  // X = Xh * B^l + floor(U / B^(2h - l)) over n + 1 limbs
  const size_t sh = 2 * h - l;
  const size_t ul = 3 * h + 1 - sh < n + 1 ? 3 * h + 1 - sh : n + 1;
  if (l) DCHK(hipMemcpyAsync(X, U + sh, l * 8, hipMemcpyDeviceToDevice, st));
  if ((rc = dev_add(X + l, Xh, U + sh + l, ul > l ? ul - l : 0, n + 1 - l, 0, 0, nullptr, st)) !=
      SBA_OK)
    return rc;
  g_dp.top = mark;
  return SBA_OK;
}

static int grow_pool(size_t limbs, size_t addmax) {
  if (g_dp.cap < limbs) {
    if (g_dp.base) DCHK(hipFree(g_dp.base));
    g_dp.base = nullptr;
    g_dp.cap = 0;
    const size_t want = limbs + limbs / 8 + 4096;
    DCHK(hipMalloc(&g_dp.base, want * 8));
    g_dp.cap = want;
  }
  const size_t lmneed = ((addmax + DLPB - 1) / DLPB + 1) * DLPB;
  if (g_dp.lmcap < lmneed) {
    if (g_dp.lm) DCHK(hipFree(g_dp.lm));
    g_dp.lm = nullptr;
    g_dp.lmcap = 0;
    const size_t nb = lmneed / DLPB;
    DCHK(hipMalloc(&g_dp.lm, lmneed + nb * 4 + 64));
    g_dp.bm = (uint32_t *)(g_dp.lm + lmneed);
    g_dp.cw = g_dp.bm + nb;
    g_dp.lmcap = lmneed;
  }
  if (!g_dp.word) DCHK(hipMalloc(&g_dp.word, 64));
  if (!g_dp.mbox) DCHK(hipHostMalloc((void **)&g_dp.mbox, 64, hipHostMallocDefault));
  g_dp.top = 0;
  return SBA_OK;
}

static int grow_pinned(size_t limbs) {
  if (g_dp.pincap >= limbs) return SBA_OK;
  if (g_dp.pin) DCHK(hipHostFree(g_dp.pin));
  g_dp.pin = nullptr;
  g_dp.pincap = 0;
  const size_t want = limbs + limbs / 8 + 64;
  DCHK(hipHostMalloc((void **)&g_dp.pin, want * 8, hipHostMallocDefault));
  g_dp.pincap = want;
  return SBA_OK;
}

static unsigned grid_for(size_t n) { return (unsigned)((n + DNT - 1) / DNT); }

// *sign = sign of a[0..n) - b[0..bn) (bn <= n, b zero-extended); synchronizes
static int dev_cmp(const u64 *a, size_t n, const u64 *b, size_t bn, int *sign, hipStream_t st) {
  DCHK(hipMemsetAsync(g_dp.word, 0, 8, st));
  hipLaunchKernelGGL(k_top_diff, dim3(grid_for(n)), dim3(DNT), 0, st, a, b, bn, n, g_dp.word);
  DCHK(hipGetLastError());
  DCHK(hipMemcpyAsync(g_dp.mbox + 2, g_dp.word, 8, hipMemcpyDeviceToHost, st));
  DCHK(hipStreamSynchronize(st));
  const u64 top = g_dp.mbox[2];
  if (top == 0) {
    *sign = 0;
    return SBA_OK;
  }
  const size_t i = (size_t)top - 1;
  g_dp.mbox[3] = g_dp.mbox[4] = 0;
  DCHK(hipMemcpyAsync(g_dp.mbox + 3, a + i, 8, hipMemcpyDeviceToHost, st));
  if (i < bn) DCHK(hipMemcpyAsync(g_dp.mbox + 4, b + i, 8, hipMemcpyDeviceToHost, st));
  DCHK(hipStreamSynchronize(st));
  *sign = g_dp.mbox[3] > g_dp.mbox[4] ? 1 : -1;
  return SBA_OK;
}

// The division's Newton branch on device buffers (host/sba_div.c decides
// when): N has nt limbs (top limb nonzero), D has dn limbs (top limb
// nonzero), s = the leading zero bits of D's top limb, qn = nt + 1 - dn.
// On the device: N' = N << s (nt + 1 limbs), D' = D << s; the top qn + 1
// limbs of D' plus one (or D' zero-extended) give Dm and its reciprocal X;
// Qest = floor((N' >> 64 (dn - 1)) X / B^(qn + 2)); R = N' - Qest D'; while
// R >= D': R -= D', Q += 1; R >>= s. Writes qp[0..min(qn, qm)) and
// rp[0..dn). Returns SBA_OK, SBA_BAND when an estimate leaves its bounds
// (the host then aborts, as its own Newton branch does), < 0 on a device
// error.
extern "C" int sba_div_dev(u64 *qp, size_t qm, u64 *rp, const u64 *np, size_t nt, const u64 *dp,
                           size_t dn, unsigned s) {
  std::lock_guard<std::mutex> lk(g_dp_mu);
  const size_t qn = nt + 1 - dn, m = qn + 1, nt1 = nt + 1;
  hipStream_t st = nullptr;
  // device buffers: raw N and D, N', D', Dm, X, P, QD / R, R >> s, recursion (about 6m)
  const size_t need = nt + dn + nt1 + (dn + 1) + m + (m + 1) + (qn + 1 + m + 1) + nt1 + dn +
                      8 * m + 64 * 64;
  int rc = grow_pool(need, nt1 + 2 * m + 2);
  if (rc != SBA_OK) return rc;
  const size_t qc = qn < qm ? qn : qm;
  if ((rc = grow_pinned(nt + dn > qc + dn ? nt + dn : qc + dn)) != SBA_OK) return rc;
  u64 *dN0 = g_dp.take(nt), *dD0 = g_dp.take(dn), *dN = g_dp.take(nt1), *dD = g_dp.take(dn + 1);
  u64 *dDm = g_dp.take(m), *dX = g_dp.take(m + 1), *dP = g_dp.take(qn + 1 + m + 1);
  u64 *dQD = g_dp.take(nt1), *dRs = g_dp.take(dn);
  if (!dN0 || !dD0 || !dN || !dD || !dDm || !dX || !dP || !dQD || !dRs) return SBA_ERR_WORKSPACE;
  // operands up through pinned staging (the second host copy overlaps the first DMA)
  sba_par_copy(g_dp.pin, np, nt * 8);
  DCHK(hipMemcpyAsync(dN0, g_dp.pin, nt * 8, hipMemcpyHostToDevice, st));
  sba_par_copy(g_dp.pin + nt, dp, dn * 8);
  DCHK(hipMemcpyAsync(dD0, g_dp.pin + nt, dn * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_shl, dim3(grid_for(nt + 1)), dim3(DNT), 0, st, dN, (const u64 *)dN0, nt, s);
  hipLaunchKernelGGL(k_shl, dim3(grid_for(dn + 1)), dim3(DNT), 0, st, dD, (const u64 *)dD0, dn, s);
  DCHK(hipGetLastError());
  // Dm = the top m limbs of D' plus one (carry out: Dm + 1 = B^m), or D' zero-extended
  int exact_pow = 0;
  if (dn >= m) {
    uint32_t cout = 0;
    if ((rc = dev_add(dDm, dD + (dn - m), nullptr, 0, m, 0, 1, &cout, st)) != SBA_OK) return rc;
    exact_pow = cout != 0;
  } else {
    DCHK(hipMemsetAsync(dDm, 0, (m - dn) * 8, st));
    DCHK(hipMemcpyAsync(dDm + (m - dn), dD, dn * 8, hipMemcpyDeviceToDevice, st));
  }
  if (exact_pow) {                   // X = B^2m / B^m = B^m
    g_dp.mbox[6] = 1;
    DCHK(hipMemsetAsync(dX, 0, m * 8, st));
    DCHK(hipMemcpyAsync(dX + m, g_dp.mbox + 6, 8, hipMemcpyHostToDevice, st));
    DCHK(hipStreamSynchronize(st));
  } else if ((rc = recip_dev(dX, dDm, m, st)) != SBA_OK) {
    return rc;
  }
  // Qest = floor(Nt * X / B^(qn + 2)), Nt = N' >> 64 (dn - 1): qn + 1 limbs
  if ((rc = dev_mul(dP, dN + (dn - 1), qn + 1, dX, m + 1, st)) != SBA_OK) return rc;
  u64 *dQ = dP + (qn + 2);
  DCHK(hipMemcpyAsync(g_dp.mbox, dQ + qn, 8, hipMemcpyDeviceToHost, st));
  DCHK(hipStreamSynchronize(st));
  if (g_dp.mbox[0] != 0) return SBA_BAND;
  // R = N' - Qest D' over nt + 1 limbs (no borrow: carry out 1)
  if ((rc = dev_mul(dQD, dQ, qn, dD, dn, st)) != SBA_OK) return rc;
  uint32_t cout = 0;
  if ((rc = dev_add(dQD, dN, dQD, nt1, nt1, 1, 1, &cout, st)) != SBA_OK) return rc;
  if (!cout) return SBA_BAND;
  // the remainder of an estimate at most a few below Q is below B^(dn+1):
  // the limbs above must be zero
  if (nt1 > dn + 1) {
    DCHK(hipMemsetAsync(g_dp.cw, 0, 4, st));
    hipLaunchKernelGGL(k_any_nonzero, dim3(grid_for(nt1 - (dn + 1))), dim3(DNT), 0, st,
                       (const u64 *)(dQD + dn + 1), nt1 - (dn + 1), g_dp.cw);
    DCHK(hipGetLastError());
    DCHK(hipMemcpyAsync(g_dp.mbox + 5, g_dp.cw, 4, hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    if (g_dp.mbox[5] & 0xFFFFFFFFu) return SBA_BAND;
  }
  // while R >= D': R -= D', Q += 1 (R over dn + 1 limbs)
  for (int fix = 0;; ++fix) {
    int sign = 0;
    if ((rc = dev_cmp(dQD, dn + 1, dD, dn, &sign, st)) != SBA_OK) return rc;
    if (sign < 0) break;
    if (fix >= 64) return SBA_BAND;
    if ((rc = dev_add(dQD, dQD, dD, dn, dn + 1, 1, 1, nullptr, st)) != SBA_OK) return rc;
    if ((rc = dev_add(dQ, dQ, nullptr, 0, qn + 1, 0, 1, nullptr, st)) != SBA_OK) return rc;
  }
  hipLaunchKernelGGL(k_shr, dim3(grid_for(dn)), dim3(DNT), 0, st, dRs, (const u64 *)dQD, dn, s);
  DCHK(hipGetLastError());
  // results down through pinned staging
  DCHK(hipMemcpyAsync(g_dp.pin, dQ, qc * 8, hipMemcpyDeviceToHost, st));
  DCHK(hipMemcpyAsync(g_dp.pin + qc, dRs, dn * 8, hipMemcpyDeviceToHost, st));
  DCHK(hipStreamSynchronize(st));
  sba_par_copy(qp, g_dp.pin, qc * 8);
  sba_par_copy(rp, g_dp.pin + qc, dn * 8);
  return SBA_OK;
}
