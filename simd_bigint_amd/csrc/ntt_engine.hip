// ntt_engine.hip -- gfx950 kernels of the multi-prime NTT multiply and the
// thin C-ABI layer that launches them (declared in include/simd_bigint_amd.h).
// The pass kernels live in ntt_tiles.h; this file holds Garner, the carry
// scan, the twiddle tables and all host code.
//
// Math (one prime p, transform length N = 2^L):
//   * Root table in the FLINT / reference convention (include/p30.h:312-329,
//     tests/p30_tft_model_test.c:66-94): w[0] = 1, w[2^(t-1) + s] =
//     zeta_t * w[s], zeta_t of order 2^t, zeta_1 = -1. The butterfly of
//     node j uses w[2j]; the device keeps the node table wn[j] = w[2j]
//     (and wni[j] = w[2j]^-1) in Montgomery form. Both are prefix-stable,
//     so one table built for the largest L serves every smaller L.
//   * Forward = the flat N-point Cooley-Tukey network: level s, block j
//     (j = i >> (L - s)) pairs i, i + N/2^(s+1): (x, y) -> (x + wn[j] y,
//     x - wn[j] y). Output position k holds the evaluation at w[k], the
//     same spectral order as the reference's forward transform.
//   * Inverse = the Gentleman-Sande mirror with wni, levels reversed; the
//     factor N is removed in the Garner stage.
//   * A PASS runs ELL consecutive levels [s0, s0 + ELL) on TILES: 2^ELL
//     elements of stride S = 2^(L - s0 - ELL) (one block of level s0) times
//     TC = 2^LOGTC adjacent such groups, staged in LDS (<= 8192 words).
//     Inside the tile, CHUNKS of up to 5 levels run in registers (radix-32
//     sub-butterflies per thread), with LDS exchanges between chunks.
//   * Garner over the five residues (ascending primes) composes each
//     coefficient into 3 u64 words; the words are summed into limbs and a
//     binary carry-lookahead (ballot masks in a block, a scan of 2-state
//     block maps across blocks) finishes the limbs bit-exactly.
#include "ntt_tiles.h"
#include "plan_table.h"

#include <hip/hip_ext.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <mutex>

// ---------------------------------------------------------------------------
// Garner: the five residues of coefficient i -> 3 u64 words
// ---------------------------------------------------------------------------
__device__ __forceinline__ void garner_coef(const GarnerArgs &ga, gc_u32 *spec,
                                            uint32_t N, uint32_t i, uint64_t &c0, uint64_t &c1,
                                            uint64_t &c2) {
  uint32_t u[SBA_K], x[SBA_K];
  // prime q's residues start at spec + q N, a uniform base kept in scalar
  // registers (readfirstlane: the compiler does not fold it into a 64-bit
  // per-lane address); coefficient i is a 32-bit byte offset (i < 2^30)
#pragma unroll
  for (int q = 0; q < SBA_K; ++q)
    u[q] = red1(ld_u32(uni_ptr(spec + (size_t)q * N), i), ga.p[q]);   // already scaled by the inverse
  // digit q: x_q = (u_q - x_0 - sum_{0<k<q} x_k c_qk) / (p_0 ... p_{q-1})
  // mod p_q. The products x_k c_qk (each < 2^62, the c_qk in Montgomery
  // form) are summed two at a time in 64 bits and reduced once per pair:
  // (s + m p) / 2^32 < (2^63 + 2^63) / 2^32, and < 2p since s < 2^32 p.
  // u_q + p - t lies in (0, 2p), a valid left operand of the last product.
  x[0] = u[0];
#pragma unroll
  for (int q = 1; q < SBA_K; ++q) {
    const uint32_t p = ga.p[q], pinv = ga.pinv[q];
    uint32_t t = 0;                                 // sum of the x_k c_qk, in [0, p)
#pragma unroll
    for (int k = 1; k < q; k += 2) {
      uint64_t s = (uint64_t)x[k] * ga.cm[q][k];
      if (k + 1 < q) s += (uint64_t)x[k + 1] * ga.cm[q][k + 1];
      const uint32_t m = (uint32_t)s * pinv;
      const uint32_t r = red1((uint32_t)(((uint64_t)m * p + s) >> 32), p);
      t = red1(t + r, p);
    }
    t = red1(t + x[0], p);                          // x_0 < p_0 < p_q
    x[q] = red1(mont_mul(u[q] + p - t, ga.im[q], p, pinv), p);
  }
  // Horner: c = (((x4 p3 + x3) p2 + x2) p1 + x1) p0 + x0  (< 2^155) on
  // 32-bit words, one 32 x 32 + 64-bit multiply-add per word and step
  uint32_t w[SBA_K];
  w[0] = x[4];
#pragma unroll
  for (int s = 0; s < SBA_K - 1; ++s) {
    const uint32_t P = ga.p[3 - s];
    uint64_t t = x[3 - s];
#pragma unroll
    for (int i = 0; i <= s; ++i) {
      t += (uint64_t)w[i] * P;                   // < 2^63 + 2^32
      w[i] = (uint32_t)t;
      t >>= 32;
    }
    w[s + 1] = (uint32_t)t;
  }
  c0 = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  c1 = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
  c2 = w[4];
}

// ---------------------------------------------------------------------------
// carry propagation. Limb j receives s_j = w0[j] + w1[j-1] + w2[j-2]
// (< 3 * 2^64) = lo_j + 2^64 hi_j, hi_j <= 2. The high part moves one limb
// up: t_j = lo_j + hi_{j-1} (mod 2^64) with g_j = [it wrapped]. What is left
// is a plain binary carry chain, c_{j+1} = g_j | (p_j & c_j) with p_j =
// [t_j = 2^64 - 1] (g_j and p_j never both hold: a wrapped t_j is < 2), and
// out_j = t_j + c_j. On a 64-limb run held by one wave the chain is an
// integer addition of ballot masks: with A = G | P and B = G, the carries
// into the limbs are (A + B + c_in) ^ A ^ B. Carries leave a run or a block
// only on a generate followed by all-ones limbs, so for most data every
// block's carry-in is 0.
//   k_garner_limbs  Garner + limb words; carries inside each 1024-limb
//                   block (block carry-in 0); the block's 2-state carry map
//   k_carry_blocks  the block carry-ins (a scan of the block maps), added
//                   where they are 1 (usually to one limb)
// ---------------------------------------------------------------------------
#ifndef SBA_GARNER_LPT
#define SBA_GARNER_LPT 4                // limbs per thread of the Garner stage (measured: 8 and 2 slower)
#endif
#ifndef SBA_GARNER_UNROLL
#define SBA_GARNER_UNROLL 2             // coefficients in flight per thread
#endif
#define CARRY_LPT SBA_GARNER_LPT
#define CARRY_LPB (CARRY_LPT * NT)      // limbs per (carry) block
#define SCAN_NT 1024                    // k_carry_blocks threads
#define CARRY_WGS 256                   // k_carry_blocks workgroups per product (at most)
// k_carry_blocks blocks per workgroup (at least): a workgroup of SCAN_NT
// threads per block of 1024 limbs would make a batch of many small products
// launch 16 waves for every few limbs that need a carry
#ifndef CARRY_MIN_SPAN
#define CARRY_MIN_SPAN 16
#endif
#define C2_ID 2u                        // 2-state map c -> c: bit c = carry out for carry in c

__device__ __forceinline__ uint32_t c2_apply(uint32_t f, uint32_t c) { return (f >> c) & 1u; }
// first f, then g
__device__ __forceinline__ uint32_t c2_then(uint32_t f, uint32_t g) {
  return c2_apply(g, f & 1u) | (c2_apply(g, (f >> 1) & 1u) << 1);
}

// The limbs [B0, B0 + LPT * NT) of a product from the coefficient words in
// LDS (slot j + 3 holds the words of coefficient B0 + j, slots 0..2 those of
// B0-3..B0-1; zero outside the product), finished with block carry-in 0;
// *bmap (if not null) gets the block's 2-state carry map. Called by the whole
// workgroup after the words are in place. W2: the high word's type (its
// values stay below 2^32 for Garner's coefficients).
template <int LPT, typename W2>
__device__ __forceinline__ void limbs_from_words(const uint64_t *ws0, const uint64_t *ws1,
                                                 const W2 *ws2, uint32_t B0, uint32_t Z,
                                                 g_u64 *rp, g_u32 *bmap) {
  constexpr int RUNS = LPT * NT / 64;             // 64-limb runs
  __shared__ uint32_t rmap[RUNS], rcin[RUNS];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t t[LPT], ra[LPT], rb[LPT];
#pragma unroll
  for (int k = 0; k < LPT; ++k) {
    const uint32_t j = threadIdx.x + k * NT;
    // s_j from slots j+3, j+2, j+1; s_{j-1} from j+2, j+1, j (its high part)
    const uint64_t a = ws0[j + 3], b = ws1[j + 2], c = ws2[j + 1];
    const uint64_t s1 = a + b;
    const uint64_t lo = s1 + c;
    const uint64_t ap = ws0[j + 2], bp = ws1[j + 1], cp = ws2[j];
    const uint64_t q1 = ap + bp;
    const uint32_t hp = (q1 < ap ? 1u : 0u) + (q1 + cp < q1 ? 1u : 0u);
    const uint64_t tt = lo + hp;
    bool g = tt < lo, pr = tt == ~(uint64_t)0;
    if (B0 + j >= Z) g = false, pr = true;        // past the product: carry passes
    t[k] = tt;
    const uint64_t G = __ballot(g), P = __ballot(pr);
    ra[k] = G | P;
    rb[k] = G;
    if (lane == 0) {
      const uint64_t sm = ra[k] + rb[k];
      const uint32_t c0 = sm < ra[k] ? 1u : 0u;
      const uint32_t c1 = c0 | (sm == ~(uint64_t)0 ? 1u : 0u);
      rmap[k * (NT / 64) + wid] = c0 | (c1 << 1);   // runs in limb order
    }
  }
  __syncthreads();
  // carry into each run (block carry-in 0), and the block's map
  if (wid == 0) {
    const uint32_t f = lane < RUNS ? rmap[lane] : C2_ID;
    uint32_t inc = f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(inc, d, 64);
      if (lane >= d) inc = c2_then(o, inc);
    }
    uint32_t exc = __shfl_up(inc, 1, 64);
    if (lane == 0) exc = C2_ID;
    if (lane < RUNS) rcin[lane] = c2_apply(exc, 0);
    if (lane == 63 && bmap) *bmap = inc;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < LPT; ++k) {
    const uint32_t j = threadIdx.x + k * NT;
    const uint64_t cin = rcin[k * (NT / 64) + wid];
    const uint64_t C = (ra[k] + rb[k] + cin) ^ ra[k] ^ rb[k];
    if (B0 + j < Z) rp[B0 + j] = t[k] + ((C >> lane) & 1u);
  }
}

__global__ __launch_bounds__(NT) void k_garner_limbs(GarnerArgs ga, JobRef jr, uint32_t N) {
  // the high word of a coefficient (< 2^155) is below 2^27
  __shared__ uint64_t ws0[CARRY_LPB + 3], ws1[CARRY_LPB + 3];
  __shared__ uint32_t ws2[CARRY_LPB + 3];
  const Job jb = get_job(jr);
  if (blockIdx.x >= jb.nb) return;                 // this job has fewer carry blocks
  gc_u32 *spec = (gc_u32 *)jb.bspec;
  const uint32_t zlen = jb.zlen, Z = jb.Z;
  const uint32_t B0 = blockIdx.x * CARRY_LPB;
#pragma unroll SBA_GARNER_UNROLL
  for (int k = 0; k < CARRY_LPT; ++k) {
    const uint32_t j = threadIdx.x + k * NT;
    const uint32_t i = B0 + j;
    uint64_t c0 = 0, c1 = 0, c2 = 0;
    if (i < zlen) garner_coef(ga, spec, N, i, c0, c1, c2);
    ws0[j + 3] = c0; ws1[j + 3] = c1; ws2[j + 3] = (uint32_t)c2;
  }
  if (threadIdx.x < 3) {
    const uint32_t j = threadIdx.x;            // coefficient B0 - 3 + j
    uint64_t c0 = 0, c1 = 0, c2 = 0;
    if (B0 + j >= 3) {
      const uint32_t i = B0 + j - 3;
      if (i < zlen) garner_coef(ga, spec, N, i, c0, c1, c2);
    }
    ws0[j] = c0; ws1[j] = c1; ws2[j] = (uint32_t)c2;
  }
  __syncthreads();
  limbs_from_words<CARRY_LPT, uint32_t>(ws0, ws1, ws2, B0, Z, (g_u64 *)jb.r,
                                        (g_u32 *)jb.bmaps + blockIdx.x);
}

// ---------------------------------------------------------------------------
// small products: schoolbook on one workgroup per product (an + bn <= 2048
// limbs, one carry block). Column k's sum of a_i b_(k-i) is kept as three
// words (< 2^(128 + 11)), which is exactly the coefficient form the Garner
// stage produces, so the same limb finishing applies. Grid x = job.
// ---------------------------------------------------------------------------
#define SCHOOL_LPT 8
#define SCHOOL_ZMAX (SCHOOL_LPT * NT)         // one workgroup's limbs
#define SCHOOL_COLS SCHOOL_LPT                // columns per thread
#ifndef SBA_SCHOOL_PMAX_DEFAULT
#define SBA_SCHOOL_PMAX_DEFAULT 32768         // schoolbook limit on an * bn at N = 1024
#endif

__global__ __launch_bounds__(NT) void k_school(JobRef jr) {
  // operands first (a at [0, an), b at [an, an + bn)), then the column words
  __shared__ uint64_t sm[3 * (SCHOOL_ZMAX + 3)];
  const Job jb = get_job(jr, blockIdx.x);
  const uint32_t an = jb.an, bn = jb.bn, Z = jb.Z;
  gc_u64 *ga_ = (gc_u64 *)jb.a;
  gc_u64 *gb_ = (gc_u64 *)jb.b;
  for (uint32_t i = threadIdx.x; i < an; i += NT) sm[i] = ga_[i];
  for (uint32_t i = threadIdx.x; i < bn; i += NT) sm[an + i] = gb_[i];
  __syncthreads();
  const uint64_t *sa = sm, *sb = sm + an;
  uint64_t w0[SCHOOL_COLS], w1[SCHOOL_COLS], w2[SCHOOL_COLS];
#pragma unroll
  for (int m = 0; m < SCHOOL_COLS; ++m) {
    const uint32_t k = threadIdx.x + m * NT;       // column (coefficient) k < Z - 1
    uint64_t lo = 0, mid = 0, hi = 0;
    if (k + 1 < Z) {
      const uint32_t i0 = k + 1 > bn ? k + 1 - bn : 0;
      const uint32_t i1 = k < an - 1 ? k : an - 1;
      for (uint32_t i = i0; i <= i1; ++i) {
        const uint64_t x = sa[i], y = sb[k - i];
        const uint64_t pl = x * y, ph = __umul64hi(x, y);
        lo += pl;
        const uint64_t c1 = lo < pl ? 1u : 0u;
        const uint64_t m1 = mid + ph;
        hi += m1 < mid ? 1u : 0u;
        mid = m1 + c1;
        hi += mid < c1 ? 1u : 0u;
      }
    }
    w0[m] = lo; w1[m] = mid; w2[m] = hi;
  }
  __syncthreads();                                 // the operands are no longer read
  uint64_t *ws0 = sm, *ws1 = sm + (SCHOOL_ZMAX + 3), *ws2 = sm + 2 * (SCHOOL_ZMAX + 3);
#pragma unroll
  for (int m = 0; m < SCHOOL_COLS; ++m) {
    const uint32_t j = threadIdx.x + m * NT;
    ws0[j + 3] = w0[m]; ws1[j + 3] = w1[m]; ws2[j + 3] = w2[m];
  }
  if (threadIdx.x < 3) ws0[threadIdx.x] = ws1[threadIdx.x] = ws2[threadIdx.x] = 0;
  __syncthreads();
  limbs_from_words<SCHOOL_LPT, uint64_t>(ws0, ws1, ws2, 0, Z, (g_u64 *)jb.r, nullptr);
}

// products with an + bn <= SCHOOL_ZMAX take the schoolbook kernel when
// an * bn <= g_school_pmax * N / 1024 in a batch (N = the NTT length they
// would get, >= 1024: the NTT's cost grows with N, the schoolbook's with
// an * bn), half that for a single call (where the NTT's few launches of
// latency dominate). Measured on MI355X: 512-job batches and single calls of
// 16x16 .. 1024x1024 and 16x2032 .. 256x1792 limbs. sba_set_school_max sets
// g_school_pmax; 0 turns the tier off.
static std::atomic<uint64_t> g_school_pmax{SBA_SCHOOL_PMAX_DEFAULT};
static bool school_job(uint64_t an, uint64_t bn, bool single) {
  if (an + bn > SCHOOL_ZMAX) return false;
  uint64_t N = 1024;
  while (N < an + bn - 1) N <<= 1;
  const uint64_t pmax = g_school_pmax.load(std::memory_order_relaxed);
  const uint64_t lim = (single ? pmax / 2 : pmax) * (N / 1024);
  return an * bn <= lim;
}

// ---------------------------------------------------------------------------
// short x long products: schoolbook with the short operand b (bn <=
// SCHOOL_SHORT_MAX limbs) against a long a, one workgroup per block of
// CARRY_LPB product limbs (grid x = block, z = job). A workgroup keeps b and
// the window of a its columns read in LDS; each thread sums 4 consecutive
// columns, register-blocked: per b limb one new a limb enters a 4-limb
// sliding window and 4 multiply-adds follow (about 2 LDS reads per 4
// products). Column sums are the Garner stage's 3-word coefficients (<
// 2^128 bn), so the same limb finishing (block carry-in 0, 2-state block
// map) and the carry scan between blocks complete the product. Used where
// an * bn is small against the NTT's N log N (very uneven shapes).
// ---------------------------------------------------------------------------
#define SCHOOL_SHORT_MAX 512
#define SCHOOL_LONG_W (CARRY_LPB + 3 + SCHOOL_SHORT_MAX)

__device__ __forceinline__ void mac128(uint64_t &lo, uint64_t &mid, uint64_t &hi, uint64_t x,
                                       uint64_t y) {
  const uint64_t pl = x * y, ph = __umul64hi(x, y);
  lo += pl;
  const uint64_t c1 = lo < pl ? 1u : 0u;
  const uint64_t m1 = mid + ph;
  hi += m1 < mid ? 1u : 0u;
  mid = m1 + c1;
  hi += mid < c1 ? 1u : 0u;
}

__global__ __launch_bounds__(NT) void k_school_long(JobRef jr) {
  __shared__ uint64_t sb[SCHOOL_SHORT_MAX];
  __shared__ uint64_t sa[SCHOOL_LONG_W];       // a[abase + x], zero outside [0, an)
  __shared__ uint64_t ws0[CARRY_LPB + 3], ws1[CARRY_LPB + 3];
  __shared__ uint32_t ws2[CARRY_LPB + 3];
  const Job jb = get_job(jr);
  if (blockIdx.x >= jb.nb) return;               // this job has fewer blocks
  const uint32_t an = jb.an, bn = jb.bn, Z = jb.Z;
  gc_u64 *ga = (gc_u64 *)jb.a;
  gc_u64 *gb = (gc_u64 *)jb.b;
  const uint32_t tid = threadIdx.x;
  const uint32_t B0 = blockIdx.x * CARRY_LPB;
  // column c = B0 - 3 + x' reads a[c - j], j < bn: window from B0 - 3 - (bn - 1)
  const int64_t abase = (int64_t)B0 - 3 - (int64_t)(bn - 1);
  const uint32_t W = CARRY_LPB + 3 + bn - 1;
  for (uint32_t x = tid; x < W; x += NT) {
    const int64_t i = abase + (int64_t)x;
    sa[x] = (i >= 0 && i < (int64_t)an) ? ga[i] : 0;
  }
  for (uint32_t x = tid; x < bn; x += NT) sb[x] = gb[x];
  __syncthreads();
  {
    // columns B0 + 4 tid + u (u < 4); a[B0 + 4 tid + u - j] sits at window
    // index x0 + u - j
    const uint32_t k0 = 4 * tid;
    const uint32_t x0 = k0 + 3 + bn - 1;
    uint64_t lo[4] = {0, 0, 0, 0}, mid[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
    uint64_t w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = sa[x0 + u];
    for (uint32_t j = 0; j < bn; ++j) {
      const uint64_t y = sb[j];
#pragma unroll
      for (int u = 0; u < 4; ++u) mac128(lo[u], mid[u], hi[u], w[u], y);
      w[3] = w[2];
      w[2] = w[1];
      w[1] = w[0];
      w[0] = sa[x0 - j - 1];                      // >= index k0 + 2 while j < bn
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ws0[k0 + u + 3] = lo[u];
      ws1[k0 + u + 3] = mid[u];
      ws2[k0 + u + 3] = (uint32_t)hi[u];
    }
  }
  if (tid < 3) {
    // columns B0 - 3 + tid (zero below column 0: the window holds zeros there)
    uint64_t lo = 0, mid = 0, hi = 0;
    for (uint32_t j = 0; j < bn; ++j) mac128(lo, mid, hi, sa[tid + bn - 1 - j], sb[j]);
    ws0[tid] = lo;
    ws1[tid] = mid;
    ws2[tid] = (uint32_t)hi;
  }
  __syncthreads();
  limbs_from_words<CARRY_LPT, uint32_t>(ws0, ws1, ws2, B0, Z, (g_u64 *)jb.r,
                                        (g_u32 *)jb.bmaps + blockIdx.x);
}

// short x long products take k_school_long when b is short and its
// multiply-adds are cheap against the NTT the product would get:
// an * bn <= g_long_ratio / 16 * N * log2 N (default measured on MI355X).
#ifndef SBA_SCHOOL_LONG_RATIO
#define SBA_SCHOOL_LONG_RATIO 48               // sixteenths
#endif
static std::atomic<uint32_t> g_long_ratio{SBA_SCHOOL_LONG_RATIO};
static bool school_long_job(uint64_t an, uint64_t bn) {
  if (an < bn) { const uint64_t t = an; an = bn; bn = t; }
  if (bn > SCHOOL_SHORT_MAX || an + bn <= SCHOOL_ZMAX) return false;
  const uint64_t zlen = an + bn - 1;
  int L = 10;
  while (((uint64_t)1 << L) < zlen) ++L;
  const uint64_t ratio = g_long_ratio.load(std::memory_order_relaxed);
  return an * bn * 16 <= ratio * ((uint64_t)L << L);
}
// the workspace of a short x long product: its carry maps
static size_t ws_school_long(uint64_t Z) { return ((Z / CARRY_LPB + 2) * 4 + 255) & ~(size_t)255; }

// workgroup-wide exclusive scan of 2-state maps in thread order; *total =
// the composition of all of them
__device__ uint32_t c2_block_scan(uint32_t f, uint32_t *total) {
  __shared__ uint32_t wt[SCAN_NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = f;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc = c2_then(o, inc);
  }
  uint32_t exc = __shfl_up(inc, 1, 64);
  if (lane == 0) exc = C2_ID;
  if (lane == 63) wt[wid] = inc;
  __syncthreads();
  uint32_t pre = C2_ID, tot = C2_ID;
  for (int w = 0; w < SCAN_NT / 64; ++w) {
    if (w < wid) pre = c2_then(pre, wt[w]);
    tot = c2_then(tot, wt[w]);
  }
  *total = tot;
  __syncthreads();
  return c2_then(pre, exc);
}

// carries between blocks (grid x = up to CARRY_WGS workgroups per product,
// grid z = job). Workgroup g owns a contiguous range of blocks: it composes
// the maps of all earlier blocks for the carry into its range, scans its
// range's maps, and adds each carry-in of 1 to its block's limbs (computed
// with carry-in 0) -- usually to the first limb only; a block whose first
// limb is all ones gets the whole workgroup (the limbs below the first one
// that is not all ones wrap to zero)
__global__ __launch_bounds__(SCAN_NT) void k_carry_blocks(JobRef jr) {
  __shared__ uint32_t list[SCAN_NT];
  __shared__ uint32_t first, nlist;
  const Job jb = get_job(jr);
  const uint32_t nb = jb.nb, Z = jb.Z;
  const uint32_t span = (nb + gridDim.x - 1) / gridDim.x;
  const uint32_t r0 = blockIdx.x * span;
  if (r0 >= nb) return;                              // uniform
  const uint32_t r1 = r0 + span < nb ? r0 + span : nb;
  gc_u32 *bm = (gc_u32 *)jb.bmaps;
  g_u64 *rp = (g_u64 *)jb.r;
  const uint32_t per = (r0 + SCAN_NT - 1) / SCAN_NT;
  const uint32_t s0 = threadIdx.x * per, s1 = s0 + per < r0 ? s0 + per : r0;
  uint32_t f = C2_ID;
  for (uint32_t i = s0; i < s1; ++i) f = c2_then(f, bm[i]);
  uint32_t tot;
  c2_block_scan(f, &tot);
  uint32_t c = c2_apply(tot, 0);                     // carry into block r0
  for (uint32_t base = r0; base < r1; base += SCAN_NT) {
    const uint32_t i = base + threadIdx.x;
    uint32_t ctot;
    const uint32_t exc = c2_block_scan(i < r1 ? bm[i] : C2_ID, &ctot);
    const uint32_t cin = c2_apply(exc, c);
    c = c2_apply(ctot, c);
    if (threadIdx.x == 0) nlist = 0;
    __syncthreads();
    if (i < r1 && cin) {
      g_u64 *x = rp + (size_t)i * CARRY_LPB;
      const uint64_t v = *x;
      if (v != ~(uint64_t)0) *x = v + 1;
      else list[atomicAdd(&nlist, 1u)] = i;
    }
    __syncthreads();
    const uint32_t nl = nlist;
    for (uint32_t e = 0; e < nl; ++e) {
      const uint32_t B0 = list[e] * CARRY_LPB;
      const uint32_t E = Z - B0 < CARRY_LPB ? Z - B0 : CARRY_LPB;
      if (threadIdx.x == 0) first = E;
      __syncthreads();
      for (uint32_t j = threadIdx.x; j < E; j += SCAN_NT)
        if (rp[B0 + j] != ~(uint64_t)0) {
          atomicMin(&first, j);
          break;
        }
      __syncthreads();
      const uint32_t fz = first;
      for (uint32_t j = threadIdx.x; j < E && j <= fz; j += SCAN_NT) {
        if (j < fz) rp[B0 + j] = 0;
        else rp[B0 + j] += 1;
      }
      __syncthreads();
    }
    __syncthreads();                                 // list and nlist are free again
  }
}

// ---------------------------------------------------------------------------
// twiddle tables: wn[j] = prod over set bits t of j of zeta_{t+2} (Montgomery
// form), i.e. w[2j] of the FLINT table; wni the same with inverse roots.
// ---------------------------------------------------------------------------
struct TwArgs {
  uint32_t p, pinv, one;          // one = R mod p
  uint32_t z[SBA_MAX_LG], zi[SBA_MAX_LG];
};

__global__ __launch_bounds__(NT) void k_twiddles(TwArgs ta, uint32_t *__restrict__ wn,
                                                 uint32_t *__restrict__ wni, uint32_t count) {
  const uint32_t j = blockIdx.x * NT + threadIdx.x;
  if (j >= count) return;
  uint32_t a = ta.one, b = ta.one;
  for (int t = 0; t < SBA_MAX_LG; ++t)
    if ((j >> t) & 1u) {
      a = red1(mont_mul(a, ta.z[t], ta.p, ta.pinv), ta.p);
      b = red1(mont_mul(b, ta.zi[t], ta.p, ta.pinv), ta.p);
    }
  wn[j] = a;
  wni[j] = b;
}

// ===========================================================================
// host side of the thin layer
// ===========================================================================
// Prime sets, each five primes in ascending order.
//   SET_W (wide): k 2^25 + 1 < 2^31 with 3 | p - 1 -- every transform length
//     up to 2^25, radix-3 lengths, CRT capacity 2^153.4.
//   SET_H: three primes below 2^30 (Harvey's 8-instruction butterflies) and
//     two below 2^31, all with 2^23 | p - 1 (none of the first two with
//     3 | p - 1) -- power-of-two lengths up to 2^23, CRT capacity 2^151.3,
//     which holds the largest coefficient such a product has,
//     min(an, bn) (2^64 - 1)^2 < 2^22 * 2^128 = 2^150.
//   SET_R: the same shape for the radix-3 lengths 3 * 2^L, L <= 22: three
//     primes below 2^30 and two below 2^31, all with 3 * 2^22 | p - 1, CRT
//     capacity 2^151.2 against min(an, bn) (2^64 - 1)^2 < 3 * 2^21 * 2^128
//     = 2^150.6. (No set of three such primes below 2^30 also has
//     2^23 | p - 1 and enough capacity, hence a set of its own.)
// The merged last pass (k_fused2, k_fused) runs Harvey's forms for the
// three small primes: its butterfly work is (3 * 8 + 2 * 10) / (5 * 10) =
// 0.88 of the set-W pass; the column passes (power-of-two and radix-3)
// choose the form chunk by chunk.
#define SET_W 0
#define SET_H 1
#define SET_R 2
#define NSETS 3
struct PrimeSetDef {
  uint32_t p[SBA_K];
  int max_lg;                          // largest power-of-two transform (row)
};
static const PrimeSetDef kSets[NSETS] = {
    {{1107296257u, 1711276033u, 1811939329u, 2013265921u, 2113929217u}, SBA_MAX_LG},
    {{880803841u, 897581057u, 998244353u, 2113929217u, 2130706433u}, 23},
    {{880803841u, 918552577u, 943718401u, 2088763393u, 2113929217u}, 22},
};
static const uint32_t *const kPrimes = kSets[SET_W].p;   // the band's (CRT cap) set

static uint64_t h_powm(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1;
  a %= p;
  for (; e; e >>= 1, a = a * a % p)
    if (e & 1) r = r * a % p;
  return r;
}
static uint64_t h_invm(uint64_t a, uint64_t p) { return h_powm(a, p - 2, p); }
static uint32_t h_mont(uint64_t a, uint64_t p) { return (uint32_t)(((a % p) << 32) % p); }
static uint64_t h_primroot(uint64_t p) {
  uint64_t f[16];
  int nf = 0;
  uint64_t n = p - 1;
  for (uint64_t d = 2; d * d <= n; d += (d == 2 ? 1 : 2))
    if (n % d == 0) {
      f[nf++] = d;
      while (n % d == 0) n /= d;
    }
  if (n > 1) f[nf++] = n;
  for (uint64_t g = 2;; ++g) {
    int ok = 1;
    for (int i = 0; i < nf && ok; ++i)
      if (h_powm(g, (p - 1) / f[i], p) == 1) ok = 0;
    if (ok) return g;
  }
}

// one prime set's constants and root tables on a device
struct SetCtx {
  int tw_lg = 0;                       // tables hold 2^(tw_lg-1) node twiddles
  uint32_t *wn[SBA_K] = {}, *wni[SBA_K] = {};
  PrimeArgs pa;
  GarnerArgs ga_base;
};

struct DevCtx {
  int init = 0;
  SetCtx set[NSETS];
  // grow-only workspace
  void *ws = nullptr;
  size_t ws_bytes = 0;
  // the allocation-free entry: this device's copy of the static scratch
  // array (its address from hipGetSymbolAddress) and whether its root
  // tables are built
  char *na_scratch = nullptr;
  int na_tables = 0;
};

#define SBA_MAX_DEV 16
static DevCtx g_ctx[SBA_MAX_DEV];
static std::mutex g_mu;

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      fprintf(stderr, "simd_bigint_amd: %s failed: %s (%s:%d)\n", #x,                 \
              hipGetErrorString(e_), __FILE__, __LINE__);                             \
      return SBA_ERR_HIP;                                                             \
    }                                                                                 \
  } while (0)

static void set_init(SetCtx &c, const PrimeSetDef &d) {
  const uint32_t *P = d.p;
  for (int q = 0; q < SBA_K; ++q) {
    const uint64_t p = P[q];
    uint32_t inv = 1;
    for (int i = 0; i < 5; ++i) inv *= 2u - (uint32_t)p * inv;   // p^-1 mod 2^32
    c.pa.p[q] = (uint32_t)p;
    c.pa.pinv[q] = 0u - inv;
    c.pa.r1[q] = (uint32_t)(((uint64_t)1 << 32) % p);
    c.ga_base.p[q] = (uint32_t)p;
    c.ga_base.pinv[q] = 0u - inv;
  }
  for (int q = 0; q < SBA_K; ++q) {
    // length-3 DFT constants: w = g^((p-1)/3) a cube root of unity (only
    // where 3 | p - 1: a set without them never runs a radix-3 class)
    const uint64_t p = P[q];
    const uint64_t inv2 = (p + 1) / 2;
    c.pa.inv2[q] = h_mont(inv2, p);
    if ((p - 1) % 3 == 0) {
      const uint64_t w = h_powm(h_primroot(p), (p - 1) / 3, p);
      const uint64_t h2 = (w + p - w * w % p) % p * inv2 % p;
      c.pa.h1[q] = h_mont(p - inv2, p);
      c.pa.h2[q] = h_mont(h2, p);
      c.pa.h2i[q] = h_mont((p - h2) % p, p);
    } else {
      c.pa.h1[q] = c.pa.h2[q] = c.pa.h2i[q] = 0;
    }
  }
  for (int q = 1; q < SBA_K; ++q) {
    const uint64_t pq = P[q];
    uint64_t pp = 1;
    for (int k = 0; k < q; ++k) {
      c.ga_base.cm[q][k] = h_mont(pp, pq);
      pp = pp * (P[k] % pq) % pq;
    }
    c.ga_base.im[q] = h_mont(h_invm(pp, pq), pq);
  }
}

static int ctx_init(DevCtx &c) {
  if (c.init) return 0;
  for (int s = 0; s < NSETS; ++s) set_init(c.set[s], kSets[s]);
  c.init = 1;
  return 0;
}

// the root-table kernel's arguments for prime q of a set
static TwArgs tw_args(const SetCtx &c, const PrimeSetDef &d, int q) {
  const uint64_t p = d.p[q];
  const uint64_t g = h_primroot(p);
  TwArgs ta;
  ta.p = (uint32_t)p;
  ta.pinv = c.pa.pinv[q];
  ta.one = h_mont(1, p);
  for (int t = 0; t < SBA_MAX_LG; ++t) {
    // bit t of a node index carries zeta_{t+2}, a root of order 2^(t+2)
    const int ord = t + 2;
    uint64_t z = ord <= d.max_lg ? h_powm(g, (p - 1) >> ord, p) : 1;
    ta.z[t] = h_mont(z, p);
    ta.zi[t] = h_mont(h_invm(z, p), p);
  }
  return ta;
}

// the prime sets with primes below 2^30 (sets H and R) are used unless
// switched off (sba_set_prime_set_h)
static std::atomic<int> g_use_set_h{1};

// one set's root tables, built once for its largest transform (SET_W: 2^24
// node twiddles per table, 640 MiB; SET_H: 2^22, 160 MiB; SET_R: 2^21,
// 80 MiB); tw_lg records the length they cover once every table is in place
static int ctx_set_tables(DevCtx &c, int s, hipStream_t st) {
  SetCtx &sc = c.set[s];
  const int Lt = kSets[s].max_lg;
  if (sc.tw_lg >= Lt) return 0;
  const uint32_t count = 1u << (Lt - 1);
  for (int q = 0; q < SBA_K; ++q) {
    if (sc.wn[q]) { HIPCHK(hipFree(sc.wn[q])); HIPCHK(hipFree(sc.wni[q])); }
    sc.wn[q] = sc.wni[q] = nullptr;
    HIPCHK(hipMalloc(&sc.wn[q], (size_t)count * 4));
    HIPCHK(hipMalloc(&sc.wni[q], (size_t)count * 4));
    const TwArgs ta = tw_args(sc, kSets[s], q);
    hipLaunchKernelGGL(k_twiddles, dim3((count + NT - 1) / NT), dim3(NT), 0, st, ta,
                       sc.wn[q], sc.wni[q], count);
    HIPCHK(hipGetLastError());
    sc.pa.wn[q] = sc.wn[q];
    sc.pa.wni[q] = sc.wni[q];
  }
  HIPCHK(hipStreamSynchronize(st));
  sc.tw_lg = Lt;
  return 0;
}

// the tables the transforms of a call (power-of-two lengths up to 2^L, or
// radix-3 rows) may use: sets H and R on the first call (240 MiB), set W
// only once a transform longer than set H's can run or the small-prime
// sets are switched off
static int ctx_twiddles(DevCtx &c, int L, hipStream_t st) {
  int rc = ctx_set_tables(c, SET_H, st);
  if (!rc) rc = ctx_set_tables(c, SET_R, st);
  if (!rc && (L > kSets[SET_H].max_lg || !g_use_set_h.load(std::memory_order_relaxed)))
    rc = ctx_set_tables(c, SET_W, st);
  return rc;
}

// a launch on set `set` for rows of 2^L points needs its tables in place;
// they are missing only if the set switch changed between a call's set-up
// and its launches (another thread): refused, nothing launched. 0 when
// ready (not SBA_OK, which is 1), else the error to return
static int set_ready(const DevCtx &c, int set, int L) {
  if (c.set[set].tw_lg >= L) return 0;
  fprintf(stderr, "simd_bigint_amd: prime set %d has no root tables for 2^%d points (the set "
          "switch changed during the call)\n", set, L);
  return SBA_ERR_HIP;
}

// the prime set of a transform (L: the row length's log for radix 3): the
// sets with primes below 2^30 where they hold the length (unless switched
// off), SET_W otherwise
extern "C" void sba_set_prime_set_h(int on) { g_use_set_h.store(on ? 1 : 0); }
extern "C" int sba_get_prime_set_h(void) { return g_use_set_h.load(); }
static int set_for(int L, int r3) {
  if (!g_use_set_h.load(std::memory_order_relaxed)) return SET_W;
  if (!r3 && L <= kSets[SET_H].max_lg) return SET_H;
  if (r3 && L <= kSets[SET_R].max_lg) return SET_R;
  return SET_W;
}

// workspace layout for transform length N: aspec, bspec (K*N u32 each),
// then one carry map (u32) per block of CARRY_LPB limbs
static size_t ws_need(size_t N) {
  const size_t nb = N / CARRY_LPB + 2;
  return 2 * SBA_K * N * 4 + nb * 4 + 4096;
}

static int lg_for(uint64_t zlen) {
  int L = 10;                           // smallest transform the kernels take
  while (((uint64_t)1 << L) < zlen) ++L;
  return L;
}


// CRT cap: largest min(an, bn) with min * (2^64-1)^2 < prod(p)
extern "C" uint64_t sba_gpu_min_operand_cap(void) {
  // prod(p) as 3 u64 words, then bisection on q * (2^64-1)^2 <= prod(p)
  uint64_t w[3] = {1, 0, 0};
  for (int i = 0; i < SBA_K; ++i) {
    unsigned __int128 cc = 0;
    for (int k = 0; k < 3; ++k) {
      cc += (unsigned __int128)w[k] * kPrimes[i];
      w[k] = (uint64_t)cc;
      cc >>= 64;
    }
  }
  uint64_t lo = 0, hi = (uint64_t)1 << 62;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo + 1) / 2;
    unsigned __int128 cc = (unsigned __int128)mid * 1u;
    uint64_t t0 = (uint64_t)cc;
    cc = (unsigned __int128)mid * ~(uint64_t)1 + (uint64_t)(cc >> 64);
    uint64_t t1 = (uint64_t)cc, t2 = (uint64_t)(cc >> 64);
    int le;
    if (t2 != w[2]) le = t2 < w[2];
    else if (t1 != w[1]) le = t1 < w[1];
    else le = t0 <= w[0];
    if (le) lo = mid; else hi = mid - 1;
  }
  return lo;    // lo * F < prod strictly: F = (2^64-1)^2 shares no factor with prod
}

extern "C" int sba_gpu_max_lg(void) { return SBA_MAX_LG; }

// ---------------------------------------------------------------------------
// launch helpers (template dispatch on the pass shape)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// plans: how a transform of length N = 2^L is cut into passes
//   last   levels of the last pass, on contiguous tiles of 2^last residues.
//          last == L: the whole transform is one tile (L <= 13). Otherwise a
//          first pass over column tiles of 2^13 residues (2^(L-last) rows x
//          2^(13-L+last) columns) takes the other L - last levels.
//   bm     widest register chunk inside a tile: radix-2^bm stages (3..5)
//   split  first forward pass with one prime per block (1), or all K primes
//          per block reusing the tile's limb loads (0)
// ---------------------------------------------------------------------------
static bool plan_ok(int L, const Plan &p) {
  if (p.bm < 3 || p.bm > 5 || (p.split != 0 && p.split != 1)) return false;
  if (p.last == L) return L >= 10 && L <= 13;
  if (p.last < 11 || p.last > 13 || L < 13) return false;
  const int ell = L - p.last;
  return ell >= 1 && ell <= 12;
}
// the hand-written rule: widest last pass, radix 32, one prime per block
static Plan heuristic_plan(int L) {
  Plan p;
  p.last = L <= 13 ? L : 13;
  p.bm = SBA_BMAX;
  p.split = SBA_FIRST_SPLIT;
  return p;
}

static Launchers launchers(int L, int sqr, const Plan &p) {
  switch (p.bm) {
    case 3: return sba_launchers_bm3(L, sqr, p.last);
    case 4: return sba_launchers_bm4(L, sqr, p.last);
    default: return sba_launchers_bm5(L, sqr, p.last);
  }
}

// plan table: (sqr, L, bucket of the class's job count) -> plan. Starts as
// the shipped table (plan_table.h: the fastest plan per key measured on
// MI355X; the rule where a key was not measured) until the autotuner or a
// caller stores another choice.
static Plan g_plan[2][SBA_MAX_LG + 1][SBA_PLAN_BUCKETS];
static Plan g_plan3[2][SBA_MAX_LG + 1][SBA_PLAN_BUCKETS];   // radix-3 classes (row length 2^L)
static int g_plan_init = 0;
static Plan r3_fix(int L, Plan p);
static bool plan_ok3(int L, const Plan &p);
// the table's entry of (kind, L, bucket), kind = square? | radix-3? << 1;
// {0, 0, 0} (not measured) or an entry the kind does not accept: none
static bool table_plan(int kind, int L, int k, Plan *out) {
  if (L < 0 || L >= SBA_PT_NL || k >= SBA_PLAN_BUCKETS) return false;
  const signed char *e = kPlanTable[kind][L][k];
  const Plan p = {e[0], e[1], e[2]};
  if (e[0] <= 0 || !((kind & 2) ? plan_ok3(L, p) : plan_ok(L, p))) return false;
  *out = p;
  return true;
}
static void plan_defaults(void) {
  for (int sq = 0; sq < 2; ++sq)
    for (int L = 0; L <= SBA_MAX_LG; ++L)
      for (int k = 0; k < SBA_PLAN_BUCKETS; ++k) {
        Plan p;
        g_plan[sq][L][k] = table_plan(sq, L, k, &p) ? p : heuristic_plan(L);
        // radix-3 classes: the measured entry, else the power-of-two plan
        // moved into the radix-3 limits
        g_plan3[sq][L][k] = table_plan(2 | sq, L, k, &p) ? p : r3_fix(L, g_plan[sq][L][k]);
      }
  g_plan_init = 1;
}
static int plan_bucket(uint32_t njobs) {
  int b = 0;
  while (njobs >= 4 && b < SBA_PLAN_BUCKETS - 1) {
    njobs >>= 2;
    ++b;
  }
  return b;
}
static Plan plan_for(int L, int sqr, uint32_t njobs) {
  if (!g_plan_init) plan_defaults();
  if (L < 0 || L > SBA_MAX_LG) return heuristic_plan(L);
  return g_plan[sqr ? 1 : 0][L][plan_bucket(njobs)];
}
static Plan plan_for3(int L, int sqr, uint32_t njobs) {
  if (!g_plan_init) plan_defaults();
  if (L < 0 || L > SBA_MAX_LG) return r3_fix(L, heuristic_plan(L));
  return g_plan3[sqr ? 1 : 0][L][plan_bucket(njobs)];
}

// ---------------------------------------------------------------------------
// optional per-kernel timing: HIP events recorded on the launch stream
// around each kernel, with the kernel's algorithmic bytes per launch
// ---------------------------------------------------------------------------
#define PROF_MAX_REC 16384
#define PROF_MAX_NAMES 32
struct ProfRec { int name; double bytes; hipEvent_t e0, e1; };
static int g_prof_on = 0;
static ProfRec g_prof_rec[PROF_MAX_REC];
static int g_prof_n = 0;
static hipEvent_t g_prof_pool[2 * PROF_MAX_REC];
static int g_prof_pool_n = 0;
static const char *g_prof_names[PROF_MAX_NAMES];
static int g_prof_nnames = 0;
static double g_prof_ms[PROF_MAX_NAMES], g_prof_bytes[PROF_MAX_NAMES];
static uint64_t g_prof_cnt[PROF_MAX_NAMES];

static int prof_name(const char *n) {
  for (int i = 0; i < g_prof_nnames; ++i)
    if (!strcmp(g_prof_names[i], n)) return i;
  if (g_prof_nnames == PROF_MAX_NAMES) return PROF_MAX_NAMES - 1;
  g_prof_names[g_prof_nnames] = n;
  return g_prof_nnames++;
}
static hipEvent_t prof_event(void) {
  const int i = g_prof_pool_n++;
  if (!g_prof_pool[i]) (void)hipEventCreate(&g_prof_pool[i]);
  return g_prof_pool[i];
}
// A PROF scope names the kernel launched next in it; with profiling on,
// that launch goes through hipExtLaunchKernelGGL with the record's start and
// stop events, which time the kernel itself without extra event packets
// between kernels (back-to-back launches keep their spacing). Launches
// outside a PROF scope, or with profiling off, are plain launches.
static int g_prof_pending = -1;                 // record awaiting its launch (under g_mu)
struct ProfScope {
  ProfScope(const char *name, double bytes) {
    if (!g_prof_on || g_prof_n >= PROF_MAX_REC) return;
    const int rec = g_prof_n++;
    ProfRec &r = g_prof_rec[rec];
    r.name = prof_name(name);
    r.bytes = bytes;
    r.e0 = prof_event();
    r.e1 = prof_event();
    g_prof_pending = rec;
  }
  ~ProfScope() { g_prof_pending = -1; }
};
#define PROF(name, bytes) ProfScope prof_scope_##__LINE__(name, (double)(bytes))
#define KLAUNCH(K, G, B, SH, ST, ...)                                                        \
  do {                                                                                      \
    if (g_prof_pending >= 0) {                                                              \
      const ProfRec &r_ = g_prof_rec[g_prof_pending];                                       \
      g_prof_pending = -1;                                                                  \
      hipExtLaunchKernelGGL(K, G, B, SH, ST, r_.e0, r_.e1, 0, __VA_ARGS__);                 \
    } else {                                                                                \
      hipLaunchKernelGGL(K, G, B, SH, ST, __VA_ARGS__);                                     \
    }                                                                                       \
  } while (0)

extern "C" void sba_gpu_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_prof_on = on;
}
// fold recorded events (the caller has synchronized) into per-kernel totals
extern "C" int sba_gpu_profile_collect(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int i = 0; i < g_prof_n; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_prof_rec[i].e0, g_prof_rec[i].e1) != hipSuccess) continue;
    const int k = g_prof_rec[i].name;
    g_prof_ms[k] += ms;
    g_prof_bytes[k] += g_prof_rec[i].bytes;
    g_prof_cnt[k] += 1;
  }
  const int n = g_prof_n;
  g_prof_n = 0;
  g_prof_pool_n = 0;
  return n;
}
extern "C" void sba_gpu_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int k = 0; k < PROF_MAX_NAMES; ++k) g_prof_ms[k] = g_prof_bytes[k] = 0, g_prof_cnt[k] = 0;
  g_prof_n = 0;
  g_prof_pool_n = 0;
}
// kernel idx: name, total ms, total algorithmic bytes, launches; 0 past the end
extern "C" int sba_gpu_profile_get(int idx, const char **name, double *ms, double *bytes,
                                   uint64_t *count) {
  if (idx < 0 || idx >= g_prof_nnames) return 0;
  *name = g_prof_names[idx];
  *ms = g_prof_ms[idx];
  *bytes = g_prof_bytes[idx];
  *count = g_prof_cnt[idx];
  return 1;
}

// ---------------------------------------------------------------------------
// VALU ceiling of the butterfly: forward butterflies on registers alone
// (eight independent pairs per thread, no memory access in the loop), so the
// rate is what the butterfly's own instruction mix issues at -- the ceiling
// the transform kernels' VALU roofline is read against.
// ---------------------------------------------------------------------------
#define PROBE_ITERS 2048
#define PROBE_BLOCKS 2048
#define PROBE_RUNS 5
__global__ __launch_bounds__(256) void k_probe_bfly(uint32_t *out, uint32_t w, uint32_t p,
                                                    uint32_t pinv) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  uint32_t x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = (t * 7u + (uint32_t)j * 977u) % p;
  for (int i = 0; i < PROBE_ITERS; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) bfly_fwd(x[2 * j], x[2 * j + 1], w, p, pinv);
  }
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += x[j];
  out[t] = s;
}

extern "C" int sba_gpu_probe_butterflies(double *per_s) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SBA_ERR_NODEV;
  const uint32_t p = kPrimes[SBA_K - 1];
  uint32_t inv = 1;
  for (int i = 0; i < 5; ++i) inv *= 2u - p * inv;
  uint32_t *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  int rc = SBA_ERR_HIP;
  if (hipMalloc((void **)&out, (size_t)PROBE_BLOCKS * 256 * sizeof(uint32_t)) == hipSuccess &&
      hipStreamCreate(&st) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
      hipEventCreate(&e1) == hipSuccess) {
    for (int r = 0; r < 2 + PROBE_RUNS; ++r) {
      if (r == 2) (void)hipEventRecord(e0, st);
      k_probe_bfly<<<PROBE_BLOCKS, 256, 0, st>>>(out, 12345u, p, 0u - inv);
    }
    if (hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipGetLastError() == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
        ms > 0.f) {
      *per_s = (double)PROBE_BLOCKS * 256 * PROBE_ITERS * 8 * PROBE_RUNS / (ms * 1e-3);
      rc = SBA_OK;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (out) (void)hipFree(out);
  return rc;
}

static int enqueue_garner(DevCtx &c, int set, uint32_t Nt, uint32_t stride, const JobRef &jr,
                          uint32_t njobs, uint32_t nbmax, double zsum, hipStream_t st);
static const AccArgs kNoAcc = {nullptr, 0, 0, 0, 0};

// the context's prime constants with the inverse's final scale for
// transform length Nt: Nt^-1 * R^4 mod p undoes Nt, the pointwise R^-1, the
// R^-1 each operand's limb residues carry and this multiply's own R^-1
static PrimeArgs prime_args(const DevCtx &c, int set, uint64_t Nt) {
  PrimeArgs pa = c.set[set].pa;
  for (int q = 0; q < SBA_K; ++q) {
    const uint64_t p = kSets[set].p[q];
    pa.scale[q] = (uint32_t)((unsigned __int128)h_invm(Nt % p, p) * h_powm(2, 128, p) % p);
  }
  return pa;
}

// carry blocks of a product of Z limbs
static uint32_t carry_blocks(uint64_t Z) { return (uint32_t)((Z + CARRY_LPB - 1) / CARRY_LPB); }
// k_carry_blocks workgroups for products of at most nb blocks
static uint32_t carry_grid(uint32_t nb) {
  const uint32_t g = (nb + CARRY_MIN_SPAN - 1) / CARRY_MIN_SPAN;
  return g < 1 ? 1 : (g > CARRY_WGS ? CARRY_WGS : g);
}

// Truncated products (two-pass plans): the merged last pass transforms only
// the tiles H < ceil(zlen / 2^last) of a job and the truncated inverse column
// pass recovers the coefficients from them. Used for a class whose shortest
// product leaves at least a fraction 1 - SBA_TRUNC_FILL_PCT / 100 of the
// tiles out; fuller products keep the full inverse (its register-chunk pass
// is faster than the truncated one's radix-2 sweeps).
// Defaults measured on MI355X (single products, 2^16 .. 2^23 at fills
// 0.51 .. 0.95 of the power-of-two length): truncation beats the full length
// by 4-6% at fills in (3/4, 0.85] from 2^21 on, breaks even near 0.85 and
// loses below 2^21 (its extra barriers and spine steps); below 3/4 the
// radix-3 length 3/4 wins, so truncation never replaces it by default.
#ifndef SBA_TRUNC_FILL_PCT
#define SBA_TRUNC_FILL_PCT 85
#endif
// products that fill at most this share (percent) of their power-of-two
// length run truncated at that length rather than at the radix-3 length
#ifndef SBA_TRUNC_R3_MIN_PCT
#define SBA_TRUNC_R3_MIN_PCT 0
#endif
#ifndef SBA_TRUNC_MIN_LG
#define SBA_TRUNC_MIN_LG 21
#endif
static std::atomic<int> g_trunc_fill{SBA_TRUNC_FILL_PCT};
static std::atomic<int> g_trunc_r3{SBA_TRUNC_R3_MIN_PCT};
static std::atomic<int> g_trunc_minlg{SBA_TRUNC_MIN_LG};
static bool use_trunc(int L, const Plan &pl, uint32_t zmin) {
  const int fill = g_trunc_fill.load(std::memory_order_relaxed);
  if (fill <= 0 || pl.last >= L || zmin == 0 || L < g_trunc_minlg.load(std::memory_order_relaxed))
    return false;
  const uint64_t R = (uint64_t)1 << (L - pl.last);
  const uint64_t M = ((uint64_t)zmin + ((uint64_t)1 << pl.last) - 1) >> pl.last;
  return M * 100 <= R * (uint64_t)fill;
}
extern "C" void sba_set_trunc(int max_fill_pct, int r3_min_pct, int min_lg) {
  g_trunc_fill.store(max_fill_pct, std::memory_order_relaxed);
  g_trunc_r3.store(r3_min_pct, std::memory_order_relaxed);
  g_trunc_minlg.store(min_lg, std::memory_order_relaxed);
}
extern "C" void sba_get_trunc(int *max_fill_pct, int *r3_min_pct, int *min_lg) {
  if (max_fill_pct) *max_fill_pct = g_trunc_fill.load(std::memory_order_relaxed);
  if (r3_min_pct) *r3_min_pct = g_trunc_r3.load(std::memory_order_relaxed);
  if (min_lg) *min_lg = g_trunc_minlg.load(std::memory_order_relaxed);
}

// Enqueue the whole multiply for `njobs` jobs sharing N = 2^L (jr.arr ==
// nullptr: the single job jr.one) under plan pl. nbmax = the largest
// carry-block count; zmin = the smallest product (coefficients) of the
// class; zsum / asum / bsum feed the profiler's byte counts.
static int enqueue_class(DevCtx &c, int set, int L, int sqr, const Plan &pl, const JobRef &jr,
                         uint32_t njobs, uint32_t nbmax, uint32_t zmin, double asum, double bsum,
                         double zsum, hipStream_t st) {
  if (!plan_ok(L, pl) || L > kSets[set].max_lg) return SBA_BAND;
  if (const int rs = set_ready(c, set, L)) return rs;
  const Launchers k = launchers(L, sqr, pl);
  const uint32_t N = 1u << L;
  const PrimeArgs pa = prime_args(c, set, N);
  const double KN4 = 4.0 * SBA_K * (double)N * njobs;
  const uint32_t gy1 = pl.split ? SBA_K : 1;
  if (pl.last == L) {
    if (!sqr) {
      PROF("first_fwd", 8.0 * asum + KN4);
      KLAUNCH(k.first, dim3(1, gy1, njobs), dim3(NT), 0, st, pa, jr, 0, L);
    }
    {
      PROF("fused_last", 8.0 * bsum + (sqr ? 1 : 2) * KN4);
      KLAUNCH(k.fused, dim3(1, SBA_K, njobs), dim3(NT), 0, st, pa, jr, L, 0);
    }
  } else {
    const int ell1 = L - pl.last;
    const uint32_t ntile1 = N >> SBA_FIRST_TILE_LG;
    const uint32_t ntile2 = N >> pl.last;
    // both operands' first passes in one launch (one kernel boundary fewer)
    if (!sqr) {
      PROF("first_fwd", 8.0 * (asum + bsum) + 2 * KN4);
      KLAUNCH(k.first, dim3(ntile1, 2 * gy1, njobs), dim3(NT), 0, st, pa, jr, 2, L);
    } else {
      PROF("first_fwd", 8.0 * bsum + KN4);
      KLAUNCH(k.first, dim3(ntile1, gy1, njobs), dim3(NT), 0, st, pa, jr, 1, L);
    }
    const bool trunc = use_trunc(L, pl, zmin);
    const int mode = trunc ? MODE_TRUNC : 0;
    // bytes of a truncated pass: the transformed tiles of the average job
    const double tf = trunc ? std::min(1.0, (zsum / njobs + (1u << pl.last)) / N) : 1.0;
    if (!sqr && SBA_MERGE_LAST) {
      PROF("fused_last", 3 * KN4 * tf);
      KLAUNCH(k.fused2, dim3(ntile2, SBA_K, njobs), dim3(NT), 0, st, pa, jr, L, mode);
    } else {
      if (!sqr) {
        PROF("fwd_last", 2 * KN4);
        KLAUNCH(k.fwd_last, dim3(ntile2, SBA_K, njobs), dim3(NT), 0, st, pa, jr, 0, L,
                           ell1, kNoAcc);
      }
      PROF("fused_last", (sqr ? 2 : 3) * KN4 * tf);
      KLAUNCH(k.fused, dim3(ntile2, SBA_K, njobs), dim3(NT), 0, st, pa, jr, L, mode);
    }
    if (trunc) {
      PROF("inv_trunc", 2 * KN4 * tf);
      KLAUNCH(sba_trunc_inv(ell1), dim3(ntile1, SBA_K, njobs), dim3(NT), 0, st, pa, jr,
                         L);
    } else {
      PROF("inv_first", 2 * KN4);
      KLAUNCH(k.inv_first, dim3(ntile1, SBA_K, njobs), dim3(NT), 0, st, pa, jr, 1, L,
                         0, kNoAcc);
    }
  }
  return enqueue_garner(c, set, N, N, jr, njobs, nbmax, zsum, st);
}

// Garner + carry scan for njobs jobs whose product residues sit in bspec
// with prime stride `stride`, left scaled by the transform length Nt by the
// unnormalised inverse
static int enqueue_garner(DevCtx &c, int set, uint32_t Nt, uint32_t stride, const JobRef &jr,
                          uint32_t njobs, uint32_t nbmax, double zsum, hipStream_t st) {
  (void)Nt;
  GarnerArgs ga = c.set[set].ga_base;
  {
    PROF("garner", 4.0 * SBA_K * zsum + 8.0 * zsum);
    KLAUNCH(k_garner_limbs, dim3(nbmax, 1, njobs), dim3(NT), 0, st, ga, jr, stride);
  }
  {
    PROF("carry_blocks", 8.0 * zsum / CARRY_LPB);
    const uint32_t g = carry_grid(nbmax);
    KLAUNCH(k_carry_blocks, dim3(g, 1, njobs), dim3(SCAN_NT), 0, st, jr);
  }
  HIPCHK(hipGetLastError());
  return SBA_OK;
}

// carve one job's buffers out of a workspace region of ws_need(N) bytes
static void job_layout(Job &j, void *ws, uint32_t N) {
  j.aspec = (uint32_t *)ws;
  j.bspec = j.aspec + (size_t)SBA_K * N;
  j.bmaps = j.bspec + (size_t)SBA_K * N;
}

static int get_ctx(DevCtx **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= SBA_MAX_DEV) return SBA_ERR_HIP;
  DevCtx &c = g_ctx[dev];
  ctx_init(c);
  *out = &c;
  return 0;
}

// ---------------------------------------------------------------------------
// chunked products: the shorter operand's spectrum (length Nc = 2^Lc) is
// computed once; the longer one is cut into chunks of C = Nc - bn + 1
// limbs, each transformed, multiplied by that spectrum and inverted straight
// into an accumulator of residues (overlapping tails added mod p); one
// Garner + carry pass over the accumulator finishes. Used when it moves
// clearly fewer transform elements than one long transform, and for
// products longer than 2^25 coefficients.
// ---------------------------------------------------------------------------
#define SBA_MAX_Z ((uint64_t)1 << 29)   // product limbs: residue offsets stay 32-bit

extern "C" uint32_t sba_gpu_carry_block_limbs(void) { return CARRY_LPB; }

extern "C" int sba_gpu_carry_blocks(uint64_t *rp, const uint32_t *bmaps, uint64_t Z,
                                    void *stream) {
  if (!rp || !bmaps || Z == 0 || Z > SBA_MAX_Z) return SBA_BAND;
  JobRef jr;
  jr.arr = nullptr;
  Job &j = jr.one;
  memset(&j, 0, sizeof j);
  j.r = rp;
  j.bmaps = (uint32_t *)bmaps;
  j.Z = (uint32_t)Z;
  j.zlen = (uint32_t)(Z - 1);
  j.nb = carry_blocks(Z);
  const uint32_t g = carry_grid(j.nb);
  KLAUNCH(k_carry_blocks, dim3(g, 1, 1), dim3(SCAN_NT), 0, (hipStream_t)stream, jr);
  HIPCHK(hipGetLastError());
  return SBA_OK;
}

#define CHUNK_MIN 8192                  // smallest chunk (limbs)
#define CHUNK_LAUNCH_EQ 150000.0        // one launch costs about this many element-transforms

struct ChunkChoice {
  int Lc;
  uint64_t C, n;                        // chunk limbs, chunks (n == 0: no chunking)
};

// an >= bn
static ChunkChoice chunk_choice(uint64_t an, uint64_t bn, int sqr) {
  ChunkChoice best = {0, 0, 0};
  const int L = lg_for(an + bn - 1);
  if (L <= SBA_MAX_LG && an < 4 * bn) return best;       // near-balanced: one transform
  // direct: forward a, forward b (not for a square), inverse
  double cbest = L <= SBA_MAX_LG ? 0.9 * (sqr ? 2.0 : 3.0) * (double)((uint64_t)1 << L) : 1e300;
  for (int Lc = 14; Lc <= SBA_MAX_LG && Lc < L; ++Lc) {
    const uint64_t Nc = (uint64_t)1 << Lc;
    if (bn + CHUNK_MIN > Nc) continue;
    const uint64_t C = Nc - bn + 1;
    const uint64_t n = (an + C - 1) / C;
    if (n < 2) continue;
    const double cost = (double)Nc * (1.0 + 2.0 * (double)n) + CHUNK_LAUNCH_EQ * 3.0 * (double)n;
    if (cost < cbest) {
      cbest = cost;
      best.Lc = Lc;
      best.C = C;
      best.n = n;
    }
  }
  return best;
}

// short spectrum, chunk spectrum, accumulator, carry maps of Z limbs
static size_t ws_chunked(int Lc, uint64_t Z) {
  const size_t Nc = (size_t)1 << Lc, Zs = (Z + 255) & ~(size_t)255, nb = Z / CARRY_LPB + 2;
  return 2 * SBA_K * Nc * 4 + SBA_K * Zs * 4 + nb * 4 + 4096;
}

// an >= bn; ws holds ws_chunked(cp.Lc, an + bn) bytes
static int enqueue_chunked(DevCtx &c, const ChunkChoice &cp, uint64_t *rp, const uint64_t *ap,
                           uint64_t an, const uint64_t *bp, uint64_t bn, void *ws,
                           hipStream_t st) {
  const int Lc = cp.Lc;
  const uint32_t Nc = 1u << Lc;
  const Plan pl = plan_for(Lc, 0, 1);
  if (!plan_ok(Lc, pl) || pl.last == Lc) return SBA_BAND;
  const Launchers k = launchers(Lc, 0, pl);
  const int set = set_for(Lc, 0);
  if (const int rs = set_ready(c, set, Lc)) return rs;
  const PrimeArgs pa = prime_args(c, set, Nc);
  const uint64_t Z = an + bn;
  const uint32_t Zs = (uint32_t)((Z + 255) & ~(uint64_t)255);
  const int ell1 = Lc - pl.last;
  const uint32_t ntile1 = Nc >> SBA_FIRST_TILE_LG, ntile2 = Nc >> pl.last;
  const uint32_t gy1 = pl.split ? SBA_K : 1;
  const double KN4 = 4.0 * SBA_K * Nc;
  JobRef jr;
  jr.arr = nullptr;
  Job &j = jr.one;
  memset(&j, 0, sizeof j);
  j.aspec = (uint32_t *)ws;
  j.bspec = j.aspec + (size_t)SBA_K * Nc;
  uint32_t *acc = j.bspec + (size_t)SBA_K * Nc;
  // the short operand's whole spectrum (the "a" side of the passes)
  j.a = bp;
  j.an = (uint32_t)bn;
  {
    PROF("first_fwd", 8.0 * bn + KN4);
    KLAUNCH(k.first, dim3(ntile1, gy1, 1), dim3(NT), 0, st, pa, jr, 0, Lc);
  }
  {
    PROF("fwd_last", 2 * KN4);
    KLAUNCH(k.fwd_last, dim3(ntile2, SBA_K, 1), dim3(NT), 0, st, pa, jr, 0, Lc, ell1,
                       kNoAcc);
  }
  for (uint64_t ci = 0; ci < cp.n; ++ci) {
    const uint64_t off = ci * cp.C;
    const uint64_t len = an - off < cp.C ? an - off : cp.C;
    j.b = ap + off;
    j.bn = (uint32_t)len;
    {
      PROF("first_fwd", 8.0 * len + KN4);
      KLAUNCH(k.first, dim3(ntile1, gy1, 1), dim3(NT), 0, st, pa, jr, 1, Lc);
    }
    {
      PROF("fused_last", 3 * KN4);
      KLAUNCH(k.fused, dim3(ntile2, SBA_K, 1), dim3(NT), 0, st, pa, jr, Lc, 0);
    }
    {
      AccArgs aa;
      aa.dst = acc;
      aa.stride = Zs;
      aa.off = (uint32_t)off;
      aa.ovl = ci ? (uint32_t)(bn - 1) : 0u;     // the previous chunk's tail
      aa.lim = (uint32_t)(len + bn - 1);          // this chunk's product
      PROF("inv_first", KN4 + 8.0 * SBA_K * (len + bn));
      KLAUNCH(k.inv_first, dim3(ntile1, SBA_K, 1), dim3(NT), 0, st, pa, jr, 1, Lc, 0,
                         aa);
    }
  }
  JobRef jg;
  jg.arr = nullptr;
  Job &g = jg.one;
  memset(&g, 0, sizeof g);
  g.bspec = acc;
  g.r = rp;
  g.zlen = (uint32_t)(Z - 1);
  g.Z = (uint32_t)Z;
  g.nb = carry_blocks(Z);
  g.bmaps = acc + (size_t)SBA_K * Zs;
  return enqueue_garner(c, set, Nc, Zs, jg, 1, g.nb, (double)Z, st);
}

// ---------------------------------------------------------------------------
// radix-3 lengths: a product with 2^(L+1) < zlen <= 3 * 2^L runs at length
// 3 * 2^L instead of 2^(L+2) (a quarter fewer residues in every pass); the
// column DFT-3 rides the first forward and the last inverse pass.
// ---------------------------------------------------------------------------
// row lengths of the radix-3 classes: 2^13 .. 2^22 (first-pass tiles keep
// >= 4 columns of 2^11 per row) in batches; single products from
// SBA_R3_SINGLE_MIN_L (measured on MI355X, one product at a time: at row
// 2^14 the radix-3 product is 37% slower than the power-of-two one of twice
// the row length -- too few blocks to hide its extra launch -- at 2^16 they
// are even, at 2^18 radix-3 wins by 15%)
#define R3_MIN_L 13
#define R3_MAX_L 22
#ifndef SBA_R3_SINGLE_MIN_L
#define SBA_R3_SINGLE_MIN_L 16
#endif
static std::atomic<int> g_r3_single_min{SBA_R3_SINGLE_MIN_L};
extern "C" void sba_set_r3_min_lg(int L) { g_r3_single_min.store(L, std::memory_order_relaxed); }
extern "C" int sba_get_r3_min_lg(void) { return g_r3_single_min.load(std::memory_order_relaxed); }

// L of the radix-3 class of zlen coefficients, 0 if none (single: a product
// on its own, else a batch job)
static int r3_lg(uint64_t zlen, bool single) {
  if (!SBA_MERGE_LAST) return 0;
  const int L2 = lg_for(zlen);
  const int L = L2 - 2;
  const int lmin = single ? std::max(R3_MIN_L, g_r3_single_min.load(std::memory_order_relaxed))
                          : R3_MIN_L;
  if (L < lmin || L > R3_MAX_L) return 0;
  if (g_trunc_fill.load(std::memory_order_relaxed) > 0 &&
      L2 >= g_trunc_minlg.load(std::memory_order_relaxed) &&
      zlen * 100 <= ((uint64_t)g_trunc_r3.load(std::memory_order_relaxed) << L2))
    return 0;
  return zlen <= (uint64_t)3 << L ? L : 0;
}

// the power-of-two plan of the same row length, with the first pass kept
// within 1..9 levels (the radix-3 column tiles hold 2^11 residues per row)
// 1: radix-3 classes default to the widest last pass (13 levels), which
// keeps the first-pass column tiles widest (fixed by measurement)
#ifndef SBA_R3_WIDE_LAST
#define SBA_R3_WIDE_LAST 1
#endif
// a power-of-two plan moved into the radix-3 limits: last pass 11..13
// levels, first pass 1..9 levels (its column tiles hold 2^11 residues per
// row). With the widest last pass the radix-32 chunks are kept: measured at
// 3 x 2^21, radix 16 with 13 last-pass levels is 45% slower.
static Plan r3_fix(int L, Plan p) {
  if (SBA_R3_WIDE_LAST) {
    p.last = 13;
    p.bm = 5;
  }
  if (p.last >= L) p.last = L - 1;
  if (p.last > 13) p.last = 13;
  if (L - p.last > 9) p.last = L - 9;
  if (p.last < 11) p.last = 11;
  return p;
}
static bool plan_ok3(int L, const Plan &p) {
  if (L < R3_MIN_L || L > R3_MAX_L || p.bm < 3 || p.bm > 5 || (p.split != 0 && p.split != 1))
    return false;
  return p.last >= 11 && p.last <= 13 && L - p.last >= 1 && L - p.last <= 9;
}
static Plan r3_plan(int L, int sqr, uint32_t njobs) { return plan_for3(L, sqr, njobs); }

// set: the prime set (set_for(L, 1), or SET_W where only its tables exist)
static int enqueue_r3(DevCtx &c, int set, int L, int sqr, const Plan &pl, const JobRef &jr,
                      uint32_t njobs, uint32_t nbmax, double asum, double bsum, double zsum,
                      hipStream_t st) {
  const int ell1 = L - pl.last;
  if (ell1 < 1 || ell1 > 9 || pl.last < 11 || pl.last > 13) return SBA_BAND;
  const Launchers k = launchers(L, sqr, pl);
  const first_fn f1 = sba_r3_first(ell1);
  const inv3_fn i1 = sba_r3_inv(ell1);
  const uint32_t M = 1u << L;
  if (L > kSets[set].max_lg) return SBA_BAND;
  if (const int rs = set_ready(c, set, L)) return rs;
  const PrimeArgs pa = prime_args(c, set, 3 * (uint64_t)M);
  const double KN4 = 4.0 * SBA_K * 3.0 * M * njobs;
  const uint32_t ntile1 = M >> 11, ntile2 = M >> pl.last;
  const uint32_t gy1 = pl.split ? SBA_K : 1;
  if (!sqr) {
    PROF("first_fwd", 8.0 * (asum + bsum) + 2 * KN4);
    KLAUNCH(f1, dim3(ntile1, 2 * gy1, njobs), dim3(NT), 0, st, pa, jr, 2, L);
  } else {
    PROF("first_fwd", 8.0 * bsum + KN4);
    KLAUNCH(f1, dim3(ntile1, gy1, njobs), dim3(NT), 0, st, pa, jr, 1, L);
  }
  {
    PROF("fused_last", (sqr ? 2 : 3) * KN4);
    if (sqr)
      KLAUNCH(k.fused, dim3(ntile2, SBA_K, 3 * njobs), dim3(NT), 0, st, pa, jr, L, 1);
    else
      KLAUNCH(k.fused2, dim3(ntile2, SBA_K, 3 * njobs), dim3(NT), 0, st, pa, jr, L,
                         1);
  }
  {
    PROF("inv_first", 2 * KN4);
    KLAUNCH(i1, dim3(ntile1, SBA_K, njobs), dim3(NT), 0, st, pa, jr, L);
  }
  return enqueue_garner(c, set, 3 * M, 3 * M, jr, njobs, nbmax, zsum, st);
}

// one transform of length <= 2^25 (batches use only this)
static int in_band_direct(int64_t an, int64_t bn) {
  if (an <= 0 || bn <= 0) return 0;
  const uint64_t mn = (uint64_t)(an < bn ? an : bn);
  if (mn > sba_gpu_min_operand_cap()) return 0;
  return lg_for((uint64_t)an + (uint64_t)bn - 1) <= SBA_MAX_LG;
}

// single products: direct, or chunked along the longer operand
static int in_band(int64_t an, int64_t bn) {
  if (in_band_direct(an, bn)) return 1;
  if (an <= 0 || bn <= 0) return 0;
  const uint64_t mx = (uint64_t)(an > bn ? an : bn), mn = (uint64_t)(an < bn ? an : bn);
  if (mn > sba_gpu_min_operand_cap() || mx + mn > SBA_MAX_Z) return 0;
  return chunk_choice(mx, mn, 0).n > 0;
}

extern "C" int sba_gpu_in_band(int64_t an, int64_t bn) { return in_band(an, bn); }

extern "C" int sba_gpu_chunking(int64_t an, int64_t bn, int sqr, int *Lc, uint64_t *chunk,
                                uint64_t *nchunks) {
  if (!in_band(an, bn)) return SBA_BAND;
  const uint64_t mx = (uint64_t)(an > bn ? an : bn), mn = (uint64_t)(an < bn ? an : bn);
  const ChunkChoice cp = (!sqr && school_long_job(mx, mn)) ? ChunkChoice{0, 0, 0}
                                                          : chunk_choice(mx, mn, sqr);
  *Lc = cp.Lc;
  *chunk = cp.C;
  *nchunks = cp.n;
  return SBA_OK;
}

extern "C" size_t sba_gpu_workspace_bytes(int64_t an, int64_t bn) {
  if (!in_band(an, bn)) return 0;
  const uint64_t mx = (uint64_t)(an > bn ? an : bn), mn = (uint64_t)(an < bn ? an : bn);
  if (school_long_job(mx, mn)) return ws_school_long(mx + mn);
  const ChunkChoice cp = chunk_choice(mx, mn, 0);
  if (cp.n) return ws_chunked(cp.Lc, mx + mn);
  const int L3 = r3_lg(mx + mn - 1, true);
  if (L3) return ws_need((size_t)3 << L3);
  return ws_need((size_t)1 << lg_for(mx + mn - 1));
}

extern "C" int sba_gpu_radix3(int64_t an, int64_t bn, int sqr) {
  if (!in_band(an, bn)) return 0;
  const uint64_t mx = (uint64_t)(an > bn ? an : bn), mn = (uint64_t)(an < bn ? an : bn);
  if ((!sqr && school_long_job(mx, mn)) || chunk_choice(mx, mn, sqr).n) return 0;
  return r3_lg(mx + mn - 1, true);
}

extern "C" int sba_gpu_school(int64_t an, int64_t bn, int sqr) {
  if (!in_band(an, bn)) return 0;
  const uint64_t mx = (uint64_t)(an > bn ? an : bn), mn = (uint64_t)(an < bn ? an : bn);
  if (school_job(mx, mn, true)) return 1;
  return (!sqr && school_long_job(mx, mn)) ? 2 : 0;
}
extern "C" void sba_set_school_long_ratio(uint32_t sixteenths) {
  g_long_ratio.store(sixteenths, std::memory_order_relaxed);
}
extern "C" uint32_t sba_get_school_long_ratio(void) {
  return g_long_ratio.load(std::memory_order_relaxed);
}

extern "C" int sba_gpu_truncated(int64_t an, int64_t bn, int sqr, int *L, int *last, int *rows) {
  if (!in_band(an, bn)) return 0;
  const uint64_t mx = (uint64_t)(an > bn ? an : bn), mn = (uint64_t)(an < bn ? an : bn);
  if (school_job(mx, mn, true) || (!sqr && school_long_job(mx, mn)) ||
      chunk_choice(mx, mn, sqr).n || r3_lg(mx + mn - 1, true))
    return 0;
  const uint64_t zlen = mx + mn - 1;
  const int Lg = lg_for(zlen);
  const Plan pl = plan_for(Lg, sqr, 1);
  if (!use_trunc(Lg, pl, (uint32_t)zlen)) return 0;
  if (L) *L = Lg;
  if (last) *last = pl.last;
  if (rows) *rows = (int)((zlen + ((uint64_t)1 << pl.last) - 1) >> pl.last);
  return 1;
}

// one short x long product (an >= bn), carry maps in bmaps
static int enqueue_school_long_one(uint64_t *rp, const uint64_t *ap, uint64_t an,
                                   const uint64_t *bp, uint64_t bn, uint32_t *bmaps,
                                   hipStream_t st) {
  JobRef jr;
  jr.arr = nullptr;
  Job &j = jr.one;
  memset(&j, 0, sizeof j);
  j.a = ap;
  j.b = bp;
  j.r = rp;
  j.an = (uint32_t)an;
  j.bn = (uint32_t)bn;
  j.zlen = (uint32_t)(an + bn - 1);
  j.Z = (uint32_t)(an + bn);
  j.nb = carry_blocks(j.Z);
  j.bmaps = bmaps;
  {
    PROF("school_long", 8.0 * (an + bn) + 8.0 * j.Z);
    KLAUNCH(k_school_long, dim3(j.nb, 1, 1), dim3(NT), 0, st, jr);
  }
  {
    PROF("carry_blocks", 8.0 * j.Z / CARRY_LPB);
    const uint32_t g = carry_grid(j.nb);
    KLAUNCH(k_carry_blocks, dim3(g, 1, 1), dim3(SCAN_NT), 0, st, jr);
  }
  HIPCHK(hipGetLastError());
  return SBA_OK;
}

static int mul_common(uint64_t *rp, const uint64_t *ap, int64_t an, const uint64_t *bp,
                      int64_t bn, int sqr, void *ws, size_t ws_bytes, void *stream) {
  if (!in_band(an, bn)) return SBA_BAND;
  if (an < bn && !sqr) {
    const uint64_t *t = ap; ap = bp; bp = t;
    int64_t tn = an; an = bn; bn = tn;
  }
  if (school_job((uint64_t)an, (uint64_t)bn, true)) {
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_mu);
    JobRef jr;
    jr.arr = nullptr;
    Job &j = jr.one;
    memset(&j, 0, sizeof j);
    j.a = ap;
    j.b = sqr ? ap : bp;
    j.r = rp;
    j.an = (uint32_t)an;
    j.bn = (uint32_t)bn;
    j.zlen = (uint32_t)(an + bn - 1);
    j.Z = (uint32_t)(an + bn);
    j.nb = 1;
    {
      PROF("school", 16.0 * (double)j.Z);
      KLAUNCH(k_school, dim3(1, 1, 1), dim3(NT), 0, st, jr);
    }
    HIPCHK(hipGetLastError());
    return SBA_OK;
  }
  if (!sqr && school_long_job((uint64_t)an, (uint64_t)bn)) {
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_mu);
    const uint64_t Z = (uint64_t)an + (uint64_t)bn;
    const size_t need = ws_school_long(Z);
    if (!ws) {
      DevCtx *c;
      int rc = get_ctx(&c);
      if (rc) return rc;
      if (c->ws_bytes < need) {
        if (c->ws) HIPCHK(hipFree(c->ws));
        c->ws = nullptr;
        c->ws_bytes = 0;
        HIPCHK(hipMalloc(&c->ws, need));
        c->ws_bytes = need;
      }
      ws = c->ws;
    } else if (ws_bytes < need) {
      return SBA_ERR_WORKSPACE;
    }
    return enqueue_school_long_one(rp, ap, (uint64_t)an, bp, (uint64_t)bn, (uint32_t *)ws, st);
  }
  const ChunkChoice cp = chunk_choice((uint64_t)an, (uint64_t)bn, sqr);
  const int L3 = cp.n ? 0 : r3_lg((uint64_t)an + (uint64_t)bn - 1, true);
  const int L = cp.n ? cp.Lc : (L3 ? L3 : lg_for((uint64_t)an + (uint64_t)bn - 1));
  const size_t Ntot = L3 ? (size_t)3 << L3 : (size_t)1 << L;
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  rc = ctx_twiddles(*c, L, st);
  if (rc) return rc;
  const size_t need = cp.n ? ws_chunked(cp.Lc, (uint64_t)an + (uint64_t)bn) : ws_need(Ntot);
  if (!ws) {
    if (c->ws_bytes < need) {
      if (c->ws) HIPCHK(hipFree(c->ws));
      c->ws = nullptr;
      c->ws_bytes = 0;
      HIPCHK(hipMalloc(&c->ws, need));
      c->ws_bytes = need;
    }
    ws = c->ws;
  } else if (ws_bytes < need) {
    return SBA_ERR_WORKSPACE;
  }
  if (cp.n)
    return enqueue_chunked(*c, cp, rp, ap, (uint64_t)an, sqr ? ap : bp, (uint64_t)bn, ws, st);
  JobRef jr;
  jr.arr = nullptr;
  Job &j = jr.one;
  j.a = ap;
  j.b = sqr ? ap : bp;
  j.r = rp;
  j.an = (uint32_t)an;
  j.bn = (uint32_t)bn;
  j.zlen = (uint32_t)(an + bn - 1);
  j.Z = (uint32_t)(an + bn);
  j.nb = carry_blocks(j.Z);
  job_layout(j, ws, (uint32_t)Ntot);
  if (L3)
    return enqueue_r3(*c, set_for(L3, 1), L3, sqr, r3_plan(L3, sqr, 1), jr, 1, j.nb,
                      (double)an, (double)bn, (double)j.Z, st);
  return enqueue_class(*c, set_for(L, 0), L, sqr, plan_for(L, sqr, 1), jr, 1, j.nb, j.zlen,
                       (double)an, (double)bn, (double)j.Z, st);
}

extern "C" int sba_gpu_mul_dev(uint64_t *rp, const uint64_t *ap, int64_t an, const uint64_t *bp,
                               int64_t bn, void *stream) {
  return mul_common(rp, ap, an, bp, bn, 0, nullptr, 0, stream);
}

extern "C" int sba_gpu_sqr_dev(uint64_t *rp, const uint64_t *ap, int64_t an, void *stream) {
  return mul_common(rp, ap, an, ap, an, 1, nullptr, 0, stream);
}

extern "C" int sba_gpu_mul_dev_ws(uint64_t *rp, const uint64_t *ap, int64_t an,
                                  const uint64_t *bp, int64_t bn, void *ws, size_t ws_bytes,
                                  void *stream) {
  if (!ws) return SBA_ERR_WORKSPACE;
  return mul_common(rp, ap, an, bp, bn, 0, ws, ws_bytes, stream);
}

extern "C" int sba_gpu_sqr_dev_ws(uint64_t *rp, const uint64_t *ap, int64_t an, void *ws,
                                  size_t ws_bytes, void *stream) {
  if (!ws) return SBA_ERR_WORKSPACE;
  return mul_common(rp, ap, an, ap, an, 1, ws, ws_bytes, stream);
}

// ---------------------------------------------------------------------------
// allocation-free multiply (for callers that cannot allocate device memory:
// code under graph capture, tight serving loops). Everything the kernels use
// beyond the operands and the product lives in one file-scope device array,
// `scratch`, of SBA_NOALLOC_SCRATCH_BYTES per device:
//   [0, NA_TW_BYTES)     root tables of the 5 primes for transforms up to
//                        2^SBA_NOALLOC_MAX_LG points (wn, wni per prime),
//                        built once per device by the first call
//   [NA_TW_BYTES, ...)   a's and b's spectra and the carry maps of one
//                        product (ws_need(2^SBA_NOALLOC_MAX_LG) bytes)
// Every buffer the kernels read is written earlier in the same call (the
// first pass writes all residues, Garner every carry map), so nothing needs
// clearing between calls. In host code `scratch` names a host-side stand-in:
// the device address comes from hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------
#define NA_TW_COUNT ((size_t)1 << (SBA_NOALLOC_MAX_LG - 1))   // node twiddles per table
#define NA_TW_BYTES (2 * SBA_K * NA_TW_COUNT * 4)
static_assert(NA_TW_BYTES + (2 * SBA_K * ((size_t)4 << SBA_NOALLOC_MAX_LG) +
                             (((size_t)1 << SBA_NOALLOC_MAX_LG) / CARRY_LPB + 2) * 4 + 4096) <=
                  SBA_NOALLOC_SCRATCH_BYTES,
              "static scratch too small for its largest product");
static __device__ __align__(256) uint8_t scratch[SBA_NOALLOC_SCRATCH_BYTES];

// this device's scratch address and its root tables (first call on the
// device: hipGetSymbolAddress, then the table kernel on the caller's stream,
// waited for so that later calls on any stream find the tables built)
static int na_setup(DevCtx &c, hipStream_t st) {
  if (!c.na_scratch) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return SBA_ERR_HIP;
    c.na_scratch = (char *)ptr;
  }
  if (!c.na_tables) {
    uint32_t *tw = (uint32_t *)c.na_scratch;
    for (int q = 0; q < SBA_K; ++q) {
      const TwArgs ta = tw_args(c.set[SET_W], kSets[SET_W], q);
      KLAUNCH(k_twiddles, dim3((uint32_t)((NA_TW_COUNT + NT - 1) / NT)), dim3(NT), 0,
                         st, ta, tw + (2 * q) * NA_TW_COUNT, tw + (2 * q + 1) * NA_TW_COUNT,
                         (uint32_t)NA_TW_COUNT);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    c.na_tables = 1;
  }
  return 0;
}

extern "C" int sba_gpu_mul_dev_noalloc(uint64_t *rp, const uint64_t *ap, int64_t an,
                                       const uint64_t *bp, int64_t bn, void *stream) {
  if (!rp || !ap || !bp || !in_band_direct(an, bn)) return SBA_BAND;
  const int sqr = ap == bp && an == bn;
  if (an < bn) {
    const uint64_t *t = ap; ap = bp; bp = t;
    int64_t tn = an; an = bn; bn = tn;
  }
  const uint64_t zlen = (uint64_t)an + (uint64_t)bn - 1;
  if (zlen > ((uint64_t)1 << SBA_NOALLOC_MAX_LG)) {
    fprintf(stderr,
            "simd_bigint_amd: sba_gpu_mul_dev_noalloc: %llu coefficients exceed the static "
            "scratch's largest transform (2^%d)\n",
            (unsigned long long)zlen, SBA_NOALLOC_MAX_LG);
    return SBA_ERR_TOO_LARGE;
  }
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  const int saved_prof = g_prof_on;
  g_prof_on = 0;                                  // the profiler creates events: off here
  struct Restore { int v; ~Restore() { g_prof_on = v; } } restore{saved_prof};
  if (school_job((uint64_t)an, (uint64_t)bn, true)) {
    JobRef jr;
    jr.arr = nullptr;
    Job &j = jr.one;
    memset(&j, 0, sizeof j);
    j.a = ap;
    j.b = sqr ? ap : bp;
    j.r = rp;
    j.an = (uint32_t)an;
    j.bn = (uint32_t)bn;
    j.zlen = (uint32_t)zlen;
    j.Z = (uint32_t)(zlen + 1);
    j.nb = 1;
    KLAUNCH(k_school, dim3(1, 1, 1), dim3(NT), 0, st, jr);
    HIPCHK(hipGetLastError());
    return SBA_OK;
  }
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  if ((rc = na_setup(*c, st)) != 0) return rc;
  // the context with the static tables in place of the allocated ones
  // (the wide prime set only: the static tables are that set's)
  DevCtx nc;
  nc.set[SET_W].pa = c->set[SET_W].pa;
  nc.set[SET_W].ga_base = c->set[SET_W].ga_base;
  const uint32_t *tw = (const uint32_t *)c->na_scratch;
  for (int q = 0; q < SBA_K; ++q) {
    nc.set[SET_W].pa.wn[q] = tw + (2 * q) * NA_TW_COUNT;
    nc.set[SET_W].pa.wni[q] = tw + (2 * q + 1) * NA_TW_COUNT;
  }
  nc.set[SET_W].tw_lg = SBA_NOALLOC_MAX_LG;     // the static tables' reach
  // always one transform (power of two or 3 * 2^L), never chunked
  const int L3 = r3_lg(zlen, true);
  const int L = L3 ? L3 : lg_for(zlen);
  const size_t Ntot = L3 ? (size_t)3 << L3 : (size_t)1 << L;
  JobRef jr;
  jr.arr = nullptr;
  Job &j = jr.one;
  memset(&j, 0, sizeof j);
  j.a = ap;
  j.b = sqr ? ap : bp;
  j.r = rp;
  j.an = (uint32_t)an;
  j.bn = (uint32_t)bn;
  j.zlen = (uint32_t)zlen;
  j.Z = (uint32_t)(zlen + 1);
  j.nb = carry_blocks(j.Z);
  job_layout(j, c->na_scratch + NA_TW_BYTES, (uint32_t)Ntot);
  if (L3)
    return enqueue_r3(nc, SET_W, L3, sqr, r3_plan(L3, sqr, 1), jr, 1, j.nb, (double)an,
                      (double)bn, (double)j.Z, st);
  return enqueue_class(nc, SET_W, L, sqr, plan_for(L, sqr, 1), jr, 1, j.nb, j.zlen,
                       (double)an, (double)bn, (double)j.Z, st);
}

// ---------------------------------------------------------------------------
// batches: many independent multiplies of uneven sizes. Jobs are grouped in
// size classes (same N = 2^L, same mul/sqr kind); one launch sequence per
// class covers all its jobs (grid z = job), under the class's plan.
// ---------------------------------------------------------------------------
struct ClassDesc {
  int L, sqr, start;
  uint32_t n, nbmax, zmin;
  double as, bs, zs;
  Plan plan;
  int r3;                          // radix-3 class: length 3 * 2^L
  int school;                      // 1: the schoolbook class (every small product),
                                   // 2: the short x long schoolbook class
};
#define SBA_MAX_CLASSES (4 * (SBA_MAX_LG + 1) + 1)

// class of a batch job: (L, radix-3?) of its transform (never chunked)
static void job_key(const sba_job &j, int *L, int *r3) {
  const uint64_t an = (uint64_t)j.an, bn = j.bp ? (uint64_t)j.bn : an;
  const int L3 = r3_lg(an + bn - 1, false);
  *r3 = L3 != 0;
  *L = L3 ? L3 : lg_for(an + bn - 1);
}

static size_t job_ws(const sba_job &j) {
  int L, r3;
  job_key(j, &L, &r3);
  return (ws_need((size_t)(r3 ? 3 : 1) << L) + 255) & ~(size_t)255;
}
static size_t table_bytes(int njobs) { return ((size_t)njobs * sizeof(Job) + 255) & ~(size_t)255; }

static bool job_school(const sba_job &j) {
  return school_job((uint64_t)j.an, j.bp ? (uint64_t)j.bn : (uint64_t)j.an, false);
}
// a short x long job (multiplies only; never a square, whose operands are equal)
static bool job_school_long(const sba_job &j) {
  return j.bp && !job_school(j) && school_long_job((uint64_t)j.an, (uint64_t)j.bn);
}
static bool job_off_ntt(const sba_job &j) { return job_school(j) || job_school_long(j); }

// Orders the jobs by class into hj[], carving each job's buffers from ws in
// that order (job_ws bytes each), and fills cls[] (at most SBA_MAX_CLASSES
// entries: the schoolbook class when school != 0, then square?, radix-3?,
// L). Every job must be in band. Returns the number of classes.
static int classify_jobs(const sba_job *jobs, int njobs, char *ws, Job *hj, ClassDesc *cls,
                         int school) {
  int ncls = 0, pos = 0;
  size_t off = 0;
  int count[2][2][SBA_MAX_LG + 1] = {};
  if (school) {
    ClassDesc k = {0, 0, pos, 0, 0, 0, 0, 0, 0, Plan{0, 0, 0}, 0, 1};
    for (int i = 0; i < njobs; ++i) {
      if (!job_school(jobs[i])) continue;
      const int isq = jobs[i].bp == nullptr;
      const uint64_t an = (uint64_t)jobs[i].an, bn = isq ? an : (uint64_t)jobs[i].bn;
      Job &j = hj[pos++];
      memset(&j, 0, sizeof j);
      j.a = jobs[i].ap;
      j.b = isq ? jobs[i].ap : jobs[i].bp;
      j.an = (uint32_t)an;
      j.bn = (uint32_t)bn;
      j.r = jobs[i].rp;
      j.zlen = (uint32_t)(an + bn - 1);
      j.Z = (uint32_t)(an + bn);
      j.nb = 1;
      off += job_ws(jobs[i]);
      k.n++;
      k.zs += j.Z;
    }
    if (k.n) cls[ncls++] = k;
    // the short x long class: carry maps in the job's workspace slot (job_ws
    // reserves the NTT's, which is larger)
    ClassDesc kl = {0, 0, pos, 0, 0, 0, 0, 0, 0, Plan{0, 0, 0}, 0, 2};
    for (int i = 0; i < njobs; ++i) {
      if (!job_school_long(jobs[i])) continue;
      const bool swap = jobs[i].an < jobs[i].bn;
      Job &j = hj[pos++];
      memset(&j, 0, sizeof j);
      j.a = swap ? jobs[i].bp : jobs[i].ap;
      j.b = swap ? jobs[i].ap : jobs[i].bp;
      j.an = (uint32_t)(swap ? jobs[i].bn : jobs[i].an);
      j.bn = (uint32_t)(swap ? jobs[i].an : jobs[i].bn);
      j.r = jobs[i].rp;
      j.zlen = j.an + j.bn - 1;
      j.Z = j.an + j.bn;
      j.nb = carry_blocks(j.Z);
      j.bmaps = (uint32_t *)(ws + off);
      off += job_ws(jobs[i]);
      kl.n++;
      if (j.nb > kl.nbmax) kl.nbmax = j.nb;
      kl.zs += j.Z;
    }
    if (kl.n) cls[ncls++] = kl;
  }
  for (int i = 0; i < njobs; ++i) {
    if (school && job_off_ntt(jobs[i])) continue;
    int L, r3;
    job_key(jobs[i], &L, &r3);
    count[jobs[i].bp == nullptr][r3][L]++;
  }
  for (int sq = 0; sq < 2; ++sq)
    for (int r3 = 0; r3 < 2; ++r3)
    for (int L = 0; L <= SBA_MAX_LG; ++L) {
      if (!count[sq][r3][L]) continue;
      ClassDesc k = {L, sq, pos, 0, 0, ~0u, 0, 0, 0, heuristic_plan(L), r3, 0};
      for (int i = 0; i < njobs; ++i) {
        if (school && job_off_ntt(jobs[i])) continue;
        const int isq = jobs[i].bp == nullptr;
        const uint64_t an = (uint64_t)jobs[i].an, bn = isq ? an : (uint64_t)jobs[i].bn;
        int jL, jr3;
        job_key(jobs[i], &jL, &jr3);
        if (isq != sq || jr3 != r3 || jL != L) continue;
        Job &j = hj[pos++];
        // larger operand first (the first pass of b then covers the shorter one)
        const bool swap = !isq && an < bn;
        j.a = swap ? jobs[i].bp : jobs[i].ap;
        j.b = isq ? jobs[i].ap : (swap ? jobs[i].ap : jobs[i].bp);
        j.an = (uint32_t)(swap ? bn : an);
        j.bn = (uint32_t)(swap ? an : bn);
        j.r = jobs[i].rp;
        j.zlen = (uint32_t)(an + bn - 1);
        j.Z = (uint32_t)(an + bn);
        j.nb = carry_blocks(j.Z);
        job_layout(j, ws + off, (r3 ? 3u : 1u) << L);
        off += job_ws(jobs[i]);
        k.n++;
        if (j.nb > k.nbmax) k.nbmax = j.nb;
        if (j.zlen < k.zmin) k.zmin = j.zlen;
        k.as += j.an;
        k.bs += j.bn;
        k.zs += j.Z;
      }
      k.plan = r3 ? plan_for3(L, sq, k.n) : plan_for(L, sq, k.n);
      cls[ncls++] = k;
    }
  return ncls;
}

static int jobs_in_band(const sba_job *jobs, int njobs) {
  if (njobs <= 0 || !jobs) return 0;
  for (int i = 0; i < njobs; ++i)
    if (!in_band_direct(jobs[i].an, jobs[i].bp ? jobs[i].bn : jobs[i].an)) return 0;
  return 1;
}

static int jobs_lmax(const sba_job *jobs, int njobs) {
  int Lmax = 0;
  for (int i = 0; i < njobs; ++i) {
    const uint64_t an = (uint64_t)jobs[i].an, bn = jobs[i].bp ? (uint64_t)jobs[i].bn : an;
    const int L = lg_for(an + bn - 1);
    if (L > Lmax) Lmax = L;
  }
  return Lmax;
}

struct sba_batch {
  int dev;
  void *ws;                        // all jobs' workspaces
  Job *djobs_all;
  int nclass;
  ClassDesc cls[SBA_MAX_CLASSES];
  hipStream_t cap_stream;          // private stream used for capture
  hipGraph_t graph;
  hipGraphExec_t exec;
  int use_graph;
};

static int enqueue_classes(DevCtx &c, Job *djobs, const ClassDesc *cls, int ncls,
                           hipStream_t st) {
  for (int k = 0; k < ncls; ++k) {
    JobRef jr;
    jr.arr = djobs + cls[k].start;
    memset(&jr.one, 0, sizeof jr.one);
    if (cls[k].school == 1) {
      {
        PROF("school", 16.0 * cls[k].zs);
        KLAUNCH(k_school, dim3(cls[k].n, 1, 1), dim3(NT), 0, st, jr);
      }
      HIPCHK(hipGetLastError());
      continue;
    }
    if (cls[k].school == 2) {
      {
        PROF("school_long", 16.0 * cls[k].zs);
        KLAUNCH(k_school_long, dim3(cls[k].nbmax, 1, cls[k].n), dim3(NT), 0, st, jr);
      }
      {
        PROF("carry_blocks", 8.0 * cls[k].zs / CARRY_LPB);
        const uint32_t g = carry_grid(cls[k].nbmax);
        KLAUNCH(k_carry_blocks, dim3(g, 1, cls[k].n), dim3(SCAN_NT), 0, st, jr);
      }
      HIPCHK(hipGetLastError());
      continue;
    }
    const int rc =
        cls[k].r3 ? enqueue_r3(c, set_for(cls[k].L, 1), cls[k].L, cls[k].sqr, cls[k].plan, jr,
                               cls[k].n, cls[k].nbmax,
                               cls[k].as, cls[k].bs, cls[k].zs, st)
                  : enqueue_class(c, set_for(cls[k].L, 0), cls[k].L, cls[k].sqr, cls[k].plan, jr,
                                  cls[k].n, cls[k].nbmax, cls[k].zmin, cls[k].as, cls[k].bs,
                                  cls[k].zs, st);
    if (rc != SBA_OK) return rc;
  }
  return SBA_OK;
}

extern "C" void sba_batch_destroy(sba_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->ws) (void)hipFree(b->ws);
  if (b->djobs_all) (void)hipFree(b->djobs_all);
  free(b);
}

// Each class takes the plan stored for (L, kind, bucket of its job count)
// when the batch is created: run the autotuner on the jobs first to have
// measured plans captured.
extern "C" sba_batch *sba_batch_create(const sba_job *jobs, int njobs, int use_graph,
                                       int *status) {
  int st_dummy;
  if (!status) status = &st_dummy;
  *status = SBA_OK;
  if (!jobs_in_band(jobs, njobs)) {
    *status = SBA_BAND;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  DevCtx *c;
  if ((*status = get_ctx(&c)) != 0) return nullptr;
  sba_batch *b = (sba_batch *)calloc(1, sizeof(sba_batch));
  if (!b) { *status = SBA_ERR_HIP; return nullptr; }
  b->use_graph = use_graph;
  (void)hipGetDevice(&b->dev);
  size_t total = 0;
  for (int i = 0; i < njobs; ++i) total += job_ws(jobs[i]);
  int rc = ctx_twiddles(*c, jobs_lmax(jobs, njobs), 0);
  if (rc) { free(b); *status = rc; return nullptr; }
  if (hipMalloc(&b->ws, total) != hipSuccess ||
      hipMalloc(&b->djobs_all, sizeof(Job) * (size_t)njobs) != hipSuccess) {
    sba_batch_destroy(b);
    *status = SBA_ERR_HIP;
    return nullptr;
  }
  Job *hj = (Job *)calloc((size_t)njobs, sizeof(Job));
  if (!hj) { sba_batch_destroy(b); *status = SBA_ERR_HIP; return nullptr; }
  b->nclass = classify_jobs(jobs, njobs, (char *)b->ws, hj, b->cls, 1);
  hipError_t e = hipMemcpy(b->djobs_all, hj, sizeof(Job) * (size_t)njobs, hipMemcpyHostToDevice);
  free(hj);
  if (e != hipSuccess) { sba_batch_destroy(b); *status = SBA_ERR_HIP; return nullptr; }
  if (use_graph) {
    if (hipStreamCreate(&b->cap_stream) != hipSuccess ||
        hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      sba_batch_destroy(b);
      *status = SBA_ERR_HIP;
      return nullptr;
    }
    const int saved = g_prof_on;
    g_prof_on = 0;                               // no event records inside the graph
    rc = enqueue_classes(*c, b->djobs_all, b->cls, b->nclass, b->cap_stream);
    g_prof_on = saved;
    hipError_t e2 = hipStreamEndCapture(b->cap_stream, &b->graph);
    if (rc != SBA_OK || e2 != hipSuccess ||
        hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
      sba_batch_destroy(b);
      *status = rc != SBA_OK ? rc : SBA_ERR_HIP;
      return nullptr;
    }
  }
  return b;
}

extern "C" int sba_batch_run(sba_batch *b, void *stream) {
  if (!b) return SBA_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  if (b->use_graph) {
    HIPCHK(hipGraphLaunch(b->exec, st));
    return SBA_OK;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  return enqueue_classes(*c, b->djobs_all, b->cls, b->nclass, st);
}

extern "C" int sba_batch_classes(const sba_batch *b) { return b ? b->nclass : 0; }

// ---------------------------------------------------------------------------
// one-shot batch launch into a caller workspace (serving: no allocation, no
// graph; the job table is staged through the workspace's tail)
// ---------------------------------------------------------------------------
extern "C" size_t sba_batch_ws_bytes(const sba_job *jobs, int njobs) {
  if (!jobs_in_band(jobs, njobs)) return 0;
  size_t t = 0;
  for (int i = 0; i < njobs; ++i) t += job_ws(jobs[i]);
  return t + table_bytes(njobs);
}

// stage the job table of `jobs` (classified) at the workspace tail; on
// success *dj is the device table and cls/ncls describe the classes
static int stage_jobs(const sba_job *jobs, int njobs, void *ws, size_t need, hipStream_t st,
                      Job **dj, ClassDesc *cls, int *ncls, int school) {
  *dj = (Job *)((char *)ws + need - table_bytes(njobs));
  Job *hj = (Job *)calloc((size_t)njobs, sizeof(Job));
  if (!hj) return SBA_ERR_HIP;
  *ncls = classify_jobs(jobs, njobs, (char *)ws, hj, cls, school);
  hipError_t e = hipMemcpyAsync(*dj, hj, sizeof(Job) * (size_t)njobs, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);   // hj is pageable and freed below
  free(hj);
  return e == hipSuccess ? SBA_OK : SBA_ERR_HIP;
}

extern "C" int sba_batch_launch_ws(const sba_job *jobs, int njobs, void *ws, size_t ws_bytes,
                                   void *stream) {
  const size_t need = sba_batch_ws_bytes(jobs, njobs);
  if (!need) return SBA_BAND;
  if (!ws || ws_bytes < need) return SBA_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  if ((rc = ctx_twiddles(*c, jobs_lmax(jobs, njobs), st)) != 0) return rc;
  Job *dj;
  ClassDesc cls[SBA_MAX_CLASSES];
  int ncls = 0;
  if ((rc = stage_jobs(jobs, njobs, ws, need, st, &dj, cls, &ncls, 1)) != SBA_OK) return rc;
  return enqueue_classes(*c, dj, cls, ncls, st);
}

// ---------------------------------------------------------------------------
// plan table access and plan timing (the autotuner's measurements)
// ---------------------------------------------------------------------------
extern "C" int sba_job_class(const sba_job *j, int *L, int *kind) {
  if (!j || !in_band_direct(j->an, j->bp ? j->bn : j->an)) return SBA_BAND;
  int r3;
  job_key(*j, L, &r3);
  *kind = (j->bp == nullptr ? 1 : 0) | (r3 ? 2 : 0);
  return SBA_OK;
}

extern "C" void sba_set_school_max(uint64_t max_product) {
  g_school_pmax.store(max_product, std::memory_order_relaxed);
}
extern "C" uint64_t sba_get_school_max(void) {
  return g_school_pmax.load(std::memory_order_relaxed);
}

extern "C" int sba_plan_bucket(int njobs) { return plan_bucket(njobs < 1 ? 1u : (uint32_t)njobs); }

static bool kind_plan_ok(int L, int kind, const Plan &p) {
  return (kind & 2) ? plan_ok3(L, p) : plan_ok(L, p);
}

extern "C" int sba_plan_set(int L, int kind, int bucket, int last, int bm, int split) {
  Plan p = {last, bm, split};
  if (L < 10 || L > SBA_MAX_LG || bucket < 0 || bucket >= SBA_PLAN_BUCKETS ||
      !kind_plan_ok(L, kind, p))
    return SBA_BAND;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_plan_init) plan_defaults();
  ((kind & 2) ? g_plan3 : g_plan)[kind & 1][L][bucket] = p;
  return SBA_OK;
}

extern "C" int sba_plan_get(int L, int kind, int bucket, int *last, int *bm, int *split) {
  if (L < 10 || L > SBA_MAX_LG || bucket < 0 || bucket >= SBA_PLAN_BUCKETS) return SBA_BAND;
  if ((kind & 2) && (L < R3_MIN_L || L > R3_MAX_L)) return SBA_BAND;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_plan_init) plan_defaults();
  const Plan &p = ((kind & 2) ? g_plan3 : g_plan)[kind & 1][L][bucket];
  *last = p.last;
  *bm = p.bm;
  *split = p.split;
  return SBA_OK;
}

extern "C" void sba_plan_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  plan_defaults();
}

extern "C" int sba_plan_heuristic(int L, int *last, int *bm, int *split) {
  if (L < 10 || L > SBA_MAX_LG) return SBA_BAND;
  const Plan p = heuristic_plan(L);
  *last = p.last;
  *bm = p.bm;
  *split = p.split;
  return SBA_OK;
}

extern "C" int sba_plan_candidates(int L, int kind, int *plans, int max) {
  int n = 0;
  for (int last = 10; last <= 13; ++last)
    for (int bm = 3; bm <= 5; ++bm)
      for (int split = 0; split <= 1; ++split) {
        const Plan p = {last, bm, split};
        if (!kind_plan_ok(L, kind, p)) continue;
        if (plans && n < max) {
          plans[3 * n] = p.last;
          plans[3 * n + 1] = p.bm;
          plans[3 * n + 2] = p.split;
        }
        ++n;
      }
  return n;
}

extern "C" int sba_class_time(const sba_job *jobs, int njobs, int last, int bm, int split,
                              int reps, void *ws, size_t ws_bytes, void *stream, float *ms) {
  const size_t need = sba_batch_ws_bytes(jobs, njobs);
  if (!need || reps < 1 || !ms) return SBA_BAND;
  if (!ws || ws_bytes < need) return SBA_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  if ((rc = ctx_twiddles(*c, jobs_lmax(jobs, njobs), st)) != 0) return rc;
  Job *dj;
  ClassDesc cls[SBA_MAX_CLASSES];
  int ncls = 0;
  // plans are NTT plans: small jobs are timed on the NTT here, not the schoolbook
  if ((rc = stage_jobs(jobs, njobs, ws, need, st, &dj, cls, &ncls, 0)) != SBA_OK) return rc;
  if (ncls != 1) return SBA_BAND;                 // one size class per measurement
  cls[0].plan = Plan{last, bm, split};
  if (!kind_plan_ok(cls[0].L, cls[0].r3 ? 2 : 0, cls[0].plan)) return SBA_BAND;
  const int saved = g_prof_on;
  g_prof_on = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = enqueue_classes(*c, dj, cls, 1, st);       // warm-up
  if (rc == SBA_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = SBA_ERR_HIP;
  if (rc == SBA_OK && hipEventRecord(e0, st) != hipSuccess) rc = SBA_ERR_HIP;
  for (int r = 0; r < reps && rc == SBA_OK; ++r) rc = enqueue_classes(*c, dj, cls, 1, st);
  if (rc == SBA_OK && hipEventRecord(e1, st) != hipSuccess) rc = SBA_ERR_HIP;
  if (rc == SBA_OK && hipEventSynchronize(e1) != hipSuccess) rc = SBA_ERR_HIP;
  float t = 0.f;
  if (rc == SBA_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = SBA_ERR_HIP;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  g_prof_on = saved;
  if (rc == SBA_OK) *ms = t / (float)reps;
  return rc;
}

// the plain per-call path over the same jobs: one sba_gpu_mul_dev /
// sba_gpu_sqr_dev call per job, in order, on one stream
extern "C" int sba_gpu_mul_loop(const sba_job *jobs, int njobs, void *stream) {
  for (int i = 0; i < njobs; ++i) {
    const sba_job &j = jobs[i];
    const int rc = j.bp ? sba_gpu_mul_dev(j.rp, j.ap, j.an, j.bp, j.bn, stream)
                        : sba_gpu_sqr_dev(j.rp, j.ap, j.an, stream);
    if (rc != SBA_OK) return rc;
  }
  return SBA_OK;
}

extern "C" int sba_gpu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
