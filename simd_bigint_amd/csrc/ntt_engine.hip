// ntt_engine.hip -- gfx950 kernels of the multi-prime NTT multiply and the
// thin C-ABI layer that launches them (declared in sba_gpu.h).
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
//     coefficient into 3 u64 words; a 3-kernel carry-lookahead scan over a
//     4-state carry map adds the words into limbs bit-exactly.
#include "gpu_arith.h"
#include "simd_bigint_amd.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>

// ---------------------------------------------------------------------------
// kernel argument block (by value)
// ---------------------------------------------------------------------------
struct PrimeArgs {
  uint32_t p[SBA_K], pinv[SBA_K], r2[SBA_K];
  const uint32_t *wn[SBA_K];
  const uint32_t *wni[SBA_K];
};

struct GarnerArgs {
  uint32_t p[SBA_K], pinv[SBA_K];
  uint32_t scale[SBA_K];          // N^-1 * R^2 mod p (undoes N and the pointwise R^-1)
  uint32_t cm[SBA_K][SBA_K];      // (prod_{j<k} p_j mod p_q) * R mod p_q
  uint32_t im[SBA_K];             // (prod_{j<q} p_j)^-1 * R mod p_q
};

#define NT 256
// occupancy hint for the transform kernels (waves per SIMD) and the widest
// register chunk (radix 2^SBA_BMAX), fixed by measurement
#ifndef SBA_WAVES
#define SBA_WAVES 1
#endif
#ifndef SBA_BMAX
#define SBA_BMAX 5
#endif
__device__ __forceinline__ uint32_t lpad(uint32_t i) { return i + (i >> 5); }

// blockIdx -> tile so that consecutive tiles (which share cache lines and
// twiddle rows) run on the same XCD: the dispatcher hands block b to XCD
// b % 8.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  if (n & 7) return b;
  return (b & 7) * (n >> 3) + (b >> 3);
}

template <int X> struct ILog2 { static constexpr int v = 1 + ILog2<X / 2>::v; };
template <> struct ILog2<1> { static constexpr int v = 0; };

// residue of a limb, lazy in [0, 2p): (hi * 2^32 mod p) + (lo mod p).
// p > 2^30 so lo < 4p: two folds bring lo into [0, p).
__device__ __forceinline__ uint32_t limb_residue(uint64_t x, uint32_t p, uint32_t pinv,
                                                 uint32_t r2) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  uint32_t h = red1(mont_mul(hi, r2, p, pinv), p);
  uint32_t l2 = lo - 2u * p;
  uint32_t l = l2 < lo ? l2 : lo;
  return h + red1(l, p);
}

// ---------------------------------------------------------------------------
// one chunk of B levels (local levels T0 .. T0+B-1 of an ELL-level pass)
//
// LDS index of tile element (k, c) is (k << LOGTC) | c, padded by one word
// per 32 (lpad) against bank conflicts. A chunk's group of 2^B elements has
// stride D = 2^(LB + LOGTC) words; when D >= 32, or when the whole group
// sits inside one 32-word row (D * 2^B <= 32), the padded addresses are
// base + m * const, so one address register serves the group (ds offsets).
// ---------------------------------------------------------------------------
template <int ELL, int LOGTC, int T0, int B> struct ChunkGeom {
  static constexpr int LB = ELL - T0 - B;
  static constexpr int LOGD = LB + LOGTC;
  static constexpr bool LINEAR = LOGD >= 5 || LOGD + B <= 5;
  static constexpr int DP = LOGD >= 5 ? (1 << LOGD) + (1 << (LOGD - 5)) : (1 << LOGD);
};

template <int ELL, int LOGTC, int T0, int B, bool INV>
__device__ __forceinline__ void tile_chunk(uint32_t *lds, const uint32_t *__restrict__ tw,
                                           uint32_t H, uint32_t p, uint32_t pinv) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int GPT = NE / ((1 << B) * NT);
  using G = ChunkGeom<ELL, LOGTC, T0, B>;
  constexpr int LB = G::LB;
  static_assert(GPT >= 1, "chunk too wide for the tile");
  __syncthreads();
#pragma unroll
  for (int gi = 0; gi < GPT; ++gi) {
    const uint32_t g = threadIdx.x + gi * NT;
    const uint32_t c = g & (TC - 1);
    const uint32_t rest = g >> LOGTC;
    const uint32_t lo = rest & ((1u << LB) - 1);
    const uint32_t hb = rest >> LB;
    const uint32_t kbase = (hb << (ELL - T0)) | lo;
    const uint32_t hj = (H << T0) | hb;
    const uint32_t i0 = (kbase << LOGTC) | c;
    uint32_t *l0 = lds + lpad(i0);
    uint32_t v[1 << B];
#pragma unroll
    for (int m = 0; m < (1 << B); ++m)
      v[m] = G::LINEAR ? l0[m * G::DP] : lds[lpad(i0 + ((uint32_t)m << G::LOGD))];
    if (!INV) {
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const int half = 1 << (B - u - 1);
#pragma unroll
        for (int m2 = 0; m2 < (1 << u); ++m2) {
          const uint32_t w = tw[(hj << u) | (uint32_t)m2];
#pragma unroll
          for (int r = 0; r < half; ++r) {
            const int m = (m2 << (B - u)) | r;
            bfly_fwd(v[m], v[m + half], w, p, pinv);
          }
        }
      }
    } else {
#pragma unroll
      for (int u = B - 1; u >= 0; --u) {
        const int half = 1 << (B - u - 1);
#pragma unroll
        for (int m2 = 0; m2 < (1 << u); ++m2) {
          const uint32_t w = tw[(hj << u) | (uint32_t)m2];
#pragma unroll
          for (int r = 0; r < half; ++r) {
            const int m = (m2 << (B - u)) | r;
            bfly_inv(v[m], v[m + half], w, p, pinv);
          }
        }
      }
    }
#pragma unroll
    for (int m = 0; m < (1 << B); ++m) {
      if (G::LINEAR) l0[m * G::DP] = v[m];
      else lds[lpad(i0 + ((uint32_t)m << G::LOGD))] = v[m];
    }
  }
}

// chunk sizes: the largest B <= BM whose group addressing is linear,
// falling back to BM when no size is
template <int ELL, int LOGTC> struct ChunkPlan {
  static constexpr int E = (1 << (ELL + LOGTC)) / NT;
  static constexpr int LE = ILog2<E>::v;
  static constexpr int BM = LE < SBA_BMAX ? LE : SBA_BMAX;
  static constexpr int pick(int t0) {
    int r = ELL - t0;
    int bmax = r < BM ? r : BM;
    for (int b = bmax; b >= 1; --b) {
      int logd = ELL - t0 - b + LOGTC;
      if (logd >= 5 || logd + b <= 5) return b;
    }
    return bmax;
  }
};

template <int ELL, int LOGTC, int T0>
__device__ __forceinline__ void tile_fwd(uint32_t *lds, const uint32_t *tw, uint32_t H,
                                         uint32_t p, uint32_t pinv) {
  if constexpr (T0 < ELL) {
    constexpr int B = ChunkPlan<ELL, LOGTC>::pick(T0);
    tile_chunk<ELL, LOGTC, T0, B, false>(lds, tw, H, p, pinv);
    tile_fwd<ELL, LOGTC, T0 + B>(lds, tw, H, p, pinv);
  }
}

template <int ELL, int LOGTC, int T0>
__device__ __forceinline__ void tile_inv(uint32_t *lds, const uint32_t *tw, uint32_t H,
                                         uint32_t p, uint32_t pinv) {
  if constexpr (T0 < ELL) {
    constexpr int B = ChunkPlan<ELL, LOGTC>::pick(T0);
    tile_inv<ELL, LOGTC, T0 + B>(lds, tw, H, p, pinv);
    tile_chunk<ELL, LOGTC, T0, B, true>(lds, tw, H, p, pinv);
  }
}

// ---------------------------------------------------------------------------
// pass kernels
// ---------------------------------------------------------------------------
// First forward pass (global levels [0, ELL)): limbs -> residues of all K
// primes -> ELL levels -> spec[q][.]. Tiles = column tiles (H = 0).
template <int ELL, int LOGTC>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_first_fwd(PrimeArgs pa, const uint64_t *__restrict__ limbs,
                                                  uint64_t nlimbs, uint32_t *__restrict__ spec,
                                                  int L) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[NE + NE / 32];
  const uint32_t N = 1u << L;             // N <= 2^25: 32-bit offsets throughout
  const uint32_t S = 1u << (L - ELL);
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t col0 = tile << LOGTC;
  const uint32_t nl = nlimbs < N ? (uint32_t)nlimbs : N;
  uint32_t *lt = lds + lpad(threadIdx.x);
  // element e = tid + i*NT sits at col0 + toff + uoff(i): a per-thread part
  // and a uniform part (scalar base + 32-bit lane offset addressing)
  const uint32_t toff = (threadIdx.x >> LOGTC) * S + (threadIdx.x & (TC - 1));
#pragma unroll 1
  for (int q = 0; q < SBA_K; ++q) {
    const uint32_t p = pa.p[q], pinv = pa.pinv[q], r2 = pa.r2[q];
    __syncthreads();
    // the tile's limbs are re-read per prime (L2-resident after the first)
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
      const uint64_t *lp = limbs + uoff;
      lt[i * (NT + NT / 32)] = limb_residue(uoff + toff < nl ? lp[toff] : 0, p, pinv, r2);
    }
    tile_fwd<ELL, LOGTC, 0>(lds, pa.wn[q], 0, p, pinv);
    __syncthreads();
    uint32_t *out = spec + q * N;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
      (out + uoff)[toff] = lt[i * (NT + NT / 32)];
    }
  }
}

// In-place pass over global levels [s0, s0 + ELL) of one prime
// (blockIdx.y), forward or inverse.
template <int ELL, int LOGTC, bool INV>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_pass(PrimeArgs pa, uint32_t *__restrict__ spec, int L,
                                             int s0) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[NE + NE / 32];
  const int q = blockIdx.y;
  const uint32_t N = 1u << L;
  const uint32_t S = 1u << (L - s0 - ELL);
  const uint32_t ctiles = S >> LOGTC;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t H = tile / ctiles;
  const uint32_t col0 = (tile % ctiles) << LOGTC;
  const uint32_t base = (H << (L - s0)) + col0;
  uint32_t *buf = spec + q * N + base;
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  uint32_t *lt = lds + lpad(threadIdx.x);
  const uint32_t toff = (threadIdx.x >> LOGTC) * S + (threadIdx.x & (TC - 1));
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
    lt[i * (NT + NT / 32)] = (buf + uoff)[toff];
  }
  if (!INV) tile_fwd<ELL, LOGTC, 0>(lds, pa.wn[q], H, p, pinv);
  else tile_inv<ELL, LOGTC, 0>(lds, pa.wni[q], H, p, pinv);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
    (buf + uoff)[toff] = lt[i * (NT + NT / 32)];
  }
}

// Last forward levels of b + pointwise product with a's spectrum (or b's own
// square) + first inverse levels, on contiguous tiles (S = 1, TC = 1).
// LIMB_IN: the pass is the whole transform (s0 = 0) and reads b's limbs.
template <int ELL, bool LIMB_IN, bool SQR>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_fused(PrimeArgs pa, const uint64_t *__restrict__ blimbs,
                                              uint64_t bn, const uint32_t *__restrict__ aspec,
                                              uint32_t *__restrict__ bspec, int L) {
  constexpr int NE = 1 << ELL;
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[NE + NE / 32];
  const int q = blockIdx.y;
  const uint32_t N = 1u << L;
  const uint32_t H = blockIdx.x;
  const uint32_t base = H << ELL;
  const uint32_t nb = bn < N ? (uint32_t)bn : N;
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  uint32_t *buf = bspec + q * N + base;
  const uint32_t *as = aspec + q * N + base;
  const uint64_t *bl = blimbs + base;
  uint32_t *lt = lds + lpad(threadIdx.x);
  const uint32_t tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    uint32_t v;
    if (LIMB_IN) {
      v = limb_residue(base + i * NT + tid < nb ? (bl + i * NT)[tid] : 0, p, pinv, pa.r2[q]);
    } else {
      v = (buf + i * NT)[tid];
    }
    lt[i * (NT + NT / 32)] = v;
  }
  tile_fwd<ELL, 0, 0>(lds, pa.wn[q], H, p, pinv);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t v = lt[i * (NT + NT / 32)];
    const uint32_t o = SQR ? red1(v, p) : red1((as + i * NT)[tid], p);
    lt[i * (NT + NT / 32)] = mont_mul(v, o, p, pinv);
  }
  tile_inv<ELL, 0, 0>(lds, pa.wni[q], H, p, pinv);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) (buf + i * NT)[tid] = lt[i * (NT + NT / 32)];
}

// ---------------------------------------------------------------------------
// Garner: the five residues of coefficient i -> 3 u64 words
// ---------------------------------------------------------------------------
__device__ __forceinline__ void garner_coef(const GarnerArgs &ga, const uint32_t *__restrict__ spec,
                                            uint32_t N, uint32_t i, uint64_t &c0, uint64_t &c1,
                                            uint64_t &c2) {
  uint32_t u[SBA_K], x[SBA_K];
#pragma unroll
  for (int q = 0; q < SBA_K; ++q)
    u[q] = red1(mont_mul(spec[q * N + i], ga.scale[q], ga.p[q], ga.pinv[q]), ga.p[q]);
  x[0] = u[0];
#pragma unroll
  for (int q = 1; q < SBA_K; ++q) {
    const uint32_t p = ga.p[q], pinv = ga.pinv[q];
    uint32_t t = x[0];
#pragma unroll
    for (int k = 1; k < q; ++k) {
      t += red1(mont_mul(x[k], ga.cm[q][k], p, pinv), p);
      t = red1(t, p);
    }
    const uint32_t d = red1(u[q] + p - t, p);
    x[q] = red1(mont_mul(d, ga.im[q], p, pinv), p);
  }
  // Horner: c = (((x4 p3 + x3) p2 + x2) p1 + x1) p0 + x0  (< 2^155)
  unsigned __int128 c = x[4];
  c = c * ga.p[3] + x[3];
  c = c * ga.p[2] + x[2];
  c = c * ga.p[1] + x[1];
  const unsigned __int128 t0 = (unsigned __int128)(uint64_t)c * ga.p[0] + x[0];
  const unsigned __int128 t1 =
      (unsigned __int128)(uint64_t)(c >> 64) * ga.p[0] + (uint64_t)(t0 >> 64);
  c0 = (uint64_t)t0;
  c1 = (uint64_t)t1;
  c2 = (uint64_t)(t1 >> 64);
}

// ---------------------------------------------------------------------------
// carry propagation. Limb j receives s_j = w0[j] + w1[j-1] + w2[j-2]
// (< 3 * 2^64) = v_j + 2^64 h_j. With incoming carry c in {0..3}:
// out_j = (v_j + c) mod 2^64 and c' = h_j + [c >= 2^64 - v_j] <= 3. A limb
// is thus a map {0..3} -> {0..3} (2 bits per entry, 8 bits); maps compose
// associatively, so the carries are an exclusive scan of maps:
//   k_garner_maps  Garner + limb sums + limb maps + block maps (LDS-staged)
//   k_carry_blocks scan of block maps -> carry into each block
//   k_carry_fix    per-limb carries from the maps, out_j = v_j + c_j
// ---------------------------------------------------------------------------
#define CARRY_LPT 8                     // limbs per thread in the map scan
#define CARRY_LPB (CARRY_LPT * NT)      // limbs per block
#define MAP_ID 0xE4u                    // c -> c

__device__ __forceinline__ uint32_t map_apply(uint32_t f, uint32_t c) { return (f >> (2 * c)) & 3u; }
// first f, then g
__device__ __forceinline__ uint32_t map_then(uint32_t f, uint32_t g) {
  uint32_t r = 0;
#pragma unroll
  for (uint32_t c = 0; c < 4; ++c) r |= map_apply(g, map_apply(f, c)) << (2 * c);
  return r;
}
__device__ __forceinline__ uint32_t limb_map(uint64_t v, uint32_t h) {
  const uint64_t room = ~v;                       // 2^64 - 1 - v
  const uint32_t z = room < 3 ? (uint32_t)room + 1 : 4u;   // overflow iff c >= z
  uint32_t f = 0;
#pragma unroll
  for (uint32_t c = 0; c < 4; ++c) f |= (h + (c >= z ? 1u : 0u)) << (2 * c);
  return f;
}

// block-wide exclusive scan of maps; returns the exclusive prefix of this
// thread and writes the block total into *total (all threads).
__device__ uint32_t block_scan_maps(uint32_t f, uint32_t *total) {
  __shared__ uint32_t wtot[NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = f;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc = map_then(o, inc);
  }
  uint32_t exc = __shfl_up(inc, 1, 64);
  if (lane == 0) exc = MAP_ID;
  if (lane == 63) wtot[wid] = inc;
  __syncthreads();
  uint32_t pre = MAP_ID;
  for (int w = 0; w < wid; ++w) pre = map_then(pre, wtot[w]);
  uint32_t tot = MAP_ID;
  for (int w = 0; w < NT / 64; ++w) tot = map_then(tot, wtot[w]);
  *total = tot;
  __syncthreads();
  return map_then(pre, exc);
}

__global__ __launch_bounds__(NT) void k_garner_maps(GarnerArgs ga, const uint32_t *__restrict__ spec,
                                                    uint32_t N, uint32_t zlen, uint32_t Z,
                                                    uint64_t *__restrict__ rp,
                                                    uint8_t *__restrict__ lmaps,
                                                    uint32_t *__restrict__ bmaps) {
  __shared__ uint64_t ws0[CARRY_LPB + 2], ws1[CARRY_LPB + 2], ws2[CARRY_LPB + 2];
  __shared__ uint64_t mp64[CARRY_LPB / 8];
  uint8_t *mp = (uint8_t *)mp64;
  const uint32_t B0 = blockIdx.x * CARRY_LPB;
  // coefficients B0-2 .. B0+CARRY_LPB-1 -> LDS slot j+2 (coalesced residue reads)
#pragma unroll 2
  for (int k = 0; k < CARRY_LPT; ++k) {
    const uint32_t j = threadIdx.x + k * NT;
    const uint32_t i = B0 + j;
    uint64_t c0 = 0, c1 = 0, c2 = 0;
    if (i < zlen) garner_coef(ga, spec, N, i, c0, c1, c2);
    ws0[j + 2] = c0; ws1[j + 2] = c1; ws2[j + 2] = c2;
  }
  if (threadIdx.x < 2) {
    const uint32_t j = threadIdx.x;            // slot j holds coefficient B0 - 2 + j
    uint64_t c0 = 0, c1 = 0, c2 = 0;
    if (B0 >= 2 - j) {
      const uint32_t i = B0 - 2 + j;
      if (i < zlen) garner_coef(ga, spec, N, i, c0, c1, c2);
    }
    ws0[j] = c0; ws1[j] = c1; ws2[j] = c2;
  }
  __syncthreads();
  // limb sums (coalesced writes of v_j into the product), limb maps in LDS
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) {
    const uint32_t j = threadIdx.x + k * NT;
    const uint32_t L = B0 + j;
    uint32_t f = MAP_ID;
    if (L < Z) {
      const uint64_t a = ws0[j + 2], b = ws1[j + 1], c = ws2[j];
      const uint64_t s1 = a + b;
      uint32_t h = s1 < a ? 1u : 0u;
      const uint64_t s2 = s1 + c;
      h += s2 < s1 ? 1u : 0u;
      rp[L] = s2;
      f = limb_map(s2, h);
    }
    mp[j] = (uint8_t)f;
  }
  __syncthreads();
  // each thread composes its 8 consecutive limb maps; block total
  const uint64_t m8 = mp64[threadIdx.x];
  *(uint64_t *)(lmaps + (size_t)B0 + threadIdx.x * CARRY_LPT) = m8;
  uint32_t f = MAP_ID;
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) f = map_then(f, (uint32_t)(m8 >> (8 * k)) & 0xFFu);
  uint32_t tot;
  block_scan_maps(f, &tot);
  if (threadIdx.x == 0) bmaps[blockIdx.x] = tot;
}

// one block: carry into every block = exclusive scan of block maps applied to 0
__global__ __launch_bounds__(NT) void k_carry_blocks(const uint32_t *__restrict__ bmaps,
                                                     uint32_t nb, uint32_t *__restrict__ bcin) {
  const uint32_t per = (nb + NT - 1) / NT;
  const uint32_t b0 = threadIdx.x * per;
  uint32_t f = MAP_ID;
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < nb) f = map_then(f, bmaps[b0 + k]);
  uint32_t tot;
  const uint32_t pre = block_scan_maps(f, &tot);
  uint32_t c = map_apply(pre, 0);
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < nb) {
      bcin[b0 + k] = c;
      c = map_apply(bmaps[b0 + k], c);
    }
}

__global__ __launch_bounds__(NT) void k_carry_fix(uint64_t *__restrict__ rp,
                                                  const uint8_t *__restrict__ lmaps, uint32_t Z,
                                                  const uint32_t *__restrict__ bcin) {
  __shared__ uint64_t cj64[CARRY_LPB / 8];
  uint8_t *cj = (uint8_t *)cj64;
  const uint32_t B0 = blockIdx.x * CARRY_LPB;
  const uint64_t m8 = *(const uint64_t *)(lmaps + (size_t)B0 + threadIdx.x * CARRY_LPT);
  uint32_t f = MAP_ID;
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) f = map_then(f, (uint32_t)(m8 >> (8 * k)) & 0xFFu);
  uint32_t tot;
  const uint32_t pre = block_scan_maps(f, &tot);
  uint32_t c = map_apply(pre, bcin[blockIdx.x]);
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) {
    cj[threadIdx.x * CARRY_LPT + k] = (uint8_t)c;
    c = map_apply((uint32_t)(m8 >> (8 * k)) & 0xFFu, c);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) {
    const uint32_t j = threadIdx.x + k * NT;
    const uint32_t L = B0 + j;
    const uint32_t cc = cj[j];
    if (L < Z && cc) rp[L] += cc;
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
static const uint32_t kPrimes[SBA_K] = {1107296257u, 1711276033u, 1811939329u, 2013265921u,
                                        2113929217u};

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

struct DevCtx {
  int init = 0;
  int tw_lg = 0;                       // tables hold 2^(tw_lg-1) node twiddles
  uint32_t *wn[SBA_K] = {}, *wni[SBA_K] = {};
  PrimeArgs pa;
  GarnerArgs ga_base;
  // grow-only workspace
  void *ws = nullptr;
  size_t ws_bytes = 0;
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

static int ctx_init(DevCtx &c) {
  if (c.init) return 0;
  for (int q = 0; q < SBA_K; ++q) {
    const uint64_t p = kPrimes[q];
    uint32_t inv = 1;
    for (int i = 0; i < 5; ++i) inv *= 2u - (uint32_t)p * inv;   // p^-1 mod 2^32
    c.pa.p[q] = (uint32_t)p;
    c.pa.pinv[q] = 0u - inv;
    c.pa.r2[q] = (uint32_t)(((unsigned __int128)1 << 64) % p);
    c.ga_base.p[q] = (uint32_t)p;
    c.ga_base.pinv[q] = 0u - inv;
  }
  for (int q = 1; q < SBA_K; ++q) {
    const uint64_t pq = kPrimes[q];
    uint64_t pp = 1;
    for (int k = 0; k < q; ++k) {
      c.ga_base.cm[q][k] = h_mont(pp, pq);
      pp = pp * (kPrimes[k] % pq) % pq;
    }
    c.ga_base.im[q] = h_mont(h_invm(pp, pq), pq);
  }
  c.init = 1;
  return 0;
}

static int ctx_twiddles(DevCtx &c, int L, hipStream_t st) {
  if (L <= c.tw_lg) return 0;
  const int Lt = SBA_MAX_LG;            // build once for the largest transform
  const uint32_t count = 1u << (Lt - 1);
  for (int q = 0; q < SBA_K; ++q) {
    if (c.wn[q]) { HIPCHK(hipFree(c.wn[q])); HIPCHK(hipFree(c.wni[q])); }
    HIPCHK(hipMalloc(&c.wn[q], (size_t)count * 4));
    HIPCHK(hipMalloc(&c.wni[q], (size_t)count * 4));
    const uint64_t p = kPrimes[q];
    const uint64_t g = h_primroot(p);
    TwArgs ta;
    ta.p = (uint32_t)p;
    ta.pinv = c.pa.pinv[q];
    ta.one = h_mont(1, p);
    for (int t = 0; t < SBA_MAX_LG; ++t) {
      // bit t of a node index carries zeta_{t+2}, a root of order 2^(t+2)
      const int ord = t + 2;
      uint64_t z = ord <= SBA_MAX_LG ? h_powm(g, (p - 1) >> ord, p) : 1;
      ta.z[t] = h_mont(z, p);
      ta.zi[t] = h_mont(h_invm(z, p), p);
    }
    hipLaunchKernelGGL(k_twiddles, dim3((count + NT - 1) / NT), dim3(NT), 0, st, ta, c.wn[q],
                       c.wni[q], count);
    HIPCHK(hipGetLastError());
    c.pa.wn[q] = c.wn[q];
    c.pa.wni[q] = c.wni[q];
  }
  HIPCHK(hipStreamSynchronize(st));
  c.tw_lg = Lt;
  return 0;
}

// workspace layout for transform length N: aspec, bspec (K*N u32 each),
// limb carry maps (one byte per limb, padded to whole blocks), block maps and
// block carry-ins (one u32 per block of CARRY_LPB limbs each)
static size_t ws_need(size_t N) {
  const size_t nb = N / CARRY_LPB + 2;
  return 2 * SBA_K * N * 4 + nb * CARRY_LPB + 2 * nb * 4 + 4096;
}

static int lg_for(uint64_t zlen) {
  int L = 10;                           // smallest transform the kernels take
  while (((uint64_t)1 << L) < zlen) ++L;
  return L;
}

extern "C" size_t sba_gpu_workspace_bytes(int64_t an, int64_t bn) {
  if (an <= 0 || bn <= 0) return 0;
  const int L = lg_for((uint64_t)an + (uint64_t)bn - 1);
  if (L > SBA_MAX_LG) return 0;
  return ws_need((size_t)1 << L);
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
typedef void (*first_fn)(PrimeArgs, const uint64_t *, uint64_t, uint32_t *, int);
typedef void (*pass_fn)(PrimeArgs, uint32_t *, int, int);
typedef void (*fused_fn)(PrimeArgs, const uint64_t *, uint64_t, const uint32_t *, uint32_t *, int);

// first pass of the two-pass plan: ELL = L - 13 levels, TC = 2^(13 - ELL)
static first_fn first_two_pass(int ell) {
  switch (ell) {
    case 1: return k_first_fwd<1, 12>;
    case 2: return k_first_fwd<2, 11>;
    case 3: return k_first_fwd<3, 10>;
    case 4: return k_first_fwd<4, 9>;
    case 5: return k_first_fwd<5, 8>;
    case 6: return k_first_fwd<6, 7>;
    case 7: return k_first_fwd<7, 6>;
    case 8: return k_first_fwd<8, 5>;
    case 9: return k_first_fwd<9, 4>;
    case 10: return k_first_fwd<10, 3>;
    case 11: return k_first_fwd<11, 2>;
    case 12: return k_first_fwd<12, 1>;
  }
  return nullptr;
}
static pass_fn inv_first_two_pass(int ell) {
  switch (ell) {
    case 1: return k_pass<1, 12, true>;
    case 2: return k_pass<2, 11, true>;
    case 3: return k_pass<3, 10, true>;
    case 4: return k_pass<4, 9, true>;
    case 5: return k_pass<5, 8, true>;
    case 6: return k_pass<6, 7, true>;
    case 7: return k_pass<7, 6, true>;
    case 8: return k_pass<8, 5, true>;
    case 9: return k_pass<9, 4, true>;
    case 10: return k_pass<10, 3, true>;
    case 11: return k_pass<11, 2, true>;
    case 12: return k_pass<12, 1, true>;
  }
  return nullptr;
}
static first_fn first_single(int L) {
  switch (L) {
    case 10: return k_first_fwd<10, 0>;
    case 11: return k_first_fwd<11, 0>;
    case 12: return k_first_fwd<12, 0>;
    case 13: return k_first_fwd<13, 0>;
  }
  return nullptr;
}
static fused_fn fused_single(int L, int sqr) {
  switch (L) {
    case 10: return sqr ? k_fused<10, true, true> : k_fused<10, true, false>;
    case 11: return sqr ? k_fused<11, true, true> : k_fused<11, true, false>;
    case 12: return sqr ? k_fused<12, true, true> : k_fused<12, true, false>;
    case 13: return sqr ? k_fused<13, true, true> : k_fused<13, true, false>;
  }
  return nullptr;
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
struct ProfScope {
  int rec = -1;
  hipStream_t st;
  ProfScope(const char *name, double bytes, hipStream_t s) : st(s) {
    if (!g_prof_on || g_prof_n >= PROF_MAX_REC) return;
    rec = g_prof_n++;
    ProfRec &r = g_prof_rec[rec];
    r.name = prof_name(name);
    r.bytes = bytes;
    r.e0 = prof_event();
    r.e1 = prof_event();
    (void)hipEventRecord(r.e0, st);
  }
  ~ProfScope() {
    if (rec >= 0) (void)hipEventRecord(g_prof_rec[rec].e1, st);
  }
};
#define PROF(name, bytes) ProfScope prof_scope_##__LINE__(name, (double)(bytes), st)

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

#define LAST_ELL 13

static int run_mul(DevCtx &c, uint64_t *rp, const uint64_t *ap, uint64_t an, const uint64_t *bp,
                   uint64_t bn, int sqr, void *ws, hipStream_t st) {
  const uint64_t zlen = an + bn - 1;
  const uint64_t Z = an + bn;
  const int L = lg_for(zlen);
  const size_t N = (size_t)1 << L;
  uint32_t *aspec = (uint32_t *)ws;
  uint32_t *bspec = aspec + SBA_K * N;
  const uint32_t nb = (uint32_t)((Z + CARRY_LPB - 1) / CARRY_LPB);
  uint8_t *lmaps = (uint8_t *)(bspec + SBA_K * N);
  uint32_t *bmaps = (uint32_t *)(lmaps + (size_t)nb * CARRY_LPB);
  uint32_t *bcin = bmaps + nb + 1;
  const PrimeArgs &pa = c.pa;

  const double KN4 = 4.0 * SBA_K * (double)N;
  if (L <= LAST_ELL) {
    if (!sqr) {
      PROF("first_fwd", 8.0 * an + KN4);
      hipLaunchKernelGGL(first_single(L), dim3(1), dim3(NT), 0, st, pa, ap, an, aspec, L);
    }
    {
      PROF("fused_last", 8.0 * bn + (sqr ? 1 : 2) * KN4);
      hipLaunchKernelGGL(fused_single(L, sqr), dim3(1, SBA_K), dim3(NT), 0, st, pa, bp, bn,
                         (const uint32_t *)aspec, bspec, L);
    }
  } else {
    const int ell1 = L - LAST_ELL;
    const uint32_t ntile1 = (uint32_t)(N >> 13);
    first_fn f1 = first_two_pass(ell1);
    if (!sqr) {
      PROF("first_fwd", 8.0 * an + KN4);
      hipLaunchKernelGGL(f1, dim3(ntile1), dim3(NT), 0, st, pa, ap, an, aspec, L);
    }
    {
      PROF("first_fwd", 8.0 * bn + KN4);
      hipLaunchKernelGGL(f1, dim3(ntile1), dim3(NT), 0, st, pa, bp, bn, bspec, L);
    }
    const uint32_t ntile2 = (uint32_t)(N >> LAST_ELL);
    if (!sqr) {
      PROF("fwd_last", 2 * KN4);
      hipLaunchKernelGGL((k_pass<LAST_ELL, 0, false>), dim3(ntile2, SBA_K), dim3(NT), 0, st, pa,
                         aspec, L, L - LAST_ELL);
    }
    {
      PROF("fused_last", (sqr ? 2 : 3) * KN4);
      if (sqr)
        hipLaunchKernelGGL((k_fused<LAST_ELL, false, true>), dim3(ntile2, SBA_K), dim3(NT), 0,
                           st, pa, (const uint64_t *)nullptr, (uint64_t)0,
                           (const uint32_t *)aspec, bspec, L);
      else
        hipLaunchKernelGGL((k_fused<LAST_ELL, false, false>), dim3(ntile2, SBA_K), dim3(NT), 0,
                           st, pa, (const uint64_t *)nullptr, (uint64_t)0,
                           (const uint32_t *)aspec, bspec, L);
    }
    {
      PROF("inv_first", 2 * KN4);
      hipLaunchKernelGGL(inv_first_two_pass(ell1), dim3(ntile1, SBA_K), dim3(NT), 0, st, pa,
                         bspec, L, 0);
    }
  }
  GarnerArgs ga = c.ga_base;
  for (int q = 0; q < SBA_K; ++q) {
    const uint64_t p = kPrimes[q];
    const uint64_t ninv = h_invm(N % p, p);
    // scale = N^-1 * R^2 mod p
    ga.scale[q] = (uint32_t)((unsigned __int128)ninv * (((unsigned __int128)1 << 64) % p) % p);
  }
  {
    PROF("garner_maps", 4.0 * SBA_K * zlen + 9.0 * Z);
    hipLaunchKernelGGL(k_garner_maps, dim3(nb), dim3(NT), 0, st, ga, (const uint32_t *)bspec,
                       (uint32_t)N, (uint32_t)zlen, (uint32_t)Z, rp, lmaps, bmaps);
  }
  {
    PROF("carry_blocks", 8.0 * nb);
    hipLaunchKernelGGL(k_carry_blocks, dim3(1), dim3(NT), 0, st, (const uint32_t *)bmaps, nb,
                       bcin);
  }
  {
    PROF("carry_fix", 17.0 * Z);
    hipLaunchKernelGGL(k_carry_fix, dim3(nb), dim3(NT), 0, st, rp, (const uint8_t *)lmaps,
                       (uint32_t)Z, (const uint32_t *)bcin);
  }
  HIPCHK(hipGetLastError());
  return SBA_OK;
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

static int mul_common(uint64_t *rp, const uint64_t *ap, int64_t an, const uint64_t *bp,
                      int64_t bn, int sqr, void *ws, size_t ws_bytes, void *stream) {
  if (an <= 0 || bn <= 0) return SBA_BAND;
  if (an < bn && !sqr) {
    const uint64_t *t = ap; ap = bp; bp = t;
    int64_t tn = an; an = bn; bn = tn;
  }
  const uint64_t mn = (uint64_t)(an < bn ? an : bn);
  if (mn > sba_gpu_min_operand_cap()) return SBA_BAND;
  const int L = lg_for((uint64_t)an + (uint64_t)bn - 1);
  if (L > SBA_MAX_LG) return SBA_BAND;
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  rc = ctx_twiddles(*c, L, st);
  if (rc) return rc;
  const size_t need = ws_need((size_t)1 << L);
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
  return run_mul(*c, rp, ap, (uint64_t)an, sqr ? ap : bp, (uint64_t)bn, sqr, ws, st);
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

extern "C" int sba_gpu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
