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
// 1: the first forward pass runs one prime per block (grid y = prime)
#ifndef SBA_FIRST_SPLIT
#define SBA_FIRST_SPLIT 1
#endif
// 1: widest chunks even where group addressing is not linear
#ifndef SBA_CHUNK_GREEDY
#define SBA_CHUNK_GREEDY 0
#endif
// LDS padding: PW words after every 32-word row. A column tile (TC columns
// of 2^ELL rows) pads by TC words for TC <= 16 -- the last chunk then puts
// the TC columns x 32/TC groups of a half-wave on 32 distinct banks -- and
// by 1 word otherwise (the contiguous tiles).
template <int LOGTC> struct Pad {
  static constexpr int PW = LOGTC <= 4 ? (1 << LOGTC) : 1;
  static constexpr int words(int ne) { return ne + (ne / 32) * PW; }
};
template <int LOGTC> __device__ __forceinline__ uint32_t lpadt(uint32_t i) {
  return i + (i >> 5) * Pad<LOGTC>::PW;
}

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
  static constexpr int DP =
      LOGD >= 5 ? (1 << LOGD) + (1 << (LOGD - 5)) * Pad<LOGTC>::PW : (1 << LOGD);
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
    uint32_t *l0 = lds + lpadt<LOGTC>(i0);
    uint32_t v[1 << B];
#pragma unroll
    for (int m = 0; m < (1 << B); ++m)
      v[m] = G::LINEAR ? l0[m * G::DP] : lds[lpadt<LOGTC>(i0 + ((uint32_t)m << G::LOGD))];
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
      else lds[lpadt<LOGTC>(i0 + ((uint32_t)m << G::LOGD))] = v[m];
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
    if (SBA_CHUNK_GREEDY) return bmax;
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
// job descriptors: every kernel serves one multiply (JobRef.one) or a whole
// size class of a batch (JobRef.arr[blockIdx.z]); all jobs of one launch
// share the transform length N = 2^L.
// ---------------------------------------------------------------------------
struct Job {
  const uint64_t *a, *b;          // operand limbs (device)
  uint64_t *r;                    // product limbs (device), an+bn of them
  uint32_t *aspec, *bspec;        // K x N residues each
  uint8_t *lmaps;                 // one carry map per limb, padded to blocks
  uint32_t *bmaps, *bcin;         // per carry block
  uint32_t an, bn, zlen, Z, nb;   // limbs, coefficients, product limbs, carry blocks
};
// global-address-space views of descriptor pointers: loads through a pointer
// read from memory are otherwise generic (flat) and serialised against LDS
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((address_space(1))) const uint32_t gc_u32;
typedef __attribute__((address_space(1))) uint64_t g_u64;
typedef __attribute__((address_space(1))) const uint64_t gc_u64;
typedef __attribute__((address_space(1))) uint8_t g_u8;

struct JobRef {
  const Job *arr;
  Job one;
};
// The descriptor is uniform across the block; readfirstlane moves every field
// into scalar registers so tile addresses stay scalar-base + lane-offset.
template <typename T> __device__ __forceinline__ T *uni_ptr(T *p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T *)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint32_t uni_u32(uint32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}
__device__ __forceinline__ Job get_job(const JobRef &jr) {
  Job j;
  if (jr.arr) {
    j = jr.arr[blockIdx.z];
  } else {
    j = jr.one;
  }
  j.a = uni_ptr(j.a); j.b = uni_ptr(j.b); j.r = uni_ptr(j.r);
  j.aspec = uni_ptr(j.aspec); j.bspec = uni_ptr(j.bspec); j.lmaps = uni_ptr(j.lmaps);
  j.bmaps = uni_ptr(j.bmaps); j.bcin = uni_ptr(j.bcin);
  j.an = uni_u32(j.an); j.bn = uni_u32(j.bn); j.zlen = uni_u32(j.zlen); j.Z = uni_u32(j.Z);
  j.nb = uni_u32(j.nb);
  return j;
}

// ---------------------------------------------------------------------------
// pass kernels
// ---------------------------------------------------------------------------
// First forward pass (global levels [0, ELL)): limbs -> residues of all K
// primes -> ELL levels -> spec[q][.]. Tiles = column tiles (H = 0).
// which = 0: operand a into aspec; 1: operand b into bspec.
template <int ELL, int LOGTC>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_first_fwd(PrimeArgs pa, JobRef jr, int which,
                                                             int L) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[Pad<LOGTC>::words(NE)];
  const Job jb = get_job(jr);
  gc_u64 *limbs = (gc_u64 *)(which ? jb.b : jb.a);
  g_u32 *spec = (g_u32 *)(which ? jb.bspec : jb.aspec);
  const uint32_t nlimbs = which ? jb.bn : jb.an;
  const uint32_t N = 1u << L;             // N <= 2^25: 32-bit offsets throughout
  const uint32_t S = 1u << (L - ELL);
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t col0 = tile << LOGTC;
  const uint32_t nl = nlimbs < N ? nlimbs : N;
  uint32_t *lt = lds + lpadt<LOGTC>(threadIdx.x);
  // element e = tid + i*NT sits at col0 + toff + uoff(i): a per-thread part
  // and a uniform part (scalar base + 32-bit lane offset addressing)
  const uint32_t toff = (threadIdx.x >> LOGTC) * S + (threadIdx.x & (TC - 1));
  // SBA_FIRST_SPLIT: one prime per block (grid y), else all primes per block
  const int q0 = SBA_FIRST_SPLIT ? (int)blockIdx.y : 0;
  const int q1 = SBA_FIRST_SPLIT ? q0 + 1 : SBA_K;
#pragma unroll 1
  for (int q = q0; q < q1; ++q) {
    const uint32_t p = pa.p[q], pinv = pa.pinv[q], r2 = pa.r2[q];
    __syncthreads();
    // the tile's limbs are re-read per prime (cache-resident after the first)
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
      gc_u64 *lp = limbs + uoff;
      lt[i * Pad<LOGTC>::words(NT)] = limb_residue(uoff + toff < nl ? lp[toff] : 0, p, pinv, r2);
    }
    tile_fwd<ELL, LOGTC, 0>(lds, pa.wn[q], 0, p, pinv);
    __syncthreads();
    g_u32 *out = spec + q * N;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
      (out + uoff)[toff] = lt[i * Pad<LOGTC>::words(NT)];
    }
  }
}

// In-place pass over global levels [s0, s0 + ELL) of one prime
// (blockIdx.y), forward or inverse, on aspec (which = 0) or bspec (1).
template <int ELL, int LOGTC, bool INV>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_pass(PrimeArgs pa, JobRef jr, int which, int L,
                                                        int s0) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[Pad<LOGTC>::words(NE)];
  const Job jb = get_job(jr);
  g_u32 *spec = (g_u32 *)(which ? jb.bspec : jb.aspec);
  const int q = blockIdx.y;
  const uint32_t N = 1u << L;
  const uint32_t S = 1u << (L - s0 - ELL);
  const uint32_t ctiles = S >> LOGTC;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t H = tile / ctiles;
  const uint32_t col0 = (tile % ctiles) << LOGTC;
  const uint32_t base = (H << (L - s0)) + col0;
  g_u32 *buf = spec + q * N + base;
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  uint32_t *lt = lds + lpadt<LOGTC>(threadIdx.x);
  const uint32_t toff = (threadIdx.x >> LOGTC) * S + (threadIdx.x & (TC - 1));
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
    lt[i * Pad<LOGTC>::words(NT)] = (buf + uoff)[toff];
  }
  if (!INV) tile_fwd<ELL, LOGTC, 0>(lds, pa.wn[q], H, p, pinv);
  else tile_inv<ELL, LOGTC, 0>(lds, pa.wni[q], H, p, pinv);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
    (buf + uoff)[toff] = lt[i * Pad<LOGTC>::words(NT)];
  }
}

// Last forward levels of b + pointwise product with a's spectrum (or b's own
// square) + first inverse levels, on contiguous tiles (S = 1, TC = 1).
// LIMB_IN: the pass is the whole transform (s0 = 0) and reads b's limbs.
template <int ELL, bool LIMB_IN, bool SQR>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_fused(PrimeArgs pa, JobRef jr, int L) {
  constexpr int NE = 1 << ELL;
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[Pad<0>::words(NE)];
  const Job jb = get_job(jr);
  const int q = blockIdx.y;
  const uint32_t N = 1u << L;
  const uint32_t H = blockIdx.x;
  const uint32_t base = H << ELL;
  const uint32_t nb = jb.bn < N ? jb.bn : N;
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  g_u32 *buf = (g_u32 *)jb.bspec + q * N + base;
  gc_u32 *as = (gc_u32 *)jb.aspec + q * N + base;
  gc_u64 *bl = (gc_u64 *)jb.b + base;
  uint32_t *lt = lds + lpadt<0>(threadIdx.x);
  const uint32_t tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    uint32_t v;
    if (LIMB_IN) {
      v = limb_residue(base + i * NT + tid < nb ? (bl + i * NT)[tid] : 0, p, pinv, pa.r2[q]);
    } else {
      v = (buf + i * NT)[tid];
    }
    lt[i * Pad<0>::words(NT)] = v;
  }
  tile_fwd<ELL, 0, 0>(lds, pa.wn[q], H, p, pinv);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t v = lt[i * Pad<0>::words(NT)];
    const uint32_t o = SQR ? red1(v, p) : red1((as + i * NT)[tid], p);
    lt[i * Pad<0>::words(NT)] = mont_mul(v, o, p, pinv);
  }
  tile_inv<ELL, 0, 0>(lds, pa.wni[q], H, p, pinv);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) (buf + i * NT)[tid] = lt[i * Pad<0>::words(NT)];
}

// ---------------------------------------------------------------------------
// Garner: the five residues of coefficient i -> 3 u64 words
// ---------------------------------------------------------------------------
__device__ __forceinline__ void garner_coef(const GarnerArgs &ga, gc_u32 *spec,
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

__global__ __launch_bounds__(NT) void k_garner_maps(GarnerArgs ga, JobRef jr, uint32_t N) {
  __shared__ uint64_t ws0[CARRY_LPB + 2], ws1[CARRY_LPB + 2], ws2[CARRY_LPB + 2];
  __shared__ uint64_t mp64[CARRY_LPB / 8];
  const Job jb = get_job(jr);
  if (blockIdx.x >= jb.nb) return;                 // this job has fewer carry blocks
  gc_u32 *spec = (gc_u32 *)jb.bspec;
  const uint32_t zlen = jb.zlen, Z = jb.Z;
  g_u64 *rp = (g_u64 *)jb.r;
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
    const uint32_t Lb = B0 + j;
    uint32_t f = MAP_ID;
    if (Lb < Z) {
      const uint64_t a = ws0[j + 2], b = ws1[j + 1], c = ws2[j];
      const uint64_t s1 = a + b;
      uint32_t h = s1 < a ? 1u : 0u;
      const uint64_t s2 = s1 + c;
      h += s2 < s1 ? 1u : 0u;
      rp[Lb] = s2;
      f = limb_map(s2, h);
    }
    mp[j] = (uint8_t)f;
  }
  __syncthreads();
  // each thread composes its 8 consecutive limb maps; block total
  const uint64_t m8 = mp64[threadIdx.x];
  *(g_u64 *)((g_u8 *)jb.lmaps + (size_t)B0 + threadIdx.x * CARRY_LPT) = m8;
  uint32_t f = MAP_ID;
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) f = map_then(f, (uint32_t)(m8 >> (8 * k)) & 0xFFu);
  uint32_t tot;
  block_scan_maps(f, &tot);
  if (threadIdx.x == 0) ((g_u32 *)jb.bmaps)[blockIdx.x] = tot;
}

// one block per job: carry into every block = exclusive scan of block maps
__global__ __launch_bounds__(NT) void k_carry_blocks(JobRef jr) {
  const Job jb = get_job(jr);
  const uint32_t nb = jb.nb;
  gc_u32 *bmaps = (gc_u32 *)jb.bmaps;
  g_u32 *bcin = (g_u32 *)jb.bcin;
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

__global__ __launch_bounds__(NT) void k_carry_fix(JobRef jr) {
  __shared__ uint64_t cj64[CARRY_LPB / 8];
  const Job jb = get_job(jr);
  if (blockIdx.x >= jb.nb) return;
  g_u64 *rp = (g_u64 *)jb.r;
  const uint32_t Z = jb.Z;
  uint8_t *cj = (uint8_t *)cj64;
  const uint32_t B0 = blockIdx.x * CARRY_LPB;
  const uint64_t m8 = *(gc_u64 *)((g_u8 *)jb.lmaps + (size_t)B0 + threadIdx.x * CARRY_LPT);
  uint32_t f = MAP_ID;
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) f = map_then(f, (uint32_t)(m8 >> (8 * k)) & 0xFFu);
  uint32_t tot;
  const uint32_t pre = block_scan_maps(f, &tot);
  uint32_t c = map_apply(pre, ((gc_u32 *)jb.bcin)[blockIdx.x]);
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) {
    cj[threadIdx.x * CARRY_LPT + k] = (uint8_t)c;
    c = map_apply((uint32_t)(m8 >> (8 * k)) & 0xFFu, c);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < CARRY_LPT; ++k) {
    const uint32_t j = threadIdx.x + k * NT;
    const uint32_t Lb = B0 + j;
    const uint32_t cc = cj[j];
    if (Lb < Z && cc) rp[Lb] += cc;
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
typedef void (*first_fn)(PrimeArgs, JobRef, int, int);
typedef void (*pass_fn)(PrimeArgs, JobRef, int, int, int);
typedef void (*fused_fn)(PrimeArgs, JobRef, int);

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

// carry blocks of a product of Z limbs
static uint32_t carry_blocks(uint64_t Z) { return (uint32_t)((Z + CARRY_LPB - 1) / CARRY_LPB); }

// Enqueue the whole multiply for `njobs` jobs sharing N = 2^L (jr.arr ==
// nullptr: the single job jr.one). nbmax = the largest carry-block count;
// zsum / asum / bsum feed the profiler's algorithmic byte counts.
static int enqueue_class(DevCtx &c, int L, int sqr, const JobRef &jr, uint32_t njobs,
                         uint32_t nbmax, double asum, double bsum, double zsum, hipStream_t st) {
  const uint32_t N = 1u << L;
  const PrimeArgs &pa = c.pa;
  const double KN4 = 4.0 * SBA_K * (double)N * njobs;
  if (L <= LAST_ELL) {
    if (!sqr) {
      PROF("first_fwd", 8.0 * asum + KN4);
      hipLaunchKernelGGL(first_single(L), dim3(1, SBA_FIRST_SPLIT ? SBA_K : 1, njobs), dim3(NT),
                         0, st, pa, jr, 0, L);
    }
    {
      PROF("fused_last", 8.0 * bsum + (sqr ? 1 : 2) * KN4);
      hipLaunchKernelGGL(fused_single(L, sqr), dim3(1, SBA_K, njobs), dim3(NT), 0, st, pa, jr,
                         L);
    }
  } else {
    const int ell1 = L - LAST_ELL;
    const uint32_t ntile1 = N >> 13;
    const uint32_t ntile2 = N >> LAST_ELL;
    first_fn f1 = first_two_pass(ell1);
    if (!sqr) {
      PROF("first_fwd", 8.0 * asum + KN4);
      hipLaunchKernelGGL(f1, dim3(ntile1, SBA_FIRST_SPLIT ? SBA_K : 1, njobs), dim3(NT), 0, st,
                         pa, jr, 0, L);
    }
    {
      PROF("first_fwd", 8.0 * bsum + KN4);
      hipLaunchKernelGGL(f1, dim3(ntile1, SBA_FIRST_SPLIT ? SBA_K : 1, njobs), dim3(NT), 0, st,
                         pa, jr, 1, L);
    }
    if (!sqr) {
      PROF("fwd_last", 2 * KN4);
      hipLaunchKernelGGL((k_pass<LAST_ELL, 0, false>), dim3(ntile2, SBA_K, njobs), dim3(NT), 0,
                         st, pa, jr, 0, L, L - LAST_ELL);
    }
    {
      PROF("fused_last", (sqr ? 2 : 3) * KN4);
      if (sqr)
        hipLaunchKernelGGL((k_fused<LAST_ELL, false, true>), dim3(ntile2, SBA_K, njobs),
                           dim3(NT), 0, st, pa, jr, L);
      else
        hipLaunchKernelGGL((k_fused<LAST_ELL, false, false>), dim3(ntile2, SBA_K, njobs),
                           dim3(NT), 0, st, pa, jr, L);
    }
    {
      PROF("inv_first", 2 * KN4);
      hipLaunchKernelGGL(inv_first_two_pass(ell1), dim3(ntile1, SBA_K, njobs), dim3(NT), 0, st,
                         pa, jr, 1, L, 0);
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
    PROF("garner_maps", 4.0 * SBA_K * zsum + 9.0 * zsum);
    hipLaunchKernelGGL(k_garner_maps, dim3(nbmax, 1, njobs), dim3(NT), 0, st, ga, jr, N);
  }
  {
    PROF("carry_blocks", 8.0 * zsum / CARRY_LPB);
    hipLaunchKernelGGL(k_carry_blocks, dim3(1, 1, njobs), dim3(NT), 0, st, jr);
  }
  {
    PROF("carry_fix", 17.0 * zsum);
    hipLaunchKernelGGL(k_carry_fix, dim3(nbmax, 1, njobs), dim3(NT), 0, st, jr);
  }
  HIPCHK(hipGetLastError());
  return SBA_OK;
}

// carve one job's buffers out of a workspace region of ws_need(N) bytes
static void job_layout(Job &j, void *ws, uint32_t N) {
  j.aspec = (uint32_t *)ws;
  j.bspec = j.aspec + (size_t)SBA_K * N;
  j.lmaps = (uint8_t *)(j.bspec + (size_t)SBA_K * N);
  const size_t nbcap = N / CARRY_LPB + 2;
  j.bmaps = (uint32_t *)(j.lmaps + nbcap * CARRY_LPB);
  j.bcin = j.bmaps + nbcap;
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

static int in_band(int64_t an, int64_t bn) {
  if (an <= 0 || bn <= 0) return 0;
  const uint64_t mn = (uint64_t)(an < bn ? an : bn);
  if (mn > sba_gpu_min_operand_cap()) return 0;
  return lg_for((uint64_t)an + (uint64_t)bn - 1) <= SBA_MAX_LG;
}

static int mul_common(uint64_t *rp, const uint64_t *ap, int64_t an, const uint64_t *bp,
                      int64_t bn, int sqr, void *ws, size_t ws_bytes, void *stream) {
  if (!in_band(an, bn)) return SBA_BAND;
  if (an < bn && !sqr) {
    const uint64_t *t = ap; ap = bp; bp = t;
    int64_t tn = an; an = bn; bn = tn;
  }
  const int L = lg_for((uint64_t)an + (uint64_t)bn - 1);
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
  job_layout(j, ws, 1u << L);
  return enqueue_class(*c, L, sqr, jr, 1, j.nb, (double)an, (double)bn, (double)j.Z, st);
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
// batches: many independent multiplies of uneven sizes. Jobs are grouped in
// size classes (same N = 2^L, same mul/sqr kind); one launch sequence per
// class covers all its jobs (grid z = job). The whole batch is captured once
// into a hipGraph and replayed by sba_batch_run.
// ---------------------------------------------------------------------------
struct BatchClass {
  int L, sqr;
  uint32_t njobs, nbmax;
  Job *djobs;                      // device array
  double asum, bsum, zsum;
};

struct sba_batch {
  int dev;
  void *ws;                        // all jobs' workspaces
  Job *djobs_all;
  int nclass;
  BatchClass cls[2 * (SBA_MAX_LG + 1)];
  hipStream_t cap_stream;          // private stream used for capture
  hipGraph_t graph;
  hipGraphExec_t exec;
  int use_graph;
};

static int batch_enqueue(sba_batch *b, hipStream_t st) {
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  for (int k = 0; k < b->nclass; ++k) {
    BatchClass &bc = b->cls[k];
    JobRef jr;
    jr.arr = bc.djobs;
    memset(&jr.one, 0, sizeof jr.one);
    rc = enqueue_class(*c, bc.L, bc.sqr, jr, bc.njobs, bc.nbmax, bc.asum, bc.bsum, bc.zsum, st);
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

extern "C" sba_batch *sba_batch_create(const sba_job *jobs, int njobs, int use_graph,
                                       int *status) {
  int st_dummy;
  if (!status) status = &st_dummy;
  *status = SBA_OK;
  if (njobs <= 0) { *status = SBA_BAND; return nullptr; }
  for (int i = 0; i < njobs; ++i)
    if (!in_band(jobs[i].an, jobs[i].bp ? jobs[i].bn : jobs[i].an)) {
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
  // classify: key (L, sqr); jobs ordered by class
  int Lmax = 0;
  size_t total = 0;
  int count[2][SBA_MAX_LG + 1] = {};
  for (int i = 0; i < njobs; ++i) {
    const int sq = jobs[i].bp == nullptr;
    const int L = lg_for((uint64_t)jobs[i].an + (uint64_t)(sq ? jobs[i].an : jobs[i].bn) - 1);
    count[sq][L]++;
    if (L > Lmax) Lmax = L;
    total += (ws_need((size_t)1 << L) + 255) & ~(size_t)255;
  }
  int rc = ctx_twiddles(*c, Lmax, 0);
  if (rc) { free(b); *status = rc; return nullptr; }
  if (hipMalloc(&b->ws, total) != hipSuccess ||
      hipMalloc(&b->djobs_all, sizeof(Job) * (size_t)njobs) != hipSuccess) {
    sba_batch_destroy(b);
    *status = SBA_ERR_HIP;
    return nullptr;
  }
  Job *hj = (Job *)calloc((size_t)njobs, sizeof(Job));
  int pos = 0;
  size_t off = 0;
  for (int sq = 0; sq < 2; ++sq)
    for (int L = 0; L <= SBA_MAX_LG; ++L) {
      if (!count[sq][L]) continue;
      BatchClass &bc = b->cls[b->nclass++];
      bc.L = L;
      bc.sqr = sq;
      bc.njobs = 0;
      bc.nbmax = 0;
      bc.djobs = b->djobs_all + pos;
      bc.asum = bc.bsum = bc.zsum = 0;
      for (int i = 0; i < njobs; ++i) {
        const int isq = jobs[i].bp == nullptr;
        const uint64_t an = (uint64_t)jobs[i].an, bn = isq ? an : (uint64_t)jobs[i].bn;
        if (isq != sq || lg_for(an + bn - 1) != L) continue;
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
        job_layout(j, (char *)b->ws + off, 1u << L);
        off += (ws_need((size_t)1 << L) + 255) & ~(size_t)255;
        bc.njobs++;
        if (j.nb > bc.nbmax) bc.nbmax = j.nb;
        bc.asum += j.an;
        bc.bsum += j.bn;
        bc.zsum += j.Z;
      }
    }
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
    rc = batch_enqueue(b, b->cap_stream);
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
  return batch_enqueue(b, st);
}

extern "C" int sba_batch_classes(const sba_batch *b) { return b ? b->nclass : 0; }

// ---------------------------------------------------------------------------
// one-shot batch launch into a caller workspace (serving: no allocation, no
// graph; the job table is staged through the workspace's tail)
// ---------------------------------------------------------------------------
static size_t job_ws(const sba_job &j) {
  const uint64_t an = (uint64_t)j.an, bn = j.bp ? (uint64_t)j.bn : an;
  return (ws_need((size_t)1 << lg_for(an + bn - 1)) + 255) & ~(size_t)255;
}

extern "C" size_t sba_batch_ws_bytes(const sba_job *jobs, int njobs) {
  size_t t = 0;
  for (int i = 0; i < njobs; ++i) {
    if (!in_band(jobs[i].an, jobs[i].bp ? jobs[i].bn : jobs[i].an)) return 0;
    t += job_ws(jobs[i]);
  }
  return t + (((size_t)njobs * sizeof(Job) + 255) & ~(size_t)255);
}

extern "C" int sba_batch_launch_ws(const sba_job *jobs, int njobs, void *ws, size_t ws_bytes,
                                   void *stream) {
  if (njobs <= 0) return SBA_BAND;
  const size_t need = sba_batch_ws_bytes(jobs, njobs);
  if (!need) return SBA_BAND;
  if (!ws || ws_bytes < need) return SBA_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  DevCtx *c;
  int rc = get_ctx(&c);
  if (rc) return rc;
  int Lmax = 0;
  for (int i = 0; i < njobs; ++i) {
    const uint64_t an = (uint64_t)jobs[i].an, bn = jobs[i].bp ? (uint64_t)jobs[i].bn : an;
    const int L = lg_for(an + bn - 1);
    if (L > Lmax) Lmax = L;
  }
  if ((rc = ctx_twiddles(*c, Lmax, st)) != 0) return rc;
  // job table at the tail of the workspace
  Job *dj = (Job *)((char *)ws + need - (((size_t)njobs * sizeof(Job) + 255) & ~(size_t)255));
  Job *hj = (Job *)calloc((size_t)njobs, sizeof(Job));
  if (!hj) return SBA_ERR_HIP;
  struct Cls { int L, sqr, start, n; uint32_t nbmax; double as, bs, zs; } cls[2 * (SBA_MAX_LG + 1)];
  int ncls = 0, pos = 0;
  size_t off = 0;
  for (int sq = 0; sq < 2; ++sq)
    for (int L = 0; L <= SBA_MAX_LG; ++L) {
      Cls k = {L, sq, pos, 0, 0, 0, 0, 0};
      for (int i = 0; i < njobs; ++i) {
        const int isq = jobs[i].bp == nullptr;
        const uint64_t an = (uint64_t)jobs[i].an, bn = isq ? an : (uint64_t)jobs[i].bn;
        if (isq != sq || lg_for(an + bn - 1) != L) continue;
        Job &j = hj[pos++];
        const bool swap = !isq && an < bn;
        j.a = swap ? jobs[i].bp : jobs[i].ap;
        j.b = isq ? jobs[i].ap : (swap ? jobs[i].ap : jobs[i].bp);
        j.an = (uint32_t)(swap ? bn : an);
        j.bn = (uint32_t)(swap ? an : bn);
        j.r = jobs[i].rp;
        j.zlen = (uint32_t)(an + bn - 1);
        j.Z = (uint32_t)(an + bn);
        j.nb = carry_blocks(j.Z);
        job_layout(j, (char *)ws + off, 1u << L);
        off += job_ws(jobs[i]);
        k.n++;
        if (j.nb > k.nbmax) k.nbmax = j.nb;
        k.as += j.an;
        k.bs += j.bn;
        k.zs += j.Z;
      }
      if (k.n) cls[ncls++] = k;
    }
  hipError_t e = hipMemcpyAsync(dj, hj, sizeof(Job) * (size_t)njobs, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);   // hj is pageable and freed below
  free(hj);
  if (e != hipSuccess) return SBA_ERR_HIP;
  for (int k = 0; k < ncls; ++k) {
    JobRef jr;
    jr.arr = dj + cls[k].start;
    memset(&jr.one, 0, sizeof jr.one);
    rc = enqueue_class(*c, cls[k].L, cls[k].sqr, jr, (uint32_t)cls[k].n, cls[k].nbmax, cls[k].as,
                       cls[k].bs, cls[k].zs, st);
    if (rc != SBA_OK) return rc;
  }
  return SBA_OK;
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
