// ntt_tiles.h -- device side of the tiled NTT passes: argument blocks, job
// descriptors, LDS tile machinery and the pass kernels (templates), plus the
// table that maps a pass plan to kernel instantiations. Included by
// ntt_engine.hip and by ntt_pass_bm.hip, which instantiates the kernels of
// one register-chunk radix per object so the three compile in parallel.
#pragma once
#include "gpu_arith.h"
#include "simd_bigint_amd.h"

// ---------------------------------------------------------------------------
// kernel argument block (by value)
// ---------------------------------------------------------------------------
struct PrimeArgs {
  uint32_t p[SBA_K], pinv[SBA_K], r1[SBA_K];   // r1 = 2^32 mod p
  // length-3 DFT constants (Montgomery, canonical): h1 = -1/2,
  // h2 = (w - w^2)/2 for the cube root w, h2i = -h2 (the inverse's)
  uint32_t h1[SBA_K], h2[SBA_K], h2i[SBA_K];
  // applied by the pass that finishes the inverse transform: Nt^-1 * R^4
  // (Nt the transform length; see the host's prime_args)
  uint32_t scale[SBA_K];
  uint32_t inv2[SBA_K];                          // 1/2 (Montgomery, canonical)
  const uint32_t *wn[SBA_K];
  const uint32_t *wni[SBA_K];
};

struct GarnerArgs {
  uint32_t p[SBA_K], pinv[SBA_K];
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
// 1: a multiply's two last forward passes and the pointwise product run as
// one kernel (k_fused2); 0: a's last pass is a separate kernel
#ifndef SBA_MERGE_LAST
#define SBA_MERGE_LAST 1
#endif
// occupancy hints (waves per SIMD) of the merged last pass (radix-32 plans;
// the radix-8/16 variants need more registers than 3 waves leave), the
// first forward pass and the inverse column pass, and the unroll of the
// first pass's limb loads, fixed by measurement
#ifndef SBA_FUSED2_WAVES
#define SBA_FUSED2_WAVES 3
#endif
// butterfly forms of the passes other than the merged last pass: HV_AUTO
// (Harvey's for primes below 2^30) or HV_W
#ifndef SBA_COL_HV
#define SBA_COL_HV HV_AUTO
#endif
// butterfly forms of k_fused (squares, chunks, single-tile products):
// SBA_FUSED_SPLIT 1 instantiates the body once per form and the block takes
// Harvey's for a prime below 2^30 (one uniform branch); 0: SBA_FUSED_HV
#ifndef SBA_FUSED_SPLIT
#define SBA_FUSED_SPLIT 1
#endif
#ifndef SBA_FUSED_HV
#define SBA_FUSED_HV HV_W
#endif
// 1: the merged last pass instantiates its whole body once per form and
// branches once; 0: one body, a branch per chunk (HV_AUTO)
#ifndef SBA_FUSED2_SPLIT
#define SBA_FUSED2_SPLIT 1
#endif
#ifndef SBA_FUSED2_REGS
#define SBA_FUSED2_REGS 1    // merged last pass: both ends and the pointwise product in registers
#endif
#ifndef SBA_COL_REGS
#define SBA_COL_REGS 1       // column passes: level-0 chunk straight from / to registers
#endif
#ifndef SBA_COL_EDGE
#define SBA_COL_EDGE 1       // column passes of >= 8 columns: the chunk at the other end too
#endif
#ifndef SBA_ZERO_LEVEL
#define SBA_ZERO_LEVEL 1     // first level of a half-empty operand as a copy
#endif
#ifndef SBA_TW_EARLY
#define SBA_TW_EARLY 1       // a chunk's twiddle loads issued before its barrier
#endif
#ifndef SBA_FUSED2_EARLY
#define SBA_FUSED2_EARLY 1   // the same in k_fused2 (which holds a's spectrum in registers)
#endif
#ifndef SBA_FIRST_WAVES
#define SBA_FIRST_WAVES SBA_WAVES
#endif
#ifndef SBA_INV_WAVES
#define SBA_INV_WAVES SBA_WAVES
#endif
#ifndef SBA_FIRST_LOAD_UNROLL
#define SBA_FIRST_LOAD_UNROLL 8
#endif
// the same for the 8-column first pass (10 levels in 13-level tiles, the
// 2^23-point transform's): 16 loads in flight, measured 3.6% faster there
// with its occupancy unchanged (3 waves/SIMD); elsewhere 16 would cost
// occupancy
#ifndef SBA_FIRST_LOAD_UNROLL_8COL
#define SBA_FIRST_LOAD_UNROLL_8COL 16
#endif
#ifndef SBA_PASS_LOAD_UNROLL
#define SBA_PASS_LOAD_UNROLL 32
#endif
// residues per first-pass (and inverse column pass) tile, log2, and the row
// padding of 16-column tiles
#ifndef SBA_FIRST_TILE_LG
#define SBA_FIRST_TILE_LG 13
#endif
#ifndef SBA_PAD16
#define SBA_PAD16 16
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
  static constexpr int PW = LOGTC < 4 ? (1 << LOGTC) : (LOGTC == 4 ? SBA_PAD16 : 1);
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

// blockIdx -> (tile, prime) of a first pass with one prime per block (grid
// x = n tiles, y = per primes of the operand). With G > 0 each XCD takes
// its tiles in groups of G neighbours and runs every prime of a group before
// the next group: the primes' re-reads of the group's limbs then come from
// that XCD's L2, and runs of G neighbouring tiles still merge their row
// segments' writes there. G = 0: prime-major order (xcd_tile per prime).
#ifndef SBA_FIRST_GROUP
#define SBA_FIRST_GROUP 32
#endif
// the radix-3 first pass reads three rows per tile: a group of 32 tiles
// holds 6 MiB of limbs, more than an XCD's L2, and measured no faster
#ifndef SBA_R3_FIRST_GROUP
#define SBA_R3_FIRST_GROUP 0
#endif
template <uint32_t G = SBA_FIRST_GROUP>
__device__ __forceinline__ void first_block(uint32_t x, uint32_t y, uint32_t n, uint32_t per,
                                            uint32_t &tile, uint32_t &q) {
  if (G == 0 || per < 2 || n % (8 * G) != 0) {
    tile = xcd_tile(x, n);
    q = y;
    return;
  }
  const uint32_t l = x + n * y;             // dispatch order: x fastest
  const uint32_t s = l >> 3;                // the block's slot on XCD l & 7
  const uint32_t grp = s / (G * per), r = s - grp * (G * per);
  q = r / G;
  tile = (l & 7) * (n >> 3) + grp * G + (r - q * G);
}


// 1: uniform twiddles come through scalar loads
#ifndef SBA_TW_SCALAR
#define SBA_TW_SCALAR 1
#endif

// A workgroup barrier that orders LDS only: every barrier of the tile
// kernels separates LDS writes from other threads' LDS reads, while each
// thread's global loads and stores concern only itself. __syncthreads()
// would also wait for all outstanding global accesses (its fence covers
// every address space), which drains loads issued ahead of a barrier to
// overlap it (next-tile prefetch, early twiddles); the LDS-scoped fences
// wait for LDS accesses (lgkmcnt) alone.
#ifndef SBA_LDS_BARRIER
#define SBA_LDS_BARRIER 1
#endif
__device__ __forceinline__ void lds_barrier() {
  if (SBA_LDS_BARRIER) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  } else {
    __syncthreads();
  }
}

// a compile-time flag passed to generic lambdas (code instantiated once per
// butterfly form)
template <bool V> struct BoolC { static constexpr bool value = V; };

template <int X> struct ILog2 { static constexpr int v = 1 + ILog2<X / 2>::v; };
template <> struct ILog2<1> { static constexpr int v = 0; };

// residue of a limb times 2^-32, lazy in [0, 2p): t = hi * (2^32 mod p) +
// lo = x mod p as an integer below 2^32 p, then one Montgomery reduction
// (t + m p < 2^64 since p < 2^31). Every operand carries the same 2^-32,
// removed with the rest in the Garner scale.
__device__ __forceinline__ uint32_t limb_residue(uint64_t x, uint32_t p, uint32_t pinv,
                                                 uint32_t r1) {
  const uint64_t t = (uint64_t)(uint32_t)(x >> 32) * r1 + (uint32_t)x;
  const uint32_t m = (uint32_t)t * pinv;
  return (uint32_t)(((uint64_t)m * p + t) >> 32);
}

// table entry i (i < 2^30) through a 32-bit byte offset from the table's
// base: the load can take the uniform base in SGPRs and one VGPR offset
__device__ __forceinline__ uint32_t tw_at(const uint32_t *tw, uint32_t i) {
  typedef __attribute__((address_space(1))) const uint8_t gc_byte;
  typedef __attribute__((address_space(1))) const uint32_t gc_word;
  return *(gc_word *)((gc_byte *)tw + (i << 2));
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

// ZL (forward only, first pass of a zero-padded operand): the levels below
// ZL pair every element with a zero and become copies (x, 0) -> (x, x).
// IO: where the chunk's groups come from and go to -- CHUNK_LDS (LDS to
// LDS), CHUNK_TO_REGS (LDS in, results left in vv), CHUNK_FROM_REGS (vv in,
// results to LDS). When a pass's level-0 chunk holds all E of a thread's
// elements (B = log2 E), its groups are exactly the elements the thread
// loads and stores (k = tid / TC + m * NT / TC, column tid % TC), so that
// chunk can start from the loaded registers or end in the stored ones.
#define CHUNK_LDS 0
#define CHUNK_TO_REGS 1
#define CHUNK_FROM_REGS 2
#define CHUNK_REGS 3          // vv in, results left in vv (a pass of one chunk)
// when a chunk loads its twiddles (SBA_TW_EARLY)
#define TW_LATE 0             // at their first use
#define TW_EARLY 1            // before the chunk's barrier
#define TW_EARLY32 2          // the same, addressed by 32-bit offsets
template <int ELL, int LOGTC, int T0, int B> struct ChunkRegs {
  static constexpr int GPT = (1 << (ELL + LOGTC)) / ((1 << B) * NT);
  typedef uint32_t type[GPT][1 << B];
};
// a chunk's twiddles: group gi, level u, sub-node m2 at [gi][(1 << u) - 1 + m2]
template <int ELL, int LOGTC, int T0, int B> struct ChunkTw {
  typedef uint32_t type[ChunkRegs<ELL, LOGTC, T0, B>::GPT][(1 << B) - 1];
};

// the twiddles of chunk (T0, B) of tile row H: a chunk whose groups all lie
// in one node (NT threads x GPT groups under 2^(LOGTC + LB)) shares them
// across the block, and a uniform index lets them come through the scalar
// cache. ZL: forward levels below ZL are copies and take none.
template <int ELL, int LOGTC, int T0, int B, bool INV, int ZL = 0, bool OFF32 = false>
__device__ __forceinline__ void chunk_twiddles(const uint32_t *__restrict__ tw, uint32_t H,
                                               typename ChunkTw<ELL, LOGTC, T0, B>::type &tv) {
  constexpr int GPT = ChunkRegs<ELL, LOGTC, T0, B>::GPT;
  constexpr int LB = ELL - T0 - B;
  constexpr bool UNI = NT * GPT <= (1 << (LOGTC + LB));
  // uniform twiddles through the constant address space: scalar loads
  // (the tables are read-only while any kernel runs), which also keeps them
  // out of the in-order count of the thread's vector memory accesses
  typedef __attribute__((address_space(4))) const uint32_t cc_u32;
  cc_u32 *twc = (cc_u32 *)tw;
#pragma unroll
  for (int gi = 0; gi < GPT; ++gi) {
    const uint32_t hj =
        UNI ? (H << T0) : (H << T0) | ((threadIdx.x + gi * NT) >> (LOGTC + LB));
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (!INV && T0 + u < ZL) continue;
#pragma unroll
      for (int m2 = 0; m2 < (1 << u); ++m2) {
        const uint32_t i = (hj << u) | (uint32_t)m2;
        tv[gi][(1 << u) - 1 + m2] =
            SBA_TW_SCALAR && UNI ? twc[__builtin_amdgcn_readfirstlane(i)] : (OFF32 ? tw_at(tw, i) : tw[i]);
      }
    }
  }
}

// GUARD (inverse only, column passes: H = 0): a butterfly runs only when its
// block of the column transform lies inside rows [0, gM) -- the complete
// blocks of a truncated inverse; the others leave their values as they are.
// HV: the butterfly forms (gpu_arith.h): HV_W the [0, 2p) forms, HV_H
// Harvey's (p < 2^30), HV_AUTO Harvey's exactly when the block's prime is
// below 2^30 (one uniform branch around the chunk's arithmetic).
template <int ELL, int LOGTC, int T0, int B, bool INV, int ZL = 0, int IO = CHUNK_LDS,
          int TE = TW_EARLY, bool GUARD = false, int HV = HV_W>
__device__ __forceinline__ void tile_chunk_io(uint32_t *lds, const uint32_t *__restrict__ tw,
                                              uint32_t H, uint32_t p, uint32_t pinv,
                                              typename ChunkRegs<ELL, LOGTC, T0, B>::type &vv,
                                              uint32_t gM = 0) {
  constexpr int TC = 1 << LOGTC;
  constexpr int GPT = ChunkRegs<ELL, LOGTC, T0, B>::GPT;
  using G = ChunkGeom<ELL, LOGTC, T0, B>;
  constexpr int LB = G::LB;
  static_assert(GPT >= 1, "chunk too wide for the tile");
  constexpr bool IN_REGS = IO == CHUNK_FROM_REGS || IO == CHUNK_REGS;
  constexpr bool OUT_REGS = IO == CHUNK_TO_REGS || IO == CHUNK_REGS;
  // twiddles of group gi, level u, sub-node m2 at tv[gi][(1 << u) - 1 + m2].
  // They depend on the tile position only, so with SBA_TW_EARLY they are
  // loaded before the barrier: the previous chunk's values are in LDS by
  // then (their registers are free) and the loads' latency overlaps the
  // wait for the other waves. TE = TW_LATE keeps them at their first use;
  // TW_EARLY32 addresses them by 32-bit offsets (fewer live address
  // registers where a kernel holds other values across its chunks).
  constexpr bool EARLY = SBA_TW_EARLY && TE != TW_LATE;
  typename ChunkTw<ELL, LOGTC, T0, B>::type tv;
  if constexpr (EARLY) {
    // not hoisted above the caller's preceding work (the pointwise product
    // before the first inverse chunk), where they would stretch live ranges
    __builtin_amdgcn_sched_barrier(0);
    chunk_twiddles<ELL, LOGTC, T0, B, INV, ZL, TE == TW_EARLY32>(tw, H, tv);
  }
  if (!IN_REGS) lds_barrier();
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
      v[m] = IN_REGS
                 ? vv[gi][m]
                 : (G::LINEAR ? l0[m * G::DP] : lds[lpadt<LOGTC>(i0 + ((uint32_t)m << G::LOGD))]);
    auto flies = [&](auto hc) {
      constexpr bool HVC = decltype(hc)::value;
      if (!INV) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const int half = 1 << (B - u - 1);
          if (T0 + u < ZL) {
#pragma unroll
            for (int m = 0; m < (1 << B); ++m)
              if ((m >> (B - u - 1)) & 1) v[m] = v[m - half];
            continue;
          }
#pragma unroll
          for (int m2 = 0; m2 < (1 << u); ++m2) {
            const uint32_t w = EARLY ? tv[gi][(1 << u) - 1 + m2] : tw[(hj << u) | (uint32_t)m2];
#pragma unroll
            for (int r = 0; r < half; ++r) {
              const int m = (m2 << (B - u)) | r;
              if (HVC) bfly_fwd_h(v[m], v[m + half], w, p, pinv);
              else bfly_fwd(v[m], v[m + half], w, p, pinv);
            }
          }
        }
      } else {
#pragma unroll
        for (int u = B - 1; u >= 0; --u) {
          const int half = 1 << (B - u - 1);
#pragma unroll
          for (int m2 = 0; m2 < (1 << u); ++m2) {
            const uint32_t w = EARLY ? tv[gi][(1 << u) - 1 + m2] : tw[(hj << u) | (uint32_t)m2];
            // the block of this sub-node: rows [b, b + 1) << (ELL - T0 - u)
            const bool run =
                !GUARD || (((((hb << u) | (uint32_t)m2) + 1u) << (ELL - T0 - u)) <= gM);
#pragma unroll
            for (int r = 0; r < half; ++r) {
              const int m = (m2 << (B - u)) | r;
              if (run) {
                if (HVC) bfly_inv_h(v[m], v[m + half], w, p, pinv);
                else bfly_inv(v[m], v[m + half], w, p, pinv);
              }
            }
          }
        }
      }
    };
    if constexpr (HV == HV_AUTO) {
      if (harvey_prime(p)) flies(BoolC<true>{});
      else flies(BoolC<false>{});
    } else {
      flies(BoolC<HV == HV_H>{});
    }
#pragma unroll
    for (int m = 0; m < (1 << B); ++m) {
      if (OUT_REGS) vv[gi][m] = v[m];
      else if (G::LINEAR) l0[m * G::DP] = v[m];
      else lds[lpadt<LOGTC>(i0 + ((uint32_t)m << G::LOGD))] = v[m];
    }
  }
}

template <int ELL, int LOGTC, int T0, int B, bool INV, int ZL = 0, int TE = TW_EARLY,
          int HV = HV_W>
__device__ __forceinline__ void tile_chunk(uint32_t *lds, const uint32_t *__restrict__ tw,
                                           uint32_t H, uint32_t p, uint32_t pinv) {
  typename ChunkRegs<ELL, LOGTC, T0, B>::type unused;
  tile_chunk_io<ELL, LOGTC, T0, B, INV, ZL, CHUNK_LDS, TE, false, HV>(lds, tw, H, p, pinv,
                                                                     unused);
}

// chunk sizes: the largest B <= BM whose group addressing is linear,
// falling back to BM when no size is. BMAX (radix 2^BMAX) is a plan choice.
template <int ELL, int LOGTC, int BMAX> struct ChunkPlan {
  static constexpr int E = (1 << (ELL + LOGTC)) / NT;
  static constexpr int LE = ILog2<E>::v;
  static constexpr int BM = LE < BMAX ? LE : BMAX;
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

template <int ELL, int LOGTC, int BMAX, int T0, int ZL = 0, int TE = TW_EARLY, int HV = HV_W>
__device__ __forceinline__ void tile_fwd(uint32_t *lds, const uint32_t *tw, uint32_t H,
                                         uint32_t p, uint32_t pinv) {
  if constexpr (T0 < ELL) {
    constexpr int B = ChunkPlan<ELL, LOGTC, BMAX>::pick(T0);
    tile_chunk<ELL, LOGTC, T0, B, false, ZL, TE, HV>(lds, tw, H, p, pinv);
    tile_fwd<ELL, LOGTC, BMAX, T0 + B, ZL, TE, HV>(lds, tw, H, p, pinv);
  }
}

// inverse chunks of the levels from T0 below TEND (highest first), LDS to LDS
template <int ELL, int LOGTC, int BMAX, int T0, int TEND, int HV = HV_W, int TE = TW_EARLY>
__device__ __forceinline__ void tile_inv_upto(uint32_t *lds, const uint32_t *tw, uint32_t H,
                                              uint32_t p, uint32_t pinv) {
  if constexpr (T0 < TEND) {
    constexpr int B = ChunkPlan<ELL, LOGTC, BMAX>::pick(T0);
    tile_inv_upto<ELL, LOGTC, BMAX, T0 + B, TEND, HV, TE>(lds, tw, H, p, pinv);
    tile_chunk<ELL, LOGTC, T0, B, true, 0, TE, HV>(lds, tw, H, p, pinv);
  }
}

// forward chunks from T0 up to (not including) level TEND, LDS to LDS
template <int ELL, int LOGTC, int BMAX, int T0, int TEND, int HV = HV_W, int TE = TW_EARLY>
__device__ __forceinline__ void tile_fwd_range(uint32_t *lds, const uint32_t *tw, uint32_t H,
                                               uint32_t p, uint32_t pinv) {
  if constexpr (T0 < TEND) {
    constexpr int B = ChunkPlan<ELL, LOGTC, BMAX>::pick(T0);
    tile_chunk<ELL, LOGTC, T0, B, false, 0, TE, HV>(lds, tw, H, p, pinv);
    tile_fwd_range<ELL, LOGTC, BMAX, T0 + B, TEND, HV, TE>(lds, tw, H, p, pinv);
  }
}

// the last chunk of a tile transform (the highest levels): its first level
template <int ELL, int LOGTC, int BMAX> struct LastChunk {
  static constexpr int t0() {
    int t = 0;
    for (;;) {
      const int b = ChunkPlan<ELL, LOGTC, BMAX>::pick(t);
      if (t + b >= ELL) return t;
      t += b;
    }
  }
  static constexpr int T0 = t0();
  static constexpr int B = ELL - T0;
};

// Global offsets of the elements of chunk (T0, B)'s groups, for a pass that
// starts or ends with that chunk straight from or to HBM: element m of the
// thread's group gi is tile row k, column c, at k * S + c. With R = NT / TC
// rows per group sweep and 2^LB | R, k splits into a per-thread part and a
// part uniform over the workgroup: off = thr + (gi * (R >> LB) * 2^(ELL -
// T0) + m * 2^LB) * S, so every access is a scalar base plus one 32-bit
// lane offset (as the passes' own loads and stores are).
template <int ELL, int LOGTC, int T0, int B> struct ChunkIO {
  using G = ChunkGeom<ELL, LOGTC, T0, B>;
  static constexpr int R = NT >> LOGTC;
  static constexpr int lg(int x) { return x > 1 ? 1 + lg(x >> 1) : 0; }
  static constexpr bool OK = LOGTC >= 3 && R >= 1 && G::LB <= lg(R);
  static __device__ __forceinline__ uint32_t thr(uint32_t S) {
    const uint32_t row = threadIdx.x >> LOGTC;
    const uint32_t k = ((row >> G::LB) << (ELL - T0)) + (row & ((1u << G::LB) - 1));
    return k * S + (threadIdx.x & ((1u << LOGTC) - 1));
  }
  static __device__ __forceinline__ uint32_t uni(int gi, int m, uint32_t S) {
    return ((uint32_t)gi * (uint32_t)(R >> G::LB) * (1u << (ELL - T0)) + ((uint32_t)m << G::LB)) *
           S;
  }
};

// the forward levels of a first pass, with its first level a copy when the
// operand fills at most half of the transform (SBA_ZERO_LEVEL; one uniform
// branch per block, each side straight-line code)
template <int ELL, int LOGTC, int BMAX, int HV = HV_W>
__device__ __forceinline__ void tile_fwd_first(uint32_t *lds, const uint32_t *tw, uint32_t p,
                                               uint32_t pinv, bool half_zero) {
  if (SBA_ZERO_LEVEL && half_zero) tile_fwd<ELL, LOGTC, BMAX, 0, 1, TW_EARLY, HV>(lds, tw, 0, p, pinv);
  else tile_fwd<ELL, LOGTC, BMAX, 0, 0, TW_EARLY, HV>(lds, tw, 0, p, pinv);
}



// the inverse of a column tile's complete blocks below row gM (GUARD above)
template <int ELL, int LOGTC, int BMAX, int T0>
__device__ __forceinline__ void tile_inv_guard(uint32_t *lds, const uint32_t *tw, uint32_t p,
                                               uint32_t pinv, uint32_t gM) {
  if constexpr (T0 < ELL) {
    constexpr int B = ChunkPlan<ELL, LOGTC, BMAX>::pick(T0);
    tile_inv_guard<ELL, LOGTC, BMAX, T0 + B>(lds, tw, p, pinv, gM);
    typename ChunkRegs<ELL, LOGTC, T0, B>::type unused;
    tile_chunk_io<ELL, LOGTC, T0, B, true, 0, CHUNK_LDS, TW_EARLY, true>(lds, tw, 0, p, pinv,
                                                                       unused, gM);
  }
}

template <int ELL, int LOGTC, int BMAX, int T0, int TE = TW_EARLY, int HV = HV_W>
__device__ __forceinline__ void tile_inv(uint32_t *lds, const uint32_t *tw, uint32_t H,
                                         uint32_t p, uint32_t pinv) {
  if constexpr (T0 < ELL) {
    constexpr int B = ChunkPlan<ELL, LOGTC, BMAX>::pick(T0);
    tile_inv<ELL, LOGTC, BMAX, T0 + B, TE, HV>(lds, tw, H, p, pinv);
    tile_chunk<ELL, LOGTC, T0, B, true, 0, TE, HV>(lds, tw, H, p, pinv);
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
  uint32_t *bmaps;                // one 2-state carry map per carry block
  uint32_t an, bn, zlen, Z, nb;   // limbs, coefficients, product limbs, carry blocks
};
// global-address-space views of descriptor pointers: loads through a pointer
// read from memory are otherwise generic (flat) and serialised against LDS
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((address_space(1))) const uint32_t gc_u32;
typedef __attribute__((address_space(1))) uint64_t g_u64;
typedef __attribute__((address_space(1))) const uint64_t gc_u64;
typedef __attribute__((address_space(1))) uint8_t g_u8;
typedef __attribute__((address_space(1))) const uint8_t gc_u8;
// element i (i < 2^30) of a uniform base through a 32-bit byte offset: the
// access takes the base in SGPRs and one VGPR offset instead of a 64-bit
// address per lane
__device__ __forceinline__ uint32_t ld_u32(const g_u32 *base, uint32_t i) {
  return *(const g_u32 *)((gc_u8 *)base + (i << 2));
}
__device__ __forceinline__ void st_u32(g_u32 *base, uint32_t i, uint32_t v) {
  *(g_u32 *)((g_u8 *)base + (i << 2)) = v;
}

// accumulator target of a chunked product's last inverse pass (dst ==
// nullptr: ordinary in-place store)
struct AccArgs {
  uint32_t *dst;                  // K x stride residues
  uint32_t stride, off, ovl, lim;
};

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
__device__ __forceinline__ Job get_job(const JobRef &jr, uint32_t idx) {
  Job j;
  if (jr.arr) {
    j = jr.arr[idx];
  } else {
    j = jr.one;
  }
  j.a = uni_ptr(j.a); j.b = uni_ptr(j.b); j.r = uni_ptr(j.r);
  j.aspec = uni_ptr(j.aspec); j.bspec = uni_ptr(j.bspec);
  j.bmaps = uni_ptr(j.bmaps);
  j.an = uni_u32(j.an); j.bn = uni_u32(j.bn); j.zlen = uni_u32(j.zlen); j.Z = uni_u32(j.Z);
  j.nb = uni_u32(j.nb);
  return j;
}
__device__ __forceinline__ Job get_job(const JobRef &jr) { return get_job(jr, blockIdx.z); }

// ---------------------------------------------------------------------------
// pass kernels
// ---------------------------------------------------------------------------
// Operand and prime of a first-pass block. which = 0 (a) or 1 (b): grid y
// = K (one prime per block) or 1 (all primes per block); which = 2: both
// operands in one launch, grid y = 2K or 2 (a in the lower half).
struct OpSel {
  int op;
  bool split;
  uint32_t q0;
};
__device__ __forceinline__ OpSel op_sel(int which) {
  const uint32_t per = which == 2 ? gridDim.y >> 1 : gridDim.y;
  const int op = which == 2 ? (int)(blockIdx.y >= per) : which;
  const uint32_t y = blockIdx.y - (op && which == 2 ? per : 0);
  return OpSel{op, per > 1, per > 1 ? y : 0u};
}

// residues of elements [I0, I1) of a first-pass thread (element i at tile
// offset col0 + (i NT / TC) S + (i NT) mod TC + toff); CHECK: limbs past nl
// read as zero
template <int E, int LOGTC, int I0, int I1, bool CHECK>
__device__ __forceinline__ void first_loads(uint32_t (&v)[E], gc_u64 *limbs, uint32_t col0,
                                            uint32_t S, uint32_t toff, uint32_t nl, uint32_t p,
                                            uint32_t pinv, uint32_t r1) {
  constexpr int TC = 1 << LOGTC;
  constexpr int UNR = LOGTC == 3 ? SBA_FIRST_LOAD_UNROLL_8COL : SBA_FIRST_LOAD_UNROLL;
#pragma unroll UNR
  for (int i = I0; i < I1; ++i) {
    const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
    gc_u64 *lp = limbs + uoff;
    v[i] = limb_residue(!CHECK || uoff + toff < nl ? lp[toff] : 0, p, pinv, r1);
  }
}

// First forward pass (global levels [0, ELL)): limbs -> residues of all K
// primes -> ELL levels -> spec[q][.]. Tiles = column tiles (H = 0).
// which = 0: operand a into aspec; 1: operand b into bspec; 2: both (op_sel).
#ifndef SBA_FIRST_HZ_SKIP
#define SBA_FIRST_HZ_SKIP 1
#endif
#ifndef SBA_FIRST_NOCHECK
#define SBA_FIRST_NOCHECK 0   // measured: unchecked loads with the zero-half skip were slower
#endif
template <int ELL, int LOGTC, int BMAX>
__global__ __launch_bounds__(NT, SBA_FIRST_WAVES) void k_first_fwd(PrimeArgs pa, JobRef jr, int which,
                                                             int L) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[Pad<LOGTC>::words(NE)];
  const Job jb = get_job(jr);
  const OpSel os = op_sel(which);
  gc_u64 *limbs = (gc_u64 *)(os.op ? jb.b : jb.a);
  g_u32 *spec = (g_u32 *)(os.op ? jb.bspec : jb.aspec);
  const uint32_t nlimbs = os.op ? jb.bn : jb.an;
  const uint32_t N = 1u << L;             // N <= 2^25: 32-bit offsets throughout
  const uint32_t S = 1u << (L - ELL);
  const bool split = os.split;
  // grouped order where it measured faster: the launch that covers both
  // operands, 8-column tiles, and a prime's sweep of at least 4 groups per
  // XCD (128 tiles, 8 MiB of limbs: twice the L2; 2^23 points). Prime-major
  // elsewhere: one operand per launch (squares at 2^22: 89 -> 99 us grouped),
  // 2^22 points (a sweep's limbs already fit the L2: no gain), 4-column
  // tiles at 2^24 (621 -> 644 us grouped).
  const bool grouped = split && which == 2 && LOGTC == 3 && gridDim.x >= 32u * SBA_FIRST_GROUP;
  uint32_t tile, q0;
  first_block(blockIdx.x, os.q0, gridDim.x, grouped ? gridDim.y >> 1 : 1u, tile, q0);
  const uint32_t col0 = tile << LOGTC;
  const uint32_t nl = nlimbs < N ? nlimbs : N;
  uint32_t *lt = lds + lpadt<LOGTC>(threadIdx.x);
  // element e = tid + i*NT sits at col0 + toff + uoff(i): a per-thread part
  // and a uniform part (scalar base + 32-bit lane offset addressing)
  const uint32_t toff = (threadIdx.x >> LOGTC) * S + (threadIdx.x & (TC - 1));
  const uint32_t q1 = split ? q0 + 1 : SBA_K;
  // an operand of at most N/2 limbs pairs every element with a zero in the
  // first level (x, 0) -> (x, x)
  const bool half_zero = nl <= N / 2;
  constexpr int B0 = ChunkPlan<ELL, LOGTC, BMAX>::pick(0);
  constexpr bool DIRECT = SBA_COL_REGS && B0 == ChunkPlan<ELL, LOGTC, BMAX>::LE;
  // EDGE: the last chunk's groups go straight to HBM as well (TC >= 8
  // columns keep 32-byte row segments in that layout too)
  using LC = LastChunk<ELL, LOGTC, BMAX>;
  constexpr bool EDGE = DIRECT && SBA_COL_EDGE && ChunkIO<ELL, LOGTC, LC::T0, LC::B>::OK;
  // one prime's residues and levels (butterfly forms chosen by the prime)
  auto prime = [&](uint32_t q) {
    constexpr int HV = SBA_COL_HV;
    const uint32_t p = pa.p[q], pinv = pa.pinv[q], r1 = pa.r1[q];
    lds_barrier();
    // the tile's limbs are re-read per prime (cache-resident after the first)
    if constexpr (DIRECT) {
      // the level-0 chunk starts from the residues in registers. Element i
      // lies in rows i NT / TC .. (i + 1) NT / TC - 1 of the tile; with a
      // half-empty operand the rows of the upper half (i >= E/2) are zero and
      // the level-0 copy overwrites them, so they are neither loaded nor
      // reduced; rows wholly inside the operand load without a bounds check
      typename ChunkRegs<ELL, LOGTC, 0, B0>::type vv;
      const bool hz = SBA_FIRST_HZ_SKIP && SBA_ZERO_LEVEL && half_zero;
      const uint32_t rows = hz ? (1u << (ELL - 1)) : (1u << ELL);
      const bool inside = SBA_FIRST_NOCHECK && col0 + (rows - 1) * S + (TC - 1) < nl;   // uniform
      if (hz) {
        if (inside) first_loads<E, LOGTC, 0, E / 2, false>(vv[0], limbs, col0, S, toff, nl, p, pinv, r1);
        else first_loads<E, LOGTC, 0, E / 2, true>(vv[0], limbs, col0, S, toff, nl, p, pinv, r1);
#pragma unroll
        for (int i = E / 2; i < E; ++i) vv[0][i] = 0;
      } else {
        if (inside) first_loads<E, LOGTC, 0, E, false>(vv[0], limbs, col0, S, toff, nl, p, pinv, r1);
        else first_loads<E, LOGTC, 0, E, true>(vv[0], limbs, col0, S, toff, nl, p, pinv, r1);
      }
      if constexpr (EDGE && LC::T0 == 0) {
        // one chunk: registers in, registers out
        if (SBA_ZERO_LEVEL && half_zero)
          tile_chunk_io<ELL, LOGTC, 0, B0, false, 1, CHUNK_REGS, TW_EARLY, false, HV>(
              lds, pa.wn[q], 0, p, pinv, vv);
        else
          tile_chunk_io<ELL, LOGTC, 0, B0, false, 0, CHUNK_REGS, TW_EARLY, false, HV>(
              lds, pa.wn[q], 0, p, pinv, vv);
        using IO = ChunkIO<ELL, LOGTC, 0, B0>;
        g_u32 *out = spec + q * N + col0;
        const uint32_t t = IO::thr(S);
#pragma unroll
        for (int m = 0; m < (1 << B0); ++m) (out + IO::uni(0, m, S))[t] = vv[0][m];
        return;
      }
      if (SBA_ZERO_LEVEL && half_zero)
        tile_chunk_io<ELL, LOGTC, 0, B0, false, 1, CHUNK_FROM_REGS, TW_EARLY, false, HV>(
            lds, pa.wn[q], 0, p, pinv, vv);
      else
        tile_chunk_io<ELL, LOGTC, 0, B0, false, 0, CHUNK_FROM_REGS, TW_EARLY, false, HV>(
            lds, pa.wn[q], 0, p, pinv, vv);
      if constexpr (EDGE) {
        tile_fwd_range<ELL, LOGTC, BMAX, B0, LC::T0, HV>(lds, pa.wn[q], 0, p, pinv);
        typename ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::type vl;
        tile_chunk_io<ELL, LOGTC, LC::T0, LC::B, false, 0, CHUNK_TO_REGS, TW_EARLY, false, HV>(
            lds, pa.wn[q], 0, p, pinv, vl);
        using IO = ChunkIO<ELL, LOGTC, LC::T0, LC::B>;
        g_u32 *out = spec + q * N + col0;
        const uint32_t t = IO::thr(S);
        constexpr int GL = ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::GPT;
#pragma unroll
        for (int gi = 0; gi < GL; ++gi)
#pragma unroll
          for (int m = 0; m < (1 << LC::B); ++m) (out + IO::uni(gi, m, S))[t] = vl[gi][m];
        return;
      }
      tile_fwd<ELL, LOGTC, BMAX, B0, 0, TW_EARLY, HV>(lds, pa.wn[q], 0, p, pinv);
    } else {
#pragma unroll SBA_FIRST_LOAD_UNROLL
      for (int i = 0; i < E; ++i) {
        const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
        gc_u64 *lp = limbs + uoff;
        lt[i * Pad<LOGTC>::words(NT)] = limb_residue(uoff + toff < nl ? lp[toff] : 0, p, pinv, r1);
      }
      tile_fwd_first<ELL, LOGTC, BMAX, HV>(lds, pa.wn[q], p, pinv, half_zero);
    }
    lds_barrier();
    g_u32 *out = spec + q * N;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
      (out + uoff)[toff] = lt[i * Pad<LOGTC>::words(NT)];
    }
  };
#pragma unroll 1
  for (uint32_t q = q0; q < q1; ++q) {
    prime(q);
  }
}

// In-place pass over global levels [s0, s0 + ELL) of one prime
// (blockIdx.y), forward or inverse, on aspec (which = 0) or bspec (1).
template <int ELL, int LOGTC, bool INV, int BMAX>
__global__ __launch_bounds__(NT, INV ? SBA_INV_WAVES : SBA_WAVES) void k_pass(PrimeArgs pa, JobRef jr, int which, int L,
                                                        int s0, AccArgs acc) {
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
  constexpr int B0 = ChunkPlan<ELL, LOGTC, BMAX>::pick(0);
  constexpr bool DIRECT = SBA_COL_REGS && B0 == ChunkPlan<ELL, LOGTC, BMAX>::LE;
  // DIRECT: the level-0 chunk starts from the loaded registers (forward) or
  // ends in the registers that are stored (inverse); EDGE (inverse, TC >= 8
  // columns): the inverse's first chunk, the tile's highest levels, also
  // loads its groups straight from HBM
  using LC = LastChunk<ELL, LOGTC, BMAX>;
  constexpr bool EDGE =
      DIRECT && INV && SBA_COL_EDGE && ChunkIO<ELL, LOGTC, LC::T0, LC::B>::OK;
  // butterfly forms chosen by the prime
  auto body = [&]() {
    constexpr int HV = SBA_COL_HV;
    typename ChunkRegs<ELL, LOGTC, 0, B0>::type vv;
    if constexpr (EDGE && LC::T0 == 0) {
      // one chunk: registers in, registers out
      using IO = ChunkIO<ELL, LOGTC, 0, B0>;
      const uint32_t t = IO::thr(S);
#pragma unroll
      for (int m = 0; m < (1 << B0); ++m) vv[0][m] = (buf + IO::uni(0, m, S))[t];
      tile_chunk_io<ELL, LOGTC, 0, B0, true, 0, CHUNK_REGS, TW_EARLY, false, HV>(lds, pa.wni[q], H, p, pinv, vv);
    } else if constexpr (EDGE) {
      using IO = ChunkIO<ELL, LOGTC, LC::T0, LC::B>;
      typename ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::type vf;
      constexpr int GL = ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::GPT;
      const uint32_t t = IO::thr(S);
#pragma unroll
      for (int gi = 0; gi < GL; ++gi)
#pragma unroll
        for (int m = 0; m < (1 << LC::B); ++m) vf[gi][m] = (buf + IO::uni(gi, m, S))[t];
      tile_chunk_io<ELL, LOGTC, LC::T0, LC::B, true, 0, CHUNK_FROM_REGS, TW_EARLY, false, HV>(lds, pa.wni[q], H, p, pinv,
                                                                       vf);
      tile_inv_upto<ELL, LOGTC, BMAX, B0, LC::T0, HV>(lds, pa.wni[q], H, p, pinv);
      tile_chunk_io<ELL, LOGTC, 0, B0, true, 0, CHUNK_TO_REGS, TW_EARLY, false, HV>(lds, pa.wni[q], H, p, pinv, vv);
    } else if (DIRECT && !INV) {
#pragma unroll SBA_PASS_LOAD_UNROLL
      for (int i = 0; i < E; ++i) {
        const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
        vv[0][i] = (buf + uoff)[toff];
      }
      tile_chunk_io<ELL, LOGTC, 0, B0, false, 0, CHUNK_FROM_REGS, TW_EARLY, false, HV>(lds, pa.wn[q], H, p, pinv, vv);
      tile_fwd<ELL, LOGTC, BMAX, B0, 0, TW_EARLY, HV>(lds, pa.wn[q], H, p, pinv);
    } else {
#pragma unroll SBA_PASS_LOAD_UNROLL
      for (int i = 0; i < E; ++i) {
        const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
        lt[i * Pad<LOGTC>::words(NT)] = (buf + uoff)[toff];
      }
      if (!INV) {
        tile_fwd<ELL, LOGTC, BMAX, 0, 0, TW_EARLY, HV>(lds, pa.wn[q], H, p, pinv);
      } else if (DIRECT) {
        tile_inv_upto<ELL, LOGTC, BMAX, B0, ELL, HV>(lds, pa.wni[q], H, p, pinv);
        tile_chunk_io<ELL, LOGTC, 0, B0, true, 0, CHUNK_TO_REGS, TW_EARLY, false, HV>(lds, pa.wni[q], H, p, pinv, vv);
      } else {
        tile_inv<ELL, LOGTC, BMAX, 0, TW_EARLY, HV>(lds, pa.wni[q], H, p, pinv);
      }
    }
    // the inverse's element i: in registers (DIRECT) or in LDS
    const bool inreg = DIRECT && INV;
    if (!inreg) lds_barrier();
    // the inverse's last levels (s0 == 0) finish the transform: scale here
    const bool fin = INV && s0 == 0;
    const uint32_t sc = pa.scale[q];
    if (INV && acc.dst) {
      // one chunk of a chunked product: coefficient x of the chunk lands at
      // acc.off + x of the accumulator; x < acc.ovl adds to the previous
      // chunk's tail, x >= acc.lim (beyond the chunk's product) is skipped
      g_u32 *d = (g_u32 *)acc.dst + q * acc.stride + acc.off + base;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
        const uint32_t x = base + uoff + toff;
        if (x < acc.lim) {
          uint32_t v = mont_mul(inreg ? vv[0][i] : lt[i * Pad<LOGTC>::words(NT)], sc, p, pinv);
          if (x < acc.ovl) v = red1(v, p) + red1((d + uoff)[toff], p);
          (d + uoff)[toff] = v;
        }
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t uoff = (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
      const uint32_t v = inreg ? vv[0][i] : lt[i * Pad<LOGTC>::words(NT)];
      (buf + uoff)[toff] = fin ? mont_mul(v, sc, p, pinv) : v;
    }
  };
  body();
}

// Last forward levels of b + pointwise product with a's spectrum (or b's own
// square) + first inverse levels, on contiguous tiles (S = 1, TC = 1).
// LIMB_IN: the pass is the whole transform (s0 = 0) and reads b's limbs.
// r3: the class has a radix-3 outer level (length 3 * 2^L): grid z runs over
// (job, row) and the row's length-2^L transform sits at row * 2^L inside
// each prime's 3 * 2^L residues (LIMB_IN is never launched with r3).
// mode: bit 0 (MODE_R3) radix-3 class; bit 1 (MODE_TRUNC) truncated product:
// only the tiles H < ceil(zlen / 2^ELL) are transformed (the others hold no
// coefficient of the product; the truncated inverse column pass treats them
// as zero coefficients)
#define MODE_R3 1
#define MODE_TRUNC 2
__device__ __forceinline__ bool tile_skipped(const Job &jb, int ell, int mode) {
  return (mode & MODE_TRUNC) && blockIdx.x >= ((jb.zlen + (1u << ell) - 1) >> ell);
}

template <int ELL, bool LIMB_IN, bool SQR, int BMAX>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_fused(PrimeArgs pa, JobRef jr, int L, int mode) {
  constexpr int NE = 1 << ELL;
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[Pad<0>::words(NE)];
  const int r3 = mode & MODE_R3;
  const uint32_t z = blockIdx.z;
  const Job jb = get_job(jr, r3 ? z / 3 : z);
  if (tile_skipped(jb, ELL, mode)) return;            // uniform per block
  const int q = blockIdx.y;
  const uint32_t N = 1u << L;
  const uint32_t pstride = r3 ? 3 * N : N;
  const uint32_t H = blockIdx.x;
  const uint32_t base = (r3 ? (z % 3) * N : 0) + (H << ELL);
  const uint32_t nb = jb.bn < N ? jb.bn : N;
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  g_u32 *buf = (g_u32 *)jb.bspec + q * pstride + base;
  gc_u32 *as = (gc_u32 *)jb.aspec + q * pstride + base;
  gc_u64 *bl = (gc_u64 *)jb.b + base;
  uint32_t *lt = lds + lpadt<0>(threadIdx.x);
  const uint32_t tid = threadIdx.x;
  // REGS (as in k_fused2): the first chunk starts from the loaded registers
  // and the inverse's last chunk ends in the stored ones; a square also
  // keeps its last forward chunk in registers through the pointwise product
  // into the inverse's first chunk (a multiply reads a's spectrum in the
  // natural layout, so its product goes through LDS)
  constexpr int B0 = ChunkPlan<ELL, 0, BMAX>::pick(0);
  using LC = LastChunk<ELL, 0, BMAX>;
  constexpr bool REGS = SBA_FUSED2_REGS && B0 == ChunkPlan<ELL, 0, BMAX>::LE && LC::T0 >= B0;
  // butterfly forms: with SBA_FUSED_SPLIT the body is instantiated for
  // Harvey's forms and for the [0, 2p) forms, and the block runs the one its
  // prime allows. Stored forward values (b's, and a's read for the product)
  // may come from a pass with Harvey's forms: fwd_in and canon<HV_AUTO>
  // take either range.
  auto body = [&](auto hv) {
    constexpr int HV = SBA_FUSED_SPLIT ? (decltype(hv)::value ? HV_H : HV_W) : SBA_FUSED_HV;
    if constexpr (REGS) {
      typedef typename ChunkRegs<ELL, 0, 0, B0>::type R0;
      R0 v0;
#pragma unroll
      for (int i = 0; i < E; ++i)
        v0[0][i] = LIMB_IN ? limb_residue(base + i * NT + tid < nb ? (bl + i * NT)[tid] : 0, p,
                                          pinv, pa.r1[q])
                           : fwd_in<HV>((buf + i * NT)[tid], p);
      // LIMB_IN: one tile (H = 0) holds the whole transform of the padded
      // limbs; a half-empty operand's first level is a copy
      if (LIMB_IN && SBA_ZERO_LEVEL && nb <= N / 2)
        tile_chunk_io<ELL, 0, 0, B0, false, 1, CHUNK_FROM_REGS, TW_EARLY, false, HV>(
            lds, pa.wn[q], H, p, pinv, v0);
      else
        tile_chunk_io<ELL, 0, 0, B0, false, 0, CHUNK_FROM_REGS, TW_EARLY, false, HV>(
            lds, pa.wn[q], H, p, pinv, v0);
      tile_fwd_range<ELL, 0, BMAX, B0, LC::T0, HV>(lds, pa.wn[q], H, p, pinv);
      if constexpr (SQR) {
        typename ChunkRegs<ELL, 0, LC::T0, LC::B>::type bv;
        constexpr int GL = ChunkRegs<ELL, 0, LC::T0, LC::B>::GPT;
        tile_chunk_io<ELL, 0, LC::T0, LC::B, false, 0, CHUNK_TO_REGS, TW_EARLY, false, HV>(
            lds, pa.wn[q], H, p, pinv, bv);
#pragma unroll
        for (int g = 0; g < GL; ++g)
#pragma unroll
          for (int m = 0; m < (1 << LC::B); ++m)
            bv[g][m] = mont_mul(bv[g][m], canon<HV>(bv[g][m], p), p, pinv);
        lds_barrier();               // the last LDS reads before the inverse's writes
        tile_chunk_io<ELL, 0, LC::T0, LC::B, true, 0, CHUNK_FROM_REGS, TW_EARLY, false, HV>(
            lds, pa.wni[q], H, p, pinv, bv);
        tile_inv_upto<ELL, 0, BMAX, B0, LC::T0, HV>(lds, pa.wni[q], H, p, pinv);
      } else {
        tile_chunk<ELL, 0, LC::T0, LC::B, false, 0, TW_EARLY, HV>(lds, pa.wn[q], H, p, pinv);
        lds_barrier();
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const uint32_t v = lt[i * Pad<0>::words(NT)];
          lt[i * Pad<0>::words(NT)] = mont_mul(v, canon<HV_AUTO>((as + i * NT)[tid], p), p, pinv);
        }
        tile_inv_upto<ELL, 0, BMAX, B0, ELL, HV>(lds, pa.wni[q], H, p, pinv);
      }
      R0 vo;
      tile_chunk_io<ELL, 0, 0, B0, true, 0, CHUNK_TO_REGS, TW_EARLY, false, HV>(lds, pa.wni[q], H,
                                                                               p, pinv, vo);
      // LIMB_IN: the single tile is the whole transform, so the inverse ends here
#pragma unroll
      for (int i = 0; i < E; ++i)
        (buf + i * NT)[tid] = LIMB_IN ? mont_mul(vo[0][i], pa.scale[q], p, pinv) : vo[0][i];
    } else {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        uint32_t v;
        if (LIMB_IN) {
          v = limb_residue(base + i * NT + tid < nb ? (bl + i * NT)[tid] : 0, p, pinv, pa.r1[q]);
        } else {
          v = fwd_in<HV>((buf + i * NT)[tid], p);
        }
        lt[i * Pad<0>::words(NT)] = v;
      }
      // LIMB_IN: one tile (H = 0) holds the whole transform of the padded limbs
      if (LIMB_IN) tile_fwd_first<ELL, 0, BMAX, HV>(lds, pa.wn[q], p, pinv, nb <= N / 2);
      else tile_fwd<ELL, 0, BMAX, 0, 0, TW_EARLY, HV>(lds, pa.wn[q], H, p, pinv);
      lds_barrier();
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const uint32_t v = lt[i * Pad<0>::words(NT)];
        const uint32_t o = SQR ? canon<HV>(v, p) : canon<HV_AUTO>((as + i * NT)[tid], p);
        lt[i * Pad<0>::words(NT)] = mont_mul(v, o, p, pinv);
      }
      tile_inv<ELL, 0, BMAX, 0, TW_EARLY, HV>(lds, pa.wni[q], H, p, pinv);
      lds_barrier();
      // LIMB_IN: the single tile is the whole transform, so the inverse ends here
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const uint32_t v = lt[i * Pad<0>::words(NT)];
        (buf + i * NT)[tid] = LIMB_IN ? mont_mul(v, pa.scale[q], p, pinv) : v;
      }
    }
  };
  if (SBA_FUSED_SPLIT && harvey_prime(p)) body(BoolC<true>{});
  else body(BoolC<false>{});
}

// Multiply, two-pass plans: the last forward levels of a AND of b on the same
// tile (one set of twiddles), the pointwise product and the first inverse
// levels. a's transformed tile is kept in registers (canonical, the
// Montgomery operand) while b's is transformed in the same LDS.
template <int ELL, int BMAX>
__global__ __launch_bounds__(NT, BMAX == 5 ? SBA_FUSED2_WAVES : SBA_WAVES) void k_fused2(PrimeArgs pa, JobRef jr, int L, int mode) {
  constexpr int NE = 1 << ELL;
  constexpr int E = NE / NT;
  __shared__ uint32_t lds[Pad<0>::words(NE)];
  // a's spectrum stays in registers from its transform to the product, so
  // early twiddles need the 32-bit-offset form to fit in 3 waves per SIMD
  constexpr int F2TE = SBA_FUSED2_EARLY ? TW_EARLY32 : TW_LATE;
  const int r3 = mode & MODE_R3;
  const uint32_t z = blockIdx.z;
  const Job jb = get_job(jr, r3 ? z / 3 : z);
  if (tile_skipped(jb, ELL, mode)) return;            // uniform per block
  const int q = blockIdx.y;
  const uint32_t N = 1u << L;
  const uint32_t pstride = r3 ? 3 * N : N;
  const uint32_t H = blockIdx.x;
  const uint32_t base = (r3 ? (z % 3) * N : 0) + (H << ELL);
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  g_u32 *buf = (g_u32 *)jb.bspec + q * pstride + base;
  gc_u32 *as = (gc_u32 *)jb.aspec + q * pstride + base;
  uint32_t *lt = lds + lpadt<0>(threadIdx.x);
  const uint32_t tid = threadIdx.x;
  // REGS: the transforms' first chunk starts from the loaded registers (its
  // groups are a thread's elements tid + i NT), the forward's last chunk
  // leaves its groups in registers, where the pointwise product and the
  // inverse's first chunk (the same groups) take them, and the inverse's last
  // chunk ends in the registers that are stored: two LDS round trips and two
  // barriers fewer per transform than staging every chunk through LDS
  constexpr int B0 = ChunkPlan<ELL, 0, BMAX>::pick(0);
  using LC = LastChunk<ELL, 0, BMAX>;
  constexpr bool REGS =
      SBA_FUSED2_REGS && B0 == ChunkPlan<ELL, 0, BMAX>::LE && LC::T0 >= B0;
  // HV: the butterfly forms (the whole body instantiated per form, or one
  // body choosing by the prime chunk by chunk)
  auto body = [&](auto hv) {
    constexpr int HV = SBA_FUSED2_SPLIT ? (decltype(hv)::value ? HV_H : HV_W) : HV_AUTO;
    if constexpr (REGS) {
      typedef typename ChunkRegs<ELL, 0, 0, B0>::type R0;
      typedef typename ChunkRegs<ELL, 0, LC::T0, LC::B>::type RL;
      constexpr int GL = ChunkRegs<ELL, 0, LC::T0, LC::B>::GPT;
      // one operand's last forward levels: HBM -> registers -> ... -> registers
      auto fwd = [&](gc_u32 *src, RL &out) {
        R0 v0;
#pragma unroll
        for (int i = 0; i < E; ++i) v0[0][i] = (src + i * NT)[tid];
        tile_chunk_io<ELL, 0, 0, B0, false, 0, CHUNK_FROM_REGS, F2TE, false, HV>(
            lds, pa.wn[q], H, p, pinv, v0);
        tile_fwd_range<ELL, 0, BMAX, B0, LC::T0, HV, F2TE>(lds, pa.wn[q], H, p, pinv);
        tile_chunk_io<ELL, 0, LC::T0, LC::B, false, 0, CHUNK_TO_REGS, F2TE, false, HV>(
            lds, pa.wn[q], H, p, pinv, out);
      };
      RL av, bv;
      fwd(as, av);
#pragma unroll
      for (int g = 0; g < GL; ++g)
#pragma unroll
        for (int m = 0; m < (1 << LC::B); ++m) av[g][m] = canon<HV>(av[g][m], p);
      lds_barrier();                 // a's last LDS reads before b's first writes
      fwd(buf, bv);
#pragma unroll
      for (int g = 0; g < GL; ++g)
#pragma unroll
        for (int m = 0; m < (1 << LC::B); ++m) bv[g][m] = mont_mul(bv[g][m], av[g][m], p, pinv);
      lds_barrier();                 // b's last LDS reads before the inverse's writes
      tile_chunk_io<ELL, 0, LC::T0, LC::B, true, 0, CHUNK_FROM_REGS, F2TE, false, HV>(
          lds, pa.wni[q], H, p, pinv, bv);
      tile_inv_upto<ELL, 0, BMAX, B0, LC::T0, HV, F2TE>(lds, pa.wni[q], H, p, pinv);
      R0 vo;
      tile_chunk_io<ELL, 0, 0, B0, true, 0, CHUNK_TO_REGS, F2TE, false, HV>(lds, pa.wni[q], H,
                                                                           p, pinv, vo);
#pragma unroll
      for (int i = 0; i < E; ++i) (buf + i * NT)[tid] = vo[0][i];
    } else {
      uint32_t av[E];
#pragma unroll
      for (int i = 0; i < E; ++i) lt[i * Pad<0>::words(NT)] = (as + i * NT)[tid];
      tile_fwd<ELL, 0, BMAX, 0, 0, F2TE, HV>(lds, pa.wn[q], H, p, pinv);
      lds_barrier();
      // each thread takes back only its own elements, then refills them with b
#pragma unroll
      for (int i = 0; i < E; ++i) av[i] = canon<HV>(lt[i * Pad<0>::words(NT)], p);
#pragma unroll
      for (int i = 0; i < E; ++i) lt[i * Pad<0>::words(NT)] = (buf + i * NT)[tid];
      tile_fwd<ELL, 0, BMAX, 0, 0, F2TE, HV>(lds, pa.wn[q], H, p, pinv);
      lds_barrier();
#pragma unroll
      for (int i = 0; i < E; ++i)
        lt[i * Pad<0>::words(NT)] = mont_mul(lt[i * Pad<0>::words(NT)], av[i], p, pinv);
      tile_inv<ELL, 0, BMAX, 0, F2TE, HV>(lds, pa.wni[q], H, p, pinv);
      lds_barrier();
#pragma unroll
      for (int i = 0; i < E; ++i) (buf + i * NT)[tid] = lt[i * Pad<0>::words(NT)];
    }
  };
  if (SBA_FUSED2_SPLIT && harvey_prime(p)) body(BoolC<true>{});
  else body(BoolC<false>{});
}

// ---------------------------------------------------------------------------
// radix-3 outer level: transform length 3M, M = 2^L, by the Good-Thomas
// index map. Coefficient i sits in row i mod 3, column i mod M of a 3 x M
// array (a ring isomorphism Z[x]/(x^3M - 1) = Z[u,v]/(u^3 - 1, v^M - 1)
// since gcd(3, M) = 1), so a length-3 DFT down every column and three
// independent length-M transforms along the rows need no twiddles between
// the two factors. Column s holds limbs s, s + M, s + 2M; limb s + tM lands
// in row (s + t*mu) mod 3, mu = M mod 3. Spectra keep the layout
// [prime][row][M]; the inverse writes coefficients back in natural order,
// and the factor 3 M is removed in the Garner stage.
// ---------------------------------------------------------------------------
// length-3 DFT of canonical (a, b, c) with h1 = -1/2 and h2 = (w - w^2)/2:
//   z0 = a + b + c, z1 = a + w b + w^2 c, z2 = a + w^2 b + w c
// (w + w^2 = -1), outputs lazy in [0, 2p)
__device__ __forceinline__ void dft3(uint32_t a, uint32_t b, uint32_t c, uint32_t &z0,
                                     uint32_t &z1, uint32_t &z2, uint32_t h1, uint32_t h2,
                                     uint32_t p, uint32_t pinv) {
  const uint32_t u = b + c;                                   // [0, 2p)
  const uint32_t v = b + p - c;                               // (0, 2p)
  z0 = a + red1(u, p);
  const uint32_t A = red1(a + red1(mont_mul(u, h1, p, pinv), p), p);
  const uint32_t B = red1(mont_mul(v, h2, p, pinv), p);
  z1 = A + B;
  z2 = A + p - B;
}

// row of limb t of column s is (s + t mu) mod 3; inversely, row r holds limb
// t(r) = ((r - s) mu) mod 3 (mu^2 = 1 mod 3)
__device__ __forceinline__ uint32_t r3_pick(uint32_t r, uint32_t sm, uint32_t mu,
                                            uint32_t x0, uint32_t x1, uint32_t x2) {
  const uint32_t t = ((r + 3 - sm) * mu) % 3;
  return t == 0 ? x0 : (t == 1 ? x1 : x2);
}

// butterfly forms of the radix-3 column passes (as SBA_COL_HV: chosen by
// the prime chunk by chunk; the column DFT-3 takes and gives [0, 2p))
#ifndef SBA_R3_COL_HV
#define SBA_R3_COL_HV SBA_COL_HV
#endif
constexpr int R3HV = SBA_R3_COL_HV;
// the radix-3 first pass at 4 waves/SIMD: at most 128 VGPRs, no spills
// (with the forms chosen by the prime it needs up to 134 unbounded)
#ifndef SBA_R3_FIRST_WAVES
#define SBA_R3_FIRST_WAVES 4
#endif

// First forward pass of a radix-3 class: limbs -> residues -> column DFT-3
// -> the first ELL levels of all three rows. Tiles: 2^ELL rows x 2^LOGTC
// columns of each of the three length-M rows (three LDS tiles).
template <int ELL, int LOGTC>
__global__ __launch_bounds__(NT, SBA_R3_FIRST_WAVES) void k_first_fwd3(PrimeArgs pa, JobRef jr, int which,
                                                              int L) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int E = NE / NT;
  constexpr int W = Pad<LOGTC>::words(NE);
  __shared__ uint32_t lds[3 * W];
  const Job jb = get_job(jr);
  const OpSel os = op_sel(which);
  gc_u64 *limbs = (gc_u64 *)(os.op ? jb.b : jb.a);
  g_u32 *spec = (g_u32 *)(os.op ? jb.bspec : jb.aspec);
  const uint32_t nlimbs = os.op ? jb.bn : jb.an;
  const uint32_t M = 1u << L;
  const uint32_t S = 1u << (L - ELL);
  const uint32_t mu = (L & 1) ? 2u : 1u;
  const bool split = os.split;
  uint32_t tile, q0;
  first_block<SBA_R3_FIRST_GROUP>(blockIdx.x, os.q0, gridDim.x,
                                  split ? (which == 2 ? gridDim.y >> 1 : gridDim.y) : 1u, tile, q0);
  const uint32_t col0 = tile << LOGTC;
  const uint32_t nl = nlimbs < 3 * M ? nlimbs : 3 * M;
  uint32_t *lt = lds + lpadt<LOGTC>(threadIdx.x);
  const uint32_t toff = (threadIdx.x >> LOGTC) * S + (threadIdx.x & (TC - 1));
  const uint32_t q1 = split ? q0 + 1 : SBA_K;
  // as in k_first_fwd: the rows' level-0 chunks start from the registers
  // and (EDGE) their last chunks go straight to HBM
  constexpr int B0 = ChunkPlan<ELL, LOGTC, 5>::pick(0);
  constexpr bool DIRECT = SBA_COL_REGS && B0 == ChunkPlan<ELL, LOGTC, 5>::LE;
  using LC = LastChunk<ELL, LOGTC, 5>;
  constexpr bool EDGE =
      DIRECT && SBA_COL_EDGE && LC::T0 > 0 && ChunkIO<ELL, LOGTC, LC::T0, LC::B>::OK;
  typedef typename ChunkRegs<ELL, LOGTC, 0, B0>::type Regs0;
#pragma unroll 1
  for (uint32_t q = q0; q < q1; ++q) {
    const uint32_t p = pa.p[q], pinv = pa.pinv[q], r1 = pa.r1[q];
    lds_barrier();
    Regs0 v0, v1, v2;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
      const uint32_t s = uoff + toff;
      uint32_t x[3];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const uint32_t g = s + (uint32_t)t * M;
        x[t] = red1(limb_residue(g < nl ? (limbs + uoff + (uint32_t)t * M)[toff] : 0, p, pinv, r1), p);
      }
      const uint32_t sm = s % 3;
      uint32_t y0, y1, y2;
      dft3(r3_pick(0, sm, mu, x[0], x[1], x[2]), r3_pick(1, sm, mu, x[0], x[1], x[2]),
           r3_pick(2, sm, mu, x[0], x[1], x[2]), y0, y1, y2, pa.h1[q], pa.h2[q], p, pinv);
      if constexpr (DIRECT) {
        v0[0][i] = y0;
        v1[0][i] = y1;
        v2[0][i] = y2;
      } else {
        lt[i * Pad<LOGTC>::words(NT)] = y0;
        lt[W + i * Pad<LOGTC>::words(NT)] = y1;
        lt[2 * W + i * Pad<LOGTC>::words(NT)] = y2;
      }
    }
    if constexpr (DIRECT) {
      auto row = [&](int k, Regs0 &vk) {
        uint32_t *lk = lds + k * W;
        tile_chunk_io<ELL, LOGTC, 0, B0, false, 0, CHUNK_FROM_REGS, TW_EARLY, false, R3HV>(lk, pa.wn[q], 0, p, pinv, vk);
        if constexpr (EDGE) {
          using IO = ChunkIO<ELL, LOGTC, LC::T0, LC::B>;
          tile_fwd_range<ELL, LOGTC, 5, B0, LC::T0, R3HV>(lk, pa.wn[q], 0, p, pinv);
          typename ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::type vl;
          tile_chunk_io<ELL, LOGTC, LC::T0, LC::B, false, 0, CHUNK_TO_REGS, TW_EARLY, false, R3HV>(
              lk, pa.wn[q], 0, p,
                                                                         pinv, vl);
          g_u32 *out = spec + q * 3 * M + (uint32_t)k * M + col0;
          const uint32_t t = IO::thr(S);
          constexpr int GL = ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::GPT;
#pragma unroll
          for (int gi = 0; gi < GL; ++gi)
#pragma unroll
            for (int m = 0; m < (1 << LC::B); ++m) (out + IO::uni(gi, m, S))[t] = vl[gi][m];
        } else {
          tile_fwd<ELL, LOGTC, 5, B0, 0, TW_EARLY, R3HV>(lk, pa.wn[q], 0, p, pinv);
        }
      };
      row(0, v0);
      row(1, v1);
      row(2, v2);
      if constexpr (EDGE) continue;
    } else {
#pragma unroll 1
      for (int k = 0; k < 3; ++k) tile_fwd<ELL, LOGTC, 5, 0, 0, TW_EARLY, R3HV>(lds + k * W, pa.wn[q], 0, p, pinv);
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      g_u32 *out = spec + q * 3 * M + (uint32_t)k * M;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
        (out + uoff)[toff] = lt[k * W + i * Pad<LOGTC>::words(NT)];
      }
    }
  }
}

// Last inverse pass of a radix-3 class (inverse levels [0, ELL) of the three
// rows of prime blockIdx.y on the same column tile), then the inverse column
// DFT-3 (root w^2, unscaled) and the map back to natural order, in place.
template <int ELL, int LOGTC>
__global__ __launch_bounds__(NT, SBA_WAVES) void k_inv_first3(PrimeArgs pa, JobRef jr, int L) {
  constexpr int TC = 1 << LOGTC;
  constexpr int NE = 1 << (ELL + LOGTC);
  constexpr int E = NE / NT;
  constexpr int W = Pad<LOGTC>::words(NE);
  __shared__ uint32_t lds[3 * W];
  const Job jb = get_job(jr);
  const int q = blockIdx.y;
  const uint32_t M = 1u << L;
  const uint32_t S = 1u << (L - ELL);
  const uint32_t mu = (L & 1) ? 2u : 1u;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t col0 = tile << LOGTC;
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  g_u32 *buf = (g_u32 *)jb.bspec + q * 3 * M;
  uint32_t *lt = lds + lpadt<LOGTC>(threadIdx.x);
  const uint32_t toff = (threadIdx.x >> LOGTC) * S + (threadIdx.x & (TC - 1));
  // EDGE: each row's first inverse chunk loads straight from HBM and its
  // level-0 chunk ends in the registers the column DFT-3 takes
  constexpr int B0 = ChunkPlan<ELL, LOGTC, 5>::pick(0);
  constexpr bool DIRECT = SBA_COL_REGS && B0 == ChunkPlan<ELL, LOGTC, 5>::LE;
  using LC = LastChunk<ELL, LOGTC, 5>;
  constexpr bool EDGE =
      DIRECT && SBA_COL_EDGE && LC::T0 > 0 && ChunkIO<ELL, LOGTC, LC::T0, LC::B>::OK;
  typedef typename ChunkRegs<ELL, LOGTC, 0, B0>::type Regs0;
  Regs0 v0, v1, v2;
  if constexpr (EDGE) {
    auto row = [&](int k, Regs0 &vk) {
      using IO = ChunkIO<ELL, LOGTC, LC::T0, LC::B>;
      uint32_t *lk = lds + k * W;
      typename ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::type vf;
      constexpr int GL = ChunkRegs<ELL, LOGTC, LC::T0, LC::B>::GPT;
      g_u32 *in = buf + (uint32_t)k * M + col0;
      const uint32_t t = IO::thr(S);
#pragma unroll
      for (int gi = 0; gi < GL; ++gi)
#pragma unroll
        for (int m = 0; m < (1 << LC::B); ++m) vf[gi][m] = (in + IO::uni(gi, m, S))[t];
      tile_chunk_io<ELL, LOGTC, LC::T0, LC::B, true, 0, CHUNK_FROM_REGS, TW_EARLY, false, R3HV>(
          lk, pa.wni[q], 0, p, pinv,
                                                                       vf);
      tile_inv_upto<ELL, LOGTC, 5, B0, LC::T0, R3HV>(lk, pa.wni[q], 0, p, pinv);
      tile_chunk_io<ELL, LOGTC, 0, B0, true, 0, CHUNK_TO_REGS, TW_EARLY, false, R3HV>(
          lk, pa.wni[q], 0, p, pinv, vk);
    };
    row(0, v0);
    row(1, v1);
    row(2, v2);
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
        lt[k * W + i * Pad<LOGTC>::words(NT)] = (buf + (uint32_t)k * M + uoff)[toff];
      }
#pragma unroll 1
    for (int k = 0; k < 3; ++k) tile_inv<ELL, LOGTC, 5, 0, TW_EARLY, R3HV>(lds + k * W, pa.wni[q], 0, p, pinv);
    lds_barrier();
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t uoff = col0 + (((uint32_t)i * NT) >> LOGTC) * S + (((uint32_t)i * NT) & (TC - 1));
    const uint32_t s = uoff + toff;
    uint32_t r0, r1, r2;
    const uint32_t sc = pa.scale[q];
    const uint32_t a0 = EDGE ? v0[0][i] : lt[i * Pad<LOGTC>::words(NT)];
    const uint32_t a1 = EDGE ? v1[0][i] : lt[W + i * Pad<LOGTC>::words(NT)];
    const uint32_t a2 = EDGE ? v2[0][i] : lt[2 * W + i * Pad<LOGTC>::words(NT)];
    dft3(red1(mont_mul(a0, sc, p, pinv), p), red1(mont_mul(a1, sc, p, pinv), p),
         red1(mont_mul(a2, sc, p, pinv), p), r0, r1, r2, pa.h1[q], pa.h2i[q], p, pinv);
    // row r holds coefficient s + t(r) M: write limb position t from row
    // r(t) = (s + t mu) mod 3
    const uint32_t sm = s % 3;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const uint32_t r = (sm + (uint32_t)t * mu) % 3;
      const uint32_t v = r == 0 ? r0 : (r == 1 ? r1 : r2);
      (buf + (uint32_t)t * M + uoff)[toff] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// truncated inverse column pass: the last ELL inverse levels (global levels
// [0, ELL), the column transform of length R = 2^ELL down each of the tile's
// TC columns) when only the rows k < M = ceil(zlen / S) were transformed
// (S = 2^(L - ELL) residues per row; MODE_TRUNC in the merged last pass).
// Per column, the values of rows k < M are spectral (the column transform's
// first M outputs) and the coefficients of rows k >= M are known to be zero
// (they lie past the product). The inverse truncated transform of van der
// Hoeven, as the oracle's tft_inv states it (tests/p30_tft_model_test.c:
// 125-181, include/p30_ref.h:556-611), recovers the coefficients of rows
// k < M, scaled by R like the full inverse. Here it runs in three phases on
// the LDS tile, each step a radix-2 sweep followed by a barrier:
//   1. the full inverse of every complete dyadic block [b*bs, (b+1)*bs) with
//      (b+1)*bs <= M, level by level from the smallest (these are the full
//      subtrees the recursion inverts before it uses them);
//   2. down the spine (the nodes that contain row M): a node [o, o+m) with
//      z = M - o rows known, h = m/2, forward root w of the node, either
//        z < h  (A) rows i in [o+z, o+h): x_i = (x_i + w x_(i+h)) / 2, go left
//        z >= h (C) rows i in [o+z-h, o+h): t = x_i - w x_(i+h), x_i += t,
//                   x_(i+h) = t, go right
//      until a node is complete or empty;
//   3. back up the spine: (B) rows i in [o, o+z): x_i = 2 x_i - w x_(i+h);
//      (D) rows i in [o, o+z-h): the plain inverse butterfly.
// Values stay lazy in [0, 2p) throughout.
// ---------------------------------------------------------------------------
template <int ELL, int LOGTC>
__global__ __launch_bounds__(NT) void k_inv_trunc(PrimeArgs pa, JobRef jr, int L) {
  constexpr int TC = 1 << LOGTC;
  constexpr int R = 1 << ELL;
  constexpr int NE = R * TC;
  constexpr int E = NE / NT;
  static_assert(E >= 1, "tile shape");
  __shared__ uint32_t lds[Pad<LOGTC>::words(NE)];
  const Job jb = get_job(jr);
  const int q = blockIdx.y;
  const uint32_t N = 1u << L;
  const uint32_t S = 1u << (L - ELL);
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t col0 = tile << LOGTC;
  g_u32 *buf = (g_u32 *)jb.bspec + q * N + col0;
  const uint32_t M0 = (jb.zlen + S - 1) >> (L - ELL);
  const uint32_t M = M0 < (uint32_t)R ? M0 : (uint32_t)R;
  const uint32_t p = pa.p[q], pinv = pa.pinv[q];
  const uint32_t *wn = pa.wn[q], *wni = pa.wni[q];
  const uint32_t tid = threadIdx.x;
  // tile element (row k, column c) in LDS; a sweep index u over rows
  // [r0, r0 + n) of all columns is row r0 + (u >> LOGTC), column u & (TC - 1)
  auto at = [&](uint32_t k, uint32_t c) -> uint32_t & {
    return lds[lpadt<LOGTC>((k << LOGTC) | c)];
  };
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t e = tid + (uint32_t)i * NT;
    const uint32_t k = e >> LOGTC, c = e & (TC - 1);
    at(k, c) = k < M ? (buf + k * S)[c] : 0u;
  }
  lds_barrier();
  // phase 1: the complete blocks, every inverse level (register chunks as
  // in the full inverse column pass, each butterfly guarded)
  tile_inv_guard<ELL, LOGTC, 5, 0>(lds, wni, p, pinv, M);
  lds_barrier();
  if (M < (uint32_t)R) {
    // phase 2: down the spine (dec bit d: the node at depth d went right)
    uint32_t dec = 0;
    int nd = 0;
    uint32_t o = 0, m = R;
    const uint32_t inv2 = pa.inv2[q];
#pragma unroll 1
    for (int d = 0; d < ELL; ++d) {
      const uint32_t zz = M - o;
      if (zz == 0 || zz >= m) break;
      const uint32_t h = m >> 1;
      const uint32_t w = tw_at(wn, o >> (ELL - d));
      if (zz < h) {
        const uint32_t r0 = o + zz, cnt = (h - zz) << LOGTC;
        for (uint32_t u = tid; u < cnt; u += NT) {
          const uint32_t r = r0 + (u >> LOGTC), c = u & (TC - 1);
          const uint32_t x = red1(at(r, c), p), y = red1(mont_mul(at(r + h, c), w, p, pinv), p);
          at(r, c) = mont_mul(x + y, inv2, p, pinv);
        }
      } else {
        const uint32_t zr = zz - h;
        const uint32_t r0 = o + zr, cnt = (h - zr) << LOGTC;
        for (uint32_t u = tid; u < cnt; u += NT) {
          const uint32_t r = r0 + (u >> LOGTC), c = u & (TC - 1);
          const uint32_t x = red1(at(r, c), p), y = red1(mont_mul(at(r + h, c), w, p, pinv), p);
          const uint32_t tt = x + p - y;            // (0, 2p)
          at(r, c) = x + red1(tt, p);
          at(r + h, c) = tt;
        }
        dec |= 1u << d;
        o += h;
      }
      m = h;
      nd = d + 1;
      lds_barrier();
    }
    // phase 3: back up the spine
#pragma unroll 1
    for (int d = nd - 1; d >= 0; --d) {
      uint32_t oo = 0;                             // the node's first row: the right turns above it
      for (int e = 0; e < d; ++e)
        if ((dec >> e) & 1u) oo += (uint32_t)R >> (e + 1);
      const uint32_t h = (uint32_t)R >> (d + 1);
      const uint32_t zz = M - oo;
      const uint32_t j = oo >> (ELL - d);
      if (!((dec >> d) & 1u)) {
        const uint32_t w = tw_at(wn, j);
        const uint32_t cnt = zz << LOGTC;
        for (uint32_t u = tid; u < cnt; u += NT) {
          const uint32_t r = oo + (u >> LOGTC), c = u & (TC - 1);
          const uint32_t x = red1(at(r, c), p);
          const uint32_t x2 = red1(x + x, p);
          const uint32_t y = red1(mont_mul(at(r + h, c), w, p, pinv), p);
          at(r, c) = x2 + p - y;
        }
      } else {
        const uint32_t wi = tw_at(wni, j);
        const uint32_t cnt = (zz - h) << LOGTC;
        for (uint32_t u = tid; u < cnt; u += NT) {
          const uint32_t r = oo + (u >> LOGTC), c = u & (TC - 1);
          uint32_t x = at(r, c), y = at(r + h, c);
          bfly_inv(x, y, wi, p, pinv);
          at(r, c) = x;
          at(r + h, c) = y;
        }
      }
      lds_barrier();
    }
  }
  // the column pass finishes the inverse: scale and store the rows k < M
  const uint32_t sc = pa.scale[q];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const uint32_t e = tid + (uint32_t)i * NT;
    const uint32_t k = e >> LOGTC, c = e & (TC - 1);
    if (k < M) (buf + k * S)[c] = mont_mul(at(k, c), sc, p, pinv);
  }
}

// ---------------------------------------------------------------------------
// launch tables
// ---------------------------------------------------------------------------
typedef void (*first_fn)(PrimeArgs, JobRef, int, int);
typedef void (*pass_fn)(PrimeArgs, JobRef, int, int, int, AccArgs);
typedef void (*fused_fn)(PrimeArgs, JobRef, int, int);
typedef void (*inv3_fn)(PrimeArgs, JobRef, int);
struct Plan { int last, bm, split; };

#define SBA_ELL_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
template <int BM> struct KTab {
  // first pass: ELL levels over column tiles of 2^(SBA_FIRST_TILE_LG - ELL) columns
  static first_fn first_two(int ell) {
    switch (ell) {
#define SBA_C(E) case E: return k_first_fwd<E, SBA_FIRST_TILE_LG - E, BM>;
      SBA_ELL_CASES(SBA_C)
#undef SBA_C
    }
    return nullptr;
  }
  static pass_fn inv_two(int ell) {
    switch (ell) {
#define SBA_C(E) case E: return k_pass<E, SBA_FIRST_TILE_LG - E, true, BM>;
      SBA_ELL_CASES(SBA_C)
#undef SBA_C
    }
    return nullptr;
  }
  static first_fn first_single(int L) {
    switch (L) {
      case 10: return k_first_fwd<10, 0, BM>;
      case 11: return k_first_fwd<11, 0, BM>;
      case 12: return k_first_fwd<12, 0, BM>;
      case 13: return k_first_fwd<13, 0, BM>;
    }
    return nullptr;
  }
  static fused_fn fused_single(int L, int sqr) {
    switch (L) {
      case 10: return sqr ? k_fused<10, true, true, BM> : k_fused<10, true, false, BM>;
      case 11: return sqr ? k_fused<11, true, true, BM> : k_fused<11, true, false, BM>;
      case 12: return sqr ? k_fused<12, true, true, BM> : k_fused<12, true, false, BM>;
      case 13: return sqr ? k_fused<13, true, true, BM> : k_fused<13, true, false, BM>;
    }
    return nullptr;
  }
  static pass_fn fwd_last(int last) {
    switch (last) {
      case 11: return k_pass<11, 0, false, BM>;
      case 12: return k_pass<12, 0, false, BM>;
      case 13: return k_pass<13, 0, false, BM>;
    }
    return nullptr;
  }
  static fused_fn fused2_last(int last) {
    switch (last) {
      case 11: return k_fused2<11, BM>;
      case 12: return k_fused2<12, BM>;
      case 13: return k_fused2<13, BM>;
    }
    return nullptr;
  }
  static fused_fn fused_last(int last, int sqr) {
    switch (last) {
      case 11: return sqr ? k_fused<11, false, true, BM> : k_fused<11, false, false, BM>;
      case 12: return sqr ? k_fused<12, false, true, BM> : k_fused<12, false, false, BM>;
      case 13: return sqr ? k_fused<13, false, true, BM> : k_fused<13, false, false, BM>;
    }
    return nullptr;
  }
};

struct Launchers {
  first_fn first;
  pass_fn fwd_last, inv_first;
  fused_fn fused, fused2;          // fused2: both operands' last pass (multiplies)
};
template <int BM> static Launchers launchers_bm(int L, int sqr, int last) {
  Launchers k = {};
  if (last == L) {
    k.first = KTab<BM>::first_single(L);
    k.fused = KTab<BM>::fused_single(L, sqr);
  } else {
    k.first = KTab<BM>::first_two(L - last);
    k.fwd_last = KTab<BM>::fwd_last(last);
    k.fused = KTab<BM>::fused_last(last, sqr);
    k.fused2 = sqr ? nullptr : KTab<BM>::fused2_last(last);
    k.inv_first = KTab<BM>::inv_two(L - last);
  }
  return k;
}
// radix-3 first / last column passes: 2^11 residues of each row per tile,
// ELL = L - last first-pass levels (defined in the radix-32 object)
first_fn sba_r3_first(int ell);
inv3_fn sba_r3_inv(int ell);
// the truncated inverse column pass of ELL = L - last levels (column tiles of
// 2^SBA_FIRST_TILE_LG residues; defined in the radix-32 object)
inv3_fn sba_trunc_inv(int ell);
// defined in ntt_pass_bm.hip, one object per radix
Launchers sba_launchers_bm3(int L, int sqr, int last);
Launchers sba_launchers_bm4(int L, int sqr, int last);
Launchers sba_launchers_bm5(int L, int sqr, int last);
