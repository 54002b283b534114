// gpu_arith.h -- mod-p arithmetic for the gfx950 NTT engine.
//
// Five primes per prime set, ascending (Garner relies on the order), each
// below 2^31. Values between butterflies are lazy in [0, 2p) (< 2^32 since
// p < 2^31); a set's leading primes below 2^30 may use Harvey's forms below,
// whose forward values are lazy in [0, 4p).
// Every multiply is a Montgomery product with R = 2^32 built from two
// v_mad_u64_u32 and one v_mul_lo_u32; the constant operand (twiddles, Garner
// constants) is stored in Montgomery form and canonical (< p), which keeps
// the product in [0, 2p) for ANY 32-bit left operand:
//   (a*b + m*p) / 2^32 < (2^32*p + 2^32*p) / 2^32 = 2p.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SBA_K 5            // primes
#define SBA_MAX_LG 25      // the largest transform, 2^25 (2^25 | p - 1 in the wide set)

__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t p,
                                             uint32_t pinv) {
  uint64_t t = (uint64_t)a * b;
  uint32_t m = (uint32_t)t * pinv;          // pinv = -p^-1 mod 2^32
  uint64_t u = (uint64_t)m * p + t;         // < 2^63 + 2^63
  return (uint32_t)(u >> 32);
}

// [0, 2p) -> [0, p): if x < p, x - p wraps above 2^31 and min keeps x.
__device__ __forceinline__ uint32_t red1(uint32_t x, uint32_t p) {
  uint32_t y = x - p;
  return y < x ? y : x;
}

// forward (Cooley-Tukey, node twiddle w): (x, y) -> (x + w y, x - w y)
__device__ __forceinline__ void bfly_fwd(uint32_t &x, uint32_t &y, uint32_t w,
                                         uint32_t p, uint32_t pinv) {
  uint32_t a = red1(x, p);
  uint32_t t = red1(mont_mul(y, w, p, pinv), p);
  x = a + t;            // [0, 2p)
  y = a + p - t;        // (0, 2p)
}

// inverse (Gentleman-Sande, inverse node twiddle wi): (x, y) -> (x + y, (x - y) wi)
__device__ __forceinline__ void bfly_inv(uint32_t &x, uint32_t &y, uint32_t wi,
                                         uint32_t p, uint32_t pinv) {
  uint32_t a = red1(x, p);
  uint32_t b = red1(y, p);
  x = a + b;                                   // [0, 2p)
  y = mont_mul(a + p - b, wi, p, pinv);        // [0, 2p)
}

// Harvey's lazy forms for the primes p < 2^30 of a prime set (4p < 2^32):
// eight VALU instructions per butterfly instead of ten.
// [0, 4p) -> [0, 2p)
__device__ __forceinline__ uint32_t red2(uint32_t x, uint32_t p2) {
  uint32_t y = x - p2;
  return y < x ? y : x;
}

// forward: x in [0, 4p), y any 32-bit value -> both outputs in [0, 4p)
__device__ __forceinline__ void bfly_fwd_h(uint32_t &x, uint32_t &y, uint32_t w, uint32_t p,
                                           uint32_t pinv) {
  const uint32_t p2 = p + p;
  const uint32_t a = red2(x, p2);                   // [0, 2p)
  const uint32_t t = mont_mul(y, w, p, pinv);       // [0, 2p)
  x = a + t;                                        // [0, 4p)
  y = a + p2 - t;                                   // (0, 4p)
}

// inverse: x, y in [0, 2p) -> both outputs in [0, 2p)
__device__ __forceinline__ void bfly_inv_h(uint32_t &x, uint32_t &y, uint32_t wi, uint32_t p,
                                           uint32_t pinv) {
  const uint32_t p2 = p + p;
  const uint32_t s = x + y;                         // [0, 4p)
  const uint32_t d = x + p2 - y;                    // (0, 4p)
  x = red2(s, p2);
  y = mont_mul(d, wi, p, pinv);
}

// which butterfly forms a pass runs: the [0, 2p) forms (HV_W), Harvey's
// (HV_H, only for p < 2^30), or Harvey's exactly where the prime allows them
// (HV_AUTO: a uniform branch on the block's prime)
#define HV_W 0
#define HV_H 1
#define HV_AUTO 2
__device__ __forceinline__ bool harvey_prime(uint32_t p) { return p < (1u << 30); }

// a stored forward value as the input of a pass running the forms HV: the
// forward values of a prime below 2^30 may lie in [0, 4p) (a pass with
// Harvey's forms wrote them); the [0, 2p) forms need them folded once
template <int HV>
__device__ __forceinline__ uint32_t fwd_in(uint32_t x, uint32_t p) {
  return (HV == HV_W && harvey_prime(p)) ? red2(x, p + p) : x;
}

// a forward output as the canonical Montgomery operand of the pointwise
// product: [0, 4p) (Harvey's forms) or [0, 2p) -> [0, p). Under HV_AUTO a
// prime below 2^30 had Harvey's forms.
template <int HV>
__device__ __forceinline__ uint32_t canon(uint32_t x, uint32_t p) {
  const bool h = HV == HV_H || (HV == HV_AUTO && harvey_prime(p));
  return h ? red1(red2(x, p + p), p) : red1(x, p);
}
