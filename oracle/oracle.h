/* oracle.h -- CPU oracle for the large-operand multiply path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the simd_bigint_amd
 * library, its kernels or its host code) may include, link or call this.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker.
 *
 * What it restates (reference paths relative to the reference repository):
 *   orc_nt_mul / orc_nt_sqr  the p30 engine's math as written in the
 *                            reference's scalar model
 *                            tests/p30_tft_model_test.c:66-321 and
 *                            include/p30.h:70-78,200-329,1847-1938:
 *                            64-bit chunking, W-lane decomposition
 *                            y = x^16, van der Hoeven truncated NTT with
 *                            the prefix-stable bit-reversed root table,
 *                            twisted cyclic convolution mod (x^16 - w[k]),
 *                            ascending-prime Garner, Horner compose,
 *                            carry into limbs. The prime set is p30's
 *                            five primes, widened to six (adding
 *                            897581057 = 107*2^23+1, the next prime of
 *                            the k*2^22+1 family) when min(an,bn)
 *                            exceeds p30's CRT cap, so the oracle covers
 *                            the sizes the reference serves with p50.
 *   orc_mul_school           u128 schoolbook, the reference's scalar
 *                            oracle tests/mpn_mul_u64_test.cpp:29-41.
 *   orc_divrem               Knuth algorithm D on u64 limbs with the
 *                            contract of include/divrem.h:212-221.
 *   orc_xorshift             the reference's generator,
 *                            bench/nt_sweep.c:15-16 (13, 7, 17).
 *
 * How each part is pinned (tests/test_oracle.py):
 *   orc_nt_mul/sqr   == products of the reference compiled from source
 *                       (tests/golden fixtures, SHA-256 + full vectors),
 *                       == orc_mul_school on random small shapes,
 *                       == the all-ones closed form, modular fingerprints.
 *   orc_mul_school   == Python big-integer multiplication (exact).
 *   orc_divrem       N == Q*D + R, 0 <= R < D, and == Python divmod.
 *   orc_residues     == Python a % p.
 *   orc_fingerprint  == Python int % q.
 */
#ifndef SBA_ORACLE_H
#define SBA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint64_t orc_xorshift(uint64_t *state);
void orc_fill_xorshift(uint64_t *p, size_t n, uint64_t *state);

void orc_mul_school(uint64_t *rp, const uint64_t *ap, size_t an,
                    const uint64_t *bp, size_t bn);

/* 1 ok, 0 beyond the oracle's band (an/bn <= 0, CRT cap of six primes,
 * or the 2^22-vector transform cap). */
int orc_nt_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
               const uint64_t *bp, ptrdiff_t bn);
int orc_nt_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an);

/* number of primes the oracle would use for this shape (5 or 6), 0 if out
 * of band */
int orc_nt_nprimes(ptrdiff_t an, ptrdiff_t bn);

void orc_residues(uint32_t *out, const uint64_t *ap, size_t an, uint32_t p);
uint64_t orc_fingerprint(const uint64_t *ap, size_t an, uint64_t q);

/* Q = floor(N/D), R = N mod D. qp holds >= nn-dn+1 limbs, rp >= dn limbs.
 * dp[dn-1] != 0. Lengths out are trimmed, >= 1. */
void orc_divrem(uint64_t *qp, uint64_t *rp, const uint64_t *np, size_t nn,
                const uint64_t *dp, size_t dn, size_t *qn_out,
                size_t *rn_out);

#ifdef __cplusplus
}
#endif
#endif
