/* sba_internal.h -- declarations shared by the host C files. */
#ifndef SBA_INTERNAL_H
#define SBA_INTERNAL_H
#include "simd_bigint_amd.h"

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* total limbs (an+bn) from which simd_mpn_mul and the division's
 * multiplies go to the GPU; below, the library's CPU Karatsuba */
#define SBA_GPU_MIN_TOTAL_DEFAULT 12288u

/* rp[0..an+bn) = a * b through the threshold dispatch (either order) */
void sba_mul(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp, size_t bn);

/* X[0..n] (n + 1 limbs) = approximate reciprocal of A[0..n) (top bit set):
 * A*X < B^2n <= A*(X+2), B = 2^64 (host Newton, sba_div.c) */
void sba_approx_recip_host(uint64_t *X, const uint64_t *A, size_t n);

/* the Newton branch of the division on device buffers (div_dev.hip): from
 * the normalised Nn (qn + dn limbs), Dn (dn limbs) and Dm (qn + 1 limbs; or
 * exact_pow when Dm + 1 = B^(qn+1)), Q[0..qn] = the quotient estimate and
 * R[0..dn] = the low limbs of Nn - Q * Dn. SBA_OK; SBA_BAND if the estimate
 * is out of its bounds; < 0 on a device error. */
int sba_div_dev(uint64_t *Q, uint64_t *R, const uint64_t *Nn, const uint64_t *Dn, size_t dn,
                const uint64_t *Dm, size_t qn, int exact_pow);

#ifdef __cplusplus
}
#endif
#endif
