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
 * multiplies go to the GPU; below, the library's CPU Karatsuba. The
 * crossover tuner (sba_tune_gpu_threshold) measured 768 on an MI355X box
 * (balanced products, host buffers, copies included: 384 x 384 limbs in
 * 36 us on either side); rounded up to 1024 (57 us CPU, 37 us GPU). */
#define SBA_GPU_MIN_TOTAL_DEFAULT 1024u

/* rp[0..an+bn) = a * b through the threshold dispatch (either order) */
void sba_mul(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp, size_t bn);

/* X[0..n] (n + 1 limbs) = approximate reciprocal of A[0..n) (top bit set):
 * A*X < B^2n <= A*(X+2), B = 2^64 (host Newton, sba_div.c) */
void sba_approx_recip_host(uint64_t *X, const uint64_t *A, size_t n);

/* a GPU failure whose product the CPU tier finishes: reported on stderr
 * and counted (sba_gpu_fallbacks) */
void sba_note_fallback(const char *what, int rc);

/* memcpy split over several host threads (sba_host.c) */
void sba_par_copy(void *dst, const void *src, size_t bytes);

/* the division's Newton branch on device buffers (div_dev.hip): N of nt
 * limbs and D of dn limbs (top limbs nonzero), s = leading zero bits of D's
 * top limb, quotient of qn = nt + 1 - dn limbs. Writes qp[0..min(qn, qm))
 * and rp[0..dn). SBA_OK; SBA_BAND if an estimate leaves its bounds; < 0 on
 * a device error. */
int sba_div_dev(uint64_t *qp, size_t qm, uint64_t *rp, const uint64_t *np, size_t nt,
                const uint64_t *dp, size_t dn, unsigned s);

#ifdef __cplusplus
}
#endif
#endif
