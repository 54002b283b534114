/* sba_internal.h -- declarations shared by the host C files. */
#ifndef SBA_INTERNAL_H
#define SBA_INTERNAL_H
#include "simd_bigint_amd.h"

#include <stddef.h>
#include <stdint.h>

/* total limbs (an+bn) from which simd_mpn_mul and the division's
 * multiplies go to the GPU; below, the library's CPU Karatsuba */
#define SBA_GPU_MIN_TOTAL_DEFAULT 12288u

/* rp[0..an+bn) = a * b through the threshold dispatch (either order) */
void sba_mul(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp, size_t bn);

#endif
