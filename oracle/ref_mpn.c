/* ref_mpn.c -- exports the reference's public multiply simd_mpn_mul
 * (include/u64cvt.h:94-132) with C linkage for golden vectors. TEST
 * INFRASTRUCTURE ONLY; includes the reference header, holds no reference
 * code. The scratch arena's thread-local storage is defined in ref_nt.c. */
#include "u64cvt.h"

void sbref_simd_mpn_mul(uint64_t *rp, const uint64_t *ap, uint64_t an,
                        const uint64_t *bp, uint64_t bn) {
  simd_mpn_mul(rp, ap, an, bp, bn);
}
