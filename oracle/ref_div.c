/* ref_div.c -- exports the reference's divisions divrem_u64
 * (include/divrem.h:218-280) and divrem_u64_dc (include/dc_div.h:154-208)
 * with C linkage for golden vectors. TEST INFRASTRUCTURE ONLY; includes the
 * reference header, holds no reference code. */
#include "dc_div.h"

void sbref_divrem_u64(uint64_t *qp, uint64_t *rp, const uint64_t *np,
                      uint64_t nn, const uint64_t *dp, uint64_t dn,
                      uint64_t *qn, uint64_t *rn) {
  divrem_u64(qp, rp, np, nn, dp, dn, qn, rn);
}
void sbref_divrem_u64_dc(uint64_t *qp, uint64_t *rp, const uint64_t *np,
                         uint64_t nn, const uint64_t *dp, uint64_t dn,
                         uint64_t *qn, uint64_t *rn) {
  divrem_u64_dc(qp, rp, np, nn, dp, dn, qn, rn);
}
