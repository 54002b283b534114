/* ref_nt.c -- exports the reference's NTT entries (include/ntmul.h:47-79,
 * include/p30.h:1847-2020, include/p50.h:1020-1178) with C linkage so the
 * tests can compute golden vectors. TEST INFRASTRUCTURE ONLY: this file holds
 * no reference code; it includes the reference headers where they lie and is
 * compiled by oracle/Makefile into oracle/_ref/. */
#define SCRATCH_IMPLEMENTATION
#include "ntmul.h"

int sbref_nt_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
                 const uint64_t *bp, ptrdiff_t bn) {
  return nt_mul(rp, ap, an, bp, bn);
}
int sbref_nt_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an) {
  return nt_sqr(rp, ap, an);
}
int sbref_p30_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
                  const uint64_t *bp, ptrdiff_t bn) {
  return p30_mul(rp, ap, an, bp, bn);
}
int sbref_p50_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
                  const uint64_t *bp, ptrdiff_t bn) {
  return p50_mul(rp, ap, an, bp, bn);
}
