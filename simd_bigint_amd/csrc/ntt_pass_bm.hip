// ntt_pass_bm.hip -- instantiates the pass kernels of one register-chunk
// radix (2^SBA_TU_BM) and exports their launch table. Compiled once per
// radix (-DSBA_TU_BM=3, 4, 5).
#include "ntt_tiles.h"

#ifndef SBA_TU_BM
#error "SBA_TU_BM must be defined"
#endif
#define SBA_CAT2(a, b) a##b
#define SBA_CAT(a, b) SBA_CAT2(a, b)

Launchers SBA_CAT(sba_launchers_bm, SBA_TU_BM)(int L, int sqr, int last) {
  return launchers_bm<SBA_TU_BM>(L, sqr, last);
}

#if SBA_TU_BM == 5
#define SBA_R3_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9)
first_fn sba_r3_first(int ell) {
  switch (ell) {
#define SBA_C(E) case E: return k_first_fwd3<E, 11 - E>;
    SBA_R3_CASES(SBA_C)
#undef SBA_C
  }
  return nullptr;
}
inv3_fn sba_r3_inv(int ell) {
  switch (ell) {
#define SBA_C(E) case E: return k_inv_first3<E, 11 - E>;
    SBA_R3_CASES(SBA_C)
#undef SBA_C
  }
  return nullptr;
}
#define SBA_TRUNC_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
inv3_fn sba_trunc_inv(int ell) {
  switch (ell) {
#define SBA_C(E) case E: return k_inv_trunc<E, SBA_FIRST_TILE_LG - E>;
    SBA_TRUNC_CASES(SBA_C)
#undef SBA_C
  }
  return nullptr;
}
#endif
