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
