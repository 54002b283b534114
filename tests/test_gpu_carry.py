"""The carry stage between limb blocks, driven with chosen carry
chains: blocks whose carry-in is 1 and whose leading limbs are all ones
(short runs, runs across most of a block, whole blocks), many such blocks
per workgroup of the stage. Random products almost never reach
this path (a carry leaves a block only on a generate followed by all-ones
limbs), so the whole-product parity tests do not cover it."""
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import xorshift_fill

pytestmark = pytest.mark.gpu
ONES = np.uint64(0xFFFFFFFFFFFFFFFF)


def expected(limbs, maps, Z, LPB):
    """Carry c into block b from the maps (bit c of maps[b] = carry out for
    carry in c); a block with carry-in 1 gains 1 as a multi-limb number."""
    out = limbs.copy()
    c = 0
    for b in range(len(maps)):
        cin, c = c, (int(maps[b]) >> c) & 1
        if cin:
            B0, E = b * LPB, min(LPB, Z - b * LPB)
            blk = out[B0:B0 + E]
            nz = np.nonzero(blk != ONES)[0]
            f = int(nz[0]) if len(nz) else E
            blk[:f] = 0
            if f < E:
                blk[f] += np.uint64(1)
    return out


@pytest.mark.parametrize("nb,tail", [(3, 5), (1500, 777), (3000, 1)])
def test_carry_blocks_chosen_chains(nb, tail):
    torch = pytest.importorskip("torch")
    LPB = sba.carry_block_limbs()
    tail = tail % LPB or 1
    print("carry_blocks", nb, tail, LPB, file=sys.stderr, flush=True)
    Z = nb * LPB - tail
    st = [nb * 31 + tail]
    limbs = xorshift_fill(Z, st)
    rng = np.random.default_rng(nb)
    # kill / propagate / generate maps: many blocks receive a carry
    maps = rng.choice(np.array([0, 2, 3], dtype=np.uint32), size=nb, p=[0.2, 0.3, 0.5])
    for b in range(nb):
        B0, E = b * LPB, min(LPB, Z - b * LPB)
        kind = b % 6
        if kind == 1:
            limbs[B0:B0 + min(E, 1 + b % 7)] = ONES                 # short run
        elif kind == 2:
            limbs[B0:B0 + max(1, E - 1 - b % 3)] = ONES              # most of the block
        elif kind == 3:
            limbs[B0:B0 + E] = ONES                                  # the whole block
        elif kind == 4:
            limbs[B0] = ONES - np.uint64(1)                          # one below all ones
    want = expected(limbs, maps, Z, LPB)
    r = torch.from_numpy(limbs.view(np.int64).copy()).cuda()
    m = torch.from_numpy(maps.view(np.int32).copy()).cuda()
    assert sba.carry_blocks(r, m) == 1
    torch.cuda.synchronize()
    got = r.cpu().numpy().view(np.uint64)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, (len(bad), bad[:5])
