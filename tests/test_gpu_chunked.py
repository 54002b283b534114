"""GPU: chunked products -- the longer operand cut into chunks multiplied
against the shorter one's spectrum, the chunks' residues accumulated with
overlapping tails before one Garner + carry pass -- are bit-exact: against
the oracle at chunk-boundary shapes, against the all-ones closed form, and
by fingerprints beyond 2^25 coefficients (the reference's band reaches
2^25 trunks of 84-88 bits)."""
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import all_ones_product, xorshift_fill
from tests import _oracle as orc

pytestmark = pytest.mark.gpu
QS = [(1 << 61) - 1, 2305843009213693921]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _check_fp(r, x, y):
    for q in QS:
        assert orc.fingerprint(r, q) == orc.fingerprint(x, q) * orc.fingerprint(y, q) % q


def _with_last_chunk(bn, rem):
    """An operand length whose chunked product against bn limbs ends with a
    last chunk of `rem` limbs (rem = 0: a full last chunk)."""
    for an0 in range(1 << 21, 1 << 23, 1 << 18):
        Lc, C, n = sba.chunking(an0, bn)
        if n < 2:
            continue
        for k in range(2, 12):
            an = k * C + rem if rem else k * C
            got = sba.chunking(an, bn)
            if got and got[1] == C and got[2] >= 2:
                return an
    raise AssertionError("no shape with last chunk %d" % rem)


def _shapes():
    return [(1 << 21, 4096), (_with_last_chunk(4096, 0), 4096),
            (_with_last_chunk(4096, 1), 4096), (_with_last_chunk(4096, 4095), 4096),
            (1 << 22, 1 << 14), (3_000_001, 777)]


@pytest.mark.parametrize("shape", range(6))
def test_chunk_boundaries_vs_oracle(shape):
    an, bn = _shapes()[shape]
    ch = sba.chunking(an, bn)
    say("an=%d bn=%d chunking %r" % (an, bn, ch))
    assert ch is not None and ch[2] >= 2
    st = [1000 + shape]
    x = xorshift_fill(an, st)
    y = xorshift_fill(bn, st)
    for a, b in ((x, y), (y, x)):                 # either operand order
        ok, r = sba.nt_mul(a, b)
        assert ok == 1
        assert (r == orc.nt_mul(a, b)).all()


def test_chunked_all_ones():
    an, bn = 1 << 22, 1 << 14
    assert sba.chunking(an, bn)[2] >= 2
    ok, r = sba.nt_mul(np.full(an, 2**64 - 1, np.uint64), np.full(bn, 2**64 - 1, np.uint64))
    assert ok == 1 and (r == all_ones_product(an, bn)).all()


@pytest.mark.parametrize("tier", ["school_long", "chunked"])
def test_chunked_short_operand_of_one_limb(tier):
    """2^23 x 1 limbs: by default the short x long schoolbook takes it; with
    that tier off, the chunked NTT (the same product either way)."""
    an = 1 << 23
    old = sba.get_school_long_ratio()
    if tier == "chunked":
        sba.set_school_long_ratio(0)
    try:
        assert (sba.school_route(an, 1) == 2) == (tier == "school_long")
        assert (sba.chunking(an, 1)[2] >= 2) == (tier == "chunked")
        st = [5]
        x = xorshift_fill(an, st)
        y = np.array([0xFFFFFFFFFFFFFFFF], np.uint64)
        ok, r = sba.nt_mul(x, y)
    finally:
        sba.set_school_long_ratio(old)
    want = np.zeros(an + 1, np.uint64)
    orc.lib().orc_mul_school(want, x, an, y, 1)
    assert ok == 1 and (r == want).all()


def test_beyond_one_transform_long_by_short():
    """2^25 + 5 by 2^20 limbs: 34.6M-limb product, over 2^25 coefficients
    (no single transform), inside the reference's band."""
    an, bn = (1 << 25) + 5, 1 << 20
    assert sba.lib().sba_ref_nt_band(an, bn) == 1
    st = [77]
    x = xorshift_fill(an, st)
    y = xorshift_fill(bn, st)
    ok, r = sba.nt_mul(x, y)
    assert ok == 1
    _check_fp(r, x, y)


def test_beyond_one_transform_balanced_band_edge():
    """The largest balanced pair the reference accepts (22020096 limbs
    each, 2^25 trunks of 84 bits): computed in two chunks; one more limb in
    each is refused."""
    n = 22020096
    assert sba.lib().sba_ref_nt_band(n, n) == 1
    assert sba.lib().sba_ref_nt_band(n + 1, n + 1) == 0
    st = [78]
    x = xorshift_fill(n, st)
    y = xorshift_fill(n, st)
    ok, r = sba.nt_mul(x, y)
    assert ok == 1
    _check_fp(r, x, y)
    ok, r = sba.nt_sqr(x)
    assert ok == 1
    _check_fp(r, x, x)
