"""GPU: radix-3 lengths. A product with 2^(L+1) < an+bn-1 <= 3 * 2^L runs at
length 3 * 2^L (single products from L = 16 by default, every class here) (Good-Thomas: a length-3 DFT down the columns of a 3 x 2^L
array, riding the first forward and last inverse passes); every shape at
the edges of each such class is bit-exact against the oracle, for
multiplies and squares, and the balanced 3 * 2^20-limb benchmark pair
(golden SHA-256 in test_gpu_parity) takes this path."""
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import all_ones_product, xorshift_fill
from tests import _oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def all_radix3_classes():
    """single products take radix-3 classes from row length 2^16 by default;
    these tests reach every class from 2^13"""
    old = sba.get_r3_min_lg()
    sba.set_r3_min_lg(13)
    yield
    sba.set_r3_min_lg(old)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _shapes(L):
    """(an, bn) at the low edge, the top edge and inside class L."""
    lo = (1 << (L + 1)) + 1          # smallest zlen of the class: 2^(L+1) + 1
    hi = 3 << L                      # largest: 3 * 2^L
    return [((lo + 2) // 2, lo + 1 - (lo + 2) // 2), ((hi + 1) // 2, hi + 1 - (hi + 1) // 2),
            (hi - hi // 4, hi // 4 + 1), ((5 << L) // 4, (5 << L) // 4 - 3)]


@pytest.mark.parametrize("L", [13, 14, 15, 17, 20])
def test_radix3_class_edges_vs_oracle(L):
    st = [300 + L]
    for an, bn in _shapes(L):
        assert sba.lib().sba_gpu_radix3(an, bn, 0) == L, (an, bn)
        a = xorshift_fill(an, st)
        b = xorshift_fill(bn, st)
        ok, r = sba.nt_mul(a, b)
        assert ok == 1
        assert (r == orc.nt_mul(a, b)).all(), (an, bn)
    say("L=%d: %d shapes exact" % (L, len(_shapes(L))))


@pytest.mark.parametrize("L", [13, 16, 21])
def test_radix3_squares_vs_oracle(L):
    st = [400 + L]
    for an in ((1 << L) + 1, (3 << L) // 2, (3 << L) // 2 - 77):
        assert sba.lib().sba_gpu_radix3(an, an, 1) == L, an
        a = xorshift_fill(an, st)
        ok, r = sba.nt_sqr(a)
        assert ok == 1
        assert (r == orc.nt_sqr(a)).all(), an


def test_radix3_all_ones_top_of_class():
    n = 3 << 19                      # zlen = 3 * 2^20 - 1: top of class L = 20
    assert sba.lib().sba_gpu_radix3(n, n, 0) == 20
    a = np.full(n, 2**64 - 1, np.uint64)
    ok, r = sba.nt_mul(a, a)
    assert ok == 1 and (r == all_ones_product(n, n)).all()
    ok, r = sba.nt_sqr(a)
    assert ok == 1 and (r == all_ones_product(n, n)).all()
