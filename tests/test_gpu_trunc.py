"""Truncated products on the GPU: a power-of-two transform whose merged last
pass runs only the tiles below ceil(zlen / 2^last), inverted by the
truncated inverse column pass (van der Hoeven's inverse TFT, as
tests/p30_tft_model_test.c:125-181 and include/p30_ref.h:556-611 state it).
Bit-exact against the oracle at every spine shape of small column
transforms, at fills across (1/2, 1] of lengths 2^14 .. 2^22, at tile
boundaries, for all-ones operands, squares and mixed-fill batch classes;
the same products with truncation off agree."""
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import all_ones_product, xorshift_fill
from tests import _oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(autouse=True)
def restore_trunc():
    """these tests truncate every two-pass product that leaves a tile out
    (the default routing truncates only where it is faster)"""
    old = sba.get_trunc()
    sba.set_trunc(100, 100, 0)
    yield
    sba.set_trunc(*old)


def _check(an, bn, st, sqr=False, want_trunc=True):
    a = xorshift_fill(an, st)
    if sqr:
        tr = sba.truncated(an, an, True)
        ok, r = sba.nt_sqr(a)
        want = orc.nt_sqr(a)
    else:
        b = xorshift_fill(bn, st)
        tr = sba.truncated(an, bn)
        ok, r = sba.nt_mul(a, b)
        want = orc.nt_mul(a, b)
    if want_trunc:
        assert tr is not None, (an, bn)
    assert ok == 1 and (r == want).all(), (an, bn, sqr, tr)
    return tr


def test_every_row_count_of_small_column_transforms():
    """L = 14, 15, 16: every tile count M of the column transform (all spine
    shapes of the truncated inverse), products ending inside the last
    transformed tile and on its boundary."""
    st = [314]
    for L in (14, 15, 16):
        seen = set()
        for zlen in range((1 << (L - 1)) + 1, (1 << L) + 1, 97):
            for z in (zlen, (zlen | 2047)):       # also on a 2^11 tile boundary
                if z > (1 << L):
                    continue
                an = z // 3 + 1
                bn = z + 1 - an
                tr = _check(an, bn, st, want_trunc=False)
                if tr:
                    seen.add((tr[1], tr[2]))
        say("L", L, "tile counts seen", len(seen))
        assert len(seen) >= 4


@pytest.mark.parametrize("L", [17, 19, 20, 22])
def test_fills_across_the_octave(L):
    st = [L * 7]
    N = 1 << L
    for f in (0.5, 0.53, 0.6, 0.63, 0.77, 0.81, 0.87, 0.9):
        zlen = int(N * f) + 1
        an = zlen // 2 + 1
        bn = zlen + 1 - an
        tr = _check(an, bn, st, want_trunc=False)
        say("L", L, "fill", f, "trunc", tr)


def test_all_ones_and_squares_truncated():
    for an, bn in ((300000, 300000), (5000, 5000), (70000, 11000)):
        assert sba.truncated(an, bn) is not None, (an, bn)
        a = np.full(an, 2**64 - 1, np.uint64)
        b = np.full(bn, 2**64 - 1, np.uint64)
        ok, r = sba.nt_mul(a, b)
        assert ok == 1 and (r == all_ones_product(an, bn)).all(), (an, bn)
    st = [27]
    for an in (150000, 4600, 600000):
        _check(an, an, st, sqr=True)


def test_truncation_off_gives_the_same_products():
    st = [5]
    for an, bn in ((300000, 300000), (1 << 19, 1 << 17), (9000, 8000)):
        a, b = xorshift_fill(an, st), xorshift_fill(bn, st)
        assert sba.truncated(an, bn) is not None
        ok, r1 = sba.nt_mul(a, b)
        sba.set_trunc(0, 0, 0)
        assert sba.truncated(an, bn) is None
        ok2, r2 = sba.nt_mul(a, b)
        sba.set_trunc(100, 100, 0)
        assert ok == 1 and ok2 == 1 and (r1 == r2).all()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_class_with_mixed_fills(use_graph):
    """One size class (length 2^16) whose jobs fill from just over half to
    all of it: each job's own tile count, one launch."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    jobs, want = [], []
    for k in range(40):
        zlen = int((1 << 15) + 1 + k * ((1 << 15) - 1) / 39)
        an = max(zlen // (2 + k % 3), 1)
        bn = zlen + 1 - an
        a = rng.integers(0, 2**64, an, dtype=np.uint64)
        b = rng.integers(0, 2**64, bn, dtype=np.uint64)
        ta = torch.from_numpy(a.view(np.int64)).cuda()
        tb = torch.from_numpy(b.view(np.int64)).cuda()
        jobs.append((torch.zeros(an + bn, dtype=torch.int64, device="cuda"), ta, tb))
        want.append(orc.nt_mul(a, b))
    bt = sba.Batch(jobs, use_graph=use_graph)
    for rep in range(2):
        for r, _, _ in jobs:
            r.zero_()
        bt.run()
        torch.cuda.synchronize()
        for k, (r, _, _) in enumerate(jobs):
            assert (r.cpu().numpy().view(np.uint64) == want[k]).all(), (rep, k)
    bt.close()


def test_default_routing_truncates_where_faster():
    """With the shipped thresholds a product filling 80% of 2^22 runs
    truncated, one filling 70% runs at the radix-3 length, one filling 95%
    at the full length; all three bit-exact."""
    sba.set_trunc(85, 0, 21)
    st = [8]
    n80 = int(0.8 * (1 << 21))
    assert sba.truncated(n80, n80) is not None
    n70 = int(0.7 * (1 << 21))
    assert sba.truncated(n70, n70) is None and sba.lib().sba_gpu_radix3(n70, n70, 0) == 20
    n95 = int(0.95 * (1 << 21))
    assert sba.truncated(n95, n95) is None and sba.lib().sba_gpu_radix3(n95, n95, 0) == 0
    for n in (n80, n70, n95):
        _check(n, n, st, want_trunc=False)
