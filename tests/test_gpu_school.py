"""The schoolbook tiers: small products (an + bn <= 2048 limbs, one
workgroup) and short x long products (the short operand <= 512 limbs, one
workgroup per 1024 product limbs, carries between blocks by the carry scan):
single calls and batches, checked against Python integers and against the NTT path
on the same operands (the tier switched off), at the tier's edges: one-limb
operands, the 2048-limb column limit, all-ones operands (the longest carry
chains), zeros and sparse limbs, squares."""
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import all_ones_product, from_int, to_int, xorshift_fill
from tests import _oracle as orc

pytestmark = pytest.mark.gpu
ONES = np.uint64(0xFFFFFFFFFFFFFFFF)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def operand(n, kind, st):
    if kind == "ones":
        return np.full(n, ONES, dtype=np.uint64)
    x = xorshift_fill(n, st)
    if kind == "sparse":
        x[np.arange(n) % 3 != 0] = 0
    if kind == "topzero" and n > 1:
        x[-1] = 0
    return x


SHAPES = [(1, 1, "rand"), (1, 2047, "rand"), (2047, 1, "ones"), (2, 3, "ones"), (64, 64, "rand"),
          (100, 327, "sparse"), (181, 181, "ones"), (37, 900, "topzero"), (1024, 1024, "ones"),
          (1000, 1048, "rand")]


def _want(a, b):
    return from_int(to_int(a) * to_int(b), len(a) + len(b))


@pytest.fixture
def wide_school():
    """Every shape above on the schoolbook kernel (up to an + bn = 2048)."""
    old = sba.get_school_max()
    sba.set_school_max(1 << 21)
    yield
    sba.set_school_max(old)


def test_school_single_calls(wide_school):
    torch = pytest.importorskip("torch")
    st = [77]
    for an, bn, kind in SHAPES:
        say("school single", an, bn, kind)
        a, b = operand(an, kind, st), operand(bn, kind, st)
        ta = torch.from_numpy(a.view(np.int64)).cuda()
        tb = torch.from_numpy(b.view(np.int64)).cuda()
        tr = torch.full((an + bn + 2,), 0x5A5A, dtype=torch.int64, device="cuda")
        assert sba.mul_dev(tr[:an + bn], ta, tb) == 1
        torch.cuda.synchronize()
        got = tr.cpu().numpy().view(np.uint64)
        assert (got[:an + bn] == _want(a, b)).all(), (an, bn, kind)
        assert (got[an + bn:] == 0x5A5A).all()             # nothing past the product
        if an == bn:
            ts = torch.zeros(2 * an, dtype=torch.int64, device="cuda")
            assert sba.sqr_dev(ts, ta) == 1
            torch.cuda.synchronize()
            assert (ts.cpu().numpy().view(np.uint64) == _want(a, a)).all(), (an, kind, "sqr")


def test_school_batch_matches_ntt_path(wide_school):
    torch = pytest.importorskip("torch")
    st = [78]
    jobs, want = [], []
    for an, bn, kind in SHAPES * 3:
        a, b = operand(an, kind, st), operand(bn, kind, st)
        ta = torch.from_numpy(a.view(np.int64)).cuda()
        tb = torch.from_numpy(b.view(np.int64)).cuda()
        jobs.append((torch.zeros(an + bn, dtype=torch.int64, device="cuda"), ta, tb))
        want.append(_want(a, b))
    for school in (1 << 21, 0):                 # the schoolbook class, then all on the NTT
        sba.set_school_max(school)
        say("school batch", school)
        bt = sba.Batch(jobs, use_graph=True)
        for rep in range(2):
            for tr, _, _ in jobs:
                tr.zero_()
            bt.run()
            torch.cuda.synchronize()
            for k, (tr, _, _) in enumerate(jobs):
                assert (tr.cpu().numpy().view(np.uint64) == want[k]).all(), (school, rep, k)
        bt.close()


@pytest.fixture
def school_long_everywhere():
    """route every short x long shape to k_school_long (the default routes
    only those measured faster than the NTT)"""
    old = sba.get_school_long_ratio()
    sba.set_school_long_ratio(1 << 30)
    yield
    sba.set_school_long_ratio(old)


def test_short_x_long_vs_oracle(school_long_everywhere):
    """The short x long schoolbook: short operands of 1 .. 512 limbs against
    long ones over many 1024-limb blocks, block-boundary lengths, both
    operand orders, random / all-ones / sparse limbs."""
    rng = np.random.default_rng(77)
    shapes = [(2049, 1), (2048, 3), (5000, 512), (4096, 511), (1023 * 7 + 2, 129),
              (65536, 64), (100003, 300), (1 << 18, 17), (3 * 1024 - 2, 2), (20000, 512)]
    for an, bn in shapes:
        assert sba.school_route(an, bn) == 2, (an, bn)
        for pat in range(3):
            a = rng.integers(0, 2**64, an, dtype=np.uint64)
            b = rng.integers(0, 2**64, bn, dtype=np.uint64)
            if pat == 1:
                a[:] = 2**64 - 1
                b[:] = 2**64 - 1
            elif pat == 2:
                a[rng.random(an) < 0.9] = 0
                b[rng.random(bn) < 0.5] = 0
            ok, r = sba.nt_mul(a, b)
            assert ok == 1 and (r == orc.nt_mul(a, b)).all(), (an, bn, pat)
            ok, r = sba.nt_mul(b, a)
            assert ok == 1 and (r == orc.nt_mul(a, b)).all(), (bn, an, pat)
    a = np.full(70000, 2**64 - 1, np.uint64)
    b = np.full(500, 2**64 - 1, np.uint64)
    ok, r = sba.nt_mul(a, b)
    assert ok == 1 and (r == all_ones_product(70000, 500)).all()


@pytest.mark.parametrize("use_graph", [False, True])
def test_short_x_long_batch_class(school_long_everywhere, use_graph):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(78)
    jobs, want = [], []
    for k in range(60):
        an = int(rng.integers(2100, 70000))
        bn = int(rng.integers(1, 513))
        if k % 2:
            an, bn = bn, an
        a = rng.integers(0, 2**64, an, dtype=np.uint64)
        b = rng.integers(0, 2**64, bn, dtype=np.uint64)
        if k % 5 == 0:
            a[:] = 2**64 - 1
            b[:] = 2**64 - 1
        jobs.append((torch.zeros(an + bn, dtype=torch.int64, device="cuda"),
                     torch.from_numpy(a.view(np.int64)).cuda(),
                     torch.from_numpy(b.view(np.int64)).cuda()))
        want.append(orc.nt_mul(a, b))
    # plus a few NTT-sized and small jobs in the same batch
    for an, bn in ((3000, 3000), (300, 200), (40000, 30000)):
        a = rng.integers(0, 2**64, an, dtype=np.uint64)
        b = rng.integers(0, 2**64, bn, dtype=np.uint64)
        jobs.append((torch.zeros(an + bn, dtype=torch.int64, device="cuda"),
                     torch.from_numpy(a.view(np.int64)).cuda(),
                     torch.from_numpy(b.view(np.int64)).cuda()))
        want.append(orc.nt_mul(a, b))
    bt = sba.Batch(jobs, use_graph=use_graph)
    assert bt.classes >= 3
    for rep in range(2):
        for r, _, _ in jobs:
            r.zero_()
        bt.run()
        torch.cuda.synchronize()
        for k, (r, _, _) in enumerate(jobs):
            assert (r.cpu().numpy().view(np.uint64) == want[k]).all(), (rep, k)
    bt.close()
