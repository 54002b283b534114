"""The schoolbook tier for small products (an + bn <= 2048 limbs): single
calls and batches, checked against Python integers and against the NTT path
on the same operands (the tier switched off), at the tier's edges: one-limb
operands, the 2048-limb column limit, all-ones operands (the longest carry
chains), zeros and sparse limbs, squares."""
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import from_int, to_int, xorshift_fill

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
