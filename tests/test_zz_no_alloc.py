"""The allocation-free entry sba_gpu_mul_dev_noalloc, through the C ABI,
against the existing device entry sba_gpu_mul_dev (bit-exact): several
calls in a row on the shared static scratch, each later one smaller, over
the radix-3, two-pass, single-tile and schoolbook routes and a square;
products beyond the static scratch are rejected with SBA_ERR_TOO_LARGE and
leave the result untouched."""
import ctypes
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import xorshift_fill

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x).view(np.int64)).cuda()


def test_noalloc_matches_mul_dev():
    torch = pytest.importorskip("torch")
    L = sba.lib()
    # largest first (radix-3 at 3 x 2^18, then 2^20 / 2^19 two-pass), then
    # smaller: every later call reuses the scratch the earlier one left
    shapes = [(400000, 300000), (600000, 400000), (1 << 19, 1 << 19), (150000, 50000),
              (300000, 300000, "sqr"), (5000, 3000), (700, 500), (100, 90)]
    st = [2024]
    for sh in shapes:
        an, bn = sh[0], sh[1]
        sq = len(sh) > 2
        say("noalloc case", an, bn, "sqr" if sq else "mul")
        a = xorshift_fill(an, st)
        b = a if sq else xorshift_fill(bn, st)
        ta = _dev(torch, a)
        tb = ta if sq else _dev(torch, b)
        want = torch.zeros(an + bn, dtype=torch.int64, device="cuda")
        got = torch.full((an + bn,), -1, dtype=torch.int64, device="cuda")
        if sq:
            assert L.sba_gpu_sqr_dev(want.data_ptr(), ta.data_ptr(), an, None) == 1
        else:
            assert L.sba_gpu_mul_dev(want.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn,
                                     None) == 1
        torch.cuda.synchronize()
        rc = L.sba_gpu_mul_dev_noalloc(got.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn, None)
        assert rc == 1, rc
        torch.cuda.synchronize()
        assert torch.equal(got, want), (an, bn)
        # the same call again (state left by the previous call of the same size)
        got.fill_(-1)
        assert L.sba_gpu_mul_dev_noalloc(got.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn,
                                         None) == 1
        torch.cuda.synchronize()
        assert torch.equal(got, want), ("repeat", an, bn)


def test_noalloc_rejects_beyond_its_scratch():
    torch = pytest.importorskip("torch")
    L = sba.lib()
    an = bn = (1 << 19) + 1                      # an + bn - 1 = 2^20 + 1 coefficients
    a = _dev(torch, xorshift_fill(an, [5]))
    r = torch.full((an + bn,), 7, dtype=torch.int64, device="cuda")
    rc = L.sba_gpu_mul_dev_noalloc(r.data_ptr(), a.data_ptr(), an, a.data_ptr(), bn, None)
    assert rc == sba.SBA_ERR_TOO_LARGE
    torch.cuda.synchronize()
    assert bool((r == 7).all())
    assert L.sba_gpu_mul_dev_noalloc(r.data_ptr(), a.data_ptr(), 0, a.data_ptr(), bn,
                                     None) == sba.SBA_BAND
