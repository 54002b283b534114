"""The host entries' error contract, on the CPU (no GPU needed): a HIP
failure inside a host entry never aborts. nt_* / p30_* / p50_* return 0
(the reference's nt_* return only 1 or 0, include/ntmul.h:19,52-57);
simd_mpn_mul and the division finish the product on the CPU tier and count
it. Failures are injected before the device is touched
(sba_test_inject_failure); a process without any HIP device still aborts
in an in-band GPU entry (subprocess test)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import to_int, xorshift_fill

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


@pytest.fixture(autouse=True)
def disarm():
    yield
    sba.inject_failure(sba.FAIL_ENTRY, 0)
    sba.inject_failure(sba.FAIL_ALLOC, 0)


def test_nt_entries_return_zero_on_a_gpu_failure():
    st = [7]
    a, b = xorshift_fill(3000, st), xorshift_fill(2000, st)
    for name, args in (("nt_mul_r", (a, b)), ("nt_sqr_r", (a, None)), ("p30_mul_r", (a, b)),
                       ("p30_sqr_r", (b, None)), ("p50_mul_r", (a, b)),
                       ("p50_sqr_r", (a, None))):
        sba.inject_failure(sba.FAIL_ENTRY, 1)
        ok, _ = sba.mul_r(name, *args)
        assert ok == 0, name
    for f, args in ((sba.nt_mul, (a, b)), (sba.nt_sqr, (a,)), (sba.p30_mul, (a, b)),
                    (sba.p50_mul, (a, b)), (sba.p30_sqr, (b,)), (sba.p50_sqr, (b,))):
        sba.inject_failure(sba.FAIL_ENTRY, 1)
        ok, _ = f(*args)
        assert ok == 0, f.__name__


def test_simd_mpn_mul_finishes_on_the_cpu_tier_and_counts_it():
    st = [11]
    a, b = xorshift_fill(1500, st), xorshift_fill(1100, st)
    assert len(a) + len(b) >= sba.get_gpu_threshold()
    before = sba.gpu_fallbacks()
    sba.inject_failure(sba.FAIL_ENTRY, 1)
    r = sba.simd_mpn_mul(a, b)
    assert to_int(r) == to_int(a) * to_int(b)
    assert sba.gpu_fallbacks() == before + 1
    sba.inject_failure(sba.FAIL_ENTRY, 1)
    r = sba.simd_mpn_mul(a, a)                     # the square route
    assert to_int(r) == to_int(a) ** 2
    assert sba.gpu_fallbacks() == before + 2


def test_division_finishes_on_the_cpu_tier():
    """The host Newton branch's multiplies fail on the GPU one after another
    and are finished by the CPU tier; Q D + R = N, 0 <= R < D."""
    st = [13]
    n = xorshift_fill(6000, st)
    d = xorshift_fill(3000, st)
    d[-1] |= np.uint64(1) << np.uint64(63)
    before = sba.gpu_fallbacks()
    sba.inject_failure(sba.FAIL_ENTRY, 1000)
    q, r = sba.divrem_u64(n, d)
    N, D, Q, R = to_int(n), to_int(d), to_int(q), to_int(r)
    assert Q * D + R == N and 0 <= R < D
    assert sba.gpu_fallbacks() > before


def test_p30_transform_cap_and_crt_cap_return_zero():
    """p30's band (include/p30.h:1851-1858): tv = ceil((an+bn-1)/16) must stay
    <= 2^18 and min(an,bn) <= 2,344,425; outside it p30_* return 0 before
    touching any device."""
    big = np.ones(16 * (1 << 18) + 1, np.uint64)
    one = np.ones(1, np.uint64)
    ok, _ = sba.p30_mul(big, one)                  # zlen = 2^22 + 1 -> tv = 2^18 + 1
    assert ok == 0
    sq = np.ones(2 * (1 << 21) + 1, np.uint64)     # zlen = 2^23 + 1
    ok, _ = sba.p30_sqr(sq)
    assert ok == 0
    for name in ("p30_mul_r", "p30_sqr_r"):
        ok, _ = sba.mul_r(name, big, one if name == "p30_mul_r" else None)
        assert ok == 0


@pytest.mark.skipif(sba.lib().sba_gpu_device_count() > 0, reason="a HIP device is present")
def test_no_device_aborts_loudly():
    code = ("import numpy as np, simd_bigint_amd as sba\n"
            "a = np.arange(1, 3001, dtype=np.uint64)\n"
            "print(sba.nt_mul(a, a)[0])\n")
    p = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode != 0
    assert "no HIP device" in p.stderr
