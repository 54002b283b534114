"""Host C logic on the CPU: the CPU tiers of simd_mpn_mul and the Newton
division, with every multiply kept below the GPU threshold (the threshold is
raised for these tests only; the GPU tests run the default dispatch)."""
import json
import os

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import to_int, xorshift_fill

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def hexarr(h):
    return np.array([int(x, 16) for x in h], dtype=np.uint64)


@pytest.fixture(autouse=True)
def cpu_only():
    old = sba.get_gpu_threshold()
    sba.set_gpu_threshold(1 << 62)
    yield
    sba.set_gpu_threshold(old)


def test_cpu_tiers_vs_python():
    rng = np.random.default_rng(1)
    shapes = [(1, 1), (6, 6), (31, 31), (32, 32), (33, 1), (64, 63), (100, 100), (257, 255),
              (1000, 31), (1000, 32), (3000, 1500), (4097, 4096)]
    for an, bn in shapes:
        for pat in range(3):
            a = rng.integers(0, 2**64, an, dtype=np.uint64)
            b = rng.integers(0, 2**64, bn, dtype=np.uint64)
            if pat == 1:
                a[:] = 2**64 - 1
                b[:] = 2**64 - 1
            if pat == 2:
                a[an // 2:] = 0
            want = to_int(a) * to_int(b)
            assert to_int(sba.cpu_mul(a, b)) == want
            assert to_int(sba.cpu_mul(b, a)) == want
            assert to_int(sba.simd_mpn_mul(a, b)) == want


def test_simd_mpn_mul_reference_vectors_and_no_overwrite():
    L = sba.lib()
    for e in GOLD["small_mpn"]:
        a, b = hexarr(e["a"]), hexarr(e["b"])
        n = e["an"] + e["bn"]
        buf = np.full(n + 2, 0xABABABABABABABAB, dtype=np.uint64)
        L.simd_mpn_mul(sba._arr(buf), sba._arr(a), e["an"], sba._arr(b), e["bn"])
        assert (buf[:n] == hexarr(e["r"])).all(), (e["an"], e["bn"])
        assert buf[n] == 0xABABABABABABABAB and buf[n + 1] == 0xABABABABABABABAB
    z = np.zeros(20, np.uint64)
    a = np.arange(1, 41, dtype=np.uint64)
    assert not sba.simd_mpn_mul(a, z).any()


def test_divrem_reference_vectors():
    for e in GOLD["divrem"]:
        if "q" not in e:
            continue
        st = [e["seed"]]
        n = xorshift_fill(e["nn"], st)
        d = xorshift_fill(e["dn"], st)
        d[-1] |= np.uint64(1 << 63)
        q, r = sba.divrem_u64(n, d, dc=e["fn"].endswith("_dc"))
        assert len(q) == e["qn"] and len(r) == e["rn"]
        assert (q == hexarr(e["q"])).all() and (r == hexarr(e["r"])).all()


def test_divrem_newton_vs_python():
    rng = np.random.default_rng(9)
    shapes = [(2, 1), (5, 3), (70, 33), (100, 34), (400, 200), (1000, 500), (2000, 999),
              (3000, 2000), (5000, 100), (600, 40), (65, 64)]
    for nn, dn in shapes:
        for pat in range(3):
            n = rng.integers(0, 2**64, nn, dtype=np.uint64)
            d = rng.integers(0, 2**64, dn, dtype=np.uint64)
            if pat == 1:
                n[:] = 2**64 - 1
                d[:] = 2**64 - 1
            if pat == 2:
                d[:] = 0
                d[-1] = 1 << 63
                d[0] = 1
            if d[-1] == 0:
                d[-1] = 1
            q, r = sba.divrem_u64(n, d)
            Q, R = divmod(to_int(n), to_int(d))
            assert to_int(q) == Q and to_int(r) == R, (nn, dn, pat)
            assert len(q) >= 1 and len(r) >= 1


def test_divrem_short_dividend_mirrors_reference_lengths():
    # bitlen(N) < bitlen(D): only qp[0] written, remainder length dn64
    # (include/divrem.h:222-227)
    n = np.array([5, 0, 0], np.uint64)
    d = np.array([1, 2, 3, 4, 5, 6, 7, 8], np.uint64)
    L = sba.lib()
    import ctypes
    q = np.full(1, 99, np.uint64)
    r = np.zeros(8, np.uint64)
    qn, rn = ctypes.c_uint64(), ctypes.c_uint64()
    L.divrem_u64(sba._arr(q), sba._arr(r), sba._arr(n), 3, sba._arr(d), 8,
                 ctypes.byref(qn), ctypes.byref(rn))
    assert qn.value == 1 and rn.value == 8 and q[0] == 0 and r[0] == 5 and not r[1:].any()


def test_nt_band_matches_reference_trunk_counting():
    """nt_mul's band is the reference's p50 band: at most 2^25 trunks in
    the product. A short operand takes 88-bit trunks (2^25 * 88/64 =
    46137344 limbs); a balanced pair past the 88-bit CRT capacity takes
    84-bit trunks (2^25 * 84/64 / 2 = 22020096 limbs each). Outside, nt_*
    return 0 before reading the operands (no device needed)."""
    import simd_bigint_amd as sba
    L = sba.lib()
    assert L.sba_ref_nt_band(46137344, 1) == 1
    assert L.sba_ref_nt_band(46137345, 1) == 0
    assert L.sba_ref_nt_band(22020096, 22020096) == 1
    assert L.sba_ref_nt_band(22020097, 22020097) == 0
    assert L.sba_ref_nt_band(0, 5) == 0
    a = np.ones(3, np.uint64)
    r = np.zeros(4, np.uint64)
    assert L.nt_mul(sba._arr(r), sba._arr(a), 46137345, sba._arr(a), 1) == 0
    assert L.nt_sqr(sba._arr(r), sba._arr(a), 22020097) == 0
    # the device band contains it
    assert L.sba_gpu_in_band(46137344, 1) == 1
    assert L.sba_gpu_in_band(22020096, 22020096) == 1


def test_staging_copy_pool_in_a_forked_child():
    """The host entries' staged transfers copy through a pool of threads
    started on first use; a forked child (which has none of those threads)
    copies on its own thread instead of waiting on them."""
    import ctypes
    L = sba.lib()
    f = L.sba_par_copy
    f.restype = None
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    n = (9 << 20) + 13
    src = np.frombuffer(np.random.default_rng(3).bytes(n), np.uint8)
    dst = np.zeros(n, np.uint8)
    f(dst.ctypes.data, src.ctypes.data, n)           # starts the pool
    assert (dst == src).all()
    pid = os.fork()
    if pid == 0:
        d2 = np.zeros(n, np.uint8)
        f(d2.ctypes.data, src.ctypes.data, n)
        os._exit(0 if (d2 == src).all() else 1)
    import signal
    import time
    t0 = time.time()
    while True:
        done, status = os.waitpid(pid, os.WNOHANG)
        if done:
            break
        if time.time() - t0 > 60:
            os.kill(pid, signal.SIGKILL)
            os.waitpid(pid, 0)
            pytest.fail("the forked child's copy did not finish within 60 s")
        time.sleep(0.01)
    assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0
