"""The oracle is pinned before it is trusted (CPU only).

Checks: the reference's own outputs (tests/golden/golden.json, written by
tests/golden/make_golden.py from the reference compiled from source), u128
schoolbook, Python big integers, the all-ones closed form, modular
fingerprints, and planted mistakes that the checks must catch.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from tests import _oracle as orc
from simd_bigint_amd.workloads import pattern_fill, to_int, from_int, all_ones_product

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def hexarr(h):
    return np.array([int(x, 16) for x in h], dtype=np.uint64)


def is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def fp_primes():
    out, q = [], (1 << 61) - 1
    while len(out) < 3:
        if is_prime(q):
            out.append(q)
        q -= 2
    return out


def test_small_against_reference_vectors():
    for e in GOLD["small_mul"]:
        st = [e["seed"]]
        a = pattern_fill(e["an"], e["pat"], st)
        if e.get("sqr"):
            r = orc.nt_sqr(a)
        else:
            b = pattern_fill(e["bn"], e["pat"], st)
            r = orc.nt_mul(a, b)
        assert (r == hexarr(e["r"])).all(), (e["an"], e["bn"], e["pat"])


@pytest.mark.parametrize("e", GOLD["large_mul"], ids=[e["name"] for e in GOLD["large_mul"]])
def test_large_against_reference_sha(e):
    st = [e["seed"]]
    a = pattern_fill(e["an"], e["pat"], st)
    if e["sqr"]:
        r = orc.nt_sqr(a)
    else:
        b = pattern_fill(e["bn"], e["pat"], st)
        r = orc.nt_mul(a, b)
    for k, v in e["samples"].items():
        assert "%016x" % int(r[int(k)]) == v, (e["name"], k)
    assert sha(r) == e["sha256"], e["name"]


def test_schoolbook_and_nt_vs_python_ints():
    rng = np.random.default_rng(11)
    for _ in range(60):
        an, bn = int(rng.integers(1, 70)), int(rng.integers(1, 70))
        a = rng.integers(0, 2**64, an, dtype=np.uint64)
        b = rng.integers(0, 2**64, bn, dtype=np.uint64)
        want = to_int(a) * to_int(b)
        assert to_int(orc.school(a, b)) == want
        assert to_int(orc.nt_mul(a, b)) == want


def test_all_ones_closed_form():
    # the limb pattern against Python integers, then the oracle against it
    for an, bn in ((1, 1), (2, 1), (1, 3), (7, 7), (40, 13), (13, 40), (1000, 999)):
        assert to_int(all_ones_product(an, bn)) == ((1 << (64 * an)) - 1) * ((1 << (64 * bn)) - 1)
    for an, bn in ((1, 1), (17, 5), (4096, 4096), (3000, 70000)):
        a = np.full(an, 2**64 - 1, np.uint64)
        b = np.full(bn, 2**64 - 1, np.uint64)
        assert (orc.nt_mul(a, b) == all_ones_product(an, bn)).all()


def test_fingerprints_and_six_prime_band():
    qs = fp_primes()
    rng = np.random.default_rng(3)
    n = 3_000_000          # above p30's CRT cap: exercises the six-prime set
    assert orc.lib().orc_nt_nprimes(2344425, 2344425) == 5
    assert orc.lib().orc_nt_nprimes(2344426, 2344426) == 6
    a = rng.integers(0, 2**64, n, dtype=np.uint64)
    b = rng.integers(0, 2**64, 200_000, dtype=np.uint64)
    b2 = rng.integers(0, 2**64, n, dtype=np.uint64)
    for x, y in ((a, b), (a, b2)):
        r = orc.nt_mul(x, y)
        for q in qs:
            assert orc.fingerprint(r, q) == orc.fingerprint(x, q) * orc.fingerprint(y, q) % q
        # fingerprint itself is checked against Python on a slice
        assert orc.fingerprint(r[:1000], qs[0]) == to_int(r[:1000]) % qs[0]


def test_residues_and_divrem_vs_python():
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2**64, 1000, dtype=np.uint64)
    out = np.zeros(1000, np.uint32)
    orc.lib().orc_residues(out, a, 1000, 998244353)
    assert (out.astype(np.uint64) == a % np.uint64(998244353)).all()
    for nn, dn in ((1, 1), (5, 2), (40, 40), (100, 33), (300, 1), (64, 70)):
        n = rng.integers(0, 2**64, nn, dtype=np.uint64)
        d = rng.integers(1, 2**64, dn, dtype=np.uint64)
        q, r = orc.divrem(n, d)
        Q, R = divmod(to_int(n), to_int(d))
        assert to_int(q) == Q and to_int(r) == R


def test_divrem_against_reference_vectors():
    from simd_bigint_amd.workloads import xorshift_fill
    for e in GOLD["divrem"]:
        if "q" not in e:
            continue
        st = [e["seed"]]
        n = xorshift_fill(e["nn"], st)
        d = xorshift_fill(e["dn"], st)
        d[-1] |= np.uint64(1 << 63)
        q, r = orc.divrem(n, d)
        assert (q == hexarr(e["q"])).all() and (r == hexarr(e["r"])).all()


def test_planted_mistakes_are_caught():
    """A wrong product (one flipped bit, a dropped carry, swapped limbs)
    must fail the checks the GPU tests rely on."""
    qs = fp_primes()
    st = [42]
    a = pattern_fill(3000, 0, st)
    b = pattern_fill(2000, 0, st)
    good = orc.nt_mul(a, b)
    bad1 = good.copy()
    bad1[1234] ^= np.uint64(1 << 17)
    bad2 = good.copy()
    bad2[2000] = bad2[2000] + np.uint64(1)     # carry landed one limb high
    bad3 = good.copy()
    bad3[[10, 11]] = bad3[[11, 10]]
    for bad in (bad1, bad2, bad3):
        assert sha(bad) != sha(good)
        assert any(orc.fingerprint(bad, q) != orc.fingerprint(a, q) * orc.fingerprint(b, q) % q
                   for q in qs)


def test_oracle_generator_matches_the_scalar_recurrence():
    """orc_fill_xorshift (the reference's xorshift64 13/7/17,
    bench/nt_sweep.c:15-16) against the recurrence written out in Python,
    and against the vectorised jump-matrix fill of the workloads module at
    a length that takes its fast path; the state carries across calls."""
    import ctypes
    from simd_bigint_amd.workloads import xorshift_fill
    M = (1 << 64) - 1

    def xs(x):
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        return x

    st = ctypes.c_uint64(42)
    a = np.zeros(1000, np.uint64)
    orc.lib().orc_fill_xorshift(a, len(a), ctypes.byref(st))
    x, want = 42, []
    for _ in range(1000):
        x = xs(x)
        want.append(x)
    assert [int(v) for v in a] == want and st.value == want[-1]
    b = np.zeros(5, np.uint64)
    orc.lib().orc_fill_xorshift(b, 5, ctypes.byref(st))
    for v in b:
        x = xs(x)
        assert int(v) == x
    n = (1 << 17) + 12345                            # the jump-matrix path
    st2 = ctypes.c_uint64(7)
    c = np.zeros(n, np.uint64)
    orc.lib().orc_fill_xorshift(c, n, ctypes.byref(st2))
    pst = [7]
    assert (xorshift_fill(n, pst) == c).all() and pst[0] == st2.value
