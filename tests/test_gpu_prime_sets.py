"""GPU parity of the prime sets.

Power-of-two transforms up to 2^23 run on set H and radix-3 transforms up
to 3 x 2^22 on set R, each with three small primes (below 2^30) that take
Harvey's butterflies in the merged last pass; every other transform runs on
the wide set. Set H holds its CRT bound only if the largest coefficient of
a 2^22 x 2^22 product, 2^22 (2^64 - 1)^2 < 2^150, stays below its primes'
product (2^151.3), and set R only if that of a 3 * 2^21 x 3 * 2^21 product
(< 2^150.6) stays below 2^151.2: all-ones operands reach those
coefficients exactly, and their products have a closed form. With the
sets switched off (sba_set_prime_set_h(0)) the same shapes run on the wide
set; both must give the reference's product."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import all_ones_product, pattern_fill, xorshift_fill
from tests import _oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


@pytest.fixture
def wide_set():
    L = sba.lib()
    assert L.sba_get_prime_set_h() == 1          # the default
    L.sba_set_prime_set_h(0)
    try:
        yield
    finally:
        L.sba_set_prime_set_h(1)


def test_set_h_crt_edge_all_ones():
    """2^22 x 2^22 all-ones: the middle coefficient is 2^22 (2^64 - 1)^2, the
    largest a set-H product can have; also the square and 2^22 x (2^22 - 1)."""
    n = 1 << 22
    a = np.full(n, 2**64 - 1, np.uint64)
    say("case all-ones", n, "x", n)
    ok, r = sba.nt_mul(a, a)
    assert ok == 1 and (r == all_ones_product(n, n)).all()
    say("case all-ones square", n)
    ok, r = sba.nt_sqr(a)
    assert ok == 1 and (r == all_ones_product(n, n)).all()
    say("case all-ones", n, "x", n - 1)
    ok, r = sba.nt_mul(a, a[: n - 1])
    assert ok == 1 and (r == all_ones_product(n, n - 1)).all()


def test_set_r_crt_edge_all_ones():
    """3 * 2^21 x 3 * 2^21 all-ones, the top of the radix-3 class of row
    length 2^22: the largest coefficient a set-R product can have; also the
    square and 3 * 2^21 x (3 * 2^21 - 1)"""
    n = 3 << 21
    L = sba.lib()
    assert L.sba_gpu_radix3(n, n, 0) == 22 and L.sba_gpu_radix3(n, n - 1, 0) == 22
    a = np.full(n, 2**64 - 1, np.uint64)
    say("case all-ones", n, "x", n)
    ok, r = sba.nt_mul(a, a)
    assert ok == 1 and (r == all_ones_product(n, n)).all()
    say("case all-ones square", n)
    ok, r = sba.nt_sqr(a)
    assert ok == 1 and (r == all_ones_product(n, n)).all()
    say("case all-ones", n, "x", n - 1)
    ok, r = sba.nt_mul(a, a[: n - 1])
    assert ok == 1 and (r == all_ones_product(n, n - 1)).all()


def test_set_h_runs_the_metric_shape():
    """the metric's golden SHA-256 (the reference's own nt_mul) on set H"""
    e = [g for g in GOLD["large_mul"] if g["name"] == "metric_4Mi"][0]
    st = [e["seed"]]
    a = pattern_fill(e["an"], e["pat"], st)
    b = pattern_fill(e["bn"], e["pat"], st)
    ok, r = sba.nt_mul(a, b)
    assert ok == 1 and sha(r) == e["sha256"]


@pytest.mark.parametrize("name", ["metric_4Mi", "sqr_2^22", "mul_2^23x2^18", "mul_16384",
                                  "mul_3x2^20"])
def test_wide_set_golden(wide_set, name):
    """the same golden products with set H switched off"""
    e = [g for g in GOLD["large_mul"] if g["name"] == name][0]
    say("case", name, "wide set")
    st = [e["seed"]]
    a = pattern_fill(e["an"], e["pat"], st)
    if e["sqr"]:
        ok, r = sba.nt_sqr(a)
    else:
        b = pattern_fill(e["bn"], e["pat"], st)
        ok, r = sba.nt_mul(a, b)
    assert ok == 1 and sha(r) == e["sha256"], name


def test_sets_agree_across_lengths():
    """every power-of-two length 2^10 .. 2^21 (each plan shape of the merged
    last pass), multiplies and squares: set H == set W == the oracle"""
    L = sba.lib()
    st = [77]
    try:
        for lg in range(10, 22):
            an = (1 << (lg - 1)) - 3
            bn = (1 << (lg - 1)) - 1
            a = xorshift_fill(an, st)
            b = xorshift_fill(bn, st)
            want = orc.nt_mul(a, b)
            want_sq = orc.nt_sqr(b)
            for on in (1, 0):
                L.sba_set_prime_set_h(on)
                ok, r = sba.nt_mul(a, b)
                assert ok == 1 and (r == want).all(), (lg, on)
                ok, r = sba.nt_sqr(b)
                assert ok == 1 and (r == want_sq).all(), (lg, on, "sqr")
            say("lg", lg, "ok")
    finally:
        L.sba_set_prime_set_h(1)



def test_radix3_sets_agree_across_lengths():
    """radix-3 single products at row lengths 2^16 .. 2^19, multiplies and
    squares: set R == set W == the oracle"""
    L = sba.lib()
    st = [91]
    try:
        for lg in range(16, 20):
            an = (3 << (lg - 1)) - 5
            bn = (3 << (lg - 1)) - 2
            assert L.sba_gpu_radix3(an, bn, 0) == lg and L.sba_gpu_radix3(bn, bn, 1) == lg
            a = xorshift_fill(an, st)
            b = xorshift_fill(bn, st)
            want = orc.nt_mul(a, b)
            want_sq = orc.nt_sqr(b)
            for on in (1, 0):
                L.sba_set_prime_set_h(on)
                ok, r = sba.nt_mul(a, b)
                assert ok == 1 and (r == want).all(), (lg, on)
                ok, r = sba.nt_sqr(b)
                assert ok == 1 and (r == want_sq).all(), (lg, on, "sqr")
            say("radix-3 row 2^%d ok" % lg)
    finally:
        L.sba_set_prime_set_h(1)
