"""GPU parity: every product, square and quotient computed on the MI355X
through the C ABI is bit-identical to the reference's (golden SHA-256 and
vectors from the reference compiled from source) and to the oracle."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import (all_ones_product, pattern_fill, to_int,
                                       xorshift_fill)
from tests import _oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def hexarr(h):
    return np.array([int(x, 16) for x in h], dtype=np.uint64)


def test_device_present():
    assert sba.lib().sba_gpu_device_count() > 0


def test_small_reference_vectors():
    for e in GOLD["small_mul"]:
        st = [e["seed"]]
        a = pattern_fill(e["an"], e["pat"], st)
        if e.get("sqr"):
            ok, r = sba.nt_sqr(a)
        else:
            b = pattern_fill(e["bn"], e["pat"], st)
            ok, r = sba.nt_mul(a, b)
        assert ok == 1 and (r == hexarr(e["r"])).all(), (e["an"], e["bn"], e["pat"])


@pytest.mark.parametrize("e", GOLD["large_mul"], ids=[e["name"] for e in GOLD["large_mul"]])
def test_reference_sha(e):
    say("case", e["name"], e["an"], e["bn"])
    st = [e["seed"]]
    a = pattern_fill(e["an"], e["pat"], st)
    if e["sqr"]:
        ok, r = sba.nt_sqr(a)
    else:
        b = pattern_fill(e["bn"], e["pat"], st)
        ok, r = sba.nt_mul(a, b)
    assert ok == 1
    for k, v in e["samples"].items():
        assert "%016x" % int(r[int(k)]) == v, (e["name"], k)
    assert sha(r) == e["sha256"], e["name"]


def test_transform_size_boundaries_vs_oracle():
    """totals straddling every transform length 2^10 .. 2^21 (the pass plan
    changes at each), rectangular shapes, both patterns (tests/p30_test.c:70-83)."""
    rng = np.random.default_rng(21)
    for lg in range(10, 22):
        for d in (-2, -1, 0, 1, 2):
            tot = (1 << lg) + 1 + d          # zlen = tot - 1 around 2^lg
            for split in (2, 5):
                an = max(tot // split, 1)
                bn = tot - an
                a = rng.integers(0, 2**64, an, dtype=np.uint64)
                b = rng.integers(0, 2**64, bn, dtype=np.uint64)
                ok, r = sba.nt_mul(a, b)
                assert ok == 1 and (r == orc.nt_mul(a, b)).all(), (lg, d, an, bn)
        say("lg", lg, "ok")


def test_exhaustive_small_shapes():
    rng = np.random.default_rng(2)
    for an in range(1, 41):
        for bn in range(1, 41, 3):
            a = rng.integers(0, 2**64, an, dtype=np.uint64)
            b = rng.integers(0, 2**64, bn, dtype=np.uint64)
            ok, r = sba.nt_mul(a, b)
            assert ok == 1 and to_int(r) == to_int(a) * to_int(b), (an, bn)


def test_all_ones_closed_form_balanced_3x2_20():
    n = 3 << 20
    a = np.full(n, 2**64 - 1, np.uint64)
    ok, r = sba.nt_mul(a, a)
    assert ok == 1
    want = np.zeros(2 * n, np.uint64)      # limb 0 = 1, limb n = 2^64-2, top n-1 all ones
    want[0] = 1
    want[n] = np.uint64(2**64 - 2)
    want[n + 1:] = np.uint64(2**64 - 1)
    assert (r == want).all()
    ok, r = sba.nt_sqr(a[: 1 << 20])
    assert ok == 1 and (r == all_ones_product(1 << 20, 1 << 20)).all()


def test_sparse_and_zero_operands():
    a = np.zeros(70000, np.uint64)
    a[::10] = np.uint64(2**64 - 1)
    b = np.zeros(50000, np.uint64)
    b[7] = 3
    ok, r = sba.nt_mul(a, b)
    assert ok == 1 and (r == orc.nt_mul(a, b)).all()
    ok, r = sba.nt_mul(np.zeros(30000, np.uint64), a)
    assert ok == 1 and not r.any()


def test_band_edges():
    cap = sba.min_operand_cap()
    assert cap > (1 << 25)
    L = sba.lib()
    a = np.ones(3, np.uint64)
    r = np.zeros(4, np.uint64)
    # beyond the reference's band (nt_mul's 0): rejected before any buffer
    # is read; 2^25 trunks of 84 bits (balanced) / 88 bits (short operand)
    assert L.nt_mul(sba._arr(r), sba._arr(a), 22020097, sba._arr(a), 22020097) == 0
    assert L.nt_mul(sba._arr(r), sba._arr(a), 46137345, sba._arr(a), 1) == 0
    assert L.nt_sqr(sba._arr(r), sba._arr(a), 22020097) == 0
    # exactly 2^25 coefficients is in band (largest transform)
    st = [3]
    x = xorshift_fill(1 << 24, st)
    y = xorshift_fill((1 << 24) + 1, st)
    ok, r = sba.nt_mul(x, y)
    assert ok == 1
    qs = [(1 << 61) - 1]
    for q in qs:
        assert orc.fingerprint(r, q) == orc.fingerprint(x, q) * orc.fingerprint(y, q) % q


def test_public_entry_default_dispatch():
    L = sba.lib()
    for e in GOLD["small_mpn"]:
        a, b = hexarr(e["a"]), hexarr(e["b"])
        assert (sba.simd_mpn_mul(a, b) == hexarr(e["r"])).all()
    st = [77]
    for an, bn in ((20000, 20000), (65536, 3000), (100000, 64)):
        a = xorshift_fill(an, st)
        b = xorshift_fill(bn, st)
        n = an + bn
        buf = np.full(n + 2, 0xABABABABABABABAB, dtype=np.uint64)
        L.simd_mpn_mul(sba._arr(buf), sba._arr(a), an, sba._arr(b), bn)
        assert (buf[:n] == orc.nt_mul(a, b)).all()
        assert buf[n] == 0xABABABABABABABAB and buf[n + 1] == 0xABABABABABABABAB


def test_host_entries_from_several_threads():
    """The host entries are reentrant across threads (the reference's are,
    include/scratch.h:339-356): four threads call nt_mul and simd_mpn_mul
    at once -- GPU-sized and CPU-tier products mixed -- and every product
    matches the oracle."""
    import threading
    L = sba.lib()
    shapes = ((20000, 20000), (300, 500), (1 << 16, 3000), (4096, 4096))
    out, errs = {}, []

    def work(k):
        try:
            st = [1000 + k]
            for rep in range(3):
                for si, (an, bn) in enumerate(shapes):
                    a = xorshift_fill(an, st)
                    b = xorshift_fill(bn, st)
                    r = np.zeros(an + bn, dtype=np.uint64)
                    if (si + rep + k) % 2:
                        rc = L.nt_mul(sba._arr(r), sba._arr(a), an, sba._arr(b), bn)
                        assert rc == 1
                    else:
                        L.simd_mpn_mul(sba._arr(r), sba._arr(a), an, sba._arr(b), bn)
                    out[(k, rep, si)] = (a, b, r)
        except Exception as e:          # reported by the main thread
            errs.append(repr(e))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    assert len(out) == 4 * 3 * len(shapes)
    for key, (a, b, r) in out.items():
        assert (r == orc.nt_mul(a, b)).all(), key


def test_divrem_reference():
    for e in GOLD["divrem"]:
        say("divrem", e["nn"], e["dn"], e["fn"])
        st = [e["seed"]]
        n = xorshift_fill(e["nn"], st)
        d = xorshift_fill(e["dn"], st)
        d[-1] |= np.uint64(1 << 63)
        q, r = sba.divrem_u64(n, d, dc=e["fn"].endswith("_dc"))
        assert len(q) == e["qn"] and len(r) == e["rn"]
        if "q" in e:
            assert (q == hexarr(e["q"])).all() and (r == hexarr(e["r"])).all()
        else:
            assert sha(q) == e["q_sha256"] and sha(r) == e["r_sha256"]


def test_divrem_invariant_large():
    st = [91]
    n = xorshift_fill(300000, st)
    d = xorshift_fill(120000, st)
    q, r = sba.divrem_u64(n, d)
    # N == Q*D + R and R < D, with the product on the GPU and checked by fingerprints
    N, D, Q, R = to_int(n), to_int(d), to_int(q), to_int(r)
    assert R < D and Q * D + R == N


def _div_case(qn, dn, st, top=None, pattern=None):
    """dividend of qn + dn - 1 limbs and divisor of dn limbs: the quotient
    has qn limbs before normalisation (m = qn + 1 Newton limbs)."""
    n = xorshift_fill(qn + dn - 1, st)
    n[-1] |= np.uint64(1 << 63)
    if pattern == "ones":          # D = B^dn - 1: the reciprocal is exact
        d = np.full(dn, ~np.uint64(0), dtype=np.uint64)
    elif pattern == "ones_m2":     # top qn+1 limbs of D plus one = B^(qn+1) - 2
        d = np.full(dn, ~np.uint64(0), dtype=np.uint64)
        d[dn - (qn + 1)] = np.uint64(0xFFFFFFFFFFFFFFFD)
    else:
        d = xorshift_fill(dn, st)
        d[-1] = np.uint64(top) if top is not None else d[-1] | np.uint64(1 << 63)
    return n, d


@pytest.mark.parametrize("qn,dn,top,pattern,dc", [
    (8191, 33, None, None, False),        # first size on device buffers
    (8190, 33, None, None, False),        # last size on the host Newton
    (40000, 40, None, None, False),       # long quotient, short divisor (chunked)
    (10001, 20000, None, "ones", False),  # Dm + 1 = B^m: no reciprocal
    (12000, 20000, None, "ones_m2", True),  # reciprocal products overflow: X corrections
    (20000, 9000, 1, None, False),        # divisor top limb 1: shift by 63
    (65536, 65536, None, None, True),
])
def test_divrem_newton_on_device(qn, dn, top, pattern, dc):
    say("divrem newton", qn, dn, top, pattern, dc)
    st = [qn * 7 + dn]
    n, d = _div_case(qn, dn, st, top, pattern)
    q, r = sba.divrem_u64(n, d, dc=dc)
    N, D, Q, R = to_int(n), to_int(d), to_int(q), to_int(r)
    assert R < D and Q * D + R == N


def test_device_resident_matches_host():
    torch = pytest.importorskip("torch")
    st = [5]
    for an, bn in ((1 << 20, 1 << 20), (1 << 18, 5000)):
        a = xorshift_fill(an, st)
        b = xorshift_fill(bn, st)
        ta = torch.from_numpy(a.view(np.int64)).cuda()
        tb = torch.from_numpy(b.view(np.int64)).cuda()
        tr = torch.empty(an + bn, dtype=torch.int64, device="cuda")
        assert sba.mul_dev(tr, ta, tb) == 1
        torch.cuda.synchronize()
        ok, r = sba.nt_mul(a, b)
        assert ok == 1 and (tr.cpu().numpy().view(np.uint64) == r).all()
        ts = torch.empty(2 * an, dtype=torch.int64, device="cuda")
        assert sba.sqr_dev(ts, ta) == 1
        torch.cuda.synchronize()
        ok, r = sba.nt_sqr(a)
        assert (ts.cpu().numpy().view(np.uint64) == r).all()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_mixed_sizes_vs_oracle(use_graph):
    torch = pytest.importorskip("torch")
    from simd_bigint_amd.workloads import batch_workload
    pairs = batch_workload(150, 123, hi=20000)
    jobs, want = [], []
    for i, (a, b) in enumerate(pairs):
        ta = torch.from_numpy(a.view(np.int64)).cuda()
        if i % 7 == 3:                       # a few squares
            tr = torch.zeros(2 * len(a), dtype=torch.int64, device="cuda")
            jobs.append((tr, ta, None))
            want.append(orc.nt_sqr(a))
        else:
            tb = torch.from_numpy(b.view(np.int64)).cuda()
            tr = torch.zeros(len(a) + len(b), dtype=torch.int64, device="cuda")
            jobs.append((tr, ta, tb))
            want.append(orc.nt_mul(a, b))
    bt = sba.Batch(jobs, use_graph=use_graph)
    assert bt.classes >= 4
    for rep in range(3):                     # replays (graph) agree every time
        for (tr, _, _) in jobs:
            tr.zero_()
        bt.run()
        torch.cuda.synchronize()
        for k, (tr, _, _) in enumerate(jobs):
            assert (tr.cpu().numpy().view(np.uint64) == want[k]).all(), (rep, k)
    bt.close()


def test_serving_results_same_alone_and_in_mix():
    """Continuous batching: each result of the served mix (Poisson arrivals,
    80% multiplies / 20% squares, admission by an HBM budget) is bit-identical
    to the same request run alone, and to the Python integer product."""
    torch = pytest.importorskip("torch")
    from simd_bigint_amd import serving
    rng = np.random.default_rng(8)
    pool_np = rng.integers(0, 2**64, 1 << 18, dtype=np.uint64)
    pool = torch.from_numpy(pool_np.view(np.int64)).cuda()
    reqs = serving.poisson_requests(120, 4000.0, pool, seed=9, lo=256, hi=1 << 15)
    budget = 48 << 20
    srv = serving.Server(budget, nstreams=3, max_batch=32)
    stats = srv.run(reqs)
    s = stats.summary()
    say("serve", s)
    assert s["requests"] == len(reqs) and stats.peak_ws_bytes <= budget
    assert stats.batches < len(reqs)          # requests were batched together
    for q in reqs:
        alone = torch.empty(q.limbs, dtype=torch.int64, device="cuda")
        if q.b is None:
            assert sba.sqr_dev(alone, q.a) == 1
        else:
            assert sba.mul_dev(alone, q.a, q.b) == 1
        torch.cuda.synchronize()
        assert torch.equal(alone, q.r), q.rid
    for q in reqs[::10]:
        a = to_int(q.a.cpu().numpy().view(np.uint64))
        b = a if q.b is None else to_int(q.b.cpu().numpy().view(np.uint64))
        assert to_int(q.r.cpu().numpy().view(np.uint64)) == a * b


def test_uneven_products_at_the_large_lengths():
    """Two-operand products at transform lengths 2^22, 2^23 and 2^24 (the
    first pass covers both operands in one launch, its blocks taken in
    groups of neighbouring tiles per XCD), operands of unequal and ragged
    lengths: one longer and one shorter than half the transform. Against
    the oracle at 2^22 and 2^23, by fingerprints mod two primes at 2^24."""
    st = [77]
    for an, bn, exact in (((1 << 21) + 777, (1 << 21) - 1000, True),
                          ((5 << 20) + 3, (3 << 20) - 11, True),
                          ((1 << 23) + 5, (1 << 22) + 3, False)):
        assert sba.lib().sba_gpu_radix3(an, bn, 0) == 0
        assert sba.chunking(an, bn)[2] == 0, (an, bn)      # one transform
        a = xorshift_fill(an, st)
        b = xorshift_fill(bn, st)
        say("uneven", an, bn)
        ok, r = sba.nt_mul(a, b)
        assert ok == 1
        if exact:
            assert (r == orc.nt_mul(a, b)).all(), (an, bn)
        else:
            for q in ((1 << 61) - 1, (1 << 59) - 55):
                assert orc.fingerprint(r, q) == orc.fingerprint(a, q) * orc.fingerprint(b, q) % q
