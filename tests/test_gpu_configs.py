"""The two multi-product workloads at their stated sizes on the GPU.

Batched: 4096 independent pairs, an and bn log-uniform in [64, 65536] limbs
(10% sparse, 5% all-ones), one hipGraph capture replayed twice; every
product checked by fingerprints mod two 61-bit primes, and a stratified
sample of at least 256 products (every size class) against the oracle.
The reference runs each such pair through simd_mpn_mul (include/u64cvt.h:
84-132) or nt_mul (include/ntmul.h:33-58).

Serving: 320 requests, Poisson arrivals, 80% multiplies / 20% squares,
operand sizes log-uniform in [256, 2^20] limbs, a fixed 1 GiB workspace
arena; every result checked by fingerprints, every result compared with the
same request run alone, and a sample against the oracle."""
import ctypes
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from tests import _oracle as orc

pytestmark = pytest.mark.gpu
QS = [(1 << 61) - 1, 2305843009213693921]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def fp_ok(r, a, b):
    return all(orc.fingerprint(r, q) == orc.fingerprint(a, q) * orc.fingerprint(b, q) % q
               for q in QS)


def test_batch_4096_stated_size():
    torch = pytest.importorskip("torch")
    from simd_bigint_amd.workloads import batch_workload
    pairs = batch_workload(4096, seed=4)
    assert len(pairs) == 4096
    sizes = [n for a, b in pairs for n in (len(a), len(b))]
    assert min(sizes) >= 64 and max(sizes) <= 65536 and max(sizes) > 60000
    dev = torch.device("cuda:0")
    jobs = []
    for a, b in pairs:
        ta = torch.from_numpy(a.view(np.int64)).to(dev)
        tb = torch.from_numpy(b.view(np.int64)).to(dev)
        jobs.append((torch.zeros(len(a) + len(b), dtype=torch.int64, device=dev), ta, tb))
    # size classes of the jobs (as the batch groups them)
    cls = {}
    for i, (r, ta, tb) in enumerate(jobs):
        Lc, kind = ctypes.c_int(), ctypes.c_int()
        arr = sba.job_array([(r, ta, tb)])
        assert sba.lib().sba_job_class(ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(Lc),
                                       ctypes.byref(kind)) == 1
        cls.setdefault((Lc.value, kind.value), []).append(i)
    say("batch 4096: %d size classes" % len(cls))
    assert len(cls) >= 8
    bt = sba.Batch(jobs, use_graph=True)
    # stratified oracle sample: >= 256 products, every class represented
    per = max(256 // len(cls) + 1, 4)
    sample = {i for ids in cls.values() for i in ids[:: max(len(ids) // per, 1)][:per]}
    for i in range(0, len(jobs), 7):             # top up across the batch to >= 256
        if len(sample) >= 256:
            break
        sample.add(i)
    sample = sorted(sample)
    assert len(sample) >= 256 and all(any(i in sample for i in ids) for ids in cls.values())
    want_fp = [[orc.fingerprint(a, q) * orc.fingerprint(b, q) % q for q in QS]
               for a, b in pairs]
    for rep in range(2):                         # two replays of the captured graph
        for (r, _, _) in jobs:
            r.zero_()
        torch.cuda.synchronize()
        bt.run()
        torch.cuda.synchronize()
        bad = [i for i in range(len(pairs))
               if [orc.fingerprint(jobs[i][0].cpu().numpy().view(np.uint64), q) for q in QS]
               != want_fp[i]]
        assert not bad, (rep, bad[:10])
        say("replay", rep, "fingerprints ok")
    for i in sample:
        a, b = pairs[i]
        got = jobs[i][0].cpu().numpy().view(np.uint64)
        assert (got == orc.nt_mul(a, b)).all(), (i, len(a), len(b))
    say("oracle sample", len(sample), "ok")
    bt.close()


def test_serving_poisson_stated_size():
    torch = pytest.importorskip("torch")
    from simd_bigint_amd import serving
    rng = np.random.default_rng(55)
    pool_np = rng.integers(0, 2**64, 1 << 21, dtype=np.uint64)
    pool = torch.from_numpy(pool_np.view(np.int64)).cuda()
    reqs = serving.poisson_requests(320, 20000.0, pool, seed=56, lo=256, hi=1 << 20)
    sizes = [q.a.numel() for q in reqs] + [q.b.numel() for q in reqs if q.b is not None]
    assert min(sizes) >= 256 and max(sizes) <= 1 << 20 and max(sizes) > (1 << 19)
    nsq = sum(q.b is None for q in reqs)
    assert 0.1 * len(reqs) < nsq < 0.3 * len(reqs)
    budget = 1 << 30
    srv = serving.Server(budget, nstreams=3, max_batch=64)
    stats = srv.run(reqs)
    s = stats.summary()
    say("serve", s)
    assert s["requests"] == len(reqs) and stats.peak_ws_bytes <= budget
    assert stats.max_batch > 3                       # requests left together
    off = {}
    for q in reqs:
        a = q.a.cpu().numpy().view(np.uint64)
        b = a if q.b is None else q.b.cpu().numpy().view(np.uint64)
        r = q.r[: q.limbs].cpu().numpy().view(np.uint64)
        assert fp_ok(r, a, b), q.rid
        off[q.rid] = (a, b, r)
    for q in reqs:
        alone = torch.empty(q.limbs, dtype=torch.int64, device="cuda")
        rc = sba.sqr_dev(alone, q.a) if q.b is None else sba.mul_dev(alone, q.a, q.b)
        assert rc == 1
        torch.cuda.synchronize()
        assert torch.equal(alone, q.r[: q.limbs]), q.rid
    order = sorted(reqs, key=lambda q: q.limbs)
    for q in order[:: len(order) // 12] + [order[-1]]:
        a, b, r = off[q.rid]
        want = orc.nt_sqr(a) if q.b is None else orc.nt_mul(a, b)
        assert (r == want).all(), q.rid
    say("serving: all fingerprints, alone == in mix, oracle sample ok")
