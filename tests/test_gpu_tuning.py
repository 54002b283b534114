"""GPU: every pass plan is bit-exact; the cost model's forward, backward and
Adam kernels agree with the CPU oracle; the autotuner stores measured plans
and batches built on them stay bit-exact."""
import ctypes
import os
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd import tuning
from simd_bigint_amd.workloads import to_int
from oracle import costmodel as ocm
from tests import _oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(autouse=True)
def fresh_plans():
    sba.lib()
    tuning.plan_reset()
    yield
    tuning.plan_reset()


def _torch():
    import torch
    return torch, torch.device("cuda:0")


def _limbs(t):
    return t.cpu().numpy().view(np.uint64)


def _expected(jobs):
    out = []
    for r, a, b in jobs:
        an = _limbs(a)
        bn = an if b is None else _limbs(b)
        want = np.zeros(len(an) + len(bn), np.uint64)
        rc = orc.lib().orc_nt_mul(want, np.ascontiguousarray(an), len(an),
                                  np.ascontiguousarray(bn), len(bn))
        assert rc == 1
        out.append(want)
    return out


@pytest.mark.parametrize("L", list(range(10, 23)))
@pytest.mark.parametrize("sqr", [0, 1])
def test_every_plan_bit_exact(L, sqr):
    torch, dev = _torch()
    g = torch.Generator(device="cpu").manual_seed(100 + L)
    pool = torch.randint(-2**63, 2**63 - 1, (1 << max(20, L),), dtype=torch.int64,
                         generator=g).to(dev)
    jobs = tuning.make_class_jobs(L, sqr, 3, 0.95, pool, dev)
    # one ragged job of the same class: a short second operand (or a square
    # filling the transform to its last coefficient)
    if sqr:
        an = 1 << (L - 1)
        r = torch.empty(2 * an, dtype=torch.int64, device=dev)
        jobs.append((r, pool[11:11 + an], None))
    else:
        an, bn = (1 << L) - 41, 37
        r = torch.empty(an + bn, dtype=torch.int64, device=dev)
        jobs.append((r, pool[7:7 + an], pool[99:99 + bn]))
    want = _expected(jobs)
    plans = tuning.plan_candidates(L)
    say("L=%d sqr=%d: %d plans x %d jobs" % (L, sqr, len(plans), len(jobs)))
    for plan in plans:
        for r, _, _ in jobs:
            r.fill_(0x5A5A)
        tuning.class_time(jobs, plan, reps=1)
        torch.cuda.synchronize()
        for k, (r, _, _) in enumerate(jobs):
            got = _limbs(r)
            assert np.array_equal(got, want[k]), "plan %r job %d differs" % (plan, k)


@pytest.mark.parametrize("L", [13, 14, 15, 16, 17, 18, 19, 20])
@pytest.mark.parametrize("sqr", [0, 1])
def test_every_radix3_plan_bit_exact(L, sqr):
    """Batch classes of length 3 * 2^L: every radix-3 plan, mixed fills,
    against the oracle."""
    torch, dev = _torch()
    g = torch.Generator(device="cpu").manual_seed(200 + L)
    pool = torch.randint(-2**63, 2**63 - 1, (1 << 21,), dtype=torch.int64, generator=g).to(dev)
    jobs = []
    for z in ((3 << L) // 2, ((1 << (L + 1)) + 3) // 2 + 1, (5 << L) // 4):
        an = z if sqr else z + 7
        bn = z if sqr else z - 7
        r = torch.empty(an + bn, dtype=torch.int64, device=dev)
        jobs.append((r, pool[13:13 + an], None if sqr else pool[501:501 + bn]))
    kinds = set()
    for r, a, b in jobs:
        Lk, kind = ctypes.c_int(), ctypes.c_int()
        arr = sba.job_array([(r, a, b)])
        assert sba.lib().sba_job_class(ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(Lk),
                                       ctypes.byref(kind)) == 1
        kinds.add((Lk.value, kind.value))
    assert kinds == {(L, 2 | sqr)}, kinds
    want = _expected(jobs)
    plans = tuning.plan_candidates(L, 2 | sqr)
    say("radix-3 L=%d sqr=%d: %d plans" % (L, sqr, len(plans)))
    assert plans
    for plan in plans:
        for r, _, _ in jobs:
            r.fill_(0x5A5A)
        tuning.class_time(jobs, plan, reps=1)
        torch.cuda.synchronize()
        for k, (r, _, _) in enumerate(jobs):
            assert np.array_equal(_limbs(r), want[k]), "plan %r job %d differs" % (plan, k)


def test_every_plan_bit_exact_large():
    """2^22 x 2^22 (L = 23): every plan reproduces the default plan's product,
    whose SHA-256 the parity tests pin to the reference's."""
    torch, dev = _torch()
    from simd_bigint_amd.workloads import xorshift_fill
    st = [42]
    n = 1 << 22
    a = torch.from_numpy(xorshift_fill(n, st).view(np.int64)).to(dev)
    b = torch.from_numpy(xorshift_fill(n, st).view(np.int64)).to(dev)
    r0 = torch.empty(2 * n, dtype=torch.int64, device=dev)
    assert sba.mul_dev(r0, a, b) == 1
    torch.cuda.synchronize()
    want = _limbs(r0)
    r = torch.empty(2 * n, dtype=torch.int64, device=dev)
    for plan in tuning.plan_candidates(23):
        r.fill_(7)
        ms = tuning.class_time([(r, a, b)], plan, reps=1)
        torch.cuda.synchronize()
        say("L=23 plan %r: %.3f ms" % (plan, ms))
        assert np.array_equal(_limbs(r), want), "plan %r differs" % (plan,)


def _cm_case(n=300, seed=4):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, ocm.N_IN)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    p = (rng.normal(0, 0.3, ocm.NPARAM)).astype(np.float32)
    return X, t, p


def test_cm_forward_matches_oracle():
    torch, dev = _torch()
    X, t, p = _cm_case()
    y = torch.empty(X.shape[0], dtype=torch.float32, device=dev)
    xs, ps = torch.from_numpy(X).to(dev), torch.from_numpy(p).to(dev)
    assert sba.lib().sba_cm_forward(ps.data_ptr(), xs.data_ptr(), X.shape[0], y.data_ptr(),
                                    None) == 1
    torch.cuda.synchronize()
    want, _ = ocm.forward(p.astype(np.float64), X.astype(np.float64))
    np.testing.assert_allclose(y.cpu().numpy(), want, rtol=1e-4, atol=1e-4)


def test_cm_gradient_matches_oracle():
    """fp32 GPU gradient vs the float64 oracle gradient (itself pinned by
    finite differences): |g - g_ref| <= 1e-4 * max|g_ref| elementwise."""
    torch, dev = _torch()
    X, t, p = _cm_case()
    n = X.shape[0]
    xs, ts, ps = (torch.from_numpy(v).to(dev) for v in (X, t, p))
    G = torch.empty(ocm.NPARAM, dtype=torch.float32, device=dev)
    ws = torch.empty(int(sba.lib().sba_cm_ws_floats(n)), dtype=torch.float32, device=dev)
    assert sba.lib().sba_cm_backward(ps.data_ptr(), xs.data_ptr(), ts.data_ptr(), n,
                                     G.data_ptr(), ws.data_ptr(), None) == 1
    torch.cuda.synchronize()
    g = G.cpu().numpy().astype(np.float64)
    want = ocm.grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    scale = np.abs(want).max()
    err = np.abs(g - want).max()
    say("gradient: max |g - g_ref| = %.3g, max |g_ref| = %.3g" % (err, scale))
    assert err <= 1e-4 * scale
    # the predictions left in ws[0:n] are the forward pass
    np.testing.assert_allclose(ws[:n].cpu().numpy(), ocm.forward(p, X)[0], rtol=1e-4, atol=1e-4)


def test_cm_adam_matches_oracle():
    torch, dev = _torch()
    rng = np.random.default_rng(9)
    p = rng.normal(size=ocm.NPARAM).astype(np.float32)
    g = rng.normal(size=ocm.NPARAM).astype(np.float32)
    m = rng.normal(0, 0.1, ocm.NPARAM).astype(np.float32)
    v = np.abs(rng.normal(0, 0.1, ocm.NPARAM)).astype(np.float32)
    P, Gd, M, V = (torch.from_numpy(x.copy()).to(dev) for x in (p, g, m, v))
    assert sba.lib().sba_cm_adam(P.data_ptr(), Gd.data_ptr(), M.data_ptr(), V.data_ptr(), 3,
                                 1e-2, None) == 1
    torch.cuda.synchronize()
    wp, wm, wv = ocm.adam_step(p.astype(np.float64), g.astype(np.float64), m.astype(np.float64),
                               v.astype(np.float64), 3, 1e-2)
    np.testing.assert_allclose(P.cpu().numpy(), wp, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(M.cpu().numpy(), wm, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(V.cpu().numpy(), wv, rtol=1e-5, atol=1e-7)


def _mixed_jobs(torch, dev):
    g = torch.Generator(device="cpu").manual_seed(77)
    pool = torch.randint(-2**63, 2**63 - 1, (1 << 20,), dtype=torch.int64, generator=g).to(dev)
    jobs = []
    for L, sqr, nj in ((11, 0, 40), (13, 1, 12), (15, 0, 6), (18, 0, 2)):
        jobs += tuning.make_class_jobs(L, sqr, nj, 0.8, pool, dev)
    # and one radix-3 class (length 3 * 2^14)
    jobs += tuning.make_class_jobs(16, 0, 3, 0.7, pool, dev)
    return jobs


def test_autotune_stores_measured_plans_and_stays_exact():
    torch, dev = _torch()
    jobs = _mixed_jobs(torch, dev)
    want = _expected(jobs)
    m = tuning.autotune(jobs, mode=tuning.EXHAUSTIVE, reps=2)
    ncand = sum(len(tuning.plan_candidates(L)) for L in (11, 13, 15, 18))
    ncand += len(tuning.plan_candidates(14, 2))
    assert m == ncand
    for L, sqr, nj in ((11, 0, 40), (13, 1, 12), (15, 0, 6), (18, 0, 2)):
        say("tuned L=%d sqr=%d n=%d: %r" % (L, sqr, nj,
                                             tuning.plan_get(L, sqr, tuning.plan_bucket(nj))))
    for use_graph in (True, False):
        for r, _, _ in jobs:
            r.fill_(3)
        bt = sba.Batch(jobs, use_graph=use_graph)
        for _ in range(2):
            bt.run()
            torch.cuda.synchronize()
            for k, (r, _, _) in enumerate(jobs):
                assert np.array_equal(_limbs(r), want[k])
        bt.close()


def test_autotune_with_model_times_fewer_plans():
    torch, dev = _torch()
    if not os.path.exists(tuning.MODEL_PATH):
        pytest.skip("no trained cost model committed")
    jobs = _mixed_jobs(torch, dev)
    want = _expected(jobs)
    tuning.CostModel.load().install()
    assert sba.lib().sba_cm_loaded() == 1
    m2 = tuning.autotune(jobs, mode=tuning.MODEL, topk=2, reps=2)
    assert m2 == 2 * 4                        # the radix-3 class keeps its plan
    m1 = tuning.autotune(jobs, mode=tuning.MODEL, topk=1, reps=2)
    assert m1 == 0
    bt = sba.Batch(jobs, use_graph=True)
    bt.run()
    torch.cuda.synchronize()
    for k, (r, _, _) in enumerate(jobs):
        assert np.array_equal(_limbs(r), want[k])
    bt.close()


def test_training_on_gpu_learns_plan_costs():
    """Train the cost model from scratch with the library's kernels on the
    committed MI355X timings (training shapes); on the held-out shapes its
    plan picks beat the hand-written rule's."""
    if not os.path.exists(tuning.TIMINGS_PATH):
        pytest.skip("no committed timing dataset")
    rows = tuning.load_timings()
    tr, te = tuning.split_shapes(rows)
    X, y = tuning.dataset(tr)
    cm = tuning.train_gpu(X, y, steps=1500, seed=1)
    say("loss %.4f -> %.5f" % (cm.history[0], cm.history[-1]))
    assert cm.history[-1] < cm.history[0] / 30      # the fit converged
    qm = tuning.pick_quality(te, tuning.model_chooser(cm.predict_log_ms_gpu))
    qh = tuning.pick_quality(te, tuning.heuristic_chooser)
    say("held-out: model %.4f heuristic %.4f time over best" %
        (qm["time_over_best"], qh["time_over_best"]))
    assert qm["time_over_best"] < qh["time_over_best"]


def test_gpu_threshold_tuner():
    """The crossover tuner times the CPU multiply against the GPU path on the
    ladder of balanced sizes; from the threshold it returns up, the GPU wins
    at every ladder size, and apply=False leaves the threshold alone."""
    before = sba.get_gpu_threshold()
    res = sba.tune_gpu_threshold(apply=False)
    print("tuner", res, file=sys.stderr, flush=True)
    assert sba.get_gpu_threshold() == before
    tot, cpu, gpu = res["totals"], res["cpu_ms"], res["gpu_ms"]
    assert len(tot) == len(cpu) == len(gpu) > 4
    assert all(x > 0 for x in cpu + gpu)
    assert gpu[-1] < cpu[-1]                    # 64Ki limbs: the GPU path wins
    i = tot.index(res["threshold"])
    assert all(g < c for g, c in zip(gpu[i:], cpu[i:]))
    if i > 0:
        assert gpu[i - 1] >= cpu[i - 1]


def test_butterfly_probe_near_issue_ceiling():
    """The register-only butterfly loop runs close to the VALU issue ceiling
    the roofline assumes: 10 operations per butterfly at 256 CUs x 4 SIMDs x
    16 lanes x 2.4 GHz, i.e. 3.93e12 butterflies per second."""
    rate = sba.probe_butterflies()
    ceiling = 256 * 4 * 16 * 2.4e9 / 10
    say("butterflies/s", rate, "of", ceiling)
    assert 0.5 * ceiling < rate < 1.2 * ceiling
