"""Cost model and plans, CPU side: the oracle MLP pinned by finite
differences and by networks with closed-form outputs; plan features and
candidate sets from their definitions; and, when the MI355X timing dataset
and trained model are committed, the trained network's gain over the
default plans on held-out shapes."""
import math
import os

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd import tuning
from oracle import costmodel as ocm


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(40, ocm.N_IN))
    t = rng.normal(size=40)
    p = rng.normal(0, 0.4, ocm.NPARAM)
    _, (_, z1, _, z2, _) = ocm.forward(p, X)
    h = 1e-7
    # no unit within a step of its ReLU kink (a step moves a pre-activation
    # by at most h * max|input| < 1e-6), so the loss is smooth here
    assert np.abs(z1).min() > 1e-5 and np.abs(z2).min() > 1e-5
    g = ocm.grad(p, X, t)
    idx = np.concatenate([rng.choice(ocm.NPARAM, 300, replace=False),
                          np.arange(ocm.NPARAM - 40, ocm.NPARAM)])
    for i in idx:
        e = np.zeros(ocm.NPARAM)
        e[i] = h
        fd = (ocm.loss(p + e, X, t) - ocm.loss(p - e, X, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), i


def test_oracle_gradient_catches_planted_mistake():
    """A dropped ReLU mask in the first layer's error (a plausible slip)
    disagrees with finite differences."""
    rng = np.random.default_rng(1)
    X = rng.normal(size=(40, ocm.N_IN))
    t = rng.normal(size=40)
    p = rng.normal(0, 0.4, ocm.NPARAM)
    W1, b1, W2, b2, w3, b3 = ocm.unpack(p)
    y, (_, z1, a1, z2, a2) = ocm.forward(p, X)
    e = 2.0 * (y - t) / len(t)
    d2 = np.outer(e, w3) * (z2 > 0)
    d1_wrong = d2 @ W2                      # mask omitted
    g_wrong = (d1_wrong.T @ X).ravel()
    h = 1e-6
    diffs = []
    for i in range(ocm.H1 * ocm.N_IN):
        ev = np.zeros(ocm.NPARAM)
        ev[i] = h
        fd = (ocm.loss(p + ev, X, t) - ocm.loss(p - ev, X, t)) / (2 * h)
        diffs.append(abs(fd - g_wrong[i]))
    assert max(diffs) > 1e-3


def _identity_net(b3=0.5):
    p = np.zeros(ocm.NPARAM)
    o = 0
    for j in range(ocm.N_IN):
        p[o + j * ocm.N_IN + j] = 1.0                     # W1 = [I; 0]
    o = ocm.H1 * ocm.N_IN + ocm.H1
    for k in range(ocm.H2):
        p[o + k * ocm.H1 + k] = 1.0                       # W2 = I
    o += ocm.H2 * ocm.H1 + ocm.H2
    p[o:o + ocm.H2] = 1.0                                 # w3 = 1
    p[-1] = b3
    return p


def test_oracle_forward_closed_form():
    """W1 = [I; 0], W2 = I, w3 = 1: f(x) = sum_i relu(x_i) + b3."""
    rng = np.random.default_rng(2)
    X = rng.normal(size=(25, ocm.N_IN))
    y, _ = ocm.forward(_identity_net(0.5), X)
    np.testing.assert_allclose(y, np.maximum(X, 0).sum(axis=1) + 0.5, rtol=0, atol=1e-12)


def test_oracle_gradient_closed_form_bias():
    """d loss / d b3 = (2/n) sum (f - t)."""
    rng = np.random.default_rng(3)
    X = rng.normal(size=(30, ocm.N_IN))
    t = rng.normal(size=30)
    p = _identity_net(0.25)
    y, _ = ocm.forward(p, X)
    assert abs(ocm.grad(p, X, t)[-1] - 2 * np.mean(y - t)) < 1e-12


def test_oracle_adam_first_step_is_sign_step():
    """Bias-corrected Adam's first step: m_hat = g, v_hat = g^2, so each
    weight moves by -lr * g / (|g| + eps), i.e. lr * sign(g) up to eps."""
    rng = np.random.default_rng(4)
    p = rng.normal(size=ocm.NPARAM)
    g = rng.normal(size=ocm.NPARAM)
    p1, m, v = ocm.adam_step(p, g, np.zeros_like(p), np.zeros_like(p), 1, 0.01)
    np.testing.assert_allclose(p1 - p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-12)


def test_features_from_definition():
    f = tuning.features(23, 0, 1, 1.0, (13, 5, 1))
    k5 = math.log2(5)
    want = [23, 0, 0, 1.0, 23, 0, 10, 13, 0, 0, 1, 1, 10 + k5, 10 + k5, 3, 5]
    np.testing.assert_allclose(f, want, rtol=1e-6)
    f = tuning.features(12, 1, 64, 0.8, (12, 4, 0))
    want = [12, 1, 6, 0.8, 18, 1, 0, 12, 0, 1, 0, 0, 6, 6 + k5, 0, 4]
    np.testing.assert_allclose(f, want, rtol=1e-6)


def test_plan_candidates_and_validation():
    sba.lib()
    counts = {L: len(tuning.plan_candidates(L)) for L in range(10, 26)}
    # single tile up to 13 (6 = 3 radices x 2 split modes); last in 11..13
    # with 1 <= L - last <= 12
    assert counts[10] == counts[11] == counts[12] == 6
    assert all(counts[L] == 18 for L in range(13, 24))
    assert counts[24] == 12 and counts[25] == 6
    for L in range(10, 26):
        assert tuning.plan_heuristic(L) in tuning.plan_candidates(L)
    with pytest.raises(ValueError):
        tuning.plan_set(25, 0, 0, (11, 5, 1))           # 14 first-pass levels
    with pytest.raises(ValueError):
        tuning.plan_set(12, 0, 0, (11, 5, 1))           # two passes below 2^13
    shipped = tuning.plan_get(20, 1, 3)
    other = [p for p in tuning.plan_candidates(20) if p != shipped][0]
    tuning.plan_set(20, 1, 3, other)
    assert tuning.plan_get(20, 1, 3) == other
    tuning.plan_reset()
    assert tuning.plan_get(20, 1, 3) == shipped
    assert [tuning.plan_bucket(n) for n in (1, 3, 4, 15, 16, 1023, 1024, 9999)] == \
        [0, 0, 1, 1, 2, 4, 5, 5]


def _committed():
    if not (os.path.exists(tuning.TIMINGS_PATH) and os.path.exists(tuning.MODEL_PATH)):
        pytest.skip("no committed MI355X timings / cost model")
    return tuning.load_timings(), tuning.CostModel.load()


def test_committed_model_beats_heuristic_on_held_out_shapes():
    """The cost model trained on the GPU (its training shapes only) picks
    plans for shapes it never saw; scored with the measured times of every
    plan of those shapes, its picks are closer to the fastest plan than the
    hand-written rule's."""
    rows, cm = _committed()
    tr, te = tuning.split_shapes(rows)
    assert cm.history[-1] < cm.history[0] / 30      # the fit converged

    def predict(X):
        y, _ = ocm.forward(cm.params.astype(np.float64), cm.standardize(X).astype(np.float64))
        return y * cm.ystd + cm.ymean
    qm = tuning.pick_quality(te, tuning.model_chooser(predict))
    qh = tuning.pick_quality(te, tuning.heuristic_chooser)
    print("held-out shapes %d: model %.4f / heuristic %.4f time over best" %
          (qm["shapes"], qm["time_over_best"], qh["time_over_best"]))
    assert qm["shapes"] >= 50
    assert qm["time_over_best"] < qh["time_over_best"]
    assert qm["geomean_slowdown"] < qh["geomean_slowdown"]


def test_shipped_plan_table_matches_dataset():
    """csrc/plan_table.h is what the generator makes of the committed
    dataset (power-of-two and radix-3 classes), and the library serves it
    after a reset."""
    rows, _ = _committed()
    rows3 = tuning.load_timings3()
    table = tuning.shipped_table(rows, rows3)
    assert any(kind & 2 for kind, _, _ in table) == bool(rows3)
    tuning.plan_reset()
    for (kind, L, b), plan in table.items():
        assert tuning.plan_get(L, kind, b) == plan, (kind, L, b)
