"""CPU tests of the surrogate's risk trees: the host trainer
(anacor_risk_fit, csrc/risk_fit.cpp, called through the C-ABI with no GPU)
against closed forms, and the oracle's restatement of the resonance features
and tree scores (oracle/mlp_oracle.py) against hand-worked values.

Pinned: the split a one-level tree must choose on a separable problem and
its threshold; the leaf values of one Newton step of the logistic loss,
recomputed in NumPy from the gradients and Hessians; base = log-odds of the
weighted label mean; duplicated rows equal to weight 2; no split without a
gain; the same trees on 1 and 4 host threads; no split without positive
labels; argument errors.  The
features: ratios and distances to p / q for hand-picked vectors, including
the round-half-even case and a zero component; the scores of a hand-built
depth-2 ensemble."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO


def fit(F, y, w=None, **kw):
    from anacor_public_amd.surrogate import risk_fit
    return risk_fit(F, y, w, **kw)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def test_one_split_separable():
    rng = np.random.default_rng(1)
    F = rng.random((200, 3)).astype(np.float32)
    y = (F[:, 1] > 0.3).astype(np.uint8)
    t = fit(F, y, n_trees=1, depth=1, learning_rate=1.0, l2=0.0, min_hess=0.0, n_bins=256)
    assert t["feat"][0, 0] == 1
    # fewer distinct values than bins: every value but the largest is a cut,
    # so the chosen threshold must separate the classes exactly
    thr = t["thr"][0, 0]
    assert np.all(F[y == 0, 1] <= thr) and np.all(F[y == 1, 1] > thr)
    assert t["base"] == pytest.approx(np.log(y.mean() / (1 - y.mean())), rel=1e-6)


def test_newton_leaves_and_base():
    rng = np.random.default_rng(2)
    n = 150                                   # fewer distinct values than bins
    F = rng.random((n, 2)).astype(np.float32)
    y = (rng.random(n) < 0.2 + 0.6 * F[:, 0]).astype(np.uint8)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    lr, l2 = 0.3, 0.7
    t = fit(F, y, w, n_trees=1, depth=1, learning_rate=lr, l2=l2, min_hess=0.0, n_bins=256)
    wd = w.astype(np.float64)
    p = np.sum(wd * y) / np.sum(wd)
    assert t["base"] == pytest.approx(np.log(p / (1 - p)), rel=1e-6)
    p0 = sigmoid(np.float64(np.float32(np.log(p / (1 - p)))))
    g = wd * (p0 - y)
    h = wd * p0 * (1 - p0)
    f, thr = int(t["feat"][0, 0]), t["thr"][0, 0]
    right = F[:, f] > thr
    exp_leaf = [-lr * g[~right].sum() / (h[~right].sum() + l2),
                -lr * g[right].sum() / (h[right].sum() + l2)]
    np.testing.assert_allclose(t["leaf"][0], exp_leaf, rtol=1e-6)
    # and the chosen split is the best one over every feature and cut
    G, H = g.sum(), h.sum()
    best = -1.0
    for ff in range(2):
        for c in np.unique(F[:, ff])[:-1]:
            r = F[:, ff] > c
            gain = (g[~r].sum() ** 2 / (h[~r].sum() + l2) + g[r].sum() ** 2 / (h[r].sum() + l2)
                    - G ** 2 / (H + l2))
            best = max(best, gain)
    got = (g[~right].sum() ** 2 / (h[~right].sum() + l2)
           + g[right].sum() ** 2 / (h[right].sum() + l2) - G ** 2 / (H + l2))
    assert got == pytest.approx(best, rel=1e-9)


def test_duplicates_equal_weight_two():
    rng = np.random.default_rng(3)
    F = rng.random((120, 4)).astype(np.float32)
    y = (F[:, 2] + 0.3 * rng.random(120) > 0.7).astype(np.uint8)
    a = fit(np.concatenate([F, F]), np.concatenate([y, y]), n_trees=5, depth=3, l2=1.0,
            min_hess=0.0)
    b = fit(F, y, np.full(120, 2.0, np.float32), n_trees=5, depth=3, l2=1.0, min_hess=0.0)
    np.testing.assert_array_equal(a["feat"], b["feat"])
    np.testing.assert_array_equal(a["thr"], b["thr"])
    np.testing.assert_allclose(a["leaf"], b["leaf"], rtol=1e-5, atol=1e-7)


def test_no_gain_no_split_and_threads():
    rng = np.random.default_rng(4)
    F = rng.random((300, 5)).astype(np.float32)
    y = np.zeros(300, np.uint8)
    y[::50] = 1
    const = fit(F[:, :1] * 0 + 1, y, n_trees=2, depth=2)     # one constant feature
    assert np.all(np.isinf(const["thr"]))
    y2 = (F[:, 3] > 0.5).astype(np.uint8) ^ (F[:, 0] > 0.8).astype(np.uint8)
    one = fit(F, y2, n_trees=20, depth=4, threads=1)
    four = fit(F, y2, n_trees=20, depth=4, threads=4)
    for k in ("feat", "thr", "leaf"):
        np.testing.assert_array_equal(one[k], four[k])
    assert one["base"] == four["base"]
    # the fitted ensemble separates the training rows it was fitted on
    s = MO.risk_scores(one, np.pad(F, ((0, 0), (0, 55))))
    assert np.mean((s > 0) == (y2 == 1)) > 0.95


def test_argument_errors():
    from anacor_public_amd import _lib
    F = np.zeros((10, 2), np.float32)
    y = np.zeros(10, np.uint8)
    with pytest.raises(_lib.AnacorError):
        fit(F, np.full(10, 2, np.uint8))
    bad = F.copy()
    bad[3, 1] = np.nan
    with pytest.raises(_lib.AnacorError):
        fit(bad, y)
    with pytest.raises(_lib.AnacorError):
        fit(F, y, depth=9)
    with pytest.raises(_lib.AnacorError):
        fit(F, y, np.full(10, -1.0, np.float32))


def test_risk_features_hand_values():
    X = np.zeros((3, 12), np.float32)
    X[0, 0:6] = [0.0, 0.6, 0.8, 0.36, 0.48, 0.8]
    X[1, 0:6] = [0.5, -0.5, 0.70710677, 0.0, 0.0, 1.0]
    X[2, 0:6] = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0]          # both components zero: t = 0
    Fr = MO.risk_features(X)
    assert Fr.dtype == np.float32 and Fr.shape == (3, 60)
    np.testing.assert_array_equal(Fr[:, :6], X[:, :6])
    f32 = np.float32
    # row 0, x: (0,1) -> 0, (0,2) -> 0, (1,2) -> 0.6/0.8; s: 0.36/0.48, 0.36/0.8, 0.48/0.8
    t = [f32(0), f32(0), f32(0.6) / f32(0.8), f32(0.36) / f32(0.48), f32(0.36) / f32(0.8),
         f32(0.48) / f32(0.8)]
    np.testing.assert_array_equal(Fr[0, 6:12], t)
    for k in range(6):
        for q in range(1, 9):
            u = f32(q) * t[k]
            exp = abs(u - np.rint(u)) * (f32(1) / f32(q))
            assert Fr[0, 12 + 8 * k + q - 1] == exp
    # t(x, (1, 2)) = 0.75 (to FP32): distance 0 at q = 4 and 8, 1/4 at q = 1,
    # 1/4 at q = 2 (1.5 rounds half to even, to 2), 1/12 at q = 3
    k = 2
    assert Fr[0, 6 + k] == pytest.approx(0.75, abs=1e-7)
    assert Fr[0, 12 + 8 * k + 3] < 1e-7 and Fr[0, 12 + 8 * k + 7] < 1e-7
    assert Fr[0, 12 + 8 * k + 0] == pytest.approx(0.25, abs=1e-7)
    assert Fr[0, 12 + 8 * k + 1] == pytest.approx(0.25, abs=1e-7)
    assert Fr[0, 12 + 8 * k + 2] == pytest.approx(1 / 12, abs=1e-7)
    # row 1: x (0, 1) ratio exactly 1 (distance 0 for every q); s (0, 1) both zero
    assert Fr[1, 6] == 1.0 and np.all(Fr[1, 12:20] == 0)
    assert Fr[1, 9] == 0.0 and np.all(Fr[1, 36:44] == 0)
    assert np.all(Fr[2, 6:9] == 0)


def test_risk_scores_hand_tree():
    # two depth-2 trees over features 0 and 1
    trees = {"depth": 2, "base": -1.5,
             "feat": np.array([[0, 1, 1], [1, 0, 0]], np.uint8),
             "thr": np.array([[0.5, 0.2, 0.8], [0.4, np.inf, 0.1]], np.float32),
             "leaf": np.array([[1, 2, 3, 4], [10, 20, 30, 40]], np.float32)}
    F = np.zeros((4, 60), np.float32)
    F[:, 0:2] = [[0.1, 0.1], [0.1, 0.9], [0.9, 0.5], [0.9, 0.95]]
    s = MO.risk_scores(trees, F)
    # tree 0: (0.1, 0.1): left, left -> 1; (0.1, 0.9): left, right -> 2;
    #         (0.9, 0.5): right, 0.5 <= 0.8 left -> 3; (0.9, 0.95): right, right -> 4
    # tree 1: f1 <= 0.4 -> node 1 (thr inf: left) -> 10; f1 > 0.4 -> node 2: f0 > 0.1 -> 40,
    #         (0.1, 0.9): 0.1 > 0.1 is false -> 30
    np.testing.assert_array_equal(s, np.float32([-1.5 + 1 + 10, -1.5 + 2 + 30, -1.5 + 3 + 40,
                                                 -1.5 + 4 + 40]))


def test_no_positive_labels():
    """A surrogate whose training errors all stay under the label threshold
    gives trees without a split and a very negative base: no row is sent to
    the tracer by them."""
    rng = np.random.default_rng(5)
    F = rng.random((500, 6)).astype(np.float32)
    t = fit(F, np.zeros(500, np.uint8), n_trees=4, depth=3)
    assert np.all(np.isinf(t["thr"])) and np.all(np.isfinite(t["leaf"]))
    assert t["base"] < -20
    s = MO.risk_scores(t, np.pad(F, ((0, 0), (0, 54))))
    assert np.all(s < -20)
