"""GPU tests of the learned surrogate (csrc/mlp.hip) against the FP64 CPU
oracle (oracle/mlp_oracle.py) and the ray tracer.

Bars: features to FP32 rounding (inputs; the resonance inputs bit for bit
from the GPU's own FP32 components) and 1e-12 (FP64 invariants); the
forward pass within 5e-5 absolute of the oracle with BF16-rounded weights
(and measurably away from FP64 weights, so the BF16 path is what runs);
gradients within 1e-4 relative L2 per parameter block; the envelope flags and
cells exactly as the oracle assigns them; the risk trees' resonance features
and scores bit for bit as the oracle's FP32 restatement computes them;
reflections outside the envelope come from the tracer bit for bit."""
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _params(seed):
    rng = np.random.default_rng(seed)
    parts = {}
    for name, (a, b, shape) in MO.offsets().items():
        fan_in = shape[1] if len(shape) == 2 else MO.WIDTH
        lim = np.sqrt(6.0 / fan_in)
        parts[name] = rng.uniform(-lim, lim, size=shape)
        if name.startswith("b"):
            parts[name] *= 0.1
    return MO.pack(parts).astype(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def refl():
    from anacor_public_amd import workload as W
    s1, om = W.make_reflections(3000, 11)
    F = W.random_rotation(3).reshape(-1)
    return s1, om, F


@pytest.mark.parametrize("feature_set", [0, 1, 2])
def test_features_match_oracle(gpu, refl, feature_set):
    from anacor_public_amd import workload as W
    from anacor_public_amd.surrogate import MLP
    s1, om, F = refl
    say(f"case: features set {feature_set}")
    with MLP(seed=1) as m:
        m.set_features(feature_set)
        X, inv, _, _ = m.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F)
    Xo, invo = MO.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, feature_set=feature_set)
    np.testing.assert_allclose(X[:, :6], Xo[:, :6], rtol=0, atol=1e-7)
    if feature_set == 0:
        np.testing.assert_allclose(X[:, 6:12], Xo[:, 6:12], rtol=0, atol=3e-7)
    else:
        np.testing.assert_array_equal(X[:, 6:12], Xo[:, 6:12])
    if feature_set == 2:
        # from the GPU's own FP32 components, bit for bit
        R = MO.resonance_inputs(X[:, :6])
        np.testing.assert_array_equal(X[:, 12:].view(np.uint32), R.view(np.uint32))
        assert np.all(X[:, 12:18] <= 1) and np.all(X[:, 18:] <= 1)
    else:
        assert np.all(X[:, 12:] == 0)
    np.testing.assert_allclose(inv[:, 0], invo[:, 0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.cos(inv[:, 1]), invo[:, 1], rtol=0, atol=1e-12)


def test_envelope_cells_and_flags(gpu, refl):
    from anacor_public_amd import workload as W
    from anacor_public_amd.surrogate import MLP, crystal_axis, make_envelope
    s1, om, F = refl
    say("case: envelope cells and flags")
    axis = crystal_axis(W.OMEGA_AXIS, F)
    env = make_envelope(axis, (-0.3, 0.4), (np.radians(10), np.radians(50)), cells=(16, 5, 12))
    rng = np.random.default_rng(2)
    trust = (rng.uniform(size=16 * 5 * 12) < 0.7).astype(np.uint8)
    with MLP(seed=1) as m:
        m.set_envelope(env, trust)
        _, inv, cells, flags = m.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, cells=True,
                                          flags=True)
    invo = MO.invariants(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, env.axis[:], env.e1[:], env.e2[:])
    np.testing.assert_allclose(inv, invo, rtol=0, atol=1e-12)
    co = MO.cell_index(invo, env.two_theta[:], env.cells[:])
    np.testing.assert_array_equal(cells, co)
    fo = ((invo[:, 0] >= -0.3) & (invo[:, 0] <= 0.4) & (invo[:, 1] >= np.radians(10)) &
          (invo[:, 1] <= np.radians(50)) & (trust[co] != 0))
    np.testing.assert_array_equal(flags, fo)
    assert 0.1 < fo.mean() < 0.9


def test_envelope_bands(gpu, refl):
    """Resonance bands on both rays: the GPU's in-envelope flags equal the
    oracle's band test (every cell trusted, ranges open)."""
    from anacor_public_amd import workload as W
    from anacor_public_amd.surrogate import MLP, crystal_axis, make_envelope
    s1, om, F = refl
    say("case: envelope resonance bands")
    axis = crystal_axis(W.OMEGA_AXIS, F)
    env = make_envelope(axis, (-2.0, 2.0), (0.0, np.pi), cells=(4, 2, 4))
    bands = [(1, 0, 1, 1.0, 0.08), (1, 0, 2, 0.5, 0.03), (0, 1, 2, 2.0 / 3.0, 0.05),
             (1, 1, 2, 0.0, 0.04)]
    with MLP(seed=1) as m:
        m.set_envelope(env, None)
        m.set_bands(bands)
        _, _, _, flags = m.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, flags=True)
        m.set_bands([])
        _, _, _, all_in = m.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, flags=True)
    x, s = MO.crystal_frame(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F)
    expect = ~MO.band_mask(x, s, bands)
    np.testing.assert_array_equal(flags, expect)
    assert all_in.all()
    assert 0.1 < expect.mean() < 0.95


def test_risk_kernel_matches_oracle(gpu, refl):
    """The risk kernel's 60 resonance features and its tree scores equal the
    oracle's FP32 restatement bit for bit, and with trees set the in-envelope
    flag is the envelope's flag and score < cut."""
    from anacor_public_amd.surrogate import MLP, risk_fit
    s1, om, F = refl
    raw = np.array([0.0, 0.0, -1.0])
    axis = np.array([1.0, 0.0, 0.0])
    say("case: risk features and trees")
    with MLP(seed=3) as m:
        X, _, _, _ = m.features(s1, om, raw, axis, F)
        # special rows: exact ratios 0, 1/2, 1, halves that round to even, zeros
        Xs = np.zeros((6, X.shape[1]), np.float32)
        Xs[0, :6] = [0.0, 0.6, 0.8, 0.6, 0.0, 0.8]
        Xs[1, :6] = [0.5, 0.5, 0.70710677, 0.25, 0.5, 0.8291562]
        Xs[2, :6] = [1.0, 0.0, 0.0, 0.0, 0.0, 1.0]
        Xs[3, :6] = [-0.3, 0.4, -0.8660254, 0.125, -0.375, 0.9185587]
        Xs[4, :6] = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
        Xs[5, :6] = [1e-30, 3e-30, 1.0, -0.0, 0.0, -1.0]
        Xa = np.concatenate([X, Xs])
        Fr, _ = m.risk(Xa, scores=False)
        Fo = MO.risk_features(Xa)
        np.testing.assert_array_equal(Fr.view(np.uint32), Fo.view(np.uint32))
        # trees fitted on a label of the features, then scored on the GPU
        lab = ((Fo[:, 6 + 2] > 0.5) ^ (Fo[:, 12 + 8 * 4 + 2] < 0.05)).astype(np.uint8)
        trees = risk_fit(Fo, lab, n_trees=30, depth=5)
        m.set_risk(trees)
        _, sc = m.risk(Xa)
        so = MO.risk_scores(trees, Fo)
        np.testing.assert_array_equal(sc.view(np.uint32), so.view(np.uint32))
        assert np.mean((so > 0) == (lab == 1)) > 0.9
        # the flag: envelope and trust as without trees, and score < cut
        from anacor_public_amd.surrogate import crystal_axis, make_envelope
        _, inv, _, _ = m.features(s1, om, raw, axis, F)
        env = make_envelope(crystal_axis(axis, F), (inv[:, 0].min() - 1e-9,
                                                    inv[:, 0].max() + 1e-9),
                            (inv[:, 1].min(), inv[:, 1].max()), (4, 4, 4))
        m.set_risk(None)
        m.set_envelope(env)
        _, _, _, f_env = m.features(s1, om, raw, axis, F, flags=True)
        cut = float(np.median(so[:len(X)]))
        m.set_risk(trees, cut)
        _, _, _, f_risk = m.features(s1, om, raw, axis, F, flags=True)
        np.testing.assert_array_equal(f_risk, f_env & (so[:len(X)] < cut))
        assert 0.3 < f_risk.mean() < 0.7
        with pytest.raises(Exception):
            bad = dict(trees)
            bad["feat"] = np.full_like(trees["feat"], 60)
            m.set_risk(bad, cut)


@pytest.mark.parametrize("n_trees,depth", [(1, 3), (7, 4), (8, 6), (17, 2)])
def test_risk_scores_any_tree_count(gpu, refl, n_trees, depth):
    """The risk kernel walks eight trees side by side and the rest one at a
    time; for tree counts below, at and past a multiple of eight its scores
    equal the oracle's bit for bit (leaves added in tree order)."""
    from anacor_public_amd.surrogate import MLP, risk_fit
    s1, om, F = refl
    raw = np.array([0.0, 0.0, -1.0])
    axis = np.array([1.0, 0.0, 0.0])
    say(f"case: risk scores, {n_trees} trees of depth {depth}")
    with MLP(seed=3) as m:
        X, _, _, _ = m.features(s1, om, raw, axis, F)
        Fo = MO.risk_features(X)
        lab = (Fo[:, 6 + 1] > np.median(Fo[:, 6 + 1])).astype(np.uint8)
        trees = risk_fit(Fo, lab, n_trees=n_trees, depth=depth)
        m.set_risk(trees)
        _, sc = m.risk(X)
        so = MO.risk_scores(trees, Fo)
        np.testing.assert_array_equal(sc.view(np.uint32), so.view(np.uint32))


def test_forward_matches_oracle(gpu, refl):
    from anacor_public_amd import workload as W
    from anacor_public_amd.surrogate import MLP
    s1, om, F = refl
    P = _params(0)
    say("case: forward n=3000")
    with MLP(seed=1) as m:
        m.set_features(2)      # every input column non-zero
        X, _, _, _ = m.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F)
        m.set_params(P)
        np.testing.assert_array_equal(m.get_params(), P.astype(np.float32))
        y = m.forward(X)
        y1 = m.forward(X[:1])
    yo, _ = MO.forward(P, X.astype(np.float64), bf16=True)
    yf, _ = MO.forward(P, X.astype(np.float64), bf16=False)
    err = np.abs(y - yo).max()
    say(f"  max |gpu - oracle(bf16)| {err:.2e}, |gpu - oracle(fp64 weights)| "
        f"{np.abs(y - yf).max():.2e}")
    assert err < 5e-5
    assert np.abs(y - yf).max() > 20 * err      # the BF16 weights are the ones used
    assert y1[0] == y[0]                        # rows are independent of the batch


@pytest.mark.parametrize("n", [777, 1024, 4099])
def test_gradients_match_oracle(gpu, refl, n):
    from anacor_public_amd import workload as W
    from anacor_public_amd.surrogate import MLP
    s1, om, F = refl
    P = _params(1)
    rng = np.random.default_rng(n)
    say(f"case: gradients n={n}")
    with MLP(seed=1) as m:
        m.set_features(2)
        X, _, _, _ = m.features(np.tile(s1, (2, 1))[:n], np.tile(om, 2)[:n], W.RAW_XRAY,
                                W.OMEGA_AXIS, F)
        m.set_params(P)
        t = rng.normal(size=n).astype(np.float32)
        g, loss = m.gradients(X, t)
    yo, cache = MO.forward(P, X.astype(np.float64), bf16=True)
    go, losso = MO.backward(cache, yo, t.astype(np.float64))
    assert loss == pytest.approx(losso, rel=1e-5)
    for name, (a, b, _) in MO.offsets().items():
        rel = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        say(f"  {name}: rel L2 {rel:.2e}")
        assert rel < 1e-4, name


def test_training_fits_a_smooth_map(gpu):
    """Adam on the GPU drives the loss of a smooth map of 12 inputs down by
    orders of magnitude, deterministically for a fixed seed (inputs [12:60]
    zero, as feature sets 0 and 1 leave them)."""
    from anacor_public_amd import _lib
    from anacor_public_amd.surrogate import MLP
    rng = np.random.default_rng(7)
    X = np.zeros((20000, _lib.MLP_INPUTS), np.float32)
    X[:, :12] = rng.uniform(-1, 1, size=(20000, 12))
    t = (np.sin(2 * X[:, 0]) * X[:, 4] + 0.5 * X[:, 7] ** 2).astype(np.float32)
    say("case: training")
    losses = []
    for rep in range(2):
        with MLP(seed=3) as m:
            losses.append(m.train(X, t, epochs=20, batch=512, lr=1e-3, lr_min=1e-5, seed=4))
            if rep == 0:
                pred = m.forward(X[:2000])
    say(f"  loss {losses[0][0]:.3e} -> {losses[0][-1]:.3e}")
    np.testing.assert_array_equal(losses[0], losses[1])
    assert losses[0][-1] < 1e-2 * losses[0][0]
    assert np.sqrt(np.mean((pred - t[:2000]) ** 2)) < 0.02


@pytest.fixture(scope="module")
def fitted():
    """The surrogate of batched sample 2 as the benchmark fits it: 3.2M traced
    training reflections, 300 epochs of batch 4096 at lr 2e-3, risk trees
    learned on 800k more."""
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    from anacor_public_amd.surrogate import Surrogate
    sw = W.surrogate_config(n_infer=200_000)
    wl = sw.sample
    say(f"case: fit surrogate of {wl.name} on {len(wl.omega)} reflections")
    tr = Tracer.for_workload(wl)
    sur = Surrogate(tr, wl.raw_xray, wl.omega_axis, wl.F)
    rep = sur.fit(wl.s1, wl.omega, epochs=300, batch=4096, lr=2e-3, risk_s1=sw.risk_s1,
                  risk_omega=sw.risk_omega)
    say(f"  trained in {rep['train_s']:.1f}s, loss {rep['final_loss']:.3e}")
    yield sw, tr, sur, rep
    sur.close()
    tr.close()


def _heldout(fitted, bounded):
    sw, tr, sur, rep = fitted
    wl = sw.sample
    A_true = tr.run(sw.heldout_s1, sw.heldout_omega, wl.raw_xray, wl.omega_axis, wl.F)
    sur.set_bounded(bounded)
    try:
        A, used = sur.predict(sw.heldout_s1, sw.heldout_omega)
    finally:
        sur.set_bounded(True)
    r = np.abs(A[used] - A_true[used]) / A_true[used]
    say(f"  held-out ({'bounded' if bounded else 'fast'}, {used.mean() * 100:.1f}% answered): "
        f"max {r.max():.2e} p99.9 {np.quantile(r, 0.999):.2e} p99 {np.quantile(r, 0.99):.2e} "
        f"mean {r.mean():.2e}")
    # the rest is the tracer's own answer (same call shape, so the same tiles)
    out = ~used
    np.testing.assert_array_equal(
        A[out], tr.run(sw.heldout_s1[out], sw.heldout_omega[out], wl.raw_xray, wl.omega_axis,
                       wl.F))
    return r, used


def test_surrogate_heldout_error_bounded(gpu, fitted):
    """Bounded mode (the default after fit(bound=1e-3)): the risk trees,
    fitted on the network's errors on a separate risk set, send the
    reflections where its error is likely to exceed the bound to the tracer
    (under 5% of them), and every held-out reflection the network answers is
    within the stated 1e-3 of the tracer."""
    sw, tr, sur, rep = fitted
    say(f"  risk trees: cut {rep['risk_cut']:.3f}, cross-fitted fallback on the "
        f"{rep['risk_rows']} risk rows {rep['risk_train_fallback'] * 100:.1f}%, fitted in "
        f"{rep['trust_fit_s']:.1f}s")
    assert sur.bounded and rep["trust"] == "trees" and sur.risk is not None
    assert rep["risk_set"] == "separate" and rep["risk_rows"] == len(sw.risk_omega)
    r, used = _heldout(fitted, True)
    assert r.max() < 1e-3
    assert used.mean() > 0.95


def test_surrogate_heldout_error_fast(gpu, fitted):
    """Fast mode (risk trees off): 99% of held-out reflections within 1e-3 and a
    mean below 3e-4.  Its tail is the tracer's own fine oscillation near
    rational direction ratios, which the training reflections do not
    resolve; it is measured and reported, not bounded."""
    r, used = _heldout(fitted, False)
    assert used.mean() > 0.99
    assert r.mean() < 3e-4
    assert np.quantile(r, 0.99) < 1e-3


def test_surrogate_fallback_and_gain(gpu, fitted):
    """Reflections outside the training 2-theta range, or from another
    goniometer, come from the tracer bit for bit; on the mixed inference set
    the fast surrogate is at least 20x faster than tracing every reflection,
    the bounded one at least 15x (on 200k rows; the benchmark's 1M-row figure
    carries fixed costs over more rows)."""
    import time
    sw, tr, sur, rep = fitted
    wl = sw.sample
    geo = (wl.raw_xray, wl.omega_axis, wl.F)
    sur.set_bounded(False)
    try:
        A, used = sur.predict(sw.infer_s1, sw.infer_omega)
        assert not used[sw.infer_outside].any()
        assert used[~sw.infer_outside].mean() > 0.99
        out = ~used
        ref_out = tr.run(sw.infer_s1[out], sw.infer_omega[out], *geo)
        np.testing.assert_array_equal(A[out], ref_out)
        # another crystal orientation: nothing is inside the envelope
        from anacor_public_amd import workload as W
        F2 = W.random_rotation(11).reshape(-1)
        sur2_s1, sur2_om = sw.heldout_s1[:300], sw.heldout_omega[:300]
        A2, used2 = sur.mlp.predict(sur2_s1, sur2_om, wl.raw_xray, wl.omega_axis, F2)
        assert not used2.any() and np.all(np.isnan(A2))
        # speed against the method without the network (trace everything)
        sur.predict(sw.infer_s1, sw.infer_omega)
        t0 = time.perf_counter()
        sur.predict(sw.infer_s1, sw.infer_omega)
        t_fast = time.perf_counter() - t0
    finally:
        sur.set_bounded(True)
    sur.predict(sw.infer_s1, sw.infer_omega)
    t0 = time.perf_counter()
    _, used_b = sur.predict(sw.infer_s1, sw.infer_omega)
    t_bounded = time.perf_counter() - t0
    t0 = time.perf_counter()
    A_tr = tr.run(sw.infer_s1, sw.infer_omega, *geo)
    t_tr = time.perf_counter() - t0
    r = np.abs(A[used] - A_tr[used]) / A_tr[used]
    say(f"  {len(A)} reflections: fast {t_fast * 1e3:.1f} ms ({t_tr / t_fast:.1f}x), bounded "
        f"{t_bounded * 1e3:.1f} ms ({t_tr / t_bounded:.1f}x, {used_b.mean() * 100:.1f}% "
        f"answered), tracer {t_tr * 1e3:.1f} ms, fast p99 error {np.quantile(r, 0.99):.2e}")
    assert t_tr / t_fast > 20
    assert t_tr / t_bounded > 15
    assert np.quantile(r, 0.99) < 1e-3


def test_surrogate_state_round_trip(gpu, fitted):
    from anacor_public_amd.surrogate import Surrogate
    sw, tr, sur, rep = fitted
    wl = sw.sample
    st = sur.state()
    with Surrogate(tr, wl.raw_xray, wl.omega_axis, wl.F, seed=99) as other:
        other.load_state(st)
        a, ua = sur.predict(sw.heldout_s1[:2000], sw.heldout_omega[:2000], fallback=False)
        b, ub = other.predict(sw.heldout_s1[:2000], sw.heldout_omega[:2000], fallback=False)
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(a[ua], b[ub])
    assert other.bounded == sur.bounded and other.bands == sur.bands
    assert other.risk_cut == sur.risk_cut
    for k in ("feat", "thr", "leaf"):
        np.testing.assert_array_equal(other.risk[k], sur.risk[k])
    assert not ua.all()                     # the trees send some rows to the tracer
