"""Pins the FP64 CPU restatement of the surrogate network (oracle/mlp_oracle.py)
before it is used to judge the GPU kernels: finite differences for the
backward pass, a library conversion and tie cases for BF16 rounding, closed
forms for SiLU, and geometric identities for the features."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = {}
    for name, (a, b, shape) in MO.offsets().items():
        fan_in = shape[1] if len(shape) == 2 else MO.WIDTH
        lim = scale * np.sqrt(6.0 / fan_in)
        parts[name] = rng.uniform(-lim, lim, size=shape)
    return MO.pack(parts)


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    P = _params(2)
    X = rng.normal(size=(6, MO.N_IN))
    t = rng.normal(size=6)
    y, cache = MO.forward(P, X, bf16=False)
    g, loss = MO.backward(cache, y, t)
    assert loss == pytest.approx(MO.loss_only(P, X, t), rel=1e-14)
    h = 1e-6
    for name, (a, b, _) in MO.offsets().items():
        for j in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            Pp, Pm = P.copy(), P.copy()
            Pp[j] += h
            Pm[j] -= h
            fd = (MO.loss_only(Pp, X, t) - MO.loss_only(Pm, X, t)) / (2 * h)
            assert g[j] == pytest.approx(fd, rel=1e-5, abs=1e-9), (name, j)


def test_bf16_rounding_ties_and_library():
    one = np.float32(1.0)
    ulp = 2.0 ** -7
    # exact halfway cases go to the even mantissa
    assert MO.bf16_round(np.float32(1 + ulp / 2))[()] == 1.0
    assert MO.bf16_round(np.float32(1 + 3 * ulp / 2))[()] == 1 + 2 * ulp
    assert MO.bf16_round(np.float32(1 + ulp / 2 + 2.0 ** -20))[()] == 1 + ulp
    assert MO.bf16_round(-one)[()] == -1.0
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32),
                        (rng.integers(1, 1 << 16, size=2000) * 2.0 ** -16 + 1).astype(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(MO.bf16_round(x), ref)


def test_silu_closed_forms():
    assert MO.silu(0.0) == 0.0
    assert MO.silu_grad(0.0) == 0.5
    assert MO.silu(40.0) == pytest.approx(40.0, rel=1e-15)
    assert MO.silu_grad(40.0) == pytest.approx(1.0, rel=1e-15)
    z = np.linspace(-8, 8, 33)
    h = 1e-6
    np.testing.assert_allclose(MO.silu_grad(z), (MO.silu(z + h) - MO.silu(z - h)) / (2 * h),
                               rtol=1e-8, atol=1e-10)


def test_features_identities():
    from anacor_public_amd import workload as W
    s1, om = W.make_reflections(500, 9)
    F = W.random_rotation(7).reshape(-1)
    X, inv = MO.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F)
    x, s = X[:, 0:3], X[:, 3:6]
    np.testing.assert_allclose(np.linalg.norm(x, axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(np.linalg.norm(s, axis=1), 1.0, rtol=1e-14)
    for v, h in ((x, X[:, 6:9]), (s, X[:, 9:12])):
        np.testing.assert_allclose(h[:, 0] ** 2 + h[:, 1] ** 2, (v[:, 0] ** 2 + v[:, 1] ** 2) ** 2,
                                   atol=1e-14)
        np.testing.assert_allclose(h[:, 2], 2 * v[:, 2] ** 2 - 1, atol=1e-15)
    # a rotation keeps the angle between the beam and s1 (2 theta) ...
    b = W.RAW_XRAY / np.linalg.norm(W.RAW_XRAY)
    np.testing.assert_allclose(inv[:, 1], s1 @ b / np.linalg.norm(s1, axis=1), atol=1e-14)
    # ... and the beam's component along the goniometer axis (0 here)
    np.testing.assert_allclose(inv[:, 0], b @ W.OMEGA_AXIS, atol=1e-14)


def test_rotation_agrees_with_c_oracle(oracle):
    rng = np.random.default_rng(4)
    for _ in range(5):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ang = rng.uniform(0, 2 * np.pi)
        np.testing.assert_allclose(MO.kp_rotation(axis, ang).reshape(-1), oracle.kp_rotation(axis, ang),
                                   rtol=0, atol=1e-15)


def test_invariants_recover_constructed_angles():
    """Build directions from chosen (alpha, two_theta, beta) and read them back."""
    rng = np.random.default_rng(5)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    e1 = np.cross(u, [1.0, 0, 0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u, e1)
    n = 200
    alpha = rng.uniform(-np.pi, np.pi, n)
    tth = rng.uniform(0.05, 1.2, n)
    beta = rng.uniform(-np.pi, np.pi, n)
    x = np.cos(alpha)[:, None] * e1 + np.sin(alpha)[:, None] * e2
    f1 = np.cross(u[None, :], x)
    f1 /= np.linalg.norm(f1, axis=1, keepdims=True)
    f2 = np.cross(x, f1)
    s = (np.cos(tth)[:, None] * x + np.sin(tth)[:, None] *
         (np.cos(beta)[:, None] * f1 + np.sin(beta)[:, None] * f2))
    # omega = 0, F = I, raw_xray = x: the crystal frame is the lab frame
    inv = np.concatenate([MO.invariants(s[i:i + 1], [0.0], x[i], u, np.eye(3).ravel(), u, e1, e2)
                          for i in range(n)])
    np.testing.assert_allclose(inv[:, 0], 0.0, atol=1e-14)
    np.testing.assert_allclose(inv[:, 1], tth, atol=1e-12)
    np.testing.assert_allclose(inv[:, 2], alpha, atol=1e-12)
    np.testing.assert_allclose(inv[:, 3], beta, atol=1e-12)
    cells = (8, 4, 6)
    c = MO.cell_index(inv, (0.05, 1.2), cells)
    ia = np.floor((alpha + np.pi) / (2 * np.pi) * 8)
    it = np.minimum(np.floor((tth - 0.05) / 1.15 * 4), 3)
    ib = np.floor((beta + np.pi) / (2 * np.pi) * 6)
    np.testing.assert_array_equal(c, (ia * 4 + it) * 6 + ib)


def test_chamber_signs():
    x = np.array([[0.0, 0.6, 0.8], [0.0, -0.8, -0.6]])
    s = np.array([[0.6, -0.48, 0.64], [-0.36, 0.48, -0.8]])
    np.testing.assert_array_equal(MO.chamber(x, s),
                                  [[1, -1, -1, 1, 1, -1], [-1, 1, -1, -1, -1, -1]])


def test_band_mask_closed_forms():
    """Directions built with a known component ratio fall inside a band of
    that ratio exactly when the width exceeds their offset; the other ray and
    the other component pairs are left alone."""
    def unit(v):
        v = np.asarray(v, dtype=np.float64)
        return v / np.linalg.norm(v)
    # s with |s0| / |s1| = 0.5 + d for a few offsets d (s2 fixed)
    d = np.array([-0.02, -0.005, 0.0, 0.004, 0.03])
    s = np.array([unit([0.5 + e, -1.0, 0.3]) for e in d])
    # x along the second axis: ratio 0 for the pairs (0, 1) and (1, 2), 0 / 0 for (0, 2)
    x = np.tile(unit([0.0, 1.0, 0.0]), (len(d), 1))
    got = MO.band_mask(x, s, [(1, 0, 1, 0.5, 0.01)])
    np.testing.assert_array_equal(got, np.abs(d) < 0.01)
    # ratio taken as min / max: |s1| / |s0| = 2 is the same band as 1/2
    s_swapped = s[:, [1, 0, 2]]
    np.testing.assert_array_equal(MO.band_mask(x, s_swapped, [(1, 0, 1, 0.5, 0.01)]),
                                  np.abs(d) < 0.01)
    # a band on the incident ray does not look at s, and vice versa
    assert not MO.band_mask(x, s, [(0, 0, 1, 0.5, 0.1)]).any()
    assert MO.band_mask(x, s, [(0, 0, 1, 0.0, 1e-9)]).all()      # x0 = 0: ratio 0 exactly
    assert MO.band_mask(x, s, [(0, 0, 2, 0.0, 1e-9)]).all()      # 0 / 0 counts as ratio 0
    assert not MO.band_mask(x, s, [(1, 0, 1, 0.5, 0.0)]).any()   # zero width: nothing


def test_resonance_inputs_hand_values():
    """Feature set 2's inputs [12:60] for hand-picked component rows: the
    ratios, and 2 |q t - rint(q t)| with its zeros at t = p / q, the
    round-half-even cases (t q = 1.5 -> 2, 4.5 -> 4) and both components 0."""
    f32 = np.float32
    V = np.zeros((3, 6), np.float32)
    V[0] = [0.0, 0.6, 0.8, 0.36, 0.48, 0.8]
    V[1] = [0.5, -0.5, 0.70710677, 0.0, 0.0, 1.0]
    V[2] = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    R = MO.resonance_inputs(V)
    assert R.dtype == np.float32 and R.shape == (3, 48)
    # row 0 ratios: x (0, 1) and (0, 2) -> 0, (1, 2) -> 0.6 / 0.8; s likewise
    t = [f32(0), f32(0), f32(0.6) / f32(0.8), f32(0.36) / f32(0.48), f32(0.36) / f32(0.8),
         f32(0.48) / f32(0.8)]
    np.testing.assert_array_equal(R[0, :6], t)
    for k in range(6):
        for q in range(2, 9):
            u = f32(q) * t[k]
            assert R[0, 6 + 7 * k + q - 2] == f32(2) * abs(u - np.rint(u))
    # t = 0.75: q t = 1.5 (rounds to 2: 1.0), 2.25 (0.5), 3 (0), 3.75 (0.5),
    # 4.5 (rounds to 4: 1.0), 5.25 (0.5), 6 (0)
    assert R[0, 2] == pytest.approx(0.75, abs=1e-7)
    np.testing.assert_allclose(R[0, 6 + 7 * 2: 6 + 7 * 3], [1.0, 0.5, 0, 0.5, 1.0, 0.5, 0],
                               atol=1e-6)
    # t = 0: every distance 0; x (0, 1) of row 1 is exactly 1: distance 0 too
    assert np.all(R[0, 6:6 + 14] == 0)
    assert R[1, 0] == 1.0 and np.all(R[1, 6:13] == 0)
    # s (0, 1) of row 1: both components zero -> t = 0
    assert R[1, 3] == 0.0 and np.all(R[1, 6 + 21:6 + 28] == 0)
    assert np.all(R[2, 0:3] == 0)


def test_feature_set_layout():
    """Sets 0 and 1 leave [12:60] zero; set 2 carries the chamber signs of set
    1 and the resonance inputs of its own FP32 components."""
    from anacor_public_amd import workload as W
    s1, om = W.make_reflections(300, 13)
    F = W.random_rotation(5).reshape(-1)
    X0, _ = MO.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, feature_set=0)
    X1, _ = MO.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, feature_set=1)
    X2, _ = MO.features(s1, om, W.RAW_XRAY, W.OMEGA_AXIS, F, feature_set=2)
    assert X0.shape == X1.shape == X2.shape == (300, MO.N_IN)
    assert np.all(X0[:, 12:] == 0) and np.all(X1[:, 12:] == 0)
    np.testing.assert_array_equal(X2[:, :12], X1[:, :12])
    np.testing.assert_array_equal(X2[:, 12:], MO.resonance_inputs(X2[:, :6]))
    # the ratio of (i, j) is symmetric in the two components and sign-blind
    flip = X2[:, :6].astype(np.float32).copy()
    flip[:, [3, 4]] = -flip[:, [4, 3]]
    R = MO.resonance_inputs(flip)
    np.testing.assert_array_equal(R[:, 3], X2[:, 12 + 3])
