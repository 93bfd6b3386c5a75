"""FP64 CPU restatement of the surrogate network (TEST INFRASTRUCTURE ONLY).

The learned surrogate has no counterpart in the reference (its nearest
analogue is the gridding approximation, src/gridding_interpolation.c:98-161);
this module is the checker for the GPU kernels of anacor_public_amd/csrc/mlp.hip:

* ``features`` -- the network inputs of a reflection: the crystal-frame
  incident and diffracted unit vectors, rotated as the tracer's setup does
  (Rodrigues matrix of utils/utils_os.py:26-43, ``(R(omega) F)^T v`` as
  ray_tracing.c:397-405 applies it), plus their second harmonics or chamber
  signs, and (``resonance_inputs``) the kinks at small rational component
  ratios;
* ``risk_features`` / ``risk_scores`` -- the resonance features of the
  surrogate's risk trees and the trees' FP32 scores, the checker of
  mlp.hip's risk_kernel;
* ``forward`` / ``backward`` -- MLP 60 -> 256 x 4 -> 1 with SiLU and the mean
  squared error, in FP64; with ``bf16=True`` the three 256x256 weight matrices
  are first rounded to BF16 (round to nearest even) exactly as the GPU uses
  them, so the GPU's remaining differences are FP32 accumulation only.

Pinned by tests/test_oracle_mlp.py: central finite differences of the loss
(FP64, ``bf16=False``) for every parameter block, BF16 rounding against
torch's conversion and hand-worked tie cases, SiLU closed forms, the
feature identities (unit norms, harmonic identities, rotation invariance of
x . s), and hand-worked values of the resonance inputs (zeros at t = p / q,
round-half-even cases, zero components).  Never imported by the product path.
"""
from __future__ import annotations

import numpy as np

N_IN = 60
WIDTH = 256
# parameter layout: W1[256][60] b1 W2[256][256] b2 W3 b3 W4 b4 w5[256] b5[1]
_SHAPES = (("W1", (WIDTH, N_IN)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
           ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH, WIDTH)), ("b4", (WIDTH,)),
           ("w5", (WIDTH,)), ("b5", (1,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in _SHAPES)
assert N_PARAMS == 213249


def offsets():
    out, o = {}, 0
    for name, shape in _SHAPES:
        n = int(np.prod(shape))
        out[name] = (o, o + n, shape)
        o += n
    return out


def unpack(flat: np.ndarray) -> dict:
    flat = np.asarray(flat, dtype=np.float64).reshape(-1)
    assert flat.size == N_PARAMS
    return {k: flat[a:b].reshape(s) for k, (a, b, s) in offsets().items()}


def pack(parts: dict) -> np.ndarray:
    out = np.empty(N_PARAMS, dtype=np.float64)
    for k, (a, b, s) in offsets().items():
        out[a:b] = np.asarray(parts[k], dtype=np.float64).reshape(-1)
    return out


def bf16_round(x) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even), returned as float64 values."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    r = ((u + np.uint64(0x7FFF) + lsb) >> np.uint64(16)) << np.uint64(16)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isfinite(f), out, f.astype(np.float64))


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, X, bf16: bool = True):
    """y[n] and the cache for ``backward``.  X: [n][60]."""
    p = unpack(params)
    W = [p["W2"], p["W3"], p["W4"]]
    if bf16:
        W = [bf16_round(w) for w in W]
    X = np.asarray(X, dtype=np.float64)
    zs, hs = [], [X]
    z = X @ p["W1"].T + p["b1"]
    zs.append(z)
    hs.append(silu(z))
    for w, b in zip(W, (p["b2"], p["b3"], p["b4"])):
        z = hs[-1] @ w.T + b
        zs.append(z)
        hs.append(silu(z))
    y = hs[-1] @ p["w5"] + p["b5"][0]
    return y, (p, W, zs, hs)


def backward(cache, y, t) -> tuple:
    """Gradient of mean((y - t)^2) with respect to every parameter (flat,
    the GPU layout) and the loss."""
    p, W, zs, hs = cache
    n = len(y)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / n
    g = {"w5": hs[4].T @ dy, "b5": np.array([dy.sum()])}
    dz = np.outer(dy, p["w5"]) * silu_grad(zs[3])
    for li in (2, 1, 0):                      # W4, W3, W2
        name = ("W2", "W3", "W4")[li]
        bname = ("b2", "b3", "b4")[li]
        g[name] = dz.T @ hs[li + 1]
        g[bname] = dz.sum(axis=0)
        dz = (dz @ W[li]) * silu_grad(zs[li])
    g["W1"] = dz.T @ hs[0]
    g["b1"] = dz.sum(axis=0)
    return pack(g), loss


def loss_only(params, X, t, bf16: bool = False) -> float:
    y, _ = forward(params, X, bf16)
    d = y - np.asarray(t, dtype=np.float64)
    return float(np.mean(d * d))


def kp_rotation(axis, angle):
    """Rodrigues matrix as utils/utils_os.py:26-43 writes it."""
    x, y, z = axis
    c, s = np.cos(angle), np.sin(angle)
    k = 1.0 - c
    return np.array([[c + x * x * k, x * y * k - z * s, y * s + x * z * k],
                     [z * s + x * y * k, c + y * y * k, -x * s + y * z * k],
                     [-y * s + x * z * k, x * s + y * z * k, c + z * z * k]])


def crystal_frame(s1, omega, raw_xray, omega_axis, F):
    """Incident and diffracted vectors of each reflection in the crystal
    frame, (R(omega) F)^T v (ray_tracing.c:397-405), as unit vectors."""
    s1 = np.asarray(s1, dtype=np.float64).reshape(-1, 3)
    Fm = np.asarray(F, dtype=np.float64).reshape(3, 3)
    xs, ss = [], []
    for v, om in zip(s1, np.asarray(omega, dtype=np.float64).reshape(-1)):
        M = kp_rotation(omega_axis, om) @ Fm
        xs.append(M.T @ np.asarray(raw_xray, dtype=np.float64))
        ss.append(M.T @ v)
    x = np.array(xs).reshape(-1, 3)
    s = np.array(ss).reshape(-1, 3)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    return x, s


def harmonics(v):
    return np.stack([2.0 * v[:, 0] * v[:, 1], v[:, 0] ** 2 - v[:, 1] ** 2,
                     2.0 * v[:, 2] ** 2 - 1.0], axis=1)


def chamber(x, s):
    """Sides (+-1, 0 on the plane) of the planes where the tracer's discrete
    rules switch: sign x2, sign(|x1| - |x2|), sign(|x0| - |x1|), sign s2,
    sign(|s0| - |s1|), sign(|s1| - |s2|).  theta > 0 of dials_2_thetaphi_22
    (ray_tracing_basic.c:13-50) is the sign of the z component, which picks
    floor or round in cal_coord (ray_tracing_basic.c:428-438)."""
    a = np.abs
    return np.stack([np.sign(x[:, 2]), np.sign(a(x[:, 1]) - a(x[:, 2])),
                     np.sign(a(x[:, 0]) - a(x[:, 1])), np.sign(s[:, 2]),
                     np.sign(a(s[:, 0]) - a(s[:, 1])), np.sign(a(s[:, 1]) - a(s[:, 2]))], axis=1)


def resonance_inputs(V) -> np.ndarray:
    """Inputs [12:60] of feature set 2 from rows of FP32 components V[n][6]
    (x, then s), in FP32: for ray r and pair (0, 1), (0, 2), (1, 2), k = 3 r +
    pair, the ratio t = min(|v_i|, |v_j|) / max(|v_i|, |v_j|) (0 when both are
    0) at [k] and 2 |q t - rint(q t)| at [6 + 7 k + q - 2] for q = 2..8."""
    V = np.asarray(V, dtype=np.float32)[:, 0:6]
    out = np.empty((V.shape[0], 48), dtype=np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        for k in range(6):
            r, p = divmod(k, 3)
            i, j = _RISK_PAIRS[p]
            a = np.abs(V[:, 3 * r + i])
            b = np.abs(V[:, 3 * r + j])
            hi = np.maximum(a, b)
            t = np.where(hi > 0, np.minimum(a, b) / np.where(hi > 0, hi, np.float32(1)),
                         np.float32(0)).astype(np.float32)
            out[:, k] = t
            for q in range(2, 9):
                u = np.float32(q) * t
                out[:, 6 + 7 * k + q - 2] = np.float32(2) * np.abs(u - np.rint(u))
    return out


def features(s1, omega, raw_xray, omega_axis, F, axis_c=None, feature_set: int = 0):
    """Network inputs [R][60] and the envelope invariants [R][2] (axial, cos2t);
    axis_c defaults to the unit F^T omega_axis.  feature_set 0: second
    harmonics, 1: chamber signs, 2: chamber signs and the resonance inputs
    of the FP32 components (zero in [12:60] for sets 0 and 1)."""
    x, s = crystal_frame(s1, omega, raw_xray, omega_axis, F)
    if axis_c is None:
        u = np.asarray(F, dtype=np.float64).reshape(3, 3).T @ np.asarray(omega_axis, float)
        axis_c = u / np.linalg.norm(u)
    extra = (np.concatenate([harmonics(x), harmonics(s)], axis=1) if feature_set == 0
             else chamber(x, s))
    comps = np.concatenate([x, s], axis=1)
    res = (resonance_inputs(comps.astype(np.float32)).astype(np.float64) if feature_set == 2
           else np.zeros((len(x), 48)))
    X = np.concatenate([comps, extra, res], axis=1)
    inv = np.stack([x @ np.asarray(axis_c, float), np.sum(x * s, axis=1)], axis=1)
    return X, inv


def invariants(s1, omega, raw_xray, omega_axis, F, axis, e1, e2):
    """(axial, two_theta, alpha, beta) per reflection: alpha = azimuth of x
    about the axis in the (e1, e2) basis, beta = azimuth of s about x in the
    basis (axis x x, x (axis x x))."""
    x, s = crystal_frame(s1, omega, raw_xray, omega_axis, F)
    u = np.asarray(axis, dtype=np.float64)
    c = np.sum(x * s, axis=1)
    tth = np.arccos(np.clip(c, -1.0, 1.0))
    alpha = np.arctan2(x @ np.asarray(e2, float), x @ np.asarray(e1, float))
    f1 = np.cross(u[None, :], x)
    f1 /= np.linalg.norm(f1, axis=1, keepdims=True)
    f2 = np.cross(x, f1)
    p = s - c[:, None] * x
    beta = np.arctan2(np.sum(p * f2, axis=1), np.sum(p * f1, axis=1))
    return np.stack([x @ u, tth, alpha, beta], axis=1)


def band_mask(x, s, bands):
    """True for the reflections inside some resonance band (ray, i, j, ratio,
    width): the band's ray (0 incident x, 1 diffracted s) has
    |min(|v_i|, |v_j|) / max(|v_i|, |v_j|) - ratio| < width (0 / 0 counts as
    ratio 0)."""
    n = len(x)
    out = np.zeros(n, dtype=bool)
    for k in range(n):
        for ray, i, j, ratio, width in bands:
            v = s[k] if ray else x[k]
            a, b = abs(v[int(i)]), abs(v[int(j)])
            hi, lo = max(a, b), min(a, b)
            t = lo / hi if hi > 0 else 0.0
            if abs(t - ratio) < width:
                out[k] = True
                break
    return out


def cell_index(inv, two_theta_range, cells):
    """Trust-grid cell of each reflection: alpha and beta over [-pi, pi],
    two_theta over its range, bins clamped to the grid."""
    def bin_(v, lo, hi, n):
        k = np.floor((v - lo) / (hi - lo) * n) if hi > lo else np.zeros_like(v)
        return np.clip(k, 0, n - 1).astype(np.int64)
    ia = bin_(inv[:, 2], -np.pi, np.pi, cells[0])
    it = bin_(inv[:, 1], two_theta_range[0], two_theta_range[1], cells[1])
    ib = bin_(inv[:, 3], -np.pi, np.pi, cells[2])
    return (ia * cells[1] + it) * cells[2] + ib


# ------------------------------------------------------------ risk trees
_RISK_PAIRS = ((0, 1), (0, 2), (1, 2))


def risk_features(X) -> np.ndarray:
    """The 60 resonance features of rows of network inputs X[n][>= 6] (x = X[:, 0:3],
    s = X[:, 3:6]) in FP32, the operation chain of mlp.hip risk_features:
    the six components, per ray and pair (i, j) the ratio
    t = min(|v_i|, |v_j|) / max(|v_i|, |v_j|), and |q t - rint(q t)| * RN(1/q)
    for q = 1..8."""
    X = np.asarray(X, dtype=np.float32)
    n = X.shape[0]
    out = np.empty((n, 60), dtype=np.float32)
    out[:, 0:6] = X[:, 0:6]
    with np.errstate(invalid="ignore", divide="ignore"):
        for k in range(6):
            r, p = divmod(k, 3)
            i, j = _RISK_PAIRS[p]
            a = np.abs(X[:, 3 * r + i])
            b = np.abs(X[:, 3 * r + j])
            hi = np.maximum(a, b)
            t = np.where(hi > 0, np.minimum(a, b) / np.where(hi > 0, hi, np.float32(1)),
                         np.float32(0)).astype(np.float32)
            out[:, 6 + k] = t
            for q in range(1, 9):
                u = np.float32(q) * t
                inv_q = np.float32(1) / np.float32(q)
                out[:, 12 + 8 * k + q - 1] = np.abs(u - np.rint(u)) * inv_q
    return out


def risk_scores(trees: dict, F) -> np.ndarray:
    """FP32 scores of complete binary trees (feat/thr [T][2^D - 1], leaf
    [T][2^D], base, depth) over feature rows F[n][60]: from node 0,
    node = 2 node + 1 + (f > thr), the leaf node - (2^D - 1); base plus the
    leaves in tree order."""
    F = np.asarray(F, dtype=np.float32)
    D = int(trees["depth"])
    n_int = (1 << D) - 1
    feat = np.asarray(trees["feat"]).reshape(-1, n_int)
    thr = np.asarray(trees["thr"], dtype=np.float32).reshape(-1, n_int)
    leaf = np.asarray(trees["leaf"], dtype=np.float32).reshape(-1, n_int + 1)
    rows = np.arange(F.shape[0])
    s = np.full(F.shape[0], np.float32(trees["base"]), dtype=np.float32)
    for t in range(feat.shape[0]):
        j = np.zeros(F.shape[0], dtype=np.int64)
        for _ in range(D):
            f = feat[t][j]
            j = 2 * j + 1 + (F[rows, f] > thr[t][j])
        s = (s + leaf[t][j - n_int]).astype(np.float32)
    return s
