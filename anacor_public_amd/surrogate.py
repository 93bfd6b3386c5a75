"""Learned surrogate of the absorption factor for one sample.

An MLP (60 -> 256 x 4 -> 1, SiLU; BF16-MFMA kernels with FP32 accumulation,
csrc/mlp.hip) maps features of the incident and diffracted directions of a
reflection in the crystal frame (by default the components, the sides of the
planes where the tracer's rounding switches, and kinks at small rational
ratios of two components) to log A_h.  It is trained on the GPU on absorption factors
the ray tracer computes for the same sample, answers only inside the
envelope it was trained on (the training 2-theta range and the training
goniometer's rotation circle), and leaves every other reflection to the
tracer, which also judges its error.

    with Tracer(volume, coords, voxel_size, coefficients) as t:
        sur = Surrogate(t, raw_xray, omega_axis, F)
        report = sur.fit(s1_train, omega_train)          # traces, then trains
        A, used = sur.predict(s1, omega)                 # float64[R], bool[R]

The reference has no surrogate; its nearest analogue is the gridding
approximation (src/gridding_interpolation.c:98-161, utils/utils_gridding.py).
There is no CPU path: the network and the fallback both run on the GPU.
"""
from __future__ import annotations

import ctypes as ct
import time
from typing import Optional

import numpy as np

from . import _lib

_vp = ct.c_void_p


def _p(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    """The GPU network (anacor_mlp_* of include/anacor_mi355x.h)."""

    n_params = _lib.MLP_PARAMS
    n_inputs = _lib.MLP_INPUTS

    def __init__(self, seed: int = 0, device: int = 0):
        self._lib = _lib.load()
        h = self._lib.anacor_mlp_create(int(seed), int(device))
        if not h:
            code, msg = _lib.last_error(self._lib)
            raise _lib.AnacorError(code, msg)
        self._h = ct.c_void_p(h)

    @property
    def handle(self):
        if not getattr(self, "_h", None):
            raise ValueError("MLP is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.anacor_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        _lib.check(rc, self._lib)

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self._lib.anacor_mlp_get_params(self.handle, _p(out), out.size))
        return out

    def set_params(self, params) -> None:
        p = _f32(params).reshape(-1)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        self._check(self._lib.anacor_mlp_set_params(self.handle, _p(p), p.size))

    def set_scale(self, mean: float, std: float) -> None:
        self._check(self._lib.anacor_mlp_set_scale(self.handle, float(mean), float(std)))

    def set_features(self, feature_set: int) -> None:
        self._check(self._lib.anacor_mlp_set_features(self.handle, int(feature_set)))

    def set_envelope(self, envelope, trust=None) -> None:
        """envelope: _lib.Envelope or None (removes it); trust: uint8 per cell or None."""
        if envelope is None:
            self._check(self._lib.anacor_mlp_set_envelope(self.handle, None, None))
            return
        n = int(np.prod(envelope.cells[:]))
        t = None
        if trust is not None:
            t = np.ascontiguousarray(trust, dtype=np.uint8).reshape(-1)
            if t.size != n:
                raise ValueError(f"trust table has {t.size} cells, envelope {n}")
        self._check(self._lib.anacor_mlp_set_envelope(self.handle, ct.byref(envelope),
                                                      _p(t) if t is not None else None))
        self._trust_keep = t

    def set_bands(self, bands) -> None:
        """bands: sequence of (ray, i, j, ratio, width) (see fit_bands); empty removes them."""
        n = len(bands)
        if n > _lib.MLP_MAX_BANDS:
            raise ValueError(f"{n} bands, at most {_lib.MLP_MAX_BANDS}")
        arr = (_lib.Band * max(n, 1))()
        for k, (ray, i, j, ratio, width) in enumerate(bands):
            arr[k].ray, arr[k].i, arr[k].j = int(ray), int(i), int(j)
            arr[k].ratio, arr[k].width = float(ratio), float(width)
        self._check(self._lib.anacor_mlp_set_bands(self.handle, arr, n))

    def set_risk(self, trees: Optional[dict], cut: float = float("inf")) -> None:
        """Risk trees (risk_fit's dict) with rows scoring >= cut sent to the
        tracer; None removes them."""
        if trees is None:
            self._check(self._lib.anacor_mlp_set_risk(self.handle, 0, 0, None, None, None,
                                                      0.0, 0.0))
            return
        feat = np.ascontiguousarray(trees["feat"], dtype=np.uint8)
        thr = _f32(trees["thr"])
        leaf = _f32(trees["leaf"])
        D = int(trees["depth"])
        T = feat.shape[0]
        if feat.shape != (T, (1 << D) - 1) or thr.shape != feat.shape or \
                leaf.shape != (T, 1 << D):
            raise ValueError("risk trees: inconsistent array shapes")
        self._check(self._lib.anacor_mlp_set_risk(self.handle, T, D, _p(feat), _p(thr),
                                                  _p(leaf), float(trees["base"]), float(cut)))

    def risk(self, X, scores: bool = True):
        """(features float32[n][60], scores float32[n] or None) of rows of
        network inputs, on the GPU; scores need trees set."""
        X = _f32(X).reshape(-1, _lib.MLP_INPUTS)
        Fr = np.empty((len(X), _lib.RISK_FEATURES), dtype=np.float32)
        sc = np.empty(len(X), dtype=np.float32) if scores else None
        self._check(self._lib.anacor_mlp_risk(self.handle, _p(X), len(X), _p(Fr),
                                              _p(sc) if sc is not None else None))
        return Fr, sc

    def features(self, s1, omega, raw_xray, omega_axis, F, cells: bool = False,
                 flags: bool = False):
        """X float32[R][MLP_INPUTS], invariants float64[R][4] (axial, two_theta, alpha,
        beta) and, with an envelope set, cell indices and in-envelope flags."""
        s1 = _f64(s1).reshape(-1, 3)
        om = _f64(omega).reshape(-1)
        R = len(om)
        if len(s1) != R:
            raise ValueError("s1 and omega lengths differ")
        X = np.empty((R, _lib.MLP_INPUTS), dtype=np.float32)
        inv = np.empty((R, 4), dtype=np.float64)
        cl = np.empty(R, dtype=np.int32) if cells else None
        fl = np.empty(R, dtype=np.uint8) if flags else None
        self._check(self._lib.anacor_mlp_features(
            self.handle, _p(s1), _p(om), R, _p(_f64(raw_xray)), _p(_f64(omega_axis)),
            _p(_f64(F)), _p(X), _p(inv), _p(cl) if cl is not None else None,
            _p(fl) if fl is not None else None))
        return X, inv, cl, (fl.astype(bool) if fl is not None else None)

    def forward(self, X) -> np.ndarray:
        X = _f32(X).reshape(-1, _lib.MLP_INPUTS)
        y = np.empty(len(X), dtype=np.float32)
        self._check(self._lib.anacor_mlp_forward(self.handle, _p(X), len(X), _p(y)))
        return y

    def gradients(self, X, t):
        """(gradient float32[n_params], loss) of the mean squared error."""
        X = _f32(X).reshape(-1, _lib.MLP_INPUTS)
        t = _f32(t).reshape(-1)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = ct.c_float(0.0)
        self._check(self._lib.anacor_mlp_gradients(self.handle, _p(X), _p(t), len(X), _p(g),
                                                   ct.byref(loss)))
        return g, float(loss.value)

    def train(self, X, t, epochs: int, batch: int = 1024, lr: float = 1e-3,
              lr_min: float = 1e-5, seed: int = 0) -> np.ndarray:
        X = _f32(X).reshape(-1, _lib.MLP_INPUTS)
        t = _f32(t).reshape(-1)
        losses = np.zeros(int(epochs), dtype=np.float32)
        self._check(self._lib.anacor_mlp_train(self.handle, _p(X), _p(t), len(X), int(epochs),
                                               int(batch), float(lr), float(lr_min), int(seed),
                                               _p(losses)))
        return losses

    def predict(self, s1, omega, raw_xray, omega_axis, F):
        s1 = _f64(s1).reshape(-1, 3)
        om = _f64(omega).reshape(-1)
        R = len(om)
        A = np.empty(R, dtype=np.float64)
        flag = np.empty(R, dtype=np.uint8)
        self._check(self._lib.anacor_mlp_predict(
            self.handle, _p(s1), _p(om), R, _p(_f64(raw_xray)), _p(_f64(omega_axis)),
            _p(_f64(F)), _p(A), _p(flag)))
        return A, flag.astype(bool)

    def stats(self) -> dict:
        st = _lib.MlpStats()
        self._check(self._lib.anacor_mlp_stats(self.handle, ct.byref(st)))
        return {"mlp_ms": st.mlp_ms, "total_ms": st.total_ms, "rows": st.rows,
                "in_envelope": st.in_envelope}


BAND_PAIRS = ((0, 1), (0, 2), (1, 2))
BAND_RATIOS = (0.0, 0.25, 1.0 / 3.0, 0.5, 2.0 / 3.0, 0.75, 1.0)


def _component_ratio(v: np.ndarray, i: int, j: int) -> np.ndarray:
    a, b = np.abs(v[:, i]), np.abs(v[:, j])
    hi = np.maximum(a, b)
    return np.where(hi > 0, np.minimum(a, b) / np.where(hi > 0, hi, 1.0), 0.0)


def fit_bands(x: np.ndarray, s: np.ndarray, err: np.ndarray, target: float,
              max_width: float = 0.06, grow: float = 1.1, max_bands: int = 32,
              max_step: float = 0.05):
    """Resonance bands that cover every training reflection whose relative
    error exceeds target.

    Near a small rational ratio t = min(|v_i|, |v_j|) / max(|v_i|, |v_j|) of
    two components of a ray's crystal-frame direction v (t = 0, 1/4, 1/3,
    1/2, 2/3, 3/4, 1), the voxel steps of all crystal voxels fall into step
    together and the traced absorption factor oscillates faster than the
    training set samples it; away from them the network's error is small.
    Greedy: take the worst uncovered reflection, widen the (ray, i, j, ratio)
    band that covers it at the least additional training fallback (a band
    wider than max_width, or one step that would take more than max_step of
    the training set, is not a candidate), repeat.  Widths are grown by
    `grow` as a margin for reflections the training set does not hold.

    x, s: [n, 3] unit vectors (incident, diffracted); err: [n] relative
    errors.  Returns (bands as (ray, i, j, ratio, width) tuples, indices of
    reflections no band can cover)."""
    cands = [(r, i, j, q) for r in (0, 1) for (i, j) in BAND_PAIRS for q in BAND_RATIOS]
    dist = np.stack([np.abs(_component_ratio(x if r == 0 else s, i, j) - q)
                     for r, i, j, q in cands])
    n = len(err)
    width = np.zeros(len(cands))
    served = np.ones(n, dtype=bool)
    left = []
    while True:
        bad = served & (err > target)
        if not bad.any():
            break
        k = int(np.argmax(np.where(bad, err, -1.0)))
        need = dist[:, k] * (1.0 + 1e-6) + 1e-12
        ok = (need <= max_width) & (dist[:, k] > 0)
        if np.count_nonzero(width) >= max_bands:
            ok &= width > 0
        best, pick = None, -1
        for c in np.flatnonzero(ok):
            add = np.count_nonzero(served & (dist[c] < max(width[c], need[c])))
            if add > max_step * n:
                continue
            if best is None or add < best:
                best, pick = add, int(c)
        if pick < 0:
            served[k] = False
            left.append(k)
            continue
        width[pick] = max(width[pick], need[pick])
        served &= ~(dist[pick] < width[pick])
    bands = [(*cands[c], float(width[c] * grow)) for c in np.flatnonzero(width)]
    return bands, np.asarray(left, dtype=np.int64)


def in_bands(bands, x: np.ndarray, s: np.ndarray) -> np.ndarray:
    """bool[n]: the reflections the bands send to the tracer."""
    out = np.zeros(len(x), dtype=bool)
    for r, i, j, q, w in bands:
        out |= np.abs(_component_ratio(x if r == 0 else s, i, j) - q) < w
    return out


def risk_fit(features: np.ndarray, labels: np.ndarray, weights: Optional[np.ndarray] = None,
             n_trees: int = 200, depth: int = 6, learning_rate: float = 0.1, l2: float = 1.0,
             min_hess: float = 1e-3, n_bins: int = 256, threads: int = 0) -> dict:
    """Risk trees fitted on the host by the library (anacor_risk_fit):
    Newton-boosted complete trees of the logistic loss of labels (0/1) over
    feature rows [n][F].  Returns feat / thr [n_trees][2^depth - 1], leaf
    [n_trees][2^depth], base and depth."""
    lib = _lib.load()
    Fm = _f32(features)
    n, nf = Fm.shape
    y = np.ascontiguousarray(labels, dtype=np.uint8).reshape(-1)
    if len(y) != n:
        raise ValueError("one label per feature row")
    w = None if weights is None else _f32(weights).reshape(-1)
    n_int = (1 << depth) - 1
    feat = np.zeros((n_trees, n_int), dtype=np.uint8)
    thr = np.zeros((n_trees, n_int), dtype=np.float32)
    leaf = np.zeros((n_trees, n_int + 1), dtype=np.float32)
    base = ct.c_float()
    _lib.check(lib.anacor_risk_fit(_p(Fm), _p(y), None if w is None else _p(w), n, nf, n_trees,
                                   depth, learning_rate, l2, min_hess, n_bins, threads,
                                   _p(feat), _p(thr), _p(leaf), ct.byref(base)), lib)
    return {"feat": feat, "thr": thr, "leaf": leaf, "base": float(base.value),
            "depth": int(depth)}


def crystal_axis(omega_axis, F) -> np.ndarray:
    """The goniometer axis in the crystal frame, unit F^T omega_axis (the
    rotation about omega_axis leaves it fixed)."""
    u = np.asarray(F, dtype=np.float64).reshape(3, 3).T @ np.asarray(omega_axis, dtype=np.float64)
    return u / np.linalg.norm(u)


def make_envelope(axis, axial_range, two_theta_range, cells=(64, 8, 32)) -> "_lib.Envelope":
    """Envelope over a training set's axial and two-theta ranges with a trust
    grid of `cells` (alpha, two_theta, beta)."""
    u = np.asarray(axis, dtype=np.float64)
    u = u / np.linalg.norm(u)
    r = np.zeros(3)
    r[int(np.argmin(np.abs(u)))] = 1.0
    e1 = np.cross(u, r)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u, e1)
    env = _lib.Envelope()
    for i in range(3):
        env.axis[i], env.e1[i], env.e2[i] = float(u[i]), float(e1[i]), float(e2[i])
        env.cells[i] = int(cells[i])
    env.axial[0], env.axial[1] = float(axial_range[0]), float(axial_range[1])
    env.two_theta[0], env.two_theta[1] = float(two_theta_range[0]), float(two_theta_range[1])
    return env


class Surrogate:
    """MLP surrogate of one sample's A_h with the tracer as fallback and judge."""

    def __init__(self, tracer, raw_xray, omega_axis, F, seed: int = 0, device: int = 0,
                 feature_set: int = _lib.MLP_FEATURES_RESONANCE):
        self.tracer = tracer
        self.raw_xray = _f64(raw_xray).reshape(3)
        self.omega_axis = _f64(omega_axis).reshape(3)
        self.F = _f64(F).reshape(9)
        self.mlp = MLP(seed=seed, device=device)
        self.feature_set = int(feature_set)
        self.mlp.set_features(self.feature_set)
        self.envelope: Optional[_lib.Envelope] = None
        self.trust: Optional[np.ndarray] = None
        self.bands: list = []
        self.risk: Optional[dict] = None
        self.risk_cut = float("inf")
        self.bounded = False
        self.y_mean = 0.0
        self.y_std = 1.0

    def close(self):
        self.mlp.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_feature_set(self, feature_set: int) -> None:
        self.mlp.set_features(feature_set)
        self.feature_set = int(feature_set)

    def _geom(self):
        return self.raw_xray, self.omega_axis, self.F

    def trace(self, s1, omega) -> np.ndarray:
        return self.tracer.run(s1, omega, *self._geom())

    def fit(self, s1, omega, epochs: int = 600, batch: int = 1024, lr: float = 1e-3,
            lr_min: float = 1e-6, seed: int = 0, targets: Optional[np.ndarray] = None,
            cells=(64, 8, 32), tau: Optional[float] = None, min_count: int = 4,
            axial_tol: float = 1e-9, bound: Optional[float] = 1e-3, trust: str = "trees",
            band_target: float = 0.85, band_grow: float = 1.1, risk_trees: int = 200,
            risk_depth: int = 6, risk_folds: int = 5, risk_label: float = 0.5,
            risk_cut: float = 0.9, risk_quantile: float = 0.01,
            risk_keep: float = 0.3, risk_s1=None, risk_omega=None) -> dict:
        """Trace the training reflections (unless targets are given), train
        with Adam, and set the envelope: the training axial and two-theta
        ranges, and trust only in the (alpha, two_theta, beta) cells holding at
        least min_count training reflections (and, with tau, whose training
        relative errors are all at most tau).

        With a bound the surrogate starts in bounded mode, in which the
        reflections where the network's error is likely to exceed the bound
        go to the tracer:
        * trust="trees" (default): risk trees (risk_fit) over the resonance
          features learn where the training errors exceed risk_label * bound;
          each training row is scored by trees fitted without its fold
          (risk_folds folds, negatives subsampled to risk_keep and weighted
          back), and the cut is the risk_quantile quantile of those scores
          over the rows whose error exceeds risk_cut * bound; the trees set
          on the GPU are refitted on every row.  With risk_s1 / risk_omega
          (reflections of the training distribution that the network does not
          train on) the trees learn from the network's errors on those, which
          are not shrunk by fitting; else from the training rows' errors.
        * trust="bands": resonance bands (fit_bands, target band_target *
          bound, widths grown by band_grow).
        set_bounded(False) lets the network answer every row of its envelope."""
        t0 = time.perf_counter()
        A = self.trace(s1, omega) if targets is None else _f64(targets)
        t_trace = time.perf_counter() - t0
        ok = np.isfinite(A) & (A > 0)
        if not np.all(ok):
            raise ValueError(f"{int((~ok).sum())} training reflections have no finite A_h > 0")
        if trust not in ("trees", "bands"):
            raise ValueError(f"trust must be 'trees' or 'bands', not {trust!r}")
        y = np.log(A)
        self.y_mean = float(y.mean())
        self.y_std = float(max(y.std(), 1e-12))
        self.mlp.set_scale(self.y_mean, self.y_std)
        self.mlp.set_envelope(None)
        self.mlp.set_bands([])
        self.mlp.set_risk(None)
        self.bands, self.risk, self.risk_cut = [], None, float("inf")
        _, inv, _, _ = self.mlp.features(s1, omega, *self._geom())
        env = make_envelope(crystal_axis(self.omega_axis, self.F),
                            (inv[:, 0].min() - axial_tol, inv[:, 0].max() + axial_tol),
                            (inv[:, 1].min(), inv[:, 1].max()), cells)
        self.mlp.set_envelope(env)
        X, inv, cell, _ = self.mlp.features(s1, omega, *self._geom(), cells=True)
        t1 = time.perf_counter()
        losses = self.mlp.train(X, (y - self.y_mean) / self.y_std, epochs=epochs, batch=batch,
                                lr=lr, lr_min=lr_min, seed=seed)
        t_train = time.perf_counter() - t1
        pred = np.exp(self.mlp.forward(X).astype(np.float64) * self.y_std + self.y_mean)
        res = np.abs(pred - A) / A
        n_cells = int(np.prod(cells))
        count = np.bincount(cell, minlength=n_cells)
        worst = np.zeros(n_cells)
        np.maximum.at(worst, cell, res)
        ok = count >= min_count
        if tau is not None:
            ok &= worst <= tau
        rep_trust = {}
        left = np.zeros(0, dtype=np.int64)
        t2 = time.perf_counter()
        if bound is not None and trust == "bands":
            x64, s64 = X[:, 0:3].astype(np.float64), X[:, 3:6].astype(np.float64)
            self.bands, left = fit_bands(x64, s64, res, band_target * bound, grow=band_grow)
            ok[cell[left]] = False   # errors no band covers: their cells go to the tracer
            band_out = (in_bands(self.bands, x64, s64) if self.bands
                        else np.zeros(len(y), dtype=bool))
            rep_trust = {"bands": [list(b) for b in self.bands],
                         "band_train_fallback": float(band_out.mean()),
                         "band_uncovered": int(len(left))}
        elif bound is not None:
            Xk, rk, n_risk = X, res, 0
            if risk_s1 is not None:
                A_r = self.trace(risk_s1, risk_omega)
                Xk, _, _, _ = self.mlp.features(risk_s1, risk_omega, *self._geom())
                pred_r = np.exp(self.mlp.forward(Xk).astype(np.float64) * self.y_std +
                                self.y_mean)
                rk = np.abs(pred_r - A_r) / A_r
                n_risk = len(rk)
            rep_trust = self._fit_risk(Xk, rk, bound, seed, risk_trees, risk_depth, risk_folds,
                                       risk_label, risk_cut, risk_quantile, risk_keep)
            rep_trust["risk_rows"] = n_risk if n_risk else int(len(res))
            rep_trust["risk_set"] = "separate" if n_risk else "training"
        t_trust = time.perf_counter() - t2
        trust_tab = ok.astype(np.uint8)
        self.mlp.set_envelope(env, trust_tab)
        self.envelope, self.trust = env, trust_tab
        self.set_bounded(bound is not None)
        served = trust_tab[cell].astype(bool)
        return {"n_train": int(len(y)), "trace_s": t_trace, "train_s": t_train,
                "trust_fit_s": t_trust, "epochs": int(epochs), "batch": int(batch), "lr": lr,
                "lr_min": lr_min, "final_loss": float(losses[-1]), "losses": losses,
                "y_mean": self.y_mean, "y_std": self.y_std,
                "two_theta_deg": (float(np.degrees(env.two_theta[0])),
                                  float(np.degrees(env.two_theta[1]))),
                "train_rel_max": float(res.max()), "train_rel_mean": float(res.mean()),
                "cells": list(cells), "tau": tau, "min_count": min_count,
                "trusted_cells": float(trust_tab.mean()), "train_served": float(served.mean()),
                "bound": bound, "trust": trust if bound is not None else None, **rep_trust}

    def _fit_risk(self, X, res, bound, seed, n_trees, depth, folds, label, cut_frac, quantile,
                  keep) -> dict:
        """Risk trees and their cut (see fit); returns the report entries."""
        Fr, _ = self.mlp.risk(X, scores=False)
        n = len(res)
        rng = np.random.default_rng([int(seed), 7919])
        fold = rng.integers(0, folds, size=n)
        drop = rng.random(n)                       # negative subsampling draw per row
        pos = res > label * bound

        def fit_on(rows):
            m = rows[pos[rows] | (drop[rows] < keep)]
            w = np.where(pos[m], 1.0, 1.0 / keep).astype(np.float32)
            return risk_fit(Fr[m], pos[m].astype(np.uint8), w, n_trees=n_trees, depth=depth)

        cv = np.empty(n, dtype=np.float32)
        all_rows = np.arange(n)
        for k in range(folds):
            self.mlp.set_risk(fit_on(all_rows[fold != k]))
            _, cv[fold == k] = self.mlp.risk(X[fold == k])
        bad = np.sort(cv[res > cut_frac * bound])
        cut = float(bad[int(quantile * len(bad))]) if len(bad) else float("inf")
        self.risk = fit_on(all_rows)
        self.risk_cut = cut
        self.mlp.set_risk(None)
        return {"risk_trees": int(n_trees), "risk_depth": int(depth), "risk_cut": cut,
                "risk_positives": int(pos.sum()),
                "risk_train_fallback": float(np.mean(cv >= cut)),
                "risk_train_missed": int(np.sum((cv < cut) & (res > bound)))}

    def set_bounded(self, bounded: bool) -> None:
        """Bounded mode: the risk trees (or resonance bands) send their
        reflections to the tracer.  Fast mode: the network answers every
        reflection of its envelope."""
        on = bool(bounded)
        self.mlp.set_bands(self.bands if on else [])
        self.mlp.set_risk(self.risk if on and self.risk is not None else None, self.risk_cut)
        self.bounded = on and (bool(self.bands) or self.risk is not None)

    def predict(self, s1, omega, fallback: bool = True):
        """(A float64[R], used bool[R]): the network's A_h where it may answer
        (used), the tracer's elsewhere (NaN there if fallback is False)."""
        if self.envelope is None:
            raise ValueError("the surrogate has not been fitted")
        A, used = self.mlp.predict(s1, omega, *self._geom())
        if fallback and not np.all(used):
            out = np.flatnonzero(~used)
            s1 = _f64(s1).reshape(-1, 3)
            A[out] = self.trace(s1[out], _f64(omega).reshape(-1)[out])
        return A, used

    def state(self) -> dict:
        """Everything needed to rebuild the fitted surrogate (numpy arrays)."""
        e = self.envelope
        return {"params": self.mlp.get_params(), "y_mean": self.y_mean, "y_std": self.y_std,
                "axis": np.array(e.axis[:]), "e1": np.array(e.e1[:]), "e2": np.array(e.e2[:]),
                "axial": np.array(e.axial[:]), "two_theta": np.array(e.two_theta[:]),
                "cells": np.array(e.cells[:]), "trust": self.trust,
                "feature_set": self.feature_set,
                "bands": np.array(self.bands, dtype=np.float64).reshape(-1, 5),
                "bounded": self.bounded, **self._risk_state()}

    def _risk_state(self) -> dict:
        if self.risk is None:
            return {"risk_depth": 0}
        r = self.risk
        return {"risk_feat": r["feat"], "risk_thr": r["thr"], "risk_leaf": r["leaf"],
                "risk_base": r["base"], "risk_depth": r["depth"], "risk_cut": self.risk_cut}

    def load_state(self, st: dict) -> None:
        self.set_feature_set(int(st["feature_set"]))
        self.mlp.set_params(st["params"])
        self.y_mean, self.y_std = float(st["y_mean"]), float(st["y_std"])
        self.mlp.set_scale(self.y_mean, self.y_std)
        env = _lib.Envelope()
        for i in range(3):
            env.axis[i], env.e1[i], env.e2[i] = (float(st["axis"][i]), float(st["e1"][i]),
                                                 float(st["e2"][i]))
            env.cells[i] = int(st["cells"][i])
        for i in range(2):
            env.axial[i] = float(st["axial"][i])
            env.two_theta[i] = float(st["two_theta"][i])
        self.envelope = env
        self.trust = np.ascontiguousarray(st["trust"], dtype=np.uint8)
        self.mlp.set_envelope(env, self.trust)
        self.bands = [(int(b[0]), int(b[1]), int(b[2]), float(b[3]), float(b[4]))
                      for b in np.asarray(st.get("bands", np.zeros((0, 5)))).reshape(-1, 5)]
        D = int(st.get("risk_depth", 0))
        if D > 0:
            self.risk = {"feat": np.asarray(st["risk_feat"], dtype=np.uint8),
                         "thr": np.asarray(st["risk_thr"], dtype=np.float32),
                         "leaf": np.asarray(st["risk_leaf"], dtype=np.float32),
                         "base": float(st["risk_base"]), "depth": D}
            self.risk_cut = float(st["risk_cut"])
        else:
            self.risk, self.risk_cut = None, float("inf")
        self.set_bounded(bool(st.get("bounded", False)))
