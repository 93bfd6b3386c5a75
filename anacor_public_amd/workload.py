"""Deterministic synthetic workloads for the absorption-correction ray tracer.

The reference takes a segmented tomography volume (``.npy``, labels 0 vacuum,
1 liquor/solvent, 2 loop, 3 crystal, 4 bubble; ``main_lite.py:93,111``), a
crystal-voxel list from ``generate_sampling`` (``utils/utils_rt.py:27-253``)
and DIALS reflections (``s1`` and the rotation angle, ``utils/utils_mp.py:151-164``).
No dataset is available offline, so this module synthesises inputs of the
same shape and dtype:

* a solvent cylinder along the volume's z axis (the rotation axis),
* a flat loop slab perpendicular to x through the volume centre, clipped to
  the cylinder,
* a crystal ellipsoid at a seeded random orientation at the centre,

with precedence crystal > loop > solvent > vacuum.  The beam travels along
-X of AnACor's frame, which is ``raw_xray = (0, -1, 0)`` in the DIALS frame
the C-ABI consumes, the goniometer axis is ``(-1, 0, 0)`` and ``F = I``
(as ``verification/main_verification.py:238-239`` sets up its phantoms).

Random numbers come from a counter-based SplitMix64 stream so the inputs do not
depend on numpy's generator implementations.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Sequence, Tuple

import numpy as np

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, stream: int, n: int) -> np.ndarray:
    """n uint64 values of the SplitMix64 output function applied to the
    states ``(seed + i) * G + stream * D`` (mod 2**64), i = 1 .. n, with G and
    D the constants below: the seed and the counter share the increment G,
    so seed s + k on a stream repeats seed s's values k counters later; sets
    meant to be independent use different streams."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        z = base + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles uniform in [0, 1) (53-bit mantissa)."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# --------------------------------------------------------------------- volume

@dataclasses.dataclass(frozen=True)
class VolumeSpec:
    shape: Tuple[int, int, int]          # (z, y, x)
    cyl_radius: float                    # solvent cylinder radius (voxels)
    slab: int                            # loop slab thickness (voxels)
    semi_axes: Tuple[float, float, float]  # crystal ellipsoid semi-axes (voxels)
    seed: int


def random_rotation(seed: int) -> np.ndarray:
    """Uniform random rotation matrix from a seeded unit quaternion."""
    u = uniform01(seed, 101, 3)
    q1 = math.sqrt(1 - u[0]) * math.sin(2 * math.pi * u[1])
    q2 = math.sqrt(1 - u[0]) * math.cos(2 * math.pi * u[1])
    q3 = math.sqrt(u[0]) * math.sin(2 * math.pi * u[2])
    q0 = math.sqrt(u[0]) * math.cos(2 * math.pi * u[2])
    w, x, y, z = q0, q1, q2, q3
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def make_volume(spec: VolumeSpec) -> np.ndarray:
    """Build the uint8 label volume [z][y][x] described by ``spec``.

    Built slice by slice so 1000x1000x800 volumes need no large temporaries.
    """
    Z, Y, X = spec.shape
    cz, cy, cx = Z // 2, Y // 2, X // 2
    vol = np.zeros((Z, Y, X), dtype=np.uint8)
    yy = (np.arange(Y, dtype=np.float64) - cy)[:, None]
    xx = (np.arange(X, dtype=np.float64) - cx)[None, :]
    cyl = (yy * yy + xx * xx) <= spec.cyl_radius ** 2
    x0 = cx - spec.slab // 2
    slab_cols = np.zeros(X, dtype=bool)
    slab_cols[max(x0, 0):max(x0 + spec.slab, 0)] = True
    loop = cyl & slab_cols[None, :]
    base = np.where(loop, np.uint8(2), np.where(cyl, np.uint8(1), np.uint8(0))).astype(np.uint8)

    # crystal: |diag(1/a) R^T (p - c)|^2 <= 1
    Rm = random_rotation(spec.seed)
    A = np.diag(1.0 / np.asarray(spec.semi_axes, dtype=np.float64)) @ Rm.T  # 3x3 acting on (z,y,x)
    rmax = max(spec.semi_axes)
    z_lo, z_hi = max(0, int(math.floor(cz - rmax)) - 1), min(Z, int(math.ceil(cz + rmax)) + 2)
    y_lo, y_hi = max(0, int(math.floor(cy - rmax)) - 1), min(Y, int(math.ceil(cy + rmax)) + 2)
    x_lo, x_hi = max(0, int(math.floor(cx - rmax)) - 1), min(X, int(math.ceil(cx + rmax)) + 2)
    py = (np.arange(y_lo, y_hi, dtype=np.float64) - cy)[:, None]
    px = (np.arange(x_lo, x_hi, dtype=np.float64) - cx)[None, :]
    for z in range(Z):
        vol[z] = base
        if z_lo <= z < z_hi:
            pz = float(z - cz)
            q0 = A[0, 0] * pz + A[0, 1] * py + A[0, 2] * px
            q1 = A[1, 0] * pz + A[1, 1] * py + A[1, 2] * px
            q2 = A[2, 0] * pz + A[2, 1] * py + A[2, 2] * px
            inside = (q0 * q0 + q1 * q1 + q2 * q2) <= 1.0
            sub = vol[z, y_lo:y_hi, x_lo:x_hi]
            sub[inside] = 3
    return vol


# ------------------------------------------------------------------- sampling

def crystal_voxels(vol: np.ndarray, label: int = 3) -> np.ndarray:
    """All voxels of ``label`` in np.where order, int64 [N,3] (z,y,x), as
    ``generate_sampling`` builds ``crystal_coordinate`` (utils_rt.py:53-54)."""
    zz, yy, xx = np.where(vol == label)
    return np.stack((zz, yy, xx), axis=1).astype(np.int64)


def sample_even(coords: np.ndarray, n: int) -> np.ndarray:
    """The reference's ``even`` rule (utils_rt.py:69-82):
    ``np.linspace(0, len, num=n, endpoint=False, dtype=int)`` indices."""
    idx = np.linspace(0, len(coords), num=n, endpoint=False, dtype=int)
    return np.ascontiguousarray(coords[idx])


def sample_stride(coords: np.ndarray, stride: int) -> np.ndarray:
    """Crystal voxels on the lattice z, y, x == 0 (mod stride)."""
    if stride <= 1:
        return np.ascontiguousarray(coords)
    keep = np.all(coords % stride == 0, axis=1)
    return np.ascontiguousarray(coords[keep])


def sample_uniform(coords: np.ndarray, n: int, seed: int) -> np.ndarray:
    """Seeded uniform draw of n voxels without replacement, kept in the
    original (np.where) order."""
    if n >= len(coords):
        return np.ascontiguousarray(coords)
    keys = splitmix64(seed, 202, len(coords))
    pick = np.sort(np.argpartition(keys, n)[:n])
    return np.ascontiguousarray(coords[pick])


# ---------------------------------------------------------------- reflections

RAW_XRAY = np.array([0.0, -1.0, 0.0])
OMEGA_AXIS = np.array([-1.0, 0.0, 0.0])
F_IDENTITY = np.eye(3).reshape(-1).astype(np.float64)


def make_reflections(R: int, seed: int, two_theta_deg=(5.0, 60.0),
                     stream: int = 303) -> Tuple[np.ndarray, np.ndarray]:
    """R reflections: rotation angle omega ~ U[0, 2pi); diffracted unit vector
    s1 at 2theta ~ U[lo, hi] degrees from the beam with uniform azimuth.
    Returns (s1 float64 [R,3] DIALS frame, omega float64 [R]).

    Seeds that differ by k on one stream give the same values shifted by k
    counters, so sets that must be independent take different streams."""
    u = uniform01(seed, stream, 3 * R).reshape(R, 3)
    omega = 2.0 * np.pi * u[:, 0]
    lo, hi = np.deg2rad(two_theta_deg[0]), np.deg2rad(two_theta_deg[1])
    tth = lo + (hi - lo) * u[:, 1]
    az = 2.0 * np.pi * u[:, 2]
    b = RAW_XRAY
    e1 = np.array([1.0, 0.0, 0.0])
    e2 = np.array([0.0, 0.0, 1.0])
    s1 = (np.cos(tth)[:, None] * b[None, :]
          + np.sin(tth)[:, None] * (np.cos(az)[:, None] * e1[None, :] + np.sin(az)[:, None] * e2[None, :]))
    return np.ascontiguousarray(s1), np.ascontiguousarray(omega)


# -------------------------------------------------------------------- configs

MU_BY_LABEL = (0.0, 0.0123, 0.0117, 0.0145)   # vacuum, solvent, loop, crystal (per voxel length)
# coefficients in the reference's [li, lo, cr, bu] order (main_lite.py:207)
COEFFICIENTS = np.array([0.0123, 0.0117, 0.0145, 0.0], dtype=np.float64)
VOXEL_SIZE = np.array([1.0, 1.0, 1.0], dtype=np.float64)


@dataclasses.dataclass
class Workload:
    name: str
    volume: np.ndarray
    coords: np.ndarray
    s1: np.ndarray
    omega: np.ndarray
    raw_xray: np.ndarray = dataclasses.field(default_factory=lambda: RAW_XRAY.copy())
    omega_axis: np.ndarray = dataclasses.field(default_factory=lambda: OMEGA_AXIS.copy())
    F: np.ndarray = dataclasses.field(default_factory=lambda: F_IDENTITY.copy())
    voxel_size: np.ndarray = dataclasses.field(default_factory=lambda: VOXEL_SIZE.copy())
    coefficients: np.ndarray = dataclasses.field(default_factory=lambda: COEFFICIENTS.copy())
    full_iteration: int = 0

    @property
    def shape(self) -> np.ndarray:
        return np.array(self.volume.shape, dtype=np.int64)

    def subset(self, R: int) -> "Workload":
        """The same workload restricted to its first R reflections."""
        return dataclasses.replace(self, s1=np.ascontiguousarray(self.s1[:R]),
                                   omega=np.ascontiguousarray(self.omega[:R]))


def small_config(seed: int = 1, R: int = 256) -> Workload:
    """64^3 volume, cylinder r=28, slab 3, crystal 12/9/7, all crystal voxels,
    256 reflections."""
    vol = make_volume(VolumeSpec((64, 64, 64), 28.0, 3, (12.0, 9.0, 7.0), seed))
    s1, om = make_reflections(R, seed)
    return Workload("plain-64^3-all-voxels", vol, crystal_voxels(vol), s1, om)


def single_gpu_config(seed: int = 2, R: int = 20000, stride: int = 2) -> Workload:
    """256^3 volume, the small construction scaled x4 (cylinder r=112, slab 12,
    crystal 48/36/28), crystal voxels on a stride-2 lattice, 20,000 reflections."""
    vol = make_volume(VolumeSpec((256, 256, 256), 112.0, 12, (48.0, 36.0, 28.0), seed))
    s1, om = make_reflections(R, seed)
    return Workload("plain-256^3-stride2", vol, sample_stride(crystal_voxels(vol), stride), s1, om)


def large_config(seed: int = 4, R: int = 100000, n_voxels: int = 100000) -> Workload:
    """1000x1000x800 (x, y, z) volume, cylinder r=420, slab 10, crystal
    300/220/150, 100k crystal voxels drawn uniformly, 100k reflections."""
    vol = make_volume(VolumeSpec((800, 1000, 1000), 420.0, 10, (300.0, 220.0, 150.0), seed))
    s1, om = make_reflections(R, seed)
    return Workload("plain-1000x1000x800-100k", vol,
                    sample_uniform(crystal_voxels(vol), n_voxels, seed), s1, om)


def batched_specs(seed: int = 3, n_samples: int = 8):
    """Eight uneven samples: volume edge 96..400 (x, y) and 96..300 (z),
    semi-axes 10..120, 5k..50k reflections (log-uniform), stride 1..4."""
    u = uniform01(seed, 404, 8 * n_samples).reshape(n_samples, 8)
    out = []
    for i in range(n_samples):
        edge = int(96 + round(u[i, 0] * (400 - 96)))
        zed = int(96 + round(u[i, 1] * (300 - 96)))
        zed = min(zed, edge)
        shape = (zed, edge, edge)
        amax = 0.3 * min(zed, edge)
        a = 10.0 + u[i, 2] * (min(120.0, amax) - 10.0)
        semi = (a, a * (0.6 + 0.3 * u[i, 3]), a * (0.45 + 0.3 * u[i, 4]))
        R = int(round(math.exp(math.log(5000) + u[i, 5] * (math.log(50000) - math.log(5000)))))
        stride = 1 + int(u[i, 6] * 4) if u[i, 6] < 1.0 else 4
        out.append(dict(shape=shape, semi_axes=semi, R=R, stride=stride,
                        cyl_radius=0.44 * edge, slab=max(3, edge // 21), seed=seed * 100 + i))
    return out


def batched_config(seed: int = 3, n_samples: int = 8, R_scale: float = 1.0):
    """The eight uneven samples as Workloads (R scaled by R_scale)."""
    wls = []
    for sp in batched_specs(seed, n_samples):
        vol = make_volume(VolumeSpec(sp["shape"], sp["cyl_radius"], sp["slab"], sp["semi_axes"], sp["seed"]))
        R = max(1, int(sp["R"] * R_scale))
        s1, om = make_reflections(R, sp["seed"])
        wls.append(Workload(f"batched-{sp['shape']}", vol,
                            sample_stride(crystal_voxels(vol), sp["stride"]), s1, om))
    return wls


@dataclasses.dataclass
class SurrogateWorkload:
    """One sample of the batched set with its training, held-out and
    inference reflection sets (disjoint seeds)."""
    sample: Workload               # volume, voxels; s1/omega = the training set
    heldout_s1: np.ndarray
    heldout_omega: np.ndarray
    infer_s1: np.ndarray
    infer_omega: np.ndarray
    infer_outside: np.ndarray      # bool: drawn outside the training 2-theta range
    audit_s1: np.ndarray = None    # a larger held-out set (its own stream) for the error tail
    audit_omega: np.ndarray = None
    risk_s1: np.ndarray = None     # reflections the risk trees learn the network's errors on
    risk_omega: np.ndarray = None


SURROGATE_SAMPLE = 2   # of batched_specs(): 100^3, crystal 23.2/19.0/10.9, all 20,089 voxels


def surrogate_config(seed: int = 5, n_train: int = 3_200_000, n_heldout: int = 10_000,
                     n_infer: int = 1_000_000, outside_fraction: float = 0.01,
                     sample: int = SURROGATE_SAMPLE, n_audit: int = 0,
                     n_risk: int = 800_000) -> SurrogateWorkload:
    """Batched sample `sample` (100^3, every crystal voxel traced) with
    n_train training reflections (2theta 5-60 deg), n_heldout held-out ones
    from the same distribution, n_infer inference reflections of which
    outside_fraction lie at 2theta 60-75 deg, outside the training range,
    n_audit more held-out reflections for the tail of the error, and n_risk
    reflections of the training distribution for the risk trees."""
    sp = batched_specs()[sample]
    vol = make_volume(VolumeSpec(sp["shape"], sp["cyl_radius"], sp["slab"], sp["semi_axes"],
                                 sp["seed"]))
    coords = sample_stride(crystal_voxels(vol), sp["stride"])
    # one stream per set: on one stream, nearby seeds only shift the values
    s1, om = make_reflections(n_train, seed * 1000 + 1)
    hs1, hom = make_reflections(n_heldout, seed * 1000 + 2, stream=311)
    n_out = int(round(n_infer * outside_fraction))
    is1, iom = make_reflections(n_infer - n_out, seed * 1000 + 3, stream=313)
    os1, oom = make_reflections(n_out, seed * 1000 + 4, two_theta_deg=(60.0, 75.0), stream=317)
    outside = np.zeros(n_infer, dtype=bool)        # n_out seeded positions
    if n_out:
        outside[np.argpartition(splitmix64(seed, 505, n_infer), n_out - 1)[:n_out]] = True
    all_s1 = np.empty((n_infer, 3))
    all_om = np.empty(n_infer)
    all_s1[outside], all_om[outside] = os1, oom
    all_s1[~outside], all_om[~outside] = is1, iom
    wl = Workload(f"surrogate-batched{sample}-{'x'.join(str(s) for s in sp['shape'])}", vol,
                  coords, s1, om)
    as1, aom = (make_reflections(n_audit, seed * 1000 + 6, stream=331) if n_audit
                else (np.zeros((0, 3)), np.zeros(0)))
    rs1, rom = (make_reflections(n_risk, seed * 1000 + 7, stream=337) if n_risk
                else (None, None))
    return SurrogateWorkload(wl, hs1, hom, np.ascontiguousarray(all_s1),
                             np.ascontiguousarray(all_om), outside, as1, aom, rs1, rom)
