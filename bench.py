"""Throughput of the MI355X absorption-correction ray tracer.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config single|small|large|batched]

A step is one pass of the plain path over the workload's whole reflection
batch: host FP64 setup of every reflection, its upload, the trace kernel,
the per-reflection reduction and the download of R absorption factors.  The
label volumes and crystal-voxel lists are resident in HBM before timing
starts.  Prints one JSON line (absorption factors per second) with the trace
kernel's roofline position and the reference C/OpenMP path timed on a bounded
sample of the same workload on this host's cores.

The reference's iterative-bisection traversal (ib_test) is timed on the same
workload in a "bisection" record of the default run, with the reference's
ib_test on this host's cores beside it.

The learned surrogate (config "surrogate", and a "surrogate" record in the
default run) is trained on 3.2M reflections the tracer computes for batched
sample 2 (100^3, 20,089 crystal voxels), its risk trees on the network's
errors on 800k more, judged by the tracer on 10k held-out reflections and a
200k audit set, and timed on 1M reflections of which 1% lie outside its
training 2-theta range and go to the tracer.  Its "named_config" record is
the same surrogate in the configuration the workload names (200k training
reflections, 12 inputs, lr 1e-3), fitted and timed on the same sets.

Workloads (anacor_public_amd/workload.py):
  single   256^3 volume, 25,323 crystal voxels (stride 2), 20,000 reflections (default)
  small    64^3 volume, all 3,169 crystal voxels, 256 reflections
  large    1000x1000x800 volume, 100,000 voxels, 100,000 reflections
  batched  8 uneven samples (100^3 .. 383x383x135), 194,281 reflections,
           per-sample trace launches behind the host setup, autotuned work shapes
  surrogate  MLP surrogate of batched sample 2, 1M inferences with tracer fallback
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
MLP_FLOPS_PER_ROW = 2 * (60 * 256 + 3 * 256 * 256 + 256)   # 2 x 212,224 MACs
# L2<->fabric bytes per voxel step of trace_kernel, by exact work shape
# (workload, layout, voxel block, tile, reflection group; tile 0 is the
# library's automatic choice): FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (the
# factors of a 1 GiB copy) over the launch's voxel steps.  The 256^3 rows are
# profiles/trace_kernel_work-shapes_plain-256cube.md (one launch per
# autotuner candidate), the 1000x1000x800 row
# profiles/trace_kernel_brick2_large-1000x1000x800.md.  A shape not listed
# here reports traffic null.
_HEAD = "plain-256^3-stride2"
_LARGE10K = "plain-1000x1000x800-100k-first10000"
TRAFFIC_BYTES_PER_STEP = {
    (_HEAD, 0, (1, 2, 64), 0, 1): 0.2101,
    (_HEAD, 0, (1, 2, 64), 512, 1): 0.0963,
    (_HEAD, 0, (1, 2, 64), 2048, 1): 0.4307,
    (_HEAD, 0, (1, 1, 128), 0, 1): 0.2105,
    (_HEAD, 0, (1, 1, 128), 512, 1): 0.0964,
    (_HEAD, 0, (1, 1, 128), 2048, 1): 0.4304,
    (_HEAD, 0, (2, 2, 32), 0, 1): 0.2105,
    (_HEAD, 0, (2, 2, 32), 512, 1): 0.0964,
    (_HEAD, 0, (2, 2, 32), 2048, 1): 0.4305,
    (_HEAD, 0, (1, 4, 32), 0, 1): 0.2216,
    (_HEAD, 0, (1, 4, 32), 512, 1): 0.1060,
    (_HEAD, 0, (1, 4, 32), 2048, 1): 0.4548,
    (_HEAD, 0, (2, 4, 16), 0, 1): 0.2245,
    (_HEAD, 0, (2, 4, 16), 512, 1): 0.1095,
    (_HEAD, 0, (2, 4, 16), 2048, 1): 0.4619,
    (_HEAD, 0, (0, 0, 0), 0, 1): 0.2040,
    (_HEAD, 0, (0, 0, 0), 512, 1): 0.1132,
    (_HEAD, 0, (0, 0, 0), 2048, 1): 0.4797,
    (_HEAD, 2, (1, 2, 64), 0, 2048): 0.0015,
    (_HEAD, 2, (1, 2, 64), 512, 2048): 0.0013,
    (_HEAD, 2, (1, 2, 64), 2048, 2048): 0.0007,
    (_HEAD, 2, (1, 1, 128), 0, 2048): 0.0015,
    (_HEAD, 2, (1, 1, 128), 512, 2048): 0.0013,
    (_HEAD, 2, (1, 1, 128), 2048, 2048): 0.0007,
    (_HEAD, 2, (2, 2, 32), 0, 2048): 0.0015,
    (_HEAD, 2, (2, 2, 32), 512, 2048): 0.0013,
    (_HEAD, 2, (2, 2, 32), 2048, 2048): 0.0007,
    (_HEAD, 2, (1, 4, 32), 0, 2048): 0.0015,
    (_HEAD, 2, (1, 4, 32), 512, 2048): 0.0013,
    (_HEAD, 2, (1, 4, 32), 2048, 2048): 0.0007,
    (_HEAD, 2, (2, 4, 16), 0, 2048): 0.0015,
    (_HEAD, 2, (2, 4, 16), 512, 2048): 0.0013,
    (_HEAD, 2, (2, 4, 16), 2048, 2048): 0.0007,
    (_HEAD, 2, (0, 0, 0), 0, 2048): 0.0015,
    (_HEAD, 2, (0, 0, 0), 512, 2048): 0.0013,
    (_HEAD, 2, (0, 0, 0), 2048, 2048): 0.0007,
    (_LARGE10K, 2, (0, 0, 0), 0, 2048): 0.6935,
}
# The same per label read of ib_kernel (2-bit bricks), over 2.207e10 reads,
# profiles/ib_kernel_bisection-256cube.md.
IB_TRAFFIC_BYTES_PER_READ = {"plain-256^3-stride2": 0.0024}
METRIC = "absorption factors/sec (rays/s) at 1 GPU, plain vs surrogate at stated rel. error"
SURROGATE_BOUND = 1e-3      # the stated relative error bound of the surrogate
SURROGATE_BATCH = 4096      # Adam mini-batch of the 3.2M-pair fit (learning rate below)
SURROGATE_LR = 2e-3
LARGE_REFLECTIONS = 10000   # the default run's 1000x1000x800 record: its first 10k reflections


def cpu_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(wls, budget_s: float = 15.0):
    """The reference's own C path (oracle/_ref) on the first reflections of
    every sample of the workload, OpenMP over reflections on this host's
    cores; A_h/s on that sample (reflections are independent, so A_h/s does
    not depend on R).  Falls back to the FP64 restatement (bit-identical
    results) when the reference build is absent."""
    from oracle import oracle as O
    cores = cpu_cores()
    kind = "reference"
    try:
        runner = O.Reference()
        run = lambda w, R: runner.overall(w, n_threads=cores, R=R)  # noqa: E731
    except (FileNotFoundError, OSError):
        kind = "port"
        runner = O.Oracle()
        run = lambda w, R: runner.trace(w, n_threads=cores, R=R)["result"]  # noqa: E731
    # calibrate on a few reflections of each sample, then share the budget
    per = []
    for w in wls:
        R0 = min(len(w.omega), cores)
        t0 = time.perf_counter()
        run(w, R0)
        per.append((time.perf_counter() - t0) / R0)
    share = budget_s / len(wls)
    total_R, total_t, outs, desc = 0, 0.0, [], []
    for w, p in zip(wls, per):
        R = int(min(len(w.omega), max(cores, share / max(p, 1e-9))))
        R = max(1, (R // cores) * cores) if R >= cores else R
        t0 = time.perf_counter()
        outs.append(run(w, R))
        total_t += time.perf_counter() - t0
        total_R += R
        desc.append(f"{R}/{len(w.omega)}")
    return dict(value=total_R / total_t, unit="absorption factors/s", cores=cores, kind=kind,
                sample=f"first reflections of each sample ({', '.join(desc)}), "
                       f"{total_t:.1f} s"), outs


def calibrate_copy(nbytes: int = 1 << 30) -> None:
    """A plain device-to-device copy of known size (HIP's blit kernel), run
    under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    a, b = ct.c_void_p(), ct.c_void_p()
    assert hip.hipMalloc(ct.byref(a), ct.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ct.byref(b), ct.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ct.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ct.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def volume_bytes(shape, layout: int) -> int:
    """Bytes of the label volume the kernel reads: one per voxel in the raw
    layout (with its one-voxel shell of zeros), 128 (2-bit) or 256 (4-bit) per
    8x8x8 brick otherwise."""
    Z, Y, X = (int(v) for v in shape)
    if layout == 0:
        return (Z + 2) * (Y + 2) * (X + 2)
    bricks = -(-Z // 8) * -(-Y // 8) * -(-X // 8)
    return bricks * (128 if layout == 2 else 256)


def gather_ceiling(nbytes: int, achieved_gbs: float) -> dict:
    """The device's one-byte gather rate with no arithmetic between loads
    (anacor_probe_gather), on a working set of the next power of two above
    the volume: every lane its own pseudo-random byte ("lanes"), and a wave's
    64 lanes one 64-byte row ("row", the shape of the walk's gathers).  At one
    byte per load, in GB/s, with the kernel's algorithmic rate as a fraction
    of each."""
    import ctypes as ct
    from anacor_public_amd import _lib
    lib = _lib.load()
    k = max(6, (int(nbytes) - 1).bit_length())
    out = {"working_set_bytes": 1 << k}
    for name, pattern in (("lanes", _lib.GATHER_LANES), ("row", _lib.GATHER_ROW)):
        rate = ct.c_double()
        _lib.check(lib.anacor_probe_gather(0, k, pattern, 1 << 34, ct.byref(rate)), lib)
        gbs = rate.value / 1e9
        out[f"{name}_GBps"] = gbs
        out[f"frac_of_{name}"] = achieved_gbs / gbs if gbs > 0 else None
    return out


def rel_stats(a, ref):
    r = np.abs(np.asarray(a) - ref) / ref
    return {"max": float(r.max()), "p99_9": float(np.quantile(r, 0.999)),
            "p99": float(np.quantile(r, 0.99)), "mean": float(r.mean())}


def gridding_baseline(tracer, sw, geo, A_true, steps: int, n_theta: int = 360,
                      n_phi: int = 180) -> dict:
    """The gridding path (absorption maps of every voxel on a 1-degree
    theta/phi grid, built on the GPU, then look-ups) on the surrogate's
    sample: error on the held-out set and rate on the inference set."""
    from anacor_public_amd.gridding import Gridding
    t0 = time.perf_counter()
    out = {"grid": [n_theta, n_phi]}
    with Gridding(tracer, n_theta, n_phi) as grid:
        out["build_s"] = time.perf_counter() - t0
        out["map_kernel_ms"] = grid.stats()["build_ms"]
        for name, method in (("nearest", Gridding.NEAREST), ("bilinear", Gridding.BILINEAR)):
            A = grid.run(sw.heldout_s1, sw.heldout_omega, *geo, method=method)
            err = rel_stats(A, A_true)
            grid.run(sw.infer_s1, sw.infer_omega, *geo, method=method)
            t1 = time.perf_counter()
            for _ in range(steps):
                grid.run(sw.infer_s1, sw.infer_omega, *geo, method=method)
            rate = len(sw.infer_omega) * steps / (time.perf_counter() - t1)
            out[name] = {"value": rate, "heldout_rel_error": err,
                         "lookup_kernel_ms": grid.stats()["lookup_ms"]}
            log(f"# gridding {n_theta}x{n_phi} {name}: {rate:.3e} A_h/s, held-out {err}")
    return out


def surrogate_mode(sur, sw, A_true, A_audit, plain_rate, steps, warmup, n_infer,
                   bounded: bool, label: str = "") -> dict:
    """One operating point of a fitted surrogate: held-out and audit errors
    against the tracer, then `steps` timed predict calls on the inference set
    (tracer fallback included)."""
    sur.set_bounded(bounded)
    A_sur, used = sur.predict(sw.heldout_s1, sw.heldout_omega)
    err = rel_stats(A_sur[used], A_true[used])
    A_a, used_a = sur.predict(sw.audit_s1, sw.audit_omega, fallback=False)
    ra = np.abs(A_a[used_a] - A_audit[used_a]) / A_audit[used_a]
    audit = {"reflections": int(len(used_a)), "answered": int(used_a.sum()),
             "rel_error": rel_stats(A_a[used_a], A_audit[used_a]),
             "above_bound": int(np.sum(ra > SURROGATE_BOUND))}
    for _ in range(warmup):
        sur.predict(sw.infer_s1, sw.infer_omega)
    mlp_ms, fb = [], 0
    t2 = time.perf_counter()
    for _ in range(steps):
        _, u = sur.predict(sw.infer_s1, sw.infer_omega)
        mlp_ms.append(sur.mlp.stats()["mlp_ms"])
        fb = int((~u).sum())
    elapsed = time.perf_counter() - t2
    value = n_infer * steps / elapsed
    out = {"value": value, "ms_per_step": 1e3 * elapsed / steps,
           "heldout_answered": int(used.sum()), "heldout_rel_error": err,
           "audit": audit, "fallback_reflections": fb,
           "fallback_fraction": fb / n_infer,
           "speedup_vs_plain_gpu": value / plain_rate,
           "mlp_ms": float(np.mean(mlp_ms))}
    name = ("bounded" if bounded else "fast") + label
    log(f"# surrogate {name}: held-out ({out['heldout_answered']}/{len(used)} answered) "
        f"{err}; audit {audit}; {value:.3e} A_h/s ({out['speedup_vs_plain_gpu']:.1f}x "
        f"the plain GPU path's {plain_rate:.3e}), fallback {fb}")
    return out


def surrogate_named_config(tr, sw, geo, A_true, A_audit, plain_rate, steps, warmup,
                           n_infer, n_train: int = 200_000, epochs: int = 600) -> dict:
    """The surrogate in the configuration the workload names: 200k training
    pairs (the first of the same training stream), 12 inputs (the two unit
    vectors and the six chamber signs), 600 epochs, risk trees learned on the
    training rows' errors; judged and timed on the same held-out, audit and
    inference sets."""
    from anacor_public_amd import _lib
    from anacor_public_amd.surrogate import Surrogate
    wl = sw.sample
    with Surrogate(tr, *geo, feature_set=_lib.MLP_FEATURES_CHAMBER) as sur:
        rep = sur.fit(wl.s1[:n_train], wl.omega[:n_train], epochs=epochs,
                      bound=SURROGATE_BOUND)
        log(f"# surrogate (named configuration) fit: train {rep['train_s']:.1f}s ({epochs} "
            f"epochs on {n_train}), risk trees {rep['trust_fit_s']:.1f}s")
        args = (sur, sw, A_true, A_audit, plain_rate, steps, warmup, n_infer)
        fast = surrogate_mode(*args, False, " (named configuration)")
        bounded = surrogate_mode(*args, True, " (named configuration)")
    keys = ("value", "ms_per_step", "heldout_answered", "heldout_rel_error", "audit",
            "fallback_reflections", "fallback_fraction", "speedup_vs_plain_gpu")
    return {"training_pairs": n_train, "network_inputs": "12: components, chamber signs",
            "epochs": epochs, "train_s": rep["train_s"],
            "trust": {k: rep[k] for k in ("trust", "risk_cut", "risk_set", "risk_rows",
                                          "risk_train_fallback")},
            "bounded": {k: bounded[k] for k in keys}, "fast": {k: fast[k] for k in keys}}


def run_surrogate(steps: int, warmup: int, n_infer: int = 1_000_000, epochs: int = 300,
                  cpu_budget: float = 0.0):
    """Fit the surrogate (risk trees on a separate risk set of 800k
    reflections), judge it on the held-out set (10k reflections) and
    on an audit set (200k more, its own stream) for the tail of the error,
    time 1M inferences (features, envelope and risk-tree tests, MLP on MFMA,
    tracer fallback) and the plain GPU path on the same sample.  Two
    operating points: bounded (risk trees fitted on the training errors send
    the reflections where the network's error is likely to exceed the 1e-3
    bound to the tracer) and fast (the network answers every reflection
    inside the training range).  Returns the surrogate record, bounded mode
    at its top level."""
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    from anacor_public_amd.surrogate import Surrogate
    t0 = time.perf_counter()
    sw = W.surrogate_config(n_infer=n_infer, n_audit=200_000)
    wl = sw.sample
    geo = (wl.raw_xray, wl.omega_axis, wl.F)
    log(f"# surrogate workload {wl.name}: N={len(wl.coords)}, train {len(wl.omega)}, "
        f"held-out {len(sw.heldout_omega)}, inference {n_infer} "
        f"({int(sw.infer_outside.sum())} outside), built in {time.perf_counter() - t0:.1f}s")
    rec = {}
    with Tracer.for_workload(wl) as tr, Surrogate(tr, *geo) as sur:
        # the plain GPU path on the same sample (20k of the inference reflections)
        ps1, pom = sw.infer_s1[:20000], sw.infer_omega[:20000]
        tr.run(ps1, pom, *geo)
        t1 = time.perf_counter()
        for _ in range(3):
            tr.run(ps1, pom, *geo)
        plain_rate = 3 * len(pom) / (time.perf_counter() - t1)
        rep = sur.fit(wl.s1, wl.omega, epochs=epochs, batch=SURROGATE_BATCH, lr=SURROGATE_LR,
                      bound=SURROGATE_BOUND, risk_s1=sw.risk_s1, risk_omega=sw.risk_omega)
        log(f"# surrogate fit: trace {rep['trace_s']:.2f}s, train {rep['train_s']:.1f}s "
            f"({epochs} epochs), final loss {rep['final_loss']:.3e}, risk trees "
            f"{rep['trust_fit_s']:.1f}s (cross-fitted fallback "
            f"{rep['risk_train_fallback'] * 100:.1f}% of the {rep['risk_rows']} risk-set "
            f"reflections)")
        A_true = tr.run(sw.heldout_s1, sw.heldout_omega, *geo)
        A_audit = tr.run(sw.audit_s1, sw.audit_omega, *geo)

        args = (sur, sw, A_true, A_audit, plain_rate, steps, warmup, n_infer)
        fast = surrogate_mode(*args, False)
        bounded = surrogate_mode(*args, True)
        mms = bounded["mlp_ms"]
        achieved = n_infer * MLP_FLOPS_PER_ROW / (mms * 1e-3) / 1e12
        rec = {
            "value": bounded["value"], "unit": "absorption factors/s",
            "ms_per_step": bounded["ms_per_step"], "mode": "bounded",
            "bound": SURROGATE_BOUND, "dtype": "bf16", "workload": wl.name,
            "n_voxels": int(len(wl.coords)), "inferences": n_infer,
            "fallback_reflections": bounded["fallback_reflections"],
            "fallback_fraction": bounded["fallback_fraction"],
            "network_inputs": "60: components, chamber signs, resonance kinks",
            "training_pairs": int(len(wl.omega)), "risk_pairs": int(rep["risk_rows"]),
            "epochs": epochs, "batch": SURROGATE_BATCH, "lr": SURROGATE_LR,
            "train_s": rep["train_s"], "train_trace_s": rep["trace_s"],
            "trust": {k: rep[k] for k in ("trust", "risk_trees", "risk_depth", "risk_cut",
                                          "risk_set", "risk_rows", "risk_train_fallback",
                                          "risk_train_missed")},
            "trust_fit_s": rep["trust_fit_s"],
            "heldout": int(len(sw.heldout_omega)),
            "heldout_answered": bounded["heldout_answered"],
            "heldout_rel_error": bounded["heldout_rel_error"],
            "audit": bounded["audit"],
            "plain_gpu_same_sample": plain_rate,
            "speedup_vs_plain_gpu": bounded["speedup_vs_plain_gpu"],
            "mlp_only_rate": n_infer / (mms * 1e-3),
            "fast": {k: fast[k] for k in ("value", "ms_per_step", "heldout_answered",
                                          "heldout_rel_error", "audit", "fallback_reflections",
                                          "fallback_fraction", "speedup_vs_plain_gpu")},
            "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFLOPS,
                         "traffic": None, "kernel": "MLP forward (HIP events)",
                         "kernel_ms": mms, "flops_per_launch": n_infer * MLP_FLOPS_PER_ROW},
        }
        rec["named_config"] = surrogate_named_config(tr, sw, geo, A_true, A_audit, plain_rate,
                                                     steps, warmup, n_infer)
        # the reference's own approximation on the same sample: gridding maps
        rec["gridding_baseline"] = gridding_baseline(tr, sw, geo, A_true, steps)
        if cpu_budget > 0:
            cpu, _ = cpu_baseline([W.Workload(wl.name, wl.volume, wl.coords, sw.infer_s1[:4096],
                                              sw.infer_omega[:4096])], cpu_budget)
            rec["cpu_baseline"] = cpu
    return rec


def bisection_cpu_baseline(wl, num_cls: int, budget_s: float):
    """The reference's ib_test (oracle/_ref), one reflection per call as its
    driver makes them (mp_worker.py:509-514), on a thread per host core (the
    calls release the GIL); A_h/s on the first reflections of the workload.
    Falls back to the FP64 restatement when the reference build is absent."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    cores = cpu_cores()
    orc = O.Oracle()
    geo = orc.reflection_setup(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F)
    try:
        ref = O.Reference()
        table, keep = O.label_table(np.ascontiguousarray(wl.volume).view(np.int8))
        kind = "reference"

        def one(i):
            return ref.ib_test(wl, wl.coords, geo["s1"][i], geo["xray"][i], num_cls, table=table)
    except (FileNotFoundError, OSError):
        keep, kind = None, "port"

        def one(i):
            return orc.ib_test(wl, wl.coords, geo["s1"][i], geo["xray"][i], num_cls)[0]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as pool:
        first = list(pool.map(one, range(cores)))
    per = (time.perf_counter() - t0) / cores
    R = int(min(len(wl.omega), max(cores, budget_s / max(per, 1e-9))))
    R = max(cores, (R // cores) * cores)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as pool:
        out = np.array(list(pool.map(one, range(R))))
    dt = time.perf_counter() - t0
    del keep, first
    return dict(value=R / dt, unit="absorption factors/s", cores=cores, kind=kind,
                sample=f"first {R}/{len(wl.omega)} reflections, one ib_test call each, "
                       f"{dt:.1f} s"), out


def run_bisection(steps: int, warmup: int, cpu_budget: float, walk_value: float):
    """The reference's iterative-bisection traversal (ib_test) over the same
    256^3 workload: every reflection in one launch; num_cls from the volume's
    labels as main_lite.py:153 counts them."""
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer, num_classes
    wl = W.single_gpu_config()
    ncls = num_classes(wl.volume)
    R = len(wl.omega)
    with Tracer.for_workload(wl) as t:
        def call():
            return t.bisection(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F, ncls)
        for _ in range(warmup):
            call()
        kms, reads = [], 0
        t0 = time.perf_counter()
        for _ in range(steps):
            A = call()
            st = t.stats()
            kms.append(st["kernel_ms"])
            reads = st["voxel_steps"]
        elapsed = time.perf_counter() - t0
        t_layout = t.layout
        layout = {0: "raw", 2: "brick2", 4: "brick4"}.get(t_layout, str(t_layout))
        outside = st["outside_reads"]
    assert np.all(np.isfinite(A)) and np.all((A > 0) & (A <= 1))
    value = R * steps / elapsed
    k_ms = float(np.mean(kms))
    achieved = reads / (k_ms * 1e-3) / 1e9      # 1 algorithmic byte per label read
    rec = {"value": value, "unit": "absorption factors/s", "workload": wl.name,
           "num_cls": ncls, "ms_per_step": 1e3 * elapsed / steps, "kernel_ms": k_ms,
           "label_reads_per_launch": int(reads), "outside_reads": int(outside),
           "label_reads_per_s": reads / (k_ms * 1e-3), "volume_layout": layout,
           "speedup_vs_plain_walk": value / walk_value,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": (IB_TRAFFIC_BYTES_PER_READ[wl.name] * reads
                                    if wl.name in IB_TRAFFIC_BYTES_PER_READ and layout == "brick2"
                                    else None),
                        "kernel": "ib_kernel", "kernel_ms": k_ms, "bytes_per_launch": int(reads),
                        "gather_ceiling": gather_ceiling(
                            volume_bytes(wl.volume.shape, t_layout), achieved)}}
    log(f"# bisection: {value:.3e} A_h/s ({rec['speedup_vs_plain_walk']:.2f}x the walk), "
        f"kernel {k_ms:.2f} ms, {reads} label reads, layout {layout}")
    if cpu_budget > 0:
        cpu, cpu_out = bisection_cpu_baseline(wl, ncls, cpu_budget)
        cpu["max_rel_diff_vs_gpu"] = float(np.max(np.abs(A[:len(cpu_out)] - cpu_out) / cpu_out))
        rec["cpu_baseline"] = cpu
        rec["speedup_vs_cpu"] = value / cpu["value"]
        log(f"# bisection cpu baseline ({cpu['kind']}, {cpu['cores']} threads): "
            f"{cpu['value']:.2f} A_h/s, max rel diff vs GPU {cpu['max_rel_diff_vs_gpu']:.2e}")
    return rec


def traffic_for(name: str, layout: int, block, tile: int, group: int, vsteps: int):
    """L2<->fabric bytes of the trace launch from the calibrated per-step
    figure measured for this exact work shape (workload, layout, voxel
    block, tile, reflection group), or None without one."""
    key = (name, int(layout), tuple(int(b) for b in block), int(tile), int(group))
    per = TRAFFIC_BYTES_PER_STEP.get(key)
    return None if per is None else per * vsteps


def measure_plain(config: str, steps: int, warmup: int, cpu_budget: float,
                  reflections: int = 0, tile: int = 0, autotune_on: bool = True,
                  gather: bool = True) -> dict:
    """Time `steps` whole plain-path steps of one workload (volumes and voxel
    lists resident): host setup, upload, trace, reduction and download of
    every reflection's A_h.  Returns the record (value in A_h/s, roofline of
    the trace kernel, CPU baseline on a bounded sample of the same
    workload)."""
    from anacor_public_amd import autotune
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Batch, Tracer
    t0 = time.perf_counter()
    if config == "batched":
        wls = W.batched_config()
    else:
        wls = [{"single": W.single_gpu_config, "small": W.small_config,
                "large": W.large_config}[config]()]
    if reflections:
        wls = [w.subset(reflections) for w in wls]
    R = sum(len(w.omega) for w in wls)
    N = sum(len(w.coords) for w in wls)
    name = wls[0].name if len(wls) == 1 else f"batched-{len(wls)}-samples"
    log(f"# workload {name}: volumes {[w.volume.shape for w in wls]}, N={N}, R={R} "
        f"(built in {time.perf_counter() - t0:.1f}s)")
    extra = {}
    if config == "batched":
        runner = Batch(wls)
        if autotune_on:
            t1 = time.perf_counter()
            picks = autotune.tune_batch(runner, wls, trial=256, use_cache=False)
            extra["autotune"] = [{"block": p["block"], "tile": p["tile"]} for p in picks]
            log(f"# autotune ({time.perf_counter() - t1:.1f}s): {extra['autotune']}")

        def call():
            return runner.run([w.s1 for w in wls], [w.omega for w in wls], wls[0].raw_xray,
                              wls[0].omega_axis, wls[0].F)
    else:
        runner = Tracer.for_workload(wls[0])
        if tile:
            runner.set_tile(tile)
        elif autotune_on:
            w = wls[0]
            t1 = time.perf_counter()
            pick = autotune.tune(runner, w.s1, w.omega, w.raw_xray, w.omega_axis, w.F, w.coords,
                                 trial=512, layouts=autotune.LAYOUTS, use_cache=False)
            extra["autotune"] = {"layout": pick["layout"], "block": pick["block"],
                                 "tile": pick["tile"]}
            log(f"# autotune ({time.perf_counter() - t1:.1f}s): {extra['autotune']}")

        def call():
            return [runner.run_workload(wls[0])]

    for _ in range(warmup):
        call()
    kernel_ms, vsteps_l, host_ms = [], [], []
    t_start = time.perf_counter()
    for _ in range(steps):
        outs = call()
        st = runner.stats()
        kernel_ms.append(st["kernel_ms"])
        vsteps_l.append(st["voxel_steps"])
        host_ms.append(st["host_ms"])
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / steps
    value = R * steps / elapsed
    k_ms = float(np.mean(kernel_ms))
    vsteps = int(vsteps_l[-1])
    achieved = vsteps * 1.0 / (k_ms * 1e-3) / 1e9   # 1 algorithmic byte per voxel step
    extra["host_setup_ms"] = float(np.mean(host_ms))
    log(f"# {name}: {steps} steps, {ms_per_step:.2f} ms/step, trace kernel {k_ms:.2f} ms, "
        f"host setup {extra['host_setup_ms']:.2f} ms, {vsteps} voxel steps per launch")
    for o in outs:
        assert np.all(np.isfinite(o)) and np.all((o > 0) & (o <= 1))
    if config == "large":
        # the FP32 compensated mean against an FP64 GPU accumulation
        f64 = runner.run_workload(wls[0], f64_mean=True)
        extra["fp32_vs_fp64_max_rel"] = float(np.max(np.abs(outs[0] - f64) / f64))
        log(f"# FP32 compensated vs FP64 mean: max rel {extra['fp32_vs_fp64_max_rel']:.2e}")
    cpu = None
    if cpu_budget > 0:
        cpu, cpu_outs = cpu_baseline(wls, cpu_budget)
        rel = max(float(np.max(np.abs(o[:len(c)] - c) / c)) for o, c in zip(outs, cpu_outs))
        cpu["max_rel_diff_vs_gpu"] = rel
        log(f"# cpu baseline ({cpu['kind']}, {cpu['cores']} cores): {cpu['value']:.2f} A_h/s, "
            f"max rel diff vs GPU {rel:.2e}")
    layout = 0 if config == "batched" else runner.layout
    traffic, shape = None, None
    if config != "batched":
        shape = runner.work_shape()
        key_name = name if not reflections else f"{name}-first{reflections}"
        traffic = traffic_for(key_name, layout, shape["block"], shape["tile"], shape["group"],
                              vsteps)
        extra["work_shape"] = shape
    pairs = sum(len(w.coords) * len(w.omega) for w in wls)
    rec = {
        "value": value,
        "unit": "absorption factors/s",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "workload": name,
        "volumes": [list(w.volume.shape) for w in wls],
        "n_voxels": N,
        "reflections": R,
        "volume_layout": {0: "raw", 2: "brick2", 4: "brick4"}[layout],
        "rays_per_s": 2.0 * pairs / (ms_per_step * 1e-3),
        "voxel_steps_per_s": vsteps / (ms_per_step * 1e-3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("trace_kernel, one launch per sample" if config == "batched"
                                else "trace_kernel"),
                     "kernel_ms": k_ms, "bytes_per_launch": vsteps},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    if config == "batched":
        rec["graph"] = runner.last_graph
        rec["launch"] = ("from the host: clear, per sample its host setup, rule upload and "
                         "trace launch (each setup overlapping the trace before it), one "
                         "finish kernel")
        rec["roofline"]["kernel_ms_covers"] = ("the whole sequence, the first sample's host "
                                               "setup included")
        # the same sequence as one hipGraph (host setups as host nodes)
        runner.set_graph(True)
        call()                                   # capture and first launch
        g_steps = max(1, min(steps, 5))
        t_g = time.perf_counter()
        g_kms = []
        for _ in range(g_steps):
            g_outs = call()
            g_kms.append(runner.stats()["kernel_ms"])
        g_el = time.perf_counter() - t_g
        same = all(np.array_equal(a.view(np.uint64), b.view(np.uint64))
                   for a, b in zip(g_outs, outs))
        rec["as_graph"] = {"graph": runner.last_graph, "steps": g_steps,
                           "ms_per_step": 1e3 * g_el / g_steps, "value": R * g_steps / g_el,
                           "kernel_ms": float(np.mean(g_kms)), "results_identical": same}
        runner.set_graph(False)
        log(f"# {name} as one hipGraph: {rec['as_graph']['ms_per_step']:.2f} ms/step "
            f"(results identical: {same})")
    if config != "batched" and gather:
        rec["roofline"]["gather_ceiling"] = gather_ceiling(
            volume_bytes(wls[0].volume.shape, layout), achieved)
        gc = rec["roofline"]["gather_ceiling"]
        log(f"# gather ceiling at {gc['working_set_bytes']} B: lanes {gc['lanes_GBps']:.0f} GB/s, "
            f"row {gc['row_GBps']:.0f} GB/s; trace kernel at {achieved:.0f} GB/s")
    rec.update(extra)
    runner.close()
    return rec


def profile_shapes() -> None:
    """One trace launch per autotuner candidate (layout, voxel block, tile)
    of the 256^3 workload in a fixed order, printing one JSON line per launch
    so that the per-dispatch counters of a rocprofv3 --pmc run can be matched
    to work shapes (the figures of TRAFFIC_BYTES_PER_STEP)."""
    from anacor_public_amd import autotune
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    wl = W.single_gpu_config()
    with Tracer.for_workload(wl) as t:
        for lay in autotune.LAYOUTS:
            t.set_layout(lay)
            for blk in autotune.BLOCKS:
                t.set_voxel_block(*blk)
                for tile in autotune.TILES:
                    t.set_tile(tile)
                    t.run_workload(wl)
                    st, ws = t.stats(), t.work_shape()
                    print(json.dumps({"workload": wl.name, "layout": t.layout,
                                      "block": ws["block"], "tile": ws["tile"],
                                      "group": ws["group"], "voxel_steps": st["voxel_steps"],
                                      "kernel_ms": st["kernel_ms"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="single",
                    choices=["single", "small", "large", "batched", "surrogate"])
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate record of the default run")
    ap.add_argument("--no-bisection", action="store_true",
                    help="skip the iterative-bisection record of the default run")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched (eight samples) record of the default run")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 1000x1000x800 record of the default run")
    ap.add_argument("--large-reflections", type=int, default=LARGE_REFLECTIONS,
                    help="reflections of the default run's 1000x1000x800 record (its first R)")
    ap.add_argument("--epochs", type=int, default=300, help="surrogate training epochs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--reflections", type=int, default=0,
                    help="trace only the first R reflections of each sample (profiling runs)")
    ap.add_argument("--autotune", action="store_true",
                    help="time every (layout, voxel block, tile) candidate on a slice of the "
                         "reflections first and keep the fastest (default: the library's "
                         "automatic work shape, whose traffic figure is on record)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="the batched record's per-sample tuning off")
    ap.add_argument("--profile-shapes", action="store_true",
                    help="one trace launch per autotuner work shape of the 256^3 workload, "
                         "in a fixed order, one JSON line each (for rocprofv3 --pmc passes)")
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.profile_shapes:
        profile_shapes()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): reflections "
              "are independent and would be sharded across GPUs, but this benchmark runs and "
              "reports N=1.", file=sys.stderr)
        sys.exit(2)
    budget = 0.0 if args.no_cpu_baseline else args.cpu_budget

    if args.config == "surrogate":
        rec = run_surrogate(args.steps, args.warmup, epochs=args.epochs, cpu_budget=budget)
        cpu = rec.pop("cpu_baseline", None)
        out = {"metric": METRIC, "value": rec["value"], "unit": rec["unit"], "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": rec["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "bf16", "data": "synthetic (seeded four-material volume and reflections)",
               "config": {"workload": rec["workload"], "path": "surrogate",
                          "network": "MLP 12-256x4-1 SiLU, BF16 weights, FP32 accumulation",
                          "inferences": rec["inferences"],
                          "training_pairs": rec["training_pairs"]},
               "roofline": rec["roofline"], "cpu_baseline": cpu,
               "speedup_vs_cpu": (rec["value"] / cpu["value"]) if cpu else None,
               "surrogate": rec}
        print(json.dumps(out), flush=True)
        return

    rec = measure_plain(args.config, args.steps, args.warmup, budget,
                        reflections=args.reflections, tile=args.tile,
                        autotune_on=(args.autotune if args.config != "batched"
                                     else not args.no_autotune))
    head = {
        "metric": METRIC,
        "value": rec.pop("value"),
        "unit": rec.pop("unit"),
        "n_gpus": 1,
        "steps": rec.pop("steps"),
        "warmup": rec.pop("warmup"),
        "ms_per_step": rec.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded four-material volumes and reflections)",
        "config": {"workload": rec.pop("workload"), "volumes": rec.pop("volumes"),
                   "n_voxels": rec.pop("n_voxels"), "reflections": rec.pop("reflections"),
                   "path": "plain", "mean": "fp32 exp + compensated sum",
                   "volume_layout": rec.pop("volume_layout")},
        "roofline": rec.pop("roofline"),
        "cpu_baseline": rec.pop("cpu_baseline"),
    }
    head.update(rec)
    default_run = args.config == "single" and not args.reflections
    if default_run and not args.no_bisection:
        head["bisection"] = run_bisection(args.steps, args.warmup, min(budget, 8.0),
                                          head["value"])
    if default_run and not args.no_surrogate:
        head["surrogate"] = run_surrogate(args.steps, args.warmup, epochs=args.epochs)
    if default_run and not args.no_batched:
        head["batched"] = measure_plain("batched", min(args.steps, 10), min(args.warmup, 2),
                                        min(budget, 10.0), autotune_on=not args.no_autotune)
    if default_run and not args.no_large:
        head["large"] = measure_plain("large", min(args.steps, 3), min(args.warmup, 1),
                                      min(budget, 6.0), reflections=args.large_reflections,
                                      autotune_on=False, gather=False)
        head["large"]["reflections_of_config"] = 100000
    print(json.dumps(head), flush=True)


if __name__ == "__main__":
    main()
