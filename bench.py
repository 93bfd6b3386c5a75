"""Throughput of the MI355X absorption-correction ray tracer.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config single|small|large|batched]

A step is one pass of the plain path over the workload's whole reflection
batch: host FP64 setup of every reflection, its upload, the trace kernel,
the per-reflection reduction and the download of R absorption factors.  The
label volumes and crystal-voxel lists are resident in HBM before timing
starts.  Prints one JSON line (absorption factors per second) with the trace
kernel's roofline position and the reference C/OpenMP path timed on a bounded
sample of the same workload on this host's cores.

Workloads (anacor_public_amd/workload.py):
  single   256^3 volume, 25,323 crystal voxels (stride 2), 20,000 reflections (default)
  small    64^3 volume, all 3,169 crystal voxels, 256 reflections
  large    1000x1000x800 volume, 100,000 voxels, 100,000 reflections
  batched  8 uneven samples (100^3 .. 383x383x135), 194,281 reflections, one
           persistent graph-launched kernel, autotuned work shapes
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
# L2<->fabric bytes per voxel step of trace_kernel on the 256^3 workload:
# FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (factors from a 1 GiB copy) over
# 1.017e11 voxel steps, profiles/trace_kernel_plain-256cube.md.
TRAFFIC_BYTES_PER_STEP = {"plain-256^3-stride2": 0.885}
METRIC = "absorption factors/sec (rays/s) at 1 GPU, plain vs surrogate at stated rel. error"


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="single", choices=["single", "small", "large", "batched"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--reflections", type=int, default=0,
                    help="trace only the first R reflections of each sample (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards "
              "reflections across GPUs by construction but this build runs and reports N=1.",
              file=sys.stderr)
        sys.exit(2)

    from anacor_public_amd import workload as W

    t0 = time.perf_counter()
    if args.config == "batched":
        wls = W.batched_config()
    else:
        wls = [{"single": W.single_gpu_config, "small": W.small_config,
                "large": W.large_config}[args.config]()]
    if args.reflections:
        wls = [w.subset(args.reflections) for w in wls]
    R = sum(len(w.omega) for w in wls)
    N = sum(len(w.coords) for w in wls)
    name = wls[0].name if len(wls) == 1 else f"batched-{len(wls)}-samples"
    log(f"# workload {name}: volumes {[w.volume.shape for w in wls]}, N={N}, R={R} "
        f"(built in {time.perf_counter() - t0:.1f}s)")

    extra = {}
    if args.config == "batched":
        from anacor_public_amd import autotune
        from anacor_public_amd.ray_tracing import Batch
        runner = Batch(wls)
        if not args.no_autotune:
            t1 = time.perf_counter()
            picks = autotune.tune_batch(runner, wls, trial=256, use_cache=False)
            extra["autotune"] = [{"block": p["block"], "tile": p["tile"]} for p in picks]
            log(f"# autotune ({time.perf_counter() - t1:.1f}s): {extra['autotune']}")

        def call():
            return runner.run([w.s1 for w in wls], [w.omega for w in wls], wls[0].raw_xray,
                              wls[0].omega_axis, wls[0].F)
    else:
        from anacor_public_amd.ray_tracing import Tracer
        runner = Tracer.for_workload(wls[0])
        if args.tile:
            runner.set_tile(args.tile)

        def call():
            return [runner.run_workload(wls[0])]

    for _ in range(args.warmup):
        call()
    kernel_ms, steps, host_ms = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        outs = call()
        st = runner.stats()
        kernel_ms.append(st["kernel_ms"])
        steps.append(st["voxel_steps"])
        host_ms.append(st["host_ms"])
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / args.steps
    value = R * args.steps / elapsed
    k_ms = float(np.mean(kernel_ms))
    vsteps = int(steps[-1])
    achieved = vsteps * 1.0 / (k_ms * 1e-3) / 1e9   # 1 algorithmic byte per voxel step
    extra["host_setup_ms"] = float(np.mean(host_ms))
    log(f"# {args.steps} steps: {ms_per_step:.2f} ms/step, trace kernel {k_ms:.2f} ms, "
        f"host setup {extra['host_setup_ms']:.2f} ms, {vsteps} voxel steps per launch")
    for o in outs:
        assert np.all(np.isfinite(o)) and np.all((o > 0) & (o <= 1))

    if args.config == "large":
        # the FP32 compensated mean against an FP64 GPU accumulation
        f64 = runner.run_workload(wls[0], f64_mean=True)
        extra["fp32_vs_fp64_max_rel"] = float(np.max(np.abs(outs[0] - f64) / f64))
        log(f"# FP32 compensated vs FP64 mean: max rel {extra['fp32_vs_fp64_max_rel']:.2e}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu, cpu_outs = cpu_baseline(wls, args.cpu_budget)
        rel = max(float(np.max(np.abs(o[:len(c)] - c) / c)) for o, c in zip(outs, cpu_outs))
        cpu["max_rel_diff_vs_gpu"] = rel
        log(f"# cpu baseline ({cpu['kind']}, {cpu['cores']} cores): {cpu['value']:.2f} A_h/s, "
            f"max rel diff vs GPU {rel:.2e}")

    traffic = None
    if name in TRAFFIC_BYTES_PER_STEP and not args.reflections:
        traffic = TRAFFIC_BYTES_PER_STEP[name] * vsteps
    pairs = sum(len(w.coords) * len(w.omega) for w in wls)
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "absorption factors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded four-material volumes and reflections)",
        "config": {"workload": name, "volumes": [list(w.volume.shape) for w in wls],
                   "n_voxels": N, "reflections": R, "path": "plain",
                   "mean": "fp32 exp + compensated sum"},
        "rays_per_s": 2.0 * pairs / (ms_per_step * 1e-3),
        "voxel_steps_per_s": vsteps / (ms_per_step * 1e-3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "batch_kernel (graph)" if args.config == "batched"
                     else "trace_kernel",
                     "kernel_ms": k_ms, "bytes_per_launch": vsteps},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    rec.update(extra)
    print(json.dumps(rec), flush=True)
    runner.close()


if __name__ == "__main__":
    main()
