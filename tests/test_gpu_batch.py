"""GPU tests of the batched path (several samples: per-sample host setup,
upload and trace launch, from the host or as one hipGraph) and of the
autotuner.

Each sample's result must equal the unbatched path on the same sample
bit for bit (same voxel order and tiles, so the same partial sums), across
repeated runs with other GPU work in between, and match the FP64 oracle to
1e-6."""
import dataclasses
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def samples():
    from anacor_public_amd import workload as W
    return W.batched_config(R_scale=0.004)


def test_batch_matches_unbatched_and_oracle(gpu, oracle, samples):
    from anacor_public_amd.ray_tracing import Batch, Tracer
    say(f"case: batch of {len(samples)} samples, R={[len(w.omega) for w in samples]}")
    args = ([w.s1 for w in samples], [w.omega for w in samples], samples[0].raw_xray,
            samples[0].omega_axis, samples[0].F)
    with Batch(samples) as b:
        outs = b.run(*args)                      # each path picks its own tiles
        st = b.stats()
        for i in range(len(samples)):
            b.sample(i).set_tile(512)
        pinned = b.run(*args)                    # the same tiles as below
    total_steps = 0
    for i, wl in enumerate(samples):
        with Tracer.for_workload(wl) as t:
            auto = t.run_workload(wl)
            total_steps += t.stats()["voxel_steps"]
            t.set_tile(512)
            single = t.run_workload(wl)
        # same voxel order and tiles: the same partial sums, bit for bit
        np.testing.assert_array_equal(bits(pinned[i]), bits(single), err_msg=f"sample {i}")
        # different tiles only regroup the compensated FP32 sum
        np.testing.assert_allclose(outs[i], auto, rtol=1e-7, atol=0, err_msg=f"sample {i}")
        ref = oracle.trace(wl.subset(3))["result"]
        np.testing.assert_allclose(outs[i][:3], ref, rtol=1e-6, atol=0, err_msg=f"sample {i}")
    assert st["voxel_steps"] == total_steps


def test_batch_with_an_empty_sample(gpu, samples):
    """A sample without reflections gives an empty array; every other sample's
    result is the same bit for bit as in the full batch."""
    from anacor_public_amd.ray_tracing import Batch
    say("case: batch with sample 1 empty")
    s1 = [w.s1 for w in samples]
    om = [w.omega for w in samples]
    geo = (samples[0].raw_xray, samples[0].omega_axis, samples[0].F)
    s1e, ome = list(s1), list(om)
    s1e[1], ome[1] = np.zeros((0, 3)), np.zeros(0)
    with Batch(samples) as b:
        full = b.run(s1, om, *geo)
        part = b.run(s1e, ome, *geo)
    assert part[1].shape == (0,)
    for i in range(len(samples)):
        if i != 1:
            np.testing.assert_array_equal(bits(part[i]), bits(full[i]), err_msg=f"sample {i}")


def test_repeated_runs_with_eager_work_between(gpu, samples):
    from anacor_public_amd.ray_tracing import Batch, Tracer
    say("case: repeated batch runs")
    wl0 = samples[0]
    args = ([w.s1 for w in samples], [w.omega for w in samples], wl0.raw_xray, wl0.omega_axis,
            wl0.F)
    with Batch(samples) as b, Tracer.for_workload(samples[2]) as eager:
        first = b.run(*args)
        for rep in range(4):
            eager.run_workload(samples[2])                  # unrelated eager launches
            again = b.run(*args)
            for i in range(len(samples)):
                np.testing.assert_array_equal(bits(again[i]), bits(first[i]),
                                              err_msg=f"run {rep} sample {i}")
        # another launch shape (5 reflections: smaller automatic tiles, so
        # other FP32 partial sums)
        short = b.run([w.s1[:5] for w in samples], [w.omega[:5] for w in samples],
                      *args[2:])
        for i in range(len(samples)):
            np.testing.assert_allclose(short[i], first[i][:5], rtol=1e-7, atol=0)
        # FP64 mean through the same launches
        f64 = b.run(*args, f64_mean=True)
        for i in range(len(samples)):
            np.testing.assert_allclose(f64[i], first[i], rtol=1e-6, atol=0)


def test_graph_replays_equal_live_launches(gpu, samples):
    """The batch graph, captured once and relaunched with eager work and
    synchronisations between the replays, gives on every replay the results
    of the same sequence launched from the host, bit for bit; new reflection
    arrays of the same sizes reuse the graph."""
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Batch, Tracer
    say("case: batch graph replays")
    wl0 = samples[0]
    geo = (wl0.raw_xray, wl0.omega_axis, wl0.F)
    args = ([w.s1 for w in samples], [w.omega for w in samples], *geo)
    other = [W.make_reflections(len(w.omega), 900 + i, stream=404) for i, w in enumerate(samples)]
    args2 = ([o[0] for o in other], [o[1] for o in other], *geo)
    with Batch(samples) as b, Tracer.for_workload(samples[3]) as eager:
        live = b.run(*args)                      # launched from the host (the default)
        live2 = b.run(*args2)
        assert not b.last_graph
        b.set_graph(True)
        for rep in range(3):
            cur = args if rep != 1 else args2
            got = b.run(*cur)
            assert b.last_graph
            ref = live if rep != 1 else live2
            for i in range(len(samples)):
                np.testing.assert_array_equal(bits(got[i]), bits(ref[i]),
                                              err_msg=f"replay {rep} sample {i}")
            eager.run_workload(samples[3])            # eager work between the replays


def test_autotuner_keeps_results(gpu, oracle, tmp_path, monkeypatch, samples):
    from anacor_public_amd import autotune
    from anacor_public_amd.ray_tracing import Batch
    say("case: autotuner")
    monkeypatch.setenv("ANACOR_TUNE_CACHE", str(tmp_path / "tune.json"))
    wls = samples[:3]
    with Batch(wls) as b:
        before = b.run([w.s1 for w in wls], [w.omega for w in wls], wls[0].raw_xray,
                       wls[0].omega_axis, wls[0].F)
        picks = autotune.tune_batch(b, wls, trial=64,
                                    blocks=((1, 2, 64), (0, 0, 0), (2, 4, 16)), tiles=(0, 1024))
        after = b.run([w.s1 for w in wls], [w.omega for w in wls], wls[0].raw_xray,
                      wls[0].omega_axis, wls[0].F)
        again = autotune.tune_batch(b, wls, trial=64)         # served from the cache
    assert all(not p["cached"] for p in picks) and all(p["cached"] for p in again)
    assert all(len(p["table"]) == 6 for p in picks)
    for i in range(len(wls)):
        np.testing.assert_allclose(after[i], before[i], rtol=1e-6, atol=0)
