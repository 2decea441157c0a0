"""Host-side logic that needs no GPU: the autotuner's coordinate descent, the
batched workload's specification, the warm-start helpers and the workload
generators."""
import os

import numpy as np
import pytest

from parallelcmacionize_amd import autotune as A
from parallelcmacionize_amd import batched as B
from parallelcmacionize_amd import warmstart as WS
from parallelcmacionize_amd import workloads as W


def test_coordinate_descent_finds_separable_minimum():
    space = {"tile": [8, 16], "block_threads": [256, 512],
             "refill_batch": [8, 16, 32, 64], "chunk": [4096, 8192, 16384]}
    best_cfg = {"tile": 16, "block_threads": 256, "refill_batch": 64, "chunk": 4096}
    calls = []

    def evaluate(cfg):
        calls.append(dict(cfg))
        return 1.0 + sum(0.1 * abs(space[k].index(cfg[k]) - space[k].index(best_cfg[k]))
                         for k in space)

    best, t, trials = A.coordinate_descent(evaluate, space, passes=3, repeats=2)
    assert best == best_cfg and t == pytest.approx(1.0)
    # every configuration is timed `repeats` times, each only once
    keys = [tuple(sorted(c.items())) for c in calls]
    assert all(keys.count(k) == 2 for k in set(keys))
    assert len(trials) == len(set(keys))


def test_coordinate_descent_rejects_failing_configurations():
    def evaluate(cfg):
        if cfg["tile"] == 32:
            raise RuntimeError("tile exceeds the LDS")
        return 2.0 if cfg["chunk"] == 8192 else 1.0

    best, t, _ = A.coordinate_descent(evaluate, {"tile": [16, 32], "block_threads": [512],
                                                 "refill_batch": [32],
                                                 "chunk": [8192, 4096]}, repeats=1)
    assert best["tile"] == 16 and best["chunk"] == 4096 and t == 1.0


def test_cache_roundtrip(tmp_path):
    p = str(tmp_path / "tuning.json")
    A.save_cache(p, {"batched64": dict(A.DEFAULT)})
    assert A.load_cache(p)["batched64"] == A.DEFAULT
    assert A.load_cache(str(tmp_path / "missing.json")) == {}


def test_batched_spec():
    spec = B.problem_set_spec(nprob=64, seed=5)
    assert len(spec) == 64
    assert {s["n"] for s in spec} <= set(B.SIZES)
    for s in spec:
        assert 0.5 <= s["sigma_ln"] <= 1.5 and 1 <= s["nsrc"] <= 16
        assert 2e4 <= s["packets"] <= 4e6
        assert s["box"][3] == pytest.approx(s["n"] * B.CELL)
    again = B.problem_set_spec(nprob=64, seed=5)             # seeded
    assert [s["n"] for s in spec] == [s["n"] for s in again]
    assert [s["packets"] for s in spec] == [s["packets"] for s in again]


def test_logit_clamps_and_inverts():
    x = np.array([0.0, 1e-9, 1e-4, 0.5, 0.99, 1.0])
    z = WS.logit(x)
    back = 1 / (1 + np.exp(-z))
    assert np.all(np.isfinite(z))
    assert back[0] == pytest.approx(WS.XN_CLAMP) and back[-1] == pytest.approx(WS.XN_HIGH)
    assert np.allclose(back[2:5], x[2:5], rtol=1e-12)


def test_crops_tile_the_lattice():
    a = np.arange(8 * 8 * 8).reshape(8, 8, 8)
    c = WS.crops(a, 4)
    assert len(c) == 8
    assert sorted(np.concatenate([x.ravel() for x in c]).tolist()) == list(range(512))


def test_lognormal_density_statistics():
    n = W.lognormal_density(64, sigma_ln=1.0, seed=2)
    ln = np.log(n / W.N_H)
    assert ln.std() == pytest.approx(1.0, rel=1e-6)          # unit-variance field
    assert n.mean() == pytest.approx(W.N_H, rel=0.15)


def test_layout_roundtrip():
    a = np.arange(8 * 12 * 16, dtype=float).reshape(8, 12, 16)
    v = W.global_to_subgrid_major(a, [2, 3, 4])
    assert np.array_equal(W.subgrid_major_to_global(v, [8, 12, 16], [2, 3, 4]), a)
    # the first subgrid block is the (0:4, 0:4, 0:4) corner, x slowest
    assert np.array_equal(v[:64].reshape(4, 4, 4), a[:4, :4, :4])


def test_stromgren_radius_closed_form():
    R = W.stromgren_radius(W.Q_STROMGREN, W.N_H, W.ALPHA_B)
    assert 4 / 3 * np.pi * R ** 3 * W.N_H ** 2 * W.ALPHA_B == pytest.approx(W.Q_STROMGREN)


def test_bench_refuses_more_than_one_gpu():
    """The measurement runs on one GPU; asking for more exits with a message
    (before any device is touched)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_traffic_record_matches_its_profile():
    """The HBM traffic the bench reports comes from a committed record of the
    calibrated counter runs: its bytes are the counters times their copy
    calibration factors, the counters are the ones in the rocprofv3 summary,
    and the profiled command is the bench's own main measurement."""
    import json
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    prof = os.path.join(root, "profiles")
    with open(os.path.join(prof, "tiles_process_kernel_256clumpy_traffic.json")) as f:
        tr = json.load(f)
    want = (tr["fetch_kib_per_launch"] * 1024 * tr["fetch_factor"] +
            tr["write_kib_per_launch"] * 1024 * tr["write_factor"])
    assert tr["hbm_bytes_per_launch"] == pytest.approx(want, abs=1.0)
    assert 0 < tr["hbm_bytes_per_launch"] < tr["algorithmic_bytes_per_launch"]
    with open(os.path.join(prof, "tiles_process_kernel_256clumpy.md")) as f:
        md = f.read()
    row = "| %s | %d | %.1f | %.1f |" % (tr["kernel"], tr["launches"],
                                         tr["fetch_kib_per_launch"],
                                         tr["write_kib_per_launch"])
    assert row in md
    # the record is per launch and keyed by workload: the profiled command
    # is the bench's main measurement with the step counts it names
    assert "--steps %d --warmup %d" % (tr["steps"], tr["warmup"]) in tr["command"]
    assert tr["command"].startswith("python3 bench.py")
