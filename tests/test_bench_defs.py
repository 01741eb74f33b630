"""The measurement definitions bench.py reports against (CPU only): the
algorithmic bytes per env-step of the roofline, the workloads' shapes and the
seeded client / group mixes of the batching and serving lines."""
import math

import numpy as np

import bench


def test_algorithmic_bytes_per_env_step():
    # obs write H*W*3 + one world read + 71 B of state, action, reward and
    # done + the world rewrite amortised over the 60-step episode
    c2 = bench.algorithmic_bytes_per_env_step(bench.WORKLOADS["c2"], 1)
    assert math.isclose(c2, 3072 + 4096 + 71 + 4096 / 60)
    assert math.isclose(c2 * 32768 / 1e6, 239.44, rel_tol=1e-4)
    c3 = bench.algorithmic_bytes_per_env_step(bench.WORKLOADS["c3"], 1)
    assert math.isclose(c3, 12288 + 32768 + 71 + 32768 / 60)
    # float32 observations write four bytes per channel
    f32 = bench.algorithmic_bytes_per_env_step(bench.WORKLOADS["c2"], 4)
    assert math.isclose(f32 - c2, 3 * 3072)


def test_headline_workload_is_the_metric_configuration():
    w = bench.WORKLOADS["c2"]
    assert (w["batch"], w["world"], w["W"], w["H"], w["seed"]) == (32768, 16, 32, 32, 1)
    assert bench.REFERENCE_C2 == 49.63e6


def test_break_heavy_action_mix():
    a = bench.make_actions(dict(bench.WORKLOADS["c3"], batch=4096), 50).numpy()
    assert a.shape == (50, 4096) and a.min() >= 0 and a.max() <= 10
    brk = (a == 6).mean()
    assert 0.51 < brk < 0.53          # 50% break plus 1/11 of the uniform 20% (0.518)
    # reproducible from the seed
    assert np.array_equal(a, bench.make_actions(dict(bench.WORKLOADS["c3"], batch=4096), 50).numpy())


def test_client_and_group_mixes():
    clients = bench.serving_clients(4)
    assert len(clients) == 64
    assert all(16 <= c["batch"] <= 4096 and c["world"] in (16, 32) and c["res"] in (32, 64)
               for c in clients)
    assert sum(c["batch"] for c in clients) == 58550
    groups = bench.batched_groups(3)
    assert len(groups) == 8
    assert all(64 <= g["batch"] <= 16384 and g["world"] in (8, 16, 32) and g["res"] in (16, 32, 64)
               for g in groups)
    assert sum(g["batch"] for g in groups) == 15412
