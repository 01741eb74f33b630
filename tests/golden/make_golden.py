"""Writes the golden fixtures of this directory from the oracle.

Each fixture is a short seeded episode run through oracle.OracleEnv (the CPU
restatement of the reference step) and stores, per step, the rewards, dones,
agent state and a SHA-256 of the observation, plus the full first/last
observation, the final voxels and the final per-pixel hit indices.  The
fixtures freeze the oracle: tests/test_golden.py re-runs it (CPU) and the HIP
env (GPU) against them.

    python tests/golden/make_golden.py [case ...]
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

CASES = {
    # name: (env kwargs, steps, action seed, break-heavy)
    "ws16_64px_u8": (dict(batch_size=8, world_size=16, render_width=64, render_height=64,
                          obs_dtype="uint8", episode_length=12), 30, 0, False),
    "ws32_64px_u8": (dict(batch_size=8, world_size=32, render_width=64, render_height=64,
                          obs_dtype="uint8", episode_length=12), 30, 1, True),
    "ws16_32px_f32": (dict(batch_size=4, world_size=16, render_width=32, render_height=32,
                           obs_dtype="float32", episode_length=None), 20, 2, True),
    # flat-colour variants (reset render shaded, step renders flat)
    "ws16_32x24_minimal": (dict(batch_size=6, world_size=16, render_width=32, render_height=24,
                                obs_dtype="uint8_minimal", episode_length=10), 24, 3, False),
    "ws32_48x32_prebasis": (dict(batch_size=6, world_size=32, render_width=48, render_height=32,
                                 obs_dtype="uint8_prebasis", episode_length=11), 24, 4, True),
    "ws32_32px_fp16": (dict(batch_size=6, world_size=32, render_width=32, render_height=32,
                            obs_dtype="uint8_fp16", episode_length=9), 20, 5, True),
}


def actions_for(B, T, seed, break_heavy):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 11, (T, B))
    if break_heavy:
        a = np.where(rng.random((T, B)) < 0.5, 6, a)
    return a.astype(np.int64)


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)


def run_case(kw, T, seed, bh):
    env = oracle.OracleEnv(**kw)
    obs0 = env.reset().copy()
    acts = actions_for(kw["batch_size"], T, seed, bh)
    rec = {k: [] for k in ("rewards", "dones", "positions", "velocities", "yaws", "pitches",
                           "step_count", "obs_sha")}
    for t in range(T):
        obs, r, d = env.step(acts[t], want_hits=(t == T - 1))
        rec["rewards"].append(r.copy())
        rec["dones"].append(d.copy())
        for k in ("positions", "velocities", "yaws", "pitches", "step_count"):
            rec[k].append(getattr(env, k).copy())
        rec["obs_sha"].append(sha(obs))
    out = {k: np.stack(v) for k, v in rec.items()}
    out.update(actions=acts, obs_first=obs0, obs_last=env.obs_buffer.copy(),
               voxels_last=env.voxels.copy(), hits_last=env.hits.copy())
    if env.basis is not None:
        out["basis_last"] = env.basis.copy()
    return out


def main(names=None):
    for name, (kw, T, seed, bh) in CASES.items():
        if names and name not in names:
            continue
        out = run_case(kw, T, seed, bh)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
