"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must keep reproducing them (CPU), and the HIP env
must match them bit for bit (GPU)."""
import hashlib
import os

import numpy as np
import pytest

from golden.make_golden import CASES, run_case, sha

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    kw, T, seed, bh = CASES[name]
    g = _load(name)
    out = run_case(kw, T, seed, bh)
    for k, v in g.items():
        assert np.array_equal(out[k], v), f"{name}: {k} differs from the golden fixture"


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_env_matches_golden(name):
    import torch
    import minecu
    kw, T, seed, bh = CASES[name]
    g = _load(name)
    env = minecu.MineEnv(device="cuda", **kw)
    obs0 = env.reset()
    assert np.array_equal(obs0.cpu().numpy(), g["obs_first"])
    acts = torch.from_numpy(g["actions"]).cuda()
    for t in range(T):
        obs, r, d = env.step(acts[t])
        assert np.array_equal(r.cpu().numpy(), g["rewards"][t]), f"{name} step {t} rewards"
        assert np.array_equal(d.cpu().numpy(), g["dones"][t]), f"{name} step {t} dones"
        for k in ("positions", "velocities", "yaws", "pitches", "step_count"):
            got = getattr(env, k).cpu().numpy()
            assert np.array_equal(got.view(np.uint8), g[k][t].view(np.uint8)), f"{name} step {t} {k}"
        assert np.array_equal(sha(obs.cpu().numpy()), g["obs_sha"][t]), f"{name} step {t} obs"
    assert np.array_equal(env.obs_buffer.cpu().numpy(), g["obs_last"])
    assert np.array_equal(env.voxels.cpu().numpy(), g["voxels_last"])
    if "basis_last" in g:
        assert np.array_equal(env.basis.cpu().numpy().view(np.uint8), g["basis_last"].view(np.uint8))
