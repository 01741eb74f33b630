"""Planar observations (obs_layout="chw", MC_OBS_U8_CHW / MC_OBS_F32_CHW): the
shaded uint8 / float32 images written [B, 3, H, W], bit-identical to the
oracle's [B, H, W, 3] images transposed, through reset, falling, landed
(constant-image), regeneration and full-height (32-high world) steps and a
state render."""
import numpy as np
import pytest
import torch

import oracle
from parity_util import Stats, compare_exact, compare_float

pytestmark = pytest.mark.gpu


def _acts(T, B, seed):
    g = torch.Generator().manual_seed(seed)
    u = torch.rand((T, B), generator=g)
    anyid = torch.randint(0, 11, (T, B), generator=g)
    return torch.where(u < 0.3, torch.full_like(anyid, 6), anyid).numpy()


def _check(env, orc, where):
    st = Stats()
    compare_exact("voxels", env.voxels.cpu().numpy(), orc.voxels, where)
    for k in ("positions", "yaws", "pitches", "cameras"):
        compare_float(k, getattr(env, k).cpu().numpy(), getattr(orc, k), st, where)
    g = env.obs_buffer.cpu().numpy()
    o = np.ascontiguousarray(orc.obs_buffer.transpose(0, 3, 1, 2))
    assert g.shape == o.shape, where
    assert np.array_equal(g.view(np.uint8), o.view(np.uint8)), f"{where}: planar obs differ"


@pytest.mark.parametrize("dtype", ["uint8", "float32"])
@pytest.mark.parametrize("ws,W,H", [(16, 32, 32), (32, 30, 20), (16, 24, 8), (16, 17, 11)])
def test_planar_observations_match_oracle(dtype, ws, W, H):
    import minecu
    B = 48
    kw = dict(batch_size=B, world_size=ws, render_width=W, render_height=H, obs_dtype=dtype,
              episode_length=12)
    env = minecu.MineEnv(device="cuda", obs_layout="chw", **kw)
    orc = oracle.OracleEnv(**kw)
    assert env.observation_shape == (3, H, W)
    assert tuple(env.obs_buffer.shape) == (B, 3, H, W)
    env.reset()
    orc.reset()
    _check(env, orc, "reset")
    acts = _acts(30, B, 21)
    for t in range(len(acts)):
        o, r, d = env.step(torch.as_tensor(acts[t]).cuda())
        oo, ro, do = orc.step(acts[t])
        assert np.array_equal(r.cpu().numpy(), ro) and np.array_equal(d.cpu().numpy(), do), t
        _check(env, orc, f"step {t}")


def test_planar_render_matches_oracle():
    import minecu
    B = 24
    kw = dict(batch_size=B, world_size=16, render_width=32, render_height=16, obs_dtype="float32",
              episode_length=None)
    env = minecu.MineEnv(device="cuda", obs_layout="chw", **kw)
    orc = oracle.OracleEnv(**kw)
    env.reset()
    orc.reset()
    acts = _acts(6, B, 5)
    for t in range(len(acts)):
        env.step(torch.as_tensor(acts[t]).cuda())
        orc.step(acts[t])
    env._render()
    orc._render()
    _check(env, orc, "render")


def test_planar_layout_is_refused_for_flat_variants():
    import minecu
    with pytest.raises(ValueError):
        minecu.MineEnv(batch_size=4, world_size=8, render_width=8, render_height=8, device="cuda",
                       obs_dtype="uint8_minimal", obs_layout="chw")
