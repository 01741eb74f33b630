"""Worlds larger than one workgroup's LDS (the reference renders any world
size from global memory, src/kernels.cu:300-470): at 64^3 a MineEnv step runs
its staged layers in LDS and the full-height work (reset renders, renders,
deferred envs) from global memory; every state tensor and image equals the
oracle's bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from parity_util import Stats, compare_env, compare_obs, run_pair

pytestmark = pytest.mark.gpu


def _envs(**kw):
    import minecu
    env = minecu.MineEnv(device="cuda", **kw)
    okw = {k: v for k, v in kw.items() if k not in ("obs_layout",)}
    orc = oracle.OracleEnv(**okw)
    return env, orc


def _acts(T, B, seed):
    g = torch.Generator().manual_seed(seed)
    u = torch.rand((T, B), generator=g)
    anyid = torch.randint(0, 11, (T, B), generator=g)
    return torch.where(u < 0.4, torch.full_like(anyid, 6), anyid).numpy()


@pytest.mark.parametrize("ws,W,H,dtype", [(64, 48, 40, "uint8"), (64, 32, 32, "float32"),
                                          (72, 40, 24, "uint8_minimal"),
                                          (64, 32, 24, "uint8_prebasis"),
                                          (64, 24, 24, "uint8_fp16")])
def test_large_world_matches_oracle(ws, W, H, dtype):
    B = 6
    env, orc = _envs(batch_size=B, world_size=ws, render_width=W, render_height=H,
                     obs_dtype=dtype, episode_length=14)
    st = Stats()
    env.reset()
    orc.reset()
    compare_env(env, orc, st, f"reset {ws}")
    run_pair(env, orc, _acts(32, B, ws), st, f"{ws}^3 {dtype}", check_every=2)
    assert st.exact, st.summary()


def test_large_world_deferred_envs_from_global_memory():
    """Solids above the staged layers and agents placed high send envs to the
    full-height pass, which at 64^3 runs from global memory."""
    B = 8
    env, orc = _envs(batch_size=B, world_size=64, render_width=32, render_height=24,
                     obs_dtype="uint8", episode_length=9)
    env.reset()
    orc.reset()
    torch.cuda.synchronize()
    vox = env.voxels.cpu().numpy()
    pos = env.positions.cpu().numpy()
    vox[0:2, 40, :, :] = 9      # ceilings far above the staged layers
    vox[2, 60, 30, 40] = 8
    pos[3:5, 1] = 50.0          # agents high up
    env.voxels.copy_(torch.from_numpy(vox))
    env.positions.copy_(torch.from_numpy(pos))
    orc.voxels[...] = vox
    orc.positions[...] = pos
    st = Stats()
    run_pair(env, orc, _acts(24, B, 7), st, "64^3 deferred")
    assert st.exact, st.summary()


def test_large_world_render_binding():
    """_C.render_uint8 / _C.render on a 96^3 world with cameras inside, outside
    and above it."""
    from minecu import _C
    B, ws = 5, 96
    rng = np.random.default_rng(8)
    vox = oracle.fresh_world(B, ws, ws, ws, 8, (48, 50), (54, 46))
    vox[rng.random(vox.shape) < 0.02] = 3
    cams = np.zeros((B, 5), np.float32)
    cams[:, :3] = rng.uniform(-8, ws + 8, (B, 3))
    cams[:, 3] = rng.uniform(-7, 7, B)
    cams[:, 4] = rng.uniform(-1.5, 1.5, B)
    vg, cg = torch.from_numpy(vox).cuda(), torch.from_numpy(cams).cuda()
    for u8 in (True, False):
        out = torch.empty((B, 20, 28, 3), dtype=torch.uint8 if u8 else torch.float32, device="cuda")
        (_C.render_uint8 if u8 else _C.render)(vg, cg, out, max_steps=100, view_distance=60.0,
                                               fov_degrees=75.0)
        ref = oracle.render(vox, cams, 28, 20, 100, 60.0, 75.0, out_u8=u8)
        st = Stats()
        compare_obs(out.cpu().numpy(), ref, st, f"render 96^3 u8={u8}")
        assert st.exact, st.summary()


@pytest.mark.parametrize("terrain,collision,layout", [("noise", False, "hwc"), ("flat", True, "hwc"),
                                                      ("noise", True, "hwc"), ("flat", False, "chw")])
def test_large_world_modes(terrain, collision, layout):
    """Seeded terrain, voxel collision and planar observations at 64^3."""
    import minecu
    B = 5
    kw = dict(batch_size=B, world_size=64, render_width=32, render_height=24, obs_dtype="uint8",
              episode_length=11, seed=5, collision=collision)
    if terrain == "noise":
        kw.update(terrain_amplitude=4)
    env = minecu.MineEnv(device="cuda", terrain=terrain, obs_layout=layout, **kw)
    orc = oracle.OracleEnv(terrain=terrain, **kw)
    env.reset()
    orc.reset()
    acts = _acts(26, B, 3)
    st = Stats()
    for t in range(len(acts)):
        _, r, d = env.step(torch.as_tensor(acts[t]).cuda())
        _, ro, do = orc.step(acts[t])
        assert np.array_equal(r.cpu().numpy(), ro) and np.array_equal(d.cpu().numpy(), do), t
        if layout == "chw":
            compare_obs(env.obs_buffer.permute(0, 2, 3, 1).cpu().numpy(), orc.obs_buffer, st,
                        f"step {t}")
        else:
            compare_env(env, orc, st, f"step {t}")
    assert st.exact, st.summary()
