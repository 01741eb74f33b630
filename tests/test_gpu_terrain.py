"""Seeded terrain (MineEnv(terrain="noise")) and voxel collision
(collision=True) on the GPU: worlds, positions,
rewards, dones and images equal the oracle's bit for bit through reset,
steps, done steps with regeneration, the staged-height fast path and the
full-height pass; amplitude 0 steps exactly like the flat world."""
import numpy as np
import pytest
import torch

import oracle
from parity_util import Stats, compare_env

pytestmark = pytest.mark.gpu


def _acts(T, B, seed):
    g = torch.Generator().manual_seed(seed)
    u = torch.rand((T, B), generator=g)
    anyid = torch.randint(0, 11, (T, B), generator=g)
    return torch.where(u < 0.4, torch.full_like(anyid, 6), anyid).numpy()


@pytest.mark.parametrize("ws,W,H,dtype,amp,seed", [
    (16, 32, 32, "uint8", 3, 42),
    (32, 32, 24, "uint8", 5, 7),
    (16, 24, 16, "float32", 2, 2 ** 31 + 11),
    (8, 16, 16, "uint8", 4, 3),
    (16, 32, 32, "uint8_minimal", 3, 11),
    (16, 32, 16, "uint8_prebasis", 2, 12),
])
def test_terrain_matches_oracle(ws, W, H, dtype, amp, seed):
    import minecu
    B = 40
    kw = dict(batch_size=B, world_size=ws, render_width=W, render_height=H, obs_dtype=dtype,
              episode_length=9, seed=seed)
    env = minecu.MineEnv(device="cuda", terrain="noise", terrain_amplitude=amp, **kw)
    orc = oracle.OracleEnv(terrain="noise", terrain_amplitude=amp, **kw)
    env.reset()
    orc.reset()
    st = Stats()
    compare_env(env, orc, st, "reset")
    acts = _acts(24, B, seed % 1000)
    for t in range(len(acts)):
        _, r, d = env.step(torch.as_tensor(acts[t]).cuda())
        _, ro, do = orc.step(acts[t])
        assert np.array_equal(r.cpu().numpy(), ro) and np.array_equal(d.cpu().numpy(), do), t
        if t % 4 == 3 or t == len(acts) - 1:
            compare_env(env, orc, st, f"step {t}")
    assert st.exact, st.summary()


@pytest.mark.parametrize("collision", [False, True])
def test_terrain_spawn_on_column_surface(collision):
    """Amplitude 8 (spawn columns from 2 to 18 high, ground_height 10): every
    env spawns and respawns on its own spawn column's surface; kernel and
    oracle agree bit for bit through reset, falls, landings and a done step."""
    import minecu
    B = 48
    kw = dict(batch_size=B, world_size=32, render_width=32, render_height=24, obs_dtype="uint8",
              episode_length=10, seed=42, ground_height=10, terrain_amplitude=8,
              collision=collision)
    env = minecu.MineEnv(device="cuda", terrain="noise", **kw)
    orc = oracle.OracleEnv(terrain="noise", **kw)
    env.reset()
    orc.reset()
    st = Stats()
    compare_env(env, orc, st, "reset")
    acts = _acts(24, B, 5)
    for t in range(len(acts)):
        _, r, d = env.step(torch.as_tensor(acts[t]).cuda())
        _, ro, do = orc.step(acts[t])
        assert np.array_equal(r.cpu().numpy(), ro) and np.array_equal(d.cpu().numpy(), do), t
        compare_env(env, orc, st, f"step {t}")
    assert st.exact, st.summary()


def test_amplitude_zero_steps_like_the_flat_world():
    import minecu
    kw = dict(batch_size=32, world_size=16, render_width=32, render_height=32, obs_dtype="uint8",
              episode_length=7, seed=9)
    flat = minecu.MineEnv(device="cuda", **kw)
    noise = minecu.MineEnv(device="cuda", terrain="noise", terrain_amplitude=0, **kw)
    flat.reset()
    noise.reset()
    acts = torch.as_tensor(_acts(20, 32, 4)).cuda()
    for t in range(20):
        flat.step(acts[t])
        noise.step(acts[t])
    for k in ("voxels", "positions", "obs_buffer", "rewards", "step_count"):
        assert torch.equal(getattr(flat, k), getattr(noise, k)), k


def test_seeds_give_different_worlds():
    import minecu
    from minecu import _lib
    kw = dict(batch_size=4, world_size=16, render_width=16, render_height=16, episode_length=5)
    a = minecu.MineEnv(device="cuda", terrain="noise", seed=1, **kw)
    b = minecu.MineEnv(device="cuda", terrain="noise", seed=2, **kw)
    a.reset()
    b.reset()
    assert not torch.equal(a.voxels, b.voxels)
    with pytest.raises(ValueError):
        minecu.MineEnv(device="cuda", terrain="noise", obs_layout="chw", **kw)
    with pytest.raises(_lib.MineCuError):
        minecu.MineEnv(device="cuda", terrain="noise", terrain_amplitude=9, **kw)


def _mixed_buttons(T, B, seed):
    rng = np.random.default_rng(seed)
    b = (rng.random((T, B, 8)) < np.array([0.6, 0.1, 0.15, 0.15, 0.3, 0.2, 0.1, 0.3])).astype(np.int8)
    look = rng.normal(0.0, 0.3, (T, B, 2)).astype(np.float32)
    return b, look


@pytest.mark.parametrize("ws,terrain,amp,seed,dtype", [
    (16, "flat", 3, 5, "uint8"),
    (16, "noise", 3, 42, "uint8"),
    (32, "noise", 5, 8, "uint8"),
    (16, "noise", 2, 13, "float32"),
    (32, "flat", 3, 14, "uint8_minimal"),
])
def test_collision_matches_oracle(ws, terrain, amp, seed, dtype):
    import minecu
    B = 48
    kw = dict(batch_size=B, world_size=ws, render_width=24, render_height=16, obs_dtype=dtype,
              episode_length=20, seed=seed)
    env = minecu.MineEnv(device="cuda", terrain=terrain, terrain_amplitude=amp, collision=True,
                         **kw)
    orc = oracle.OracleEnv(terrain=terrain, terrain_amplitude=amp, collision=True, **kw)
    env.reset()
    orc.reset()
    bts, looks = _mixed_buttons(45, B, seed)
    st = Stats()
    for t in range(45):
        _, r, d = env.step(torch.as_tensor(bts[t]).cuda(), torch.as_tensor(looks[t]).cuda())
        _, ro, do = orc.step(bts[t], looks[t])
        assert np.array_equal(r.cpu().numpy(), ro) and np.array_equal(d.cpu().numpy(), do), t
        if t % 5 == 4:
            compare_env(env, orc, st, f"step {t}")
    assert st.exact, st.summary()
    if terrain == "flat":
        # the agents stand on the grass (y = 8), not on the plane y = Y / 4
        assert float(env.positions[:, 1].min()) > ws // 4


@pytest.mark.parametrize("lift", [0, 1])
def test_terrain_collision_48_full_height_pass(lift):
    """48^3 seeded terrain with voxel collision: envs with agents high up or
    solids above the staged layers take the full-height pass, whose whole
    world (110,592 B) needs the 512-thread kernels (k_env_fb_col<512, ...>),
    with the upward-ray exit off and on.  Bit-exact against the oracle through
    falls, landings and a done step."""
    import minecu
    B = 12
    kw = dict(batch_size=B, world_size=48, render_width=32, render_height=24, obs_dtype="uint8",
              episode_length=10, seed=9, terrain_amplitude=3, collision=True)
    env = minecu.MineEnv(device="cuda", terrain="noise", **kw)
    orc = oracle.OracleEnv(terrain="noise", **kw)
    env.set_shape(0, 0, lift)
    env.reset()
    orc.reset()
    torch.cuda.synchronize()
    vox = env.voxels.cpu().numpy()
    pos = env.positions.cpu().numpy()
    vox[0:3, 40, :, :] = 9          # ceilings above the staged layers
    pos[3:6, 1] = 36.0              # agents high up
    env.voxels.copy_(torch.from_numpy(vox))
    env.positions.copy_(torch.from_numpy(pos))
    orc.voxels[...] = vox
    orc.positions[...] = pos
    st = Stats()
    acts = _acts(22, B, 17)
    for t in range(len(acts)):
        _, r, d = env.step(torch.as_tensor(acts[t]).cuda())
        _, ro, do = orc.step(acts[t])
        assert np.array_equal(r.cpu().numpy(), ro) and np.array_equal(d.cpu().numpy(), do), t
        compare_env(env, orc, st, f"step {t}")
    assert st.exact, st.summary()
