"""Batched env groups in one persistent launch per mode family: every group's
results equal stepping it alone (bit for bit), whatever the launch shape, and
the oracle."""
import numpy as np
import pytest
import torch

import oracle
from parity_util import Stats, compare_env

pytestmark = pytest.mark.gpu

# Uneven groups: batch, world, resolution, obs, episode length.
GROUPS = [
    dict(batch_size=37, world_size=8, render_width=16, render_height=16, obs_dtype="uint8",
         episode_length=7),
    dict(batch_size=130, world_size=16, render_width=32, render_height=32, obs_dtype="uint8",
         episode_length=20),
    dict(batch_size=16, world_size=32, render_width=64, render_height=64, obs_dtype="uint8",
         episode_length=15),
    dict(batch_size=250, world_size=16, render_width=16, render_height=16, obs_dtype="uint8",
         episode_length=None),
    dict(batch_size=24, world_size=32, render_width=32, render_height=32, obs_dtype="float32",
         episode_length=11),
]


def _make(kw):
    import minecu
    e = minecu.MineEnv(device="cuda", **kw)
    e.reset()
    return e


def _actions(T, B, seed):
    g = torch.Generator().manual_seed(seed)
    a = torch.randint(0, 11, (T, B), generator=g)
    brk = torch.rand((T, B), generator=g) < 0.3
    return torch.where(brk, torch.full_like(a, 6), a).cuda()


def _state(e):
    return {k: getattr(e, k).cpu().numpy().copy() for k in
            ("voxels", "positions", "velocities", "yaws", "pitches", "on_ground", "step_count",
             "cameras", "rewards", "obs_buffer")}


def _assert_same(a, b, where):
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), f"{where}: {k}"


@pytest.mark.parametrize("persistent", [True, False])
def test_batch_equals_alone(persistent):
    import minecu
    T = 30
    alone = [_make(kw) for kw in GROUPS]
    mixed = [_make(kw) for kw in GROUPS]
    acts = [_actions(T, kw["batch_size"], 100 + i) for i, kw in enumerate(GROUPS)]
    batch = minecu.EnvBatch(persistent=persistent)
    for t in range(T):
        outs = batch.step(mixed, [a[t] for a in acts])
        for i, e in enumerate(alone):
            o, r, d = e.step(acts[i][t])
            assert np.array_equal(d.cpu().numpy(), outs[i][2].cpu().numpy()), f"t{t} g{i} dones"
            assert np.array_equal(r.cpu().numpy(), outs[i][1].cpu().numpy()), f"t{t} g{i} rewards"
        if t % 5 == 4 or t == T - 1:
            for i in range(len(GROUPS)):
                _assert_same(_state(alone[i]), _state(mixed[i]), f"t{t} group {i}")
    batch.close()


def test_batch_matches_oracle():
    import minecu
    T = 25
    kws = [GROUPS[0], GROUPS[2], GROUPS[4]]
    envs = [_make(kw) for kw in kws]
    orcs = []
    for kw in kws:
        o = oracle.OracleEnv(**kw)
        o.reset()
        orcs.append(o)
    acts = [_actions(T, kw["batch_size"], 7 + i) for i, kw in enumerate(kws)]
    batch = minecu.EnvBatch()
    st = Stats()
    for t in range(T):
        outs = batch.step(envs, [a[t] for a in acts])
        for i, o in enumerate(orcs):
            _, ro, do = o.step(acts[i][t].cpu().numpy())
            assert np.array_equal(outs[i][1].cpu().numpy(), ro)
            assert np.array_equal(outs[i][2].cpu().numpy(), do)
    for i in range(len(kws)):
        compare_env(envs[i], orcs[i], st, f"group {i}")
    assert st.exact, st.summary()


@pytest.mark.parametrize("shape", [(64, 1, 0), (256, 1, 0), (256, 2, 1), (256, 4, 0),
                                   (512, 8, 1), (64, 1, 1)])
@pytest.mark.parametrize("ws", [16, 32])
def test_launch_shape_does_not_change_results(shape, ws):
    kw = dict(batch_size=40, world_size=ws, render_width=32, render_height=24, obs_dtype="uint8",
              episode_length=9)
    ref = _make(kw)
    e = _make(kw)
    e.set_shape(*shape)
    acts = _actions(20, 40, 5)
    for t in range(20):
        ref.step(acts[t])
        e.step(acts[t])
    _assert_same(_state(ref), _state(e), f"shape {shape}")


def test_autotune_then_batch_still_exact():
    import minecu
    T = 12
    kws = GROUPS[:4]
    alone = [_make(kw) for kw in kws]
    mixed = [_make(kw) for kw in kws]
    batch = minecu.EnvBatch()
    shapes = batch.autotune(mixed, steps=6)
    assert len(shapes) == len(kws)
    for t_, w_, l_ in shapes:
        assert t_ in (64, 256, 512) and w_ in (1, 2, 4, 8) and w_ * 64 <= t_ and l_ in (0, 1)
    acts = [_actions(T, kw["batch_size"], 40 + i) for i, kw in enumerate(kws)]
    for t in range(T):
        batch.step(mixed, [a[t] for a in acts])
        for i, e in enumerate(alone):
            e.step(acts[i][t])
    for i in range(len(kws)):
        _assert_same(_state(alone[i]), _state(mixed[i]), f"group {i}")


def test_batch_with_buttons_and_look():
    import minecu
    kws = [GROUPS[1], GROUPS[2]]
    alone = [_make(kw) for kw in kws]
    mixed = [_make(kw) for kw in kws]
    rng = np.random.default_rng(2)
    batch = minecu.EnvBatch()
    for t in range(15):
        bts = [torch.from_numpy(rng.integers(0, 2, (kw["batch_size"], 8)).astype(np.int8)).cuda()
               for kw in kws]
        lks = [torch.from_numpy(rng.normal(0, 0.4, (kw["batch_size"], 2)).astype(np.float32)).cuda()
               for kw in kws]
        batch.step(mixed, bts, lks)
        for i, e in enumerate(alone):
            e.step(bts[i], lks[i])
    for i in range(len(kws)):
        _assert_same(_state(alone[i]), _state(mixed[i]), f"group {i}")


# One group of every mode family: flat interleaved, planar (uint8 and float32),
# seeded terrain, voxel collision on the flat world and on seeded terrain, and
# a 48^3 terrain world with collision whose staged groups defer envs to the
# full-height pass.
FAMILIES = [
    dict(batch_size=30, world_size=16, render_width=32, render_height=32, obs_dtype="uint8",
         episode_length=9),
    dict(batch_size=20, world_size=16, render_width=32, render_height=24, obs_dtype="uint8",
         obs_layout="chw", episode_length=8),
    dict(batch_size=12, world_size=32, render_width=24, render_height=16, obs_dtype="float32",
         obs_layout="chw", episode_length=11),
    dict(batch_size=24, world_size=32, render_width=32, render_height=24, obs_dtype="uint8",
         terrain="noise", terrain_amplitude=5, seed=7, episode_length=10),
    dict(batch_size=16, world_size=16, render_width=24, render_height=16, obs_dtype="uint8",
         collision=True, episode_length=12),
    dict(batch_size=20, world_size=32, render_width=24, render_height=16, obs_dtype="float32",
         terrain="noise", terrain_amplitude=4, collision=True, seed=3, episode_length=9),
    dict(batch_size=12, world_size=48, render_width=32, render_height=24, obs_dtype="uint8",
         terrain="noise", terrain_amplitude=3, collision=True, seed=9, episode_length=10),
]


def _plant_high(e):
    """Ceilings above the staged layers (envs 0-2) and agents high up (envs
    3-5) of the 48^3 group, so its step defers envs to the full-height pass."""
    torch.cuda.synchronize()
    vox = e.voxels.cpu().numpy()
    pos = e.positions.cpu().numpy()
    vox[0:3, 40, :, :] = 9
    pos[3:6, 1] = 36.0
    e.voxels.copy_(torch.from_numpy(vox))
    e.positions.copy_(torch.from_numpy(pos))
    torch.cuda.synchronize()


def _make_families():
    envs = [_make(kw) for kw in FAMILIES]
    _plant_high(envs[-1])
    return envs


def _family_actions(T, seed):
    """Discrete ids for the groups without collision, buttons and look for the
    collision groups (both kinds in one launch)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, kw in enumerate(FAMILIES):
        B = kw["batch_size"]
        if kw.get("collision"):
            p = np.array([0.6, 0.1, 0.15, 0.15, 0.3, 0.2, 0.1, 0.3])
            bt = (rng.random((T, B, 8)) < p).astype(np.int8)
            lk = rng.normal(0.0, 0.3, (T, B, 2)).astype(np.float32)
            out.append((torch.from_numpy(bt).cuda(), torch.from_numpy(lk).cuda()))
        else:
            out.append((_actions(T, B, seed + i), None))
    return out


@pytest.mark.parametrize("persistent", [True, False])
def test_mixed_families_batch_equals_alone(persistent):
    """Every mode family in one batched step (one launch per family present,
    plus the full-height passes): each group's results equal stepping it
    alone, bit for bit, through done steps and regeneration."""
    import minecu
    T = 24
    alone = _make_families()
    mixed = _make_families()
    acts = _family_actions(T, 60)
    batch = minecu.EnvBatch(persistent=persistent)
    for t in range(T):
        outs = batch.step(mixed, [a[t] for a, _ in acts],
                          [None if l is None else l[t] for _, l in acts])
        for i, e in enumerate(alone):
            a, l = acts[i]
            _, r, d = e.step(a[t], None if l is None else l[t])
            assert np.array_equal(d.cpu().numpy(), outs[i][2].cpu().numpy()), f"t{t} g{i} dones"
            assert np.array_equal(r.cpu().numpy(), outs[i][1].cpu().numpy()), f"t{t} g{i} rewards"
        if t % 6 == 5 or t == T - 1:
            for i in range(len(FAMILIES)):
                _assert_same(_state(alone[i]), _state(mixed[i]), f"t{t} group {i}")
    batch.close()


def test_mixed_families_batch_matches_oracle():
    import minecu
    T = 16
    envs = _make_families()
    orcs = []
    for kw, e in zip(FAMILIES, envs):
        o = oracle.OracleEnv(**{k: v for k, v in kw.items() if k != "obs_layout"})
        o.reset()
        o.voxels[...] = e.voxels.cpu().numpy()
        o.positions[...] = e.positions.cpu().numpy()
        orcs.append(o)
    acts = _family_actions(T, 90)
    batch = minecu.EnvBatch()
    st = Stats()
    for t in range(T):
        outs = batch.step(envs, [a[t] for a, _ in acts],
                          [None if l is None else l[t] for _, l in acts])
        for i, o in enumerate(orcs):
            a, l = acts[i]
            if l is None:
                _, ro, do = o.step(a[t].cpu().numpy())
            else:
                _, ro, do = o.step(a[t].cpu().numpy(), l[t].cpu().numpy())
            assert np.array_equal(outs[i][1].cpu().numpy(), ro), f"t{t} g{i} rewards"
            assert np.array_equal(outs[i][2].cpu().numpy(), do), f"t{t} g{i} dones"
    for i, kw in enumerate(FAMILIES):
        planar = kw.get("obs_layout") == "chw"
        compare_env(envs[i], orcs[i], st, f"group {i}", check_obs=not planar)
        if planar:
            g = envs[i].obs_buffer.cpu().numpy()
            o = np.ascontiguousarray(orcs[i].obs_buffer.transpose(0, 3, 1, 2))
            assert np.array_equal(g.view(np.uint8), o.view(np.uint8)), f"group {i}: planar obs"
    assert st.exact, st.summary()
    batch.close()


def test_autotune_mixed_families_then_batch_still_exact():
    import minecu
    T = 10
    alone = _make_families()
    mixed = _make_families()
    batch = minecu.EnvBatch()
    shapes = batch.autotune(mixed, steps=4)
    assert len(shapes) == len(FAMILIES)
    acts = _family_actions(T, 120)
    for t in range(T):
        batch.step(mixed, [a[t] for a, _ in acts], [None if l is None else l[t] for _, l in acts])
        for i, e in enumerate(alone):
            a, l = acts[i]
            e.step(a[t], None if l is None else l[t])
    for i in range(len(FAMILIES)):
        _assert_same(_state(alone[i]), _state(mixed[i]), f"group {i}")
    batch.close()
