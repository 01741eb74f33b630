"""Batched env groups in one persistent launch: every group's results equal
stepping it alone (bit for bit), whatever the launch shape, and the oracle."""
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
