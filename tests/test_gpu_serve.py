"""Continuous-batching server: every group's result is the same alone and in
the served mix (bit for bit), HBM-budget admission, request timestamps."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(batch_size=48, world_size=16, render_width=32, render_height=32, obs_dtype="uint8",
         episode_length=9),
    dict(batch_size=16, world_size=32, render_width=64, render_height=64, obs_dtype="uint8",
         episode_length=13),
    dict(batch_size=100, world_size=16, render_width=64, render_height=64, obs_dtype="uint8",
         episode_length=None),
    dict(batch_size=20, world_size=32, render_width=32, render_height=32, obs_dtype="uint8",
         episode_length=6),
]


def _make(kw):
    import minecu
    e = minecu.MineEnv(device="cuda", **kw)
    e.reset()
    return e


def _state(e):
    return {k: getattr(e, k).cpu().numpy().copy() for k in e._STATE_KEYS}


def _same(a, b, where):
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), f"{where}: {k}"


def _acts(T, B, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randint(0, 11, (T, B), generator=g).cuda()


def test_served_equals_alone_interleaved_submits():
    from minecu.serve import StepServer
    T = 14
    served = [_make(kw) for kw in SHAPES]
    alone = [_make(kw) for kw in SHAPES]
    acts = [_acts(T, kw["batch_size"], 10 + i) for i, kw in enumerate(SHAPES)]
    torch.cuda.synchronize()
    srv = StepServer(n_streams=2)
    gids = [srv.add_group(e) for e in served]
    # interleave: group i's step t is submitted in a shuffled global order
    # that keeps each group's own steps in sequence
    rng = np.random.default_rng(0)
    order = np.concatenate([np.full(T, i) for i in range(len(SHAPES))])
    rng.shuffle(order)
    nxt = [0] * len(SHAPES)
    for i in order:
        srv.submit(gids[i], acts[i][nxt[i]])
        nxt[i] += 1
    srv.drain()
    rec = srv.records()
    assert len(rec) == T * len(SHAPES)
    assert (rec["t_done_ns"] >= rec["t_launch_ns"]).all()
    assert (rec["t_launch_ns"] >= rec["t_submit_ns"]).all()
    st = srv.stats()
    assert st["requests"] == T * len(SHAPES) and st["launches"] <= st["requests"]
    for i, e in enumerate(alone):
        for t in range(T):
            e.step(acts[i][t])
    torch.cuda.synchronize()
    for i in range(len(SHAPES)):
        _same(_state(alone[i]), _state(served[i]), f"group {i}")
    srv.close()


def test_served_closed_loop_replay_equals_alone():
    from minecu.serve import StepServer
    T = 20
    served = [_make(kw) for kw in SHAPES]
    alone = [_make(kw) for kw in SHAPES]
    acts = [_acts(T, kw["batch_size"], 30 + i) for i, kw in enumerate(SHAPES)]
    torch.cuda.synchronize()
    srv = StepServer(n_streams=2)
    gids = [srv.add_group(e) for e in served]
    for g, a in zip(gids, acts):
        srv.replay(g, a)
    srv.drain()
    assert len(srv.records()) == T * len(SHAPES)
    for i, e in enumerate(alone):
        for t in range(T):
            e.step(acts[i][t])
    torch.cuda.synchronize()
    for i in range(len(SHAPES)):
        _same(_state(alone[i]), _state(served[i]), f"group {i}")
    srv.close()


def test_admission_by_hbm_budget():
    from minecu.serve import ADMITTED, REMOVED, WAITING, StepServer, env_bytes
    envs = [_make(SHAPES[0]) for _ in range(3)]
    nb = env_bytes(envs[0])
    srv = StepServer(hbm_budget_bytes=2 * nb + nb // 2, n_streams=1)
    g = [srv.add_group(e) for e in envs]
    assert [srv.state(x) for x in g] == [ADMITTED, ADMITTED, WAITING]
    a = _acts(1, envs[0].batch_size, 0)[0]
    with pytest.raises(Exception):
        srv.submit(g[2], a)
    srv.submit(g[0], a)
    srv.drain()
    srv.remove_group(g[0])
    assert [srv.state(x) for x in g] == [REMOVED, ADMITTED, ADMITTED]
    assert srv.stats()["used_bytes"] == 2 * nb
    srv.submit(g[2], a)
    srv.drain()
    assert len(srv.records()) == 2
    srv.close()


def test_scheduled_open_loop_equals_alone():
    """Planned Poisson arrivals released by the server's loop thread: every
    group ends where stepping it alone ends, each request's recorded submit
    time is its planned arrival, and malformed schedules are refused."""
    from minecu import _lib
    from minecu.serve import StepServer
    T = 12
    served = [_make(kw) for kw in SHAPES]
    alone = [_make(kw) for kw in SHAPES]
    acts = [_acts(T, kw["batch_size"], 50 + i) for i, kw in enumerate(SHAPES)]
    torch.cuda.synchronize()
    srv = StepServer(n_streams=2)
    gids = [srv.add_group(e) for e in served]
    rng = np.random.default_rng(7)
    times, who = [], []
    for i in range(len(SHAPES)):
        t = np.cumsum(rng.exponential(1.0 / 2000.0, T))  # about 2000 requests/s per group
        times.append(t)
        who.append(np.full(T, i))
    times, who = np.concatenate(times), np.concatenate(who)
    order = np.argsort(times, kind="stable")
    times, who = times[order], who[order]
    step_of = np.zeros(len(who), dtype=np.int64)
    seen = [0] * len(SHAPES)
    for k, i in enumerate(who.tolist()):
        step_of[k] = seen[i]
        seen[i] += 1
    a = [acts[i][s] for i, s in zip(who.tolist(), step_of.tolist())]
    with pytest.raises(_lib.MineCuError):
        srv.schedule([gids[0], gids[0]], [acts[0][0], acts[0][1]], [0.002, 0.001])  # decreasing
    with pytest.raises(ValueError):
        srv.schedule([gids[0]], [acts[1][0]], [0.0])  # another group's batch size
    srv.schedule([gids[i] for i in who.tolist()], a, times)
    srv.drain()
    rec = srv.records()
    assert len(rec) == len(who)
    assert (rec["t_done_ns"] >= rec["t_launch_ns"]).all()
    assert (rec["t_launch_ns"] >= rec["t_submit_ns"]).all()
    # submit times are the planned arrivals: their gaps match the plan
    sub = np.sort(rec["t_submit_ns"]) - rec["t_submit_ns"].min()
    plan = np.round((times - times[0]) * 1e9).astype(np.int64)
    assert np.abs(sub - plan).max() <= 1000
    for i, e in enumerate(alone):
        for t in range(T):
            e.step(acts[i][t])
    torch.cuda.synchronize()
    for i in range(len(SHAPES)):
        _same(_state(alone[i]), _state(served[i]), f"group {i}")
    srv.close()


def test_served_mode_families_equal_alone():
    """Planar, seeded-terrain and collision groups served beside a flat
    interleaved group (closed-loop replay of discrete ids): each equals
    stepping it alone, bit for bit."""
    from minecu.serve import StepServer
    kws = [
        SHAPES[0],
        dict(batch_size=24, world_size=16, render_width=32, render_height=24, obs_dtype="float32",
             obs_layout="chw", episode_length=7),
        dict(batch_size=32, world_size=32, render_width=32, render_height=32, obs_dtype="uint8",
             terrain="noise", terrain_amplitude=5, seed=5, episode_length=8),
        dict(batch_size=20, world_size=32, render_width=24, render_height=16, obs_dtype="uint8",
             terrain="noise", terrain_amplitude=3, collision=True, seed=6, episode_length=10),
    ]
    T = 16
    served = [_make(kw) for kw in kws]
    alone = [_make(kw) for kw in kws]
    acts = [_acts(T, kw["batch_size"], 70 + i) for i, kw in enumerate(kws)]
    torch.cuda.synchronize()
    srv = StepServer(n_streams=2)
    gids = [srv.add_group(e) for e in served]
    for g, a in zip(gids, acts):
        srv.replay(g, a)
    srv.drain()
    assert len(srv.records()) == T * len(kws)
    for i, e in enumerate(alone):
        for t in range(T):
            e.step(acts[i][t])
    torch.cuda.synchronize()
    for i in range(len(kws)):
        _same(_state(alone[i]), _state(served[i]), f"group {i}")
    srv.close()
