"""The benchmark configurations at their real sizes on the GPU:

* the plain step at batch 32768 (16^3, 32x32) and 4096 (32^3, 64x64, the
  break-heavy action mix) through a done step with regeneration, 256 sampled
  envs compared with the oracle fed those envs' actions, and over each
  workload's configured length (1000 and 500 steps) with 64 sampled envs;
* the C2 sizes on the seeded noise terrain, with and without voxel
  collision, and with float32 and planar observations, over 200 steps with
  64 sampled envs;
* the batched workload (8 uneven groups, seed 3) stepped in one launch per
  step after autotuning: every group equals stepping it alone, bit for bit,
  through a done step, and 64 sampled envs per group equal the oracle;
* the serving workload (64 clients, seed 4) on a Poisson schedule through the
  continuous-batching server: every client's group equals stepping it alone,
  and sampled envs equal the oracle.

Envs are independent and episodes run in lockstep from one reset, so a
sample of envs stepped by the oracle alone sees the same group-wide
regeneration steps as the full batch."""
import os
import sys

import numpy as np
import pytest
import torch

import oracle
from parity_util import Stats, compare_sampled

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the workload definitions)

pytestmark = pytest.mark.gpu

L = 60          # episode length of every workload: done on step index 60
T = 62          # steps: the done step, its regeneration and one step after


def _sample(B, n, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(B, size=min(n, B), replace=False))


def _plain(workload, n_sample=256):
    import minecu
    w = bench.WORKLOADS[workload]
    B = w["batch"]
    kw = dict(world_size=w["world"], render_width=w["W"], render_height=w["H"],
              obs_dtype="uint8", episode_length=w["episode_length"])
    env = minecu.MineEnv(batch_size=B, device="cuda", **kw)
    acts = bench.make_actions(w, T)                 # CPU generator, the bench's seed
    idx = _sample(B, n_sample, 100 + w["seed"])
    orc = oracle.OracleEnv(batch_size=len(idx), **kw)
    env.reset()
    orc.reset()
    st = Stats()
    compare_sampled(env, idx, orc, st, f"{workload} reset")
    a_dev = acts.cuda()
    for t in range(T):
        _, r, d = env.step(a_dev[t])
        orc.step(acts[t].numpy()[idx])
        if t == L:
            assert bool(d.all()), "every env is done on step index 60"
        if t in (L - 1, L, T - 1):
            compare_sampled(env, idx, orc, st, f"{workload} step {t}")
            assert np.array_equal(r.cpu().numpy()[idx], orc.rewards)
    print(workload, st.summary())
    assert st.exact, st.summary()


def test_c2_full_batch_sampled_against_oracle():
    _plain("c2")


def test_c3_full_batch_sampled_against_oracle():
    _plain("c3")


def _long_run(workload, n_steps, n_sample=64, **extra):
    """The plain step for n_steps, n_sample sampled envs against the oracle
    stepping those env indices: rewards and dones every step; state, decode
    buffers and images on every step either side of a done step, on the done
    step itself and every 50th step."""
    import minecu
    w = bench.WORKLOADS[workload]
    B, ep = w["batch"], w["episode_length"]
    kw = dict(world_size=w["world"], render_width=w["W"], render_height=w["H"],
              obs_dtype="uint8", episode_length=ep)
    kw.update(extra)
    layout = kw.pop("obs_layout", "hwc")
    env = minecu.MineEnv(batch_size=B, device="cuda", obs_layout=layout, **kw)
    acts = bench.make_actions(w, n_steps)           # CPU generator, the bench's seed
    idx = _sample(B, n_sample, 200 + w["seed"])
    orc = oracle.OracleEnv(batch_size=len(idx), env_ids=idx, **kw)
    env.reset()
    orc.reset()
    a_dev = acts.cuda()
    ti = torch.as_tensor(idx, device="cuda")
    st = Stats()
    n_done = 0
    for t in range(n_steps):
        _, r, d = env.step(a_dev[t])
        _, o_r, o_d = orc.step(acts[t].numpy()[idx])
        where = f"{workload} {extra} step {t}"
        assert np.array_equal(r.index_select(0, ti).cpu().numpy(), o_r), f"{where}: rewards"
        g_d = d.index_select(0, ti).cpu().numpy().astype(bool)
        assert np.array_equal(g_d, o_d), f"{where}: dones"
        n_done += int(g_d.all())
        if t % ep in (ep - 1, 0, 1) or t % 50 == 0:
            compare_sampled(env, idx, orc, st, where, chw=layout == "chw")
    assert n_done == n_steps // ep, "a done step every episode_length steps"
    print(workload, extra, n_steps, "steps", st.summary())
    assert st.exact, st.summary()


@pytest.mark.parametrize("workload", ["c2", "c3"])
def test_configured_length_sampled_against_oracle(workload):
    """C2 over its configured 1000 steps (about 16 episodes), C3 over 500."""
    _long_run(workload, bench.WORKLOADS[workload]["config_steps"])


@pytest.mark.parametrize("collision", [False, True])
def test_c2_noise_terrain_sampled_against_oracle(collision):
    """The C2 sizes on the seeded noise terrain (amplitude 3), with and
    without voxel collision, over 200 steps: three done steps, each
    regenerating every world as a fresh height map."""
    _long_run("c2", 200, terrain="noise", terrain_amplitude=3, seed=42, collision=collision)


@pytest.mark.parametrize("obs_dtype,layout", [("float32", "hwc"), ("uint8", "chw"),
                                              ("float32", "chw")])
def test_c2_obs_modes_sampled_against_oracle(obs_dtype, layout):
    """The C2 sizes with float32 and planar observations over 200 steps."""
    _long_run("c2", 200, obs_dtype=obs_dtype, obs_layout=layout)


def _group_env(c, **extra):
    import minecu
    e = minecu.MineEnv(batch_size=c["batch"], world_size=c["world"], render_width=c["res"],
                       render_height=c["res"], device="cuda", obs_dtype="uint8",
                       episode_length=L, **extra)
    e.reset()
    return e


def _group_acts(groups, seed):
    g = torch.Generator().manual_seed(seed)
    return [torch.randint(0, 11, (T, c["batch"]), generator=g) for c in groups]


def _same_state(a, b, where):
    for k in a._STATE_KEYS:
        x, y = getattr(a, k), getattr(b, k)
        if x.dtype.is_floating_point:
            x, y = x.view(torch.int32), y.view(torch.int32)
        assert torch.equal(x, y), f"{where}: {k} differs"


def _oracle_sample(c, acts, idx):
    o = oracle.OracleEnv(batch_size=len(idx), world_size=c["world"], render_width=c["res"],
                         render_height=c["res"], obs_dtype="uint8", episode_length=L)
    o.reset()
    for t in range(T):
        o.step(acts[t].numpy()[idx])
    return o


def test_c4_uneven_groups_batched_equals_alone_and_oracle():
    from minecu.batch import EnvBatch
    groups = bench.batched_groups(3)
    acts = _group_acts(groups, 3)
    mixed = [_group_env(c) for c in groups]
    alone = [_group_env(c) for c in groups]
    batch = EnvBatch(persistent=True)
    batch.autotune(mixed)
    dev_acts = [a.cuda() for a in acts]
    for t in range(T):
        outs = batch.step(mixed, [a[t] for a in dev_acts])
        for i, e in enumerate(alone):
            _, r, d = e.step(dev_acts[i][t])
            assert torch.equal(outs[i][2], d), f"group {i} step {t}: dones"
    torch.cuda.synchronize()
    batch.close()
    st = Stats()
    for i, c in enumerate(groups):
        _same_state(mixed[i], alone[i], f"group {i} {c}")
        idx = _sample(c["batch"], 64, 300 + i)
        compare_sampled(mixed[i], idx, _oracle_sample(c, acts[i], idx), st, f"group {i} {c}")
    print("C4", st.summary())
    assert st.exact, st.summary()


def test_c5_serving_mix_poisson_equals_alone_and_oracle():
    from minecu.serve import StepServer
    clients = bench.serving_clients(4)
    acts = _group_acts(clients, 4)
    served = [_group_env(c) for c in clients]
    alone = [_group_env(c) for c in clients]
    dev_acts = [a.cuda() for a in acts]
    torch.cuda.synchronize()
    srv = StepServer(hbm_budget_bytes=64 << 30, n_streams=2)
    gids = [srv.add_group(e) for e in served]
    # Poisson arrivals, 1000 requests/s per client, T requests each.
    rng = np.random.default_rng(4)
    times, who = [], []
    for i in range(len(clients)):
        times.append(np.cumsum(rng.exponential(1.0 / 1000.0, T)))
        who.append(np.full(T, i))
    times, who = np.concatenate(times), np.concatenate(who)
    order = np.argsort(times, kind="stable")
    times, who = times[order], who[order]
    nxt = [0] * len(clients)
    req = []
    for i in who.tolist():
        req.append(dev_acts[i][nxt[i]])
        nxt[i] += 1
    srv.schedule([gids[i] for i in who.tolist()], req, times)
    srv.drain()
    rec = srv.records()
    assert len(rec) == T * len(clients)
    st_srv = srv.stats()
    srv.close()
    for i, e in enumerate(alone):
        for t in range(T):
            e.step(dev_acts[i][t])
    torch.cuda.synchronize()
    st = Stats()
    for i, c in enumerate(clients):
        _same_state(served[i], alone[i], f"client {i} {c}")
    # Oracle: 16 sampled envs of the first client of each (world, image) kind.
    seen = set()
    for i, c in enumerate(clients):
        kind = (c["world"], c["res"])
        if kind in seen:
            continue
        seen.add(kind)
        idx = _sample(c["batch"], 16, 400 + i)
        compare_sampled(served[i], idx, _oracle_sample(c, acts[i], idx), st, f"client {i} {c}")
    assert len(seen) == 4
    print("C5", st.summary(), "launches", st_srv["launches"], "requests", st_srv["requests"])
    assert st.exact, st.summary()
