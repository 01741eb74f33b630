"""A step captured into a CUDA/HIP graph (MineEnv.capture_step) replays as
eager steps would: every replay equals the eager env's step bit for bit,
through done steps and world regeneration, with eager work (including eager
steps of the graph env itself) and synchronisations between replays."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _state(e):
    return {k: getattr(e, k).cpu().numpy().copy() for k in e._state_keys()}


def _same(a, b, where):
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), f"{where}: {k}"


@pytest.mark.parametrize("obs_dtype,ws,extra", [
    ("uint8", 16, {}), ("uint8_prebasis", 32, {}),
    ("uint8", 16, {"terrain": "noise", "collision": True}), ("float32", 32, {"terrain": "noise"}),
])
def test_graph_replays_equal_eager_steps(obs_dtype, ws, extra):
    import minecu
    B, T = 64, 26
    kw = dict(batch_size=B, world_size=ws, render_width=32, render_height=32,
              obs_dtype=obs_dtype, episode_length=5, device="cuda", **extra)
    eager, graphed = minecu.MineEnv(**kw), minecu.MineEnv(**kw)
    eager.reset()
    graphed.reset()
    g = torch.Generator().manual_seed(3)
    acts = torch.randint(0, 11, (T, B), generator=g)
    acts = torch.where(torch.rand((T, B), generator=g) < 0.4, torch.full_like(acts, 6), acts).cuda()
    static = torch.zeros(B, dtype=torch.int64, device="cuda")
    graph, obs, rew, dones = graphed.capture_step(static)
    for t in range(T):
        if t == 11:  # an eager step of the graph-mode env between replays
            graphed.step(acts[t])
        else:
            static.copy_(acts[t])
            graph.replay()
        _, r_e, d_e = eager.step(acts[t])
        torch.cuda.synchronize()
        if t != 11:
            assert np.array_equal(rew.cpu().numpy(), r_e.cpu().numpy()), f"step {t} rewards"
            assert np.array_equal(dones.cpu().numpy(), d_e.cpu().numpy()), f"step {t} dones"
            assert np.array_equal(obs.cpu().numpy(), eager.obs_buffer.cpu().numpy()), f"step {t} obs"
        _same(_state(eager), _state(graphed), f"step {t}")
    # back to host-tracked slots: eager stepping continues in lock-step
    graphed.set_graph_mode(False)
    for t in range(4):
        graphed.step(acts[t])
        eager.step(acts[t])
    _same(_state(eager), _state(graphed), "after graph mode")


def test_graph_mode_groups_are_refused_by_batches():
    import minecu
    e = minecu.MineEnv(batch_size=4, world_size=8, render_width=8, render_height=8,
                       device="cuda", obs_dtype="uint8")
    e.reset()
    e.set_graph_mode(True)
    with pytest.raises(Exception, match="graph mode"):
        minecu.EnvBatch().step([e], [torch.zeros(4, dtype=torch.int64, device="cuda")])
