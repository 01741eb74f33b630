"""The ray-step counting build of the step kernel (a measurement hook): the
same observations and state as the production kernel, bit for bit, and
counters that add up (every pixel counted once, a marched ray looks up at
least its start voxel, a filled pixel exactly one)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw", [
    dict(batch_size=64, world_size=16, render_width=32, render_height=32, obs_dtype="uint8",
         episode_length=20),
    dict(batch_size=16, world_size=32, render_width=48, render_height=40, obs_dtype="float32",
         episode_length=None),
    dict(batch_size=24, world_size=16, render_width=32, render_height=24, obs_dtype="uint8_fp16",
         episode_length=7),
])
def test_counting_build_matches_production(kw):
    import minecu
    a = minecu.MineEnv(device="cuda", **kw)
    b = minecu.MineEnv(device="cuda", **kw)
    a.reset()
    b.reset()
    c = torch.zeros(3, dtype=torch.int64, device="cuda")
    b.count_ray_steps(c)
    B, T = kw["batch_size"], 30
    g = torch.Generator().manual_seed(1)
    acts = torch.randint(0, 11, (T, B), generator=g).cuda()
    for t in range(T):
        a.step(acts[t])
        b.step(acts[t])
    torch.cuda.synchronize()
    for k in a._STATE_KEYS:
        x, y = getattr(a, k), getattr(b, k)
        if x.dtype.is_floating_point:
            x, y = x.view(torch.int32), y.view(torch.int32)
        assert torch.equal(x, y), k
    looks, marched, pixels = c.cpu().tolist()
    assert pixels == T * B * kw["render_width"] * kw["render_height"]
    assert 0 < marched <= pixels
    assert looks >= pixels            # in-world eyes: every pixel looks up >= 1 voxel
    # after counting is switched off the counters stay put
    b.count_ray_steps(None)
    b.step(acts[0])
    torch.cuda.synchronize()
    assert c.cpu().tolist() == [looks, marched, pixels]


def test_counting_refused_for_terrain():
    import minecu
    e = minecu.MineEnv(batch_size=4, world_size=16, render_width=8, render_height=8,
                       device="cuda", obs_dtype="uint8", terrain="noise")
    with pytest.raises(Exception):
        e.count_ray_steps(torch.zeros(3, dtype=torch.int64, device="cuda"))
