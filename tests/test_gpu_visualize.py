"""Demo frames from live envs: every collected frame is the tiled observation
batch of that step, and the demo run writes one PNG per step."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", ["hwc", "chw"])
def test_frames_are_the_tiled_observations(layout):
    import minecu
    from minecu import visualize as vz
    env = minecu.MineEnv(batch_size=5, world_size=16, render_width=16, render_height=12,
                         device="cuda", obs_dtype="uint8", episode_length=4, obs_layout=layout)
    g = torch.Generator(device="cuda").manual_seed(3)
    frames = vz.collect(env, 6, generator=g)
    assert len(frames) == 6
    last = env.obs_buffer if layout == "hwc" else env.obs_buffer.permute(0, 2, 3, 1)
    last = last.cpu().numpy()
    for i in range(5):
        y, x = (i // 3) * 14, (i % 3) * 18
        assert np.array_equal(frames[-1][y:y + 12, x:x + 16], last[i]), i
    env.close()


def test_demo_run_writes_every_step(tmp_path):
    from minecu import visualize as vz
    out = vz.create_demo_video(str(tmp_path / "demo"), num_envs=3, steps=4, render_size=16)
    assert sorted(os.listdir(out)) == [f"frame_{i:05d}.png" for i in range(4)]
