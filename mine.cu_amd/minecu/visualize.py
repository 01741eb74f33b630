"""Demo frames of a batch of envs: every step's observations tiled into one
image, written as an animated GIF or a folder of PNG frames.

Mirrors ``create_demo_video`` of examples/visualize.py:22-118 (same
arguments and the same random-policy run: world 32^3, episode length 60,
target OAKLOG, uniform action ids).  That script draws through matplotlib
into an MP4 with ffmpeg; here the grid is assembled on the GPU with one copy
to the host per step, and Pillow writes the file (no ffmpeg needed).

    python -m minecu.visualize --output demo.gif --steps 200 --fps 20
"""
import argparse
import math
import os
from typing import Callable, List, Optional

import numpy as np
import torch


def to_u8(obs: torch.Tensor) -> torch.Tensor:
    """uint8 view of an observation batch: float32 images (values in [0, 1])
    are scaled by 255 and truncated, as the uint8 renders are."""
    if obs.dtype == torch.uint8:
        return obs
    return (obs.clamp(0.0, 1.0) * 255.0).to(torch.uint8)


def tile(obs: torch.Tensor, n_cols: Optional[int] = None, pad: int = 2,
         layout: str = "hwc") -> torch.Tensor:
    """Tile B images into one [rows*(H+pad)-pad, cols*(W+pad)-pad, 3] uint8
    image, row-major (env i at row i // cols, column i % cols), with black
    gaps of ``pad`` pixels.  ``obs`` is [B, H, W, 3] (layout "hwc") or
    [B, 3, H, W] ("chw"); the result stays on obs's device."""
    if obs.dim() != 4:
        raise ValueError("obs must be a [B, H, W, 3] or [B, 3, H, W] batch")
    x = to_u8(obs)
    if layout == "chw":
        x = x.permute(0, 2, 3, 1)
    elif layout != "hwc":
        raise ValueError("layout must be 'hwc' or 'chw'")
    B, H, W, C = x.shape
    if C != 3:
        raise ValueError("images must have 3 channels")
    cols = n_cols or int(math.ceil(math.sqrt(B)))
    rows = (B + cols - 1) // cols
    grid = torch.zeros((rows * cols, H + pad, W + pad, 3), dtype=torch.uint8, device=x.device)
    grid[:B, :H, :W] = x
    grid = grid.view(rows, cols, H + pad, W + pad, 3).permute(0, 2, 1, 3, 4)
    grid = grid.reshape(rows * (H + pad), cols * (W + pad), 3)
    return grid[: rows * (H + pad) - pad, : cols * (W + pad) - pad]


def collect(env, steps: int, policy: Optional[Callable] = None, n_cols: Optional[int] = None,
            generator: Optional[torch.Generator] = None) -> List[np.ndarray]:
    """Reset ``env``, then step it ``steps`` times and keep each step's tiled
    observations.  ``policy(obs) -> actions`` defaults to uniform random ids
    in [0, env.action_dim) (examples/visualize.py:50-58)."""
    layout = getattr(env, "obs_layout", "hwc")
    obs = env.reset()
    frames = []
    for _ in range(steps):
        if policy is None:
            actions = torch.randint(0, env.action_dim, (env.batch_size,), device=env.device,
                                    generator=generator)
        else:
            actions = policy(obs)
        obs, _, _ = env.step(actions)
        frames.append(tile(obs, n_cols, layout=layout).cpu().numpy())
    return frames


def save(frames: List[np.ndarray], path: str, fps: int = 20) -> str:
    """Write ``frames`` ([h, w, 3] uint8 each): ``*.gif`` as one animated GIF
    (256-colour palette per frame), anything else as a folder of lossless
    ``frame_00000.png`` files.  Returns the path written."""
    from PIL import Image
    if not frames:
        raise ValueError("no frames to write")
    images = [Image.fromarray(np.ascontiguousarray(f)) for f in frames]
    if path.lower().endswith(".gif"):
        images[0].save(path, save_all=True, append_images=images[1:],
                       duration=max(1, int(round(1000.0 / fps))), loop=0)
        return path
    os.makedirs(path, exist_ok=True)
    for i, im in enumerate(images):
        im.save(os.path.join(path, f"frame_{i:05d}.png"))
    return path


def create_demo_video(output_path: str = "demo.gif", num_envs: int = 8, steps: int = 200,
                      fps: int = 20, render_size: int = 64, random_policy: bool = True,
                      device: str = "cuda", seed: int = 0) -> str:
    """examples/visualize.py:22-118: ``num_envs`` envs of a 32^3 world,
    episode length 60, target OAKLOG, driven by uniform random actions (the
    reference's non-random branch is random as well)."""
    from . import BlockType, MineEnv
    env = MineEnv(batch_size=num_envs, world_size=32, render_width=render_size,
                  render_height=render_size, episode_length=60, target_block=BlockType.OAKLOG,
                  device=device)
    g = torch.Generator(device=env.device).manual_seed(seed)
    frames = collect(env, steps, generator=g)
    out = save(frames, output_path, fps)
    env.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Tiled demo frames of random-policy envs")
    ap.add_argument("--output", default="demo.gif", help="*.gif, or a folder for PNG frames")
    ap.add_argument("--num-envs", type=int, default=8)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--fps", type=int, default=20)
    ap.add_argument("--render-size", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    out = create_demo_video(a.output, a.num_envs, a.steps, a.fps, a.render_size, seed=a.seed)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
