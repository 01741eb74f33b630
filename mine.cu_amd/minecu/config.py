"""Configuration dataclasses and presets.

Same fields, defaults and presets as the reference's minecu/config.py:12-101
(EnvConfig, TrainConfig, DEFAULT_ENV, FAST_ENV, QUALITY_ENV), so
``MineEnv.from_config`` accepts the same objects.  ``memory_bytes`` counts the
observation buffer at the configured dtype (the reference always counts it as
float32, config.py:58).
"""
from dataclasses import dataclass
from typing import Optional


@dataclass
class EnvConfig:
    """Environment configuration (minecu/config.py:12-65)."""

    batch_size: int = 4096
    world_size: int = 32

    render_width: int = 64
    render_height: int = 64
    max_steps: int = 64
    view_distance: float = 32.0
    fov_degrees: float = 70.0

    dt: float = 0.05
    gravity: float = -20.0
    walk_speed: float = 4.0
    jump_vel: float = 8.0

    ground_height: int = 8
    seed: int = 42

    target_block: int = 3
    reward_value: float = 1.0

    episode_length: Optional[int] = 60

    device: str = "cuda"

    def memory_bytes(self, obs_bytes_per_channel: int = 4) -> int:
        """Estimated device bytes of the env state (minecu/config.py:47-61)."""
        B = self.batch_size
        W = self.world_size
        H, WW = self.render_height, self.render_width
        voxels = B * W * W * W
        positions = B * 3 * 4
        velocities = B * 3 * 4
        yaws = B * 4
        pitches = B * 4
        on_ground = B
        obs_buffer = B * H * WW * 3 * obs_bytes_per_channel
        cameras = B * 5 * 4
        action_buffers = B * 4 * 5 + B * 2
        return (voxels + positions + velocities + yaws + pitches + on_ground + obs_buffer
                + cameras + action_buffers)

    def memory_mb(self) -> float:
        return self.memory_bytes() / (1024 * 1024)


@dataclass
class TrainConfig:
    """Training configuration (minecu/config.py:68-83)."""

    lr: float = 3e-4
    gamma: float = 0.99
    total_steps: int = 10_000_000
    log_interval: int = 100_000
    hidden_dim: int = 256


DEFAULT_ENV = EnvConfig()

FAST_ENV = EnvConfig(
    batch_size=32768,
    render_width=32,
    render_height=32,
    max_steps=32,
)

QUALITY_ENV = EnvConfig(
    batch_size=4096,
    render_width=128,
    render_height=128,
    max_steps=128,
    view_distance=64.0,
)
