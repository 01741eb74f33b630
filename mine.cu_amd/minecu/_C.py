"""Tensor-level functions with the reference's ``minecu._C`` names and
argument lists (src/bindings.cpp:985-1269), implemented over the C ABI.

The reference's bindings check device and dtype only (e.g.
bindings.cpp:537-543); these also check shapes and contiguity, so a mismatch
raises RuntimeError instead of reading out of bounds on the device.  Every
launch goes to torch's current stream (the reference uses the legacy default
stream).  Not implemented: the ground-accelerated step_uint8_fast and the
[Y,X,Z,B]-layout render_uint8_coalesced, neither of which MineEnv calls.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

_lib.load()


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _need(t: torch.Tensor, name: str, dtype: torch.dtype, shape=None, device=None):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be CUDA tensor")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise RuntimeError(f"{name} must have shape {tuple(shape)} (got {tuple(t.shape)})")
    if device is not None and t.device != device:
        raise RuntimeError(f"{name} must be on {device} (got {t.device})")


def generate_world(voxels: torch.Tensor, ground_height: int, seed: int = 42) -> None:
    """bindings.cpp:357-371 -> launch_generate_world."""
    _need(voxels, "voxels", torch.int8)
    B, Y, X, Z = voxels.shape
    _lib.check(_lib.load().mc_generate_world(_ptr(voxels), B, X, Y, Z, ground_height,
                                             seed & 0xFFFFFFFF, _stream(voxels)), "generate_world")


def place_tree(voxels: torch.Tensor, tree_x: int, tree_z: int, ground_height: int) -> None:
    """bindings.cpp:373-388 -> launch_place_tree."""
    _need(voxels, "voxels", torch.int8)
    B, Y, X, Z = voxels.shape
    _lib.check(_lib.load().mc_place_tree(_ptr(voxels), B, X, Y, Z, tree_x, tree_z, ground_height,
                                         _stream(voxels)), "place_tree")


def place_house(voxels: torch.Tensor, house_x: int, house_z: int, ground_height: int) -> None:
    """bindings.cpp:390-406 -> launch_place_house."""
    _need(voxels, "voxels", torch.int8)
    B, Y, X, Z = voxels.shape
    _lib.check(_lib.load().mc_place_house(_ptr(voxels), B, X, Y, Z, house_x, house_z,
                                          ground_height, _stream(voxels)), "place_house")


def reset(positions, velocities, yaws, pitches, on_ground, do_reset, spawn_x, spawn_y,
          spawn_z) -> None:
    """bindings.cpp:333-355 -> launch_reset."""
    B = positions.shape[0]
    dev = positions.device
    _need(positions, "positions", torch.float32, (B, 3), dev)
    _need(velocities, "velocities", torch.float32, (B, 3), dev)
    _need(yaws, "yaws", torch.float32, (B,), dev)
    _need(pitches, "pitches", torch.float32, (B,), dev)
    _need(on_ground, "on_ground", torch.bool, (B,), dev)
    _need(do_reset, "do_reset", torch.bool, (B,), dev)
    _lib.check(_lib.load().mc_reset(_ptr(positions), _ptr(velocities), _ptr(yaws), _ptr(pitches),
                                    _ptr(on_ground), _ptr(do_reset), B, spawn_x, spawn_y, spawn_z,
                                    _stream(positions)), "reset")


def _render(voxels, cameras, output, max_steps, view_distance, fov_degrees, mode, dtype):
    _need(voxels, "voxels", torch.int8)
    B, Y, X, Z = voxels.shape
    dev = voxels.device
    _need(cameras, "cameras", torch.float32, (B, 5), dev)
    _need(output, "output", dtype, None, dev)
    if output.dim() != 4 or output.shape[0] != B or output.shape[3] != 3:
        raise RuntimeError(f"output must have shape [B, H, W, 3] (got {tuple(output.shape)})")
    H, W = output.shape[1], output.shape[2]
    _lib.check(_lib.load().mc_render(_ptr(voxels), _ptr(cameras), _ptr(output), mode, B, X, Y, Z,
                                     W, H, max_steps, view_distance, fov_degrees,
                                     _stream(voxels)), "render")


def render(voxels, cameras, output, max_steps=64, view_distance=32.0, fov_degrees=70.0) -> None:
    """bindings.cpp:215-242 -> launch_render (float32 output)."""
    _render(voxels, cameras, output, max_steps, view_distance, fov_degrees, _lib.MC_OBS_F32,
            torch.float32)


def render_uint8(voxels, cameras, output, max_steps=64, view_distance=32.0,
                 fov_degrees=70.0) -> None:
    """bindings.cpp:244-271 -> launch_render_uint8."""
    _render(voxels, cameras, output, max_steps, view_distance, fov_degrees, _lib.MC_OBS_U8,
            torch.uint8)


def _step(fn_name, obs_dtype, voxels, positions, velocities, yaws, pitches, on_ground, buttons,
          look, do_reset, forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in, do_break,
          do_place, speed_mult, cameras, obs_buffer, rewards, width, height, max_steps,
          view_distance, fov_degrees, dt, gravity, walk_speed, sprint_mult, jump_vel,
          target_block, reward_value, spawn_x, spawn_y, spawn_z, basis=None, step_count=None,
          episode_length=None, per_axis=False):
    """Checks every tensor, then calls the C launcher.  basis / step_count /
    episode_length: the prebasis variants' extra arguments; per_axis: the
    launcher takes world_x, world_y, world_z (else the cubic world_size)."""
    _need(voxels, "voxels", torch.int8)
    B, Y, X, Z = voxels.shape
    if not per_axis and not (X == Y == Z):
        raise RuntimeError("this step binding takes a cubic world (world_size = voxels.size(1))")
    dev = voxels.device
    if basis is not None:
        _need(basis, "basis", torch.float32, (B, 14), dev)
    if step_count is not None:
        _need(step_count, "step_count", torch.int32, (B,), dev)
    for t, n, d, s in ((positions, "positions", torch.float32, (B, 3)),
                       (velocities, "velocities", torch.float32, (B, 3)),
                       (yaws, "yaws", torch.float32, (B,)), (pitches, "pitches", torch.float32, (B,)),
                       (on_ground, "on_ground", torch.bool, (B,)),
                       (buttons, "buttons", torch.int8, (B, 8)), (look, "look", torch.float32, (B, 2)),
                       (do_reset, "do_reset", torch.bool, (B,)),
                       (forward_in, "forward_in", torch.float32, (B,)),
                       (strafe_in, "strafe_in", torch.float32, (B,)),
                       (delta_yaw_in, "delta_yaw_in", torch.float32, (B,)),
                       (delta_pitch_in, "delta_pitch_in", torch.float32, (B,)),
                       (jump_in, "jump_in", torch.bool, (B,)), (do_break, "do_break", torch.bool, (B,)),
                       (do_place, "do_place", torch.bool, (B,)),
                       (speed_mult, "speed_mult", torch.float32, (B,)),
                       (cameras, "cameras", torch.float32, (B, 5)),
                       (obs_buffer, "obs_buffer", obs_dtype, (B, height, width, 3)),
                       (rewards, "rewards", torch.float32, (B,))):
        _need(t, n, d, s, dev)
    fn = getattr(_lib.load(), fn_name)
    ptrs = [voxels, positions, velocities, yaws, pitches, on_ground, buttons, look, do_reset]
    if step_count is not None:
        ptrs.append(step_count)
    ptrs += [forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in, do_break, do_place,
             speed_mult, cameras]
    if basis is not None:
        ptrs.append(basis)
    ptrs += [obs_buffer, rewards]
    dims = (B, X, Y, Z) if per_axis else (B, Y)
    tail = [] if episode_length is None else [int(episode_length)]
    _lib.check(fn(*(_ptr(t) for t in ptrs), *dims, width, height, max_steps, view_distance,
                  fov_degrees, dt, gravity, walk_speed, sprint_mult, jump_vel, int(target_block),
                  reward_value, spawn_x, spawn_y, spawn_z, *tail, _stream(voxels)), fn_name)


def step_uint8(voxels, positions, velocities, yaws, pitches, on_ground, buttons, look, do_reset,
               forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in, do_break, do_place,
               speed_mult, cameras, obs_buffer, rewards, width, height, max_steps=64,
               view_distance=32.0, fov_degrees=70.0, dt=0.05, gravity=-20.0, walk_speed=4.0,
               sprint_mult=1.5, jump_vel=8.0, target_block=-1, reward_value=1.0, spawn_x=16.0,
               spawn_y=9.0, spawn_z=16.0) -> None:
    """bindings.cpp:492-575 -> launch_step_uint8 (defaults as bindings.cpp:1062-1096)."""
    _step("mc_launch_step_uint8", torch.uint8, voxels, positions, velocities, yaws, pitches,
          on_ground, buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
          delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, obs_buffer, rewards,
          width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed,
          sprint_mult, jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z)


def step(voxels, positions, velocities, yaws, pitches, on_ground, buttons, look, do_reset,
         forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in, do_break, do_place,
         speed_mult, cameras, obs_buffer, rewards, width, height, max_steps=64,
         view_distance=32.0, fov_degrees=70.0, dt=0.05, gravity=-20.0, walk_speed=4.0,
         sprint_mult=1.5, jump_vel=8.0, target_block=-1, reward_value=1.0, spawn_x=16.0,
         spawn_y=9.0, spawn_z=16.0) -> None:
    """bindings.cpp:408-490 -> launch_step (float32 observations)."""
    _step("mc_launch_step", torch.float32, voxels, positions, velocities, yaws, pitches,
          on_ground, buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
          delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, obs_buffer, rewards,
          width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed,
          sprint_mult, jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z)



def step_uint8_minimal(voxels, positions, velocities, yaws, pitches, on_ground, buttons, look,
                       do_reset, forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in,
                       do_break, do_place, speed_mult, cameras, obs_buffer, rewards, width, height,
                       max_steps=64, view_distance=32.0, fov_degrees=70.0, dt=0.05, gravity=-20.0,
                       walk_speed=4.0, sprint_mult=1.5, jump_vel=8.0, target_block=-1,
                       reward_value=1.0, spawn_x=16.0, spawn_y=9.0, spawn_z=16.0) -> None:
    """bindings.cpp:632-682 -> launch_step_uint8_minimal (flat colours)."""
    _step("mc_launch_step_uint8_minimal", torch.uint8, voxels, positions, velocities, yaws,
          pitches, on_ground, buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
          delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, obs_buffer, rewards,
          width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed,
          sprint_mult, jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z)


def step_uint8_fp16(voxels, positions, velocities, yaws, pitches, on_ground, buttons, look,
                    do_reset, forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in,
                    do_break, do_place, speed_mult, cameras, obs_buffer, rewards, width, height,
                    max_steps=64, view_distance=32.0, fov_degrees=70.0, dt=0.05, gravity=-20.0,
                    walk_speed=4.0, sprint_mult=1.5, jump_vel=8.0, target_block=-1,
                    reward_value=1.0, spawn_x=16.0, spawn_y=9.0, spawn_z=16.0) -> None:
    """bindings.cpp:684-734 -> launch_step_uint8_fp16 (flat colours, binary16 DDA)."""
    _step("mc_launch_step_uint8_fp16", torch.uint8, voxels, positions, velocities, yaws,
          pitches, on_ground, buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
          delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, obs_buffer, rewards,
          width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed,
          sprint_mult, jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z)


def step_uint8_prebasis(voxels, positions, velocities, yaws, pitches, on_ground, buttons, look,
                        do_reset, forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in,
                        do_break, do_place, speed_mult, cameras, basis, obs_buffer, rewards, width,
                        height, max_steps=64, view_distance=32.0, fov_degrees=70.0, dt=0.05,
                        gravity=-20.0, walk_speed=4.0, sprint_mult=1.5, jump_vel=8.0,
                        target_block=-1, reward_value=1.0, spawn_x=16.0, spawn_y=9.0,
                        spawn_z=16.0) -> None:
    """bindings.cpp:736-792 -> launch_step_uint8_prebasis (writes basis [B,14])."""
    _step("mc_launch_step_uint8_prebasis", torch.uint8, voxels, positions, velocities, yaws,
          pitches, on_ground, buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
          delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, obs_buffer, rewards,
          width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed,
          sprint_mult, jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z,
          basis=basis, per_axis=True)


def step_uint8_prebasis_full(voxels, positions, velocities, yaws, pitches, on_ground, buttons,
                             look, do_reset, step_count, forward_in, strafe_in, delta_yaw_in,
                             delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras,
                             basis, obs_buffer, rewards, width, height, max_steps=64,
                             view_distance=32.0, fov_degrees=70.0, dt=0.05, gravity=-20.0,
                             walk_speed=4.0, sprint_mult=1.5, jump_vel=8.0, target_block=-1,
                             reward_value=1.0, spawn_x=16.0, spawn_y=9.0, spawn_z=16.0,
                             episode_length=0) -> None:
    """bindings.cpp:794-875 -> launch_step_uint8_prebasis_full: do_reset =
    step_count >= episode_length (0: never) and the step_count update happen
    in the kernel; world regeneration stays with the caller."""
    _step("mc_launch_step_uint8_prebasis_full", torch.uint8, voxels, positions, velocities, yaws,
          pitches, on_ground, buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
          delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, obs_buffer, rewards,
          width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed,
          sprint_mult, jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z,
          basis=basis, step_count=step_count, episode_length=episode_length, per_axis=True)


def step_uint8_smem(voxels, positions, velocities, yaws, pitches, on_ground, buttons, look,
                    do_reset, forward_in, strafe_in, delta_yaw_in, delta_pitch_in, jump_in,
                    do_break, do_place, speed_mult, cameras, basis, obs_buffer, rewards, width,
                    height, max_steps=64, view_distance=32.0, fov_degrees=70.0, dt=0.05,
                    gravity=-20.0, walk_speed=4.0, sprint_mult=1.5, jump_vel=8.0,
                    target_block=-1, reward_value=1.0, spawn_x=16.0, spawn_y=9.0,
                    spawn_z=16.0) -> None:
    """bindings.cpp:930-981 -> launch_step_uint8_smem (cubic world; the image
    equals the prebasis step's)."""
    _step("mc_launch_step_uint8_smem", torch.uint8, voxels, positions, velocities, yaws,
          pitches, on_ground, buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
          delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, obs_buffer, rewards,
          width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed,
          sprint_mult, jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z,
          basis=basis)


def precompute_camera_basis(cameras, basis, width, height, fov_degrees=70.0) -> None:
    """bindings.cpp:857-875 -> launch_precompute_camera_basis."""
    _need(cameras, "cameras", torch.float32)
    if cameras.dim() != 2 or cameras.shape[1] != 5:
        raise RuntimeError(f"cameras must have shape [B, 5] (got {tuple(cameras.shape)})")
    B = cameras.shape[0]
    _need(basis, "basis", torch.float32, (B, 14), cameras.device)
    _lib.check(_lib.load().mc_launch_precompute_camera_basis(
        _ptr(cameras), _ptr(basis), B, width, height, fov_degrees, _stream(cameras)),
        "precompute_camera_basis")


def render_uint8_prebasis(voxels, basis, output, max_steps=64) -> None:
    """bindings.cpp:877-899 -> launch_render_uint8_prebasis (cubic world_size =
    voxels.size(1); flat colours from a stored basis)."""
    _need(voxels, "voxels", torch.int8)
    B, Y, X, Z = voxels.shape
    if not (X == Y == Z):
        raise RuntimeError("render_uint8_prebasis takes a cubic world (world_size = voxels.size(1))")
    dev = voxels.device
    _need(basis, "basis", torch.float32, (B, 14), dev)
    _need(output, "output", torch.uint8, None, dev)
    if output.dim() != 4 or output.shape[0] != B or output.shape[3] != 3:
        raise RuntimeError(f"output must have shape [B, H, W, 3] (got {tuple(output.shape)})")
    H, W = output.shape[1], output.shape[2]
    _lib.check(_lib.load().mc_launch_render_uint8_prebasis(
        _ptr(voxels), _ptr(basis), _ptr(output), B, Y, W, H, max_steps, _stream(voxels)),
        "render_uint8_prebasis")
