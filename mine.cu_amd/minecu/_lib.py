"""ctypes binding of libminecu_hip.so (C ABI in include/minecu_hip.h).

The library is built in-tree (``make -C mine.cu_amd``) for gfx950 and linked
against the HIP runtime inside the installed torch.  Importing this module
without the library raises ImportError -- there is no CPU fallback, matching
the reference, which refuses to import without its extension
(minecu/__init__.py:33-39).
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libminecu_hip.so")

MC_OK, MC_E_INVALID, MC_E_HIP, MC_E_UNSUPPORTED = 0, -1, -2, -3
MC_OBS_U8, MC_OBS_F32, MC_OBS_U8_MINIMAL, MC_OBS_U8_PREBASIS, MC_OBS_U8_FP16 = 0, 1, 2, 3, 4
MC_OBS_U8_CHW, MC_OBS_F32_CHW = 5, 6
MC_TERRAIN_FLAT, MC_TERRAIN_NOISE = 0, 1
MC_ACT_BUTTONS, MC_ACT_DISCRETE = 0, 1
MC_WORKSPACE_BYTES = 256
MC_STEP_STATIC_MAX_BATCH = 262142

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_U = ctypes.c_uint


class McConfig(ctypes.Structure):
    _fields_ = [
        ("batch_size", _I), ("world_x", _I), ("world_y", _I), ("world_z", _I),
        ("width", _I), ("height", _I), ("max_steps", _I),
        ("view_distance", _F), ("fov_degrees", _F),
        ("dt", _F), ("gravity", _F), ("walk_speed", _F), ("sprint_mult", _F), ("jump_vel", _F),
        ("target_block", _I), ("reward_value", _F),
        ("spawn_x", _F), ("spawn_y", _F), ("spawn_z", _F),
        ("ground_height", _I), ("seed", _U),
        ("tree_x", _I), ("tree_z", _I), ("house_x", _I), ("house_z", _I),
        ("episode_length", _I), ("look_delta", _F), ("obs_mode", _I),
        ("terrain", _I), ("terrain_amplitude", _I), ("collision", _I),
    ]


class McState(ctypes.Structure):
    _fields_ = [(n, _P) for n in (
        "voxels", "positions", "velocities", "yaws", "pitches", "on_ground", "step_count",
        "do_reset", "cameras", "obs", "rewards", "basis", "forward_in", "strafe_in",
        "delta_yaw_in", "delta_pitch_in", "jump_in", "do_break", "do_place", "speed_mult")]


# Exported symbols and their argument types.
SIGNATURES = {
    "mc_last_error": ([], ctypes.c_char_p),
    "mc_version": ([], _I),
    "mc_generate_world": ([_P, _I, _I, _I, _I, _I, _U, _P], _I),
    "mc_place_tree": ([_P, _I, _I, _I, _I, _I, _I, _I, _P], _I),
    "mc_place_house": ([_P, _I, _I, _I, _I, _I, _I, _I, _P], _I),
    "mc_regen_worlds": ([_P, _P, _P, _P], _I),
    "mc_reset": ([_P, _P, _P, _P, _P, _P, _I, _F, _F, _F, _P], _I),
    "mc_render": ([_P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _F, _F, _P], _I),
    "mc_step": ([_P, _P, _P, _I, _P, _P, _P, _P], _I),
    "mc_step_static": ([_P, _P, _P, _I, _P, _P, _P], _I),
    "mc_env_reset_state": ([_P, _P, _P, _P, _P], _I),
    "mc_render_state": ([_P, _P, _P], _I),
    "mc_launch_step_uint8": ([_P] * 20 + [_I] * 5 + [_F] * 7 + [ctypes.c_int8, _F, _F, _F, _F, _P], _I),
    "mc_launch_step": ([_P] * 20 + [_I] * 5 + [_F] * 7 + [ctypes.c_int8, _F, _F, _F, _F, _P], _I),
    "mc_launch_render_uint8": ([_P, _P, _P, _I, _I, _I, _I, _I, _F, _F, _P], _I),
    "mc_launch_render": ([_P, _P, _P, _I, _I, _I, _I, _I, _F, _F, _P], _I),
    "mc_launch_step_uint8_minimal": ([_P] * 20 + [_I] * 5 + [_F] * 7 + [ctypes.c_int8, _F, _F, _F, _F, _P], _I),
    "mc_launch_step_uint8_fp16": ([_P] * 20 + [_I] * 5 + [_F] * 7 + [ctypes.c_int8, _F, _F, _F, _F, _P], _I),
    "mc_launch_render_uint8_fp16": ([_P, _P, _P, _I, _I, _I, _I, _I, _F, _F, _P], _I),
    "mc_launch_step_uint8_prebasis": ([_P] * 21 + [_I] * 7 + [_F] * 7 + [ctypes.c_int8, _F, _F, _F, _F, _P], _I),
    "mc_launch_step_uint8_prebasis_full": ([_P] * 22 + [_I] * 7 + [_F] * 7 + [ctypes.c_int8, _F, _F, _F, _F, _I, _P], _I),
    "mc_launch_step_uint8_smem": ([_P] * 21 + [_I] * 5 + [_F] * 7 + [ctypes.c_int8, _F, _F, _F, _F, _P], _I),
    "mc_launch_precompute_camera_basis": ([_P, _P, _I, _I, _I, _F, _P], _I),
    "mc_launch_render_uint8_minimal": ([_P, _P, _P, _I, _I, _I, _I, _I, _F, _F, _P], _I),
    "mc_launch_render_uint8_prebasis": ([_P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "mc_launch_render_uint8_smem": ([_P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "mc_env_create": ([_P, _P, _P], _I),
    "mc_env_destroy": ([_P], _I),
    "mc_env_sync": ([_P, _P], _I),
    "mc_env_step": ([_P, _P, _I, _P, _P, _P], _I),
    "mc_env_reset": ([_P, _P, _P], _I),
    "mc_env_render": ([_P, _P], _I),
    "mc_env_set_shape": ([_P, _I, _I, _I], _I),
    "mc_env_get_shape": ([_P, _P, _P, _P], _I),
    "mc_env_get_config": ([_P, _P], _I),
    "mc_batch_create": ([_P], _I),
    "mc_batch_destroy": ([_P], _I),
    "mc_batch_set_persistent": ([_P, _I], _I),
    "mc_batch_step": ([_P, _P, _P, _P, _P, _P, _I, _P], _I),
    "mc_batch_autotune": ([_P, _P, _I, _I, _P], _I),
    "mc_batch_tuned_shape": ([_P, _P, _P, _P, _P], _I),
    "mc_server_create": ([ctypes.c_int64, _I, _P], _I),
    "mc_server_destroy": ([_P], _I),
    "mc_server_last_error": ([_P], ctypes.c_char_p),
    "mc_server_add_group": ([_P, _P, ctypes.c_int64, _I, _P], _I),
    "mc_server_group_state": ([_P, _I, _P], _I),
    "mc_server_remove_group": ([_P, _I], _I),
    "mc_server_submit": ([_P, _I, _P, _I, _P, _P, _P], _I),
    "mc_server_set_replay": ([_P, _I, _P, ctypes.c_int64, ctypes.c_int64, _I, _P, ctypes.c_int64, _P], _I),
    "mc_server_start_replay": ([_P, _I], _I),
    "mc_server_schedule": ([_P, _P, _P, _I, _P, ctypes.c_int64], _I),
    "mc_server_drain": ([_P], _I),
    "mc_server_take_records": ([_P, _P, _I, _P], _I),
    "mc_server_stats": ([_P, _P, _P, _P, _P], _I),
    "mc_server_autotune": ([_P, _I, _P], _I),
    "mc_debug_sincosf": ([_P, _P, _P, _I, _P], _I),
    "mc_env_set_graph_mode": ([_P, _I, _P], _I),
    "mc_env_set_ray_counter": ([_P, _P], _I),
    "mc_debug_f16": ([_P, _P, _P, _P, _P, _I, _P], _I),
    "mc_debug_fast_math": ([_I, _U, _U, _P, _P], _I),
}

_lib = None


def load():
    """Load the library once; raise ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build it with `make -C mine.cu_amd` "
                "(or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


class MineCuError(RuntimeError):
    """A libminecu_hip call returned an error status."""


def check(status: int, what: str = "") -> None:
    if status != MC_OK:
        msg = load().mc_last_error().decode(errors="replace")
        kind = {MC_E_INVALID: "invalid argument", MC_E_HIP: "HIP error",
                MC_E_UNSUPPORTED: "unsupported"}.get(status, f"status {status}")
        raise MineCuError(f"{what}: {kind}: {msg}" if what else f"{kind}: {msg}")
