"""minecu -- batched voxel RL environment, MI355X (gfx950) native.

Drop-in for Infatoshi/mine.cu's Python API (minecu/__init__.py): same
``MineEnv`` constructor, ``reset(mask)``, ``step(buttons, look)``, tensor
attributes and layouts, ``BlockType``, ``EnvConfig``/``TrainConfig``.  Every
step runs as hand-written HIP kernels for CDNA4 through libminecu_hip.so:
one fused kernel launch per call (decode, physics, break/place raycast,
respawn, render from an LDS-staged world, regeneration), with no host
synchronisation.  ``EnvBatch`` steps several groups in one launch.

Example::

    import torch
    from minecu import MineEnv

    env = MineEnv(batch_size=4096, world_size=32, device="cuda", obs_dtype="uint8")
    obs = env.reset()
    buttons = torch.zeros(4096, 8, dtype=torch.int8, device="cuda")
    buttons[:, 0] = 1                       # forward
    look = torch.zeros(4096, 2, device="cuda")
    obs, rewards, dones = env.step(buttons, look)
    obs, rewards, dones = env.step(torch.randint(0, 11, (4096,), device="cuda"))  # discrete
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Tuple

import torch

from . import _lib
from .config import EnvConfig, TrainConfig

_lib.load()  # fail loudly at import when the HIP library is missing
from . import _C  # noqa: E402  (reference-named tensor functions)

__version__ = "0.1.0"
__all__ = ["MineEnv", "BlockType", "EnvConfig", "TrainConfig", "EnvBatch"]


class BlockType:
    """Block ids (minecu/__init__.py:47-62 / src/kernels.cu:11-24)."""
    AIR = -1
    GRASS = 0
    DIRT = 1
    STONE = 2
    OAKLOG = 3
    LEAVES = 4
    SAND = 5
    WATER = 6
    GLASS = 7
    BRICK = 8
    COBBLESTONE = 9
    PLANKS = 10
    SNOW = 11
    BEDROCK = 12


# obs_dtype strings the reference accepts (minecu/__init__.py:167) and the
# encoding each maps to here.  "uint8_smem" renders the same image as
# "uint8_prebasis" (src/kernels.cu:835-982 only moves the world to shared
# memory, which every mode here does); "uint8_prebasis_full" adds its episode
# convention (0 = no limit).
_OBS_MODES = {"uint8": _lib.MC_OBS_U8, "float32": _lib.MC_OBS_F32,
              "uint8_minimal": _lib.MC_OBS_U8_MINIMAL,
              "uint8_prebasis": _lib.MC_OBS_U8_PREBASIS,
              "uint8_prebasis_full": _lib.MC_OBS_U8_PREBASIS,
              "uint8_smem": _lib.MC_OBS_U8_PREBASIS,
              "uint8_fp16": _lib.MC_OBS_U8_FP16}
_BASIS_OBS = ("uint8_prebasis", "uint8_prebasis_full", "uint8_smem")
_REFERENCE_OBS = ("float32", "uint8", "uint8_minimal", "uint8_fp16", "uint8_prebasis",
                  "uint8_prebasis_full", "uint8_smem")

# Discrete action table (docs/environment.md:3-17).
ACTION_NAMES = ("noop", "forward", "backward", "strafe_left", "strafe_right", "jump", "break",
                "look_left", "look_right", "look_up", "look_down")


class MineEnv:
    """Batched voxel environment; every op runs as gfx950 HIP kernels.

    Arguments match the reference (minecu/__init__.py:141-165) plus
    ``look_delta``: the yaw/pitch change in radians of the discrete look
    actions 7-10, which the reference never defines.

    Actions: ``step(buttons int8[B,8], look f32[B,2])`` exactly as the
    reference, or ``step(actions)`` with integer ids in [0, 11)
    (docs/environment.md:3-17; ids outside that range are no-ops).
    """

    button_dim = 8
    look_dim = 2
    action_dim = 11

    @classmethod
    def from_config(cls, config: "EnvConfig", **overrides) -> "MineEnv":
        """minecu/__init__.py:115-139."""
        kw = dict(
            batch_size=config.batch_size,
            world_size=getattr(config, "world_size", 32),
            world_x=getattr(config, "world_x", None),
            world_y=getattr(config, "world_y", None),
            world_z=getattr(config, "world_z", None),
            render_width=config.render_width,
            render_height=config.render_height,
            device=config.device,
            max_steps=config.max_steps,
            view_distance=config.view_distance,
            fov_degrees=config.fov_degrees,
            dt=config.dt,
            gravity=config.gravity,
            walk_speed=config.walk_speed,
            sprint_mult=getattr(config, "sprint_mult", 1.5),
            jump_vel=config.jump_vel,
            ground_height=config.ground_height,
            target_block=config.target_block,
            reward_value=config.reward_value,
            episode_length=config.episode_length,
            seed=config.seed,
        )
        kw.update(overrides)
        return cls(**kw)

    def __init__(
        self,
        batch_size: int = 4096,
        world_size: int = 32,
        world_x: Optional[int] = None,
        world_y: Optional[int] = None,
        world_z: Optional[int] = None,
        render_width: int = 64,
        render_height: int = 64,
        device: str = "cuda",
        max_steps: int = 64,
        view_distance: float = 32.0,
        fov_degrees: float = 70.0,
        dt: float = 0.05,
        gravity: float = -20.0,
        walk_speed: float = 4.0,
        sprint_mult: float = 1.5,
        jump_vel: float = 8.0,
        ground_height: int = 8,
        target_block: int = BlockType.OAKLOG,
        reward_value: float = 1.0,
        episode_length: Optional[int] = None,
        seed: int = 42,
        obs_dtype: str = "float32",
        look_delta: float = 0.1,
        obs_layout: str = "hwc",
        terrain: str = "flat",
        terrain_amplitude: int = 3,
        collision: bool = False,
    ):
        assert "cuda" in str(device), "MineEnv requires CUDA device"
        assert obs_dtype in _REFERENCE_OBS, \
            "obs_dtype must be one of " + ", ".join(repr(o) for o in _REFERENCE_OBS)
        # "chw": the shaded uint8 / float32 observations written planar,
        # [B, 3, H, W] -- what a convolutional policy reads (the reference's
        # training example permutes every observation, examples/train_wood.py:53).
        if obs_layout not in ("hwc", "chw"):
            raise ValueError("obs_layout must be 'hwc' or 'chw'")
        if obs_layout == "chw" and obs_dtype not in ("uint8", "float32"):
            raise ValueError("obs_layout 'chw' is available for obs_dtype 'uint8' and 'float32'")
        self.obs_layout = obs_layout
        # "noise": each env gets a seeded height map (column heights
        # ground_height + [-terrain_amplitude, terrain_amplitude]) with the tree
        # and house on it; "flat" is the reference's world, whose seed is unused.
        if terrain not in ("flat", "noise"):
            raise ValueError("terrain must be 'flat' or 'noise'")
        if terrain == "noise" and obs_layout == "chw":
            raise ValueError("terrain 'noise' renders the 'hwc' layout only")
        self.terrain, self.terrain_amplitude = terrain, int(terrain_amplitude)
        # collision=True: the agent stands on, and is blocked by, solid voxels
        # instead of the reference's ground plane at world_y / 4.
        if collision and obs_layout == "chw":
            raise ValueError("collision renders the 'hwc' layout only")
        self.collision = bool(collision)
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.batch_size = batch_size
        self.world_x = world_x if world_x is not None else world_size
        self.world_y = world_y if world_y is not None else world_size
        self.world_z = world_z if world_z is not None else world_size
        self.world_size = world_size
        self.render_width = render_width
        self.render_height = render_height
        self.max_steps = max_steps
        self.view_distance = view_distance
        self.fov_degrees = fov_degrees
        self.dt = dt
        self.gravity = gravity
        self.walk_speed = walk_speed
        self.sprint_mult = sprint_mult
        self.jump_vel = jump_vel
        self.ground_height = ground_height
        self.target_block = target_block
        self.reward_value = reward_value
        self.episode_length = episode_length
        self.seed = seed
        self.obs_dtype = obs_dtype
        self.look_delta = look_delta

        # Spawn / tree / house (minecu/__init__.py:196-206).
        self.spawn_x = self.world_x / 2.0
        self.spawn_y = float(ground_height + 2)
        self.spawn_z = self.world_z / 2.0
        self.tree_x = int(self.spawn_x)
        self.tree_z = int(self.spawn_z) + 2
        self.house_x = int(self.spawn_x) + 6
        self.house_z = int(self.spawn_z) - 2

        self._allocate_state()
        self._cfg, self._st = self._make_descriptors()
        handle = ctypes.c_void_p()
        _lib.check(_lib.load().mc_env_create(ctypes.byref(self._cfg), ctypes.byref(self._st),
                                             ctypes.byref(handle)), "MineEnv")
        self._handle = handle
        self._generate_world()
        self._sync()

    # -- state -------------------------------------------------------------
    def _allocate_state(self):
        """minecu/__init__.py:214-256, with the same tensors, names, dtypes
        and shapes, the eight decode buffers included."""
        B = self.batch_size
        X, Y, Z = self.world_x, self.world_y, self.world_z
        H, W = self.render_height, self.render_width
        dev = self.device
        self.voxels = torch.zeros((B, Y, X, Z), dtype=torch.int8, device=dev)
        self.positions = torch.zeros((B, 3), dtype=torch.float32, device=dev)
        self.velocities = torch.zeros((B, 3), dtype=torch.float32, device=dev)
        self.yaws = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.pitches = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.on_ground = torch.ones((B,), dtype=torch.bool, device=dev)
        obs_t = torch.uint8 if self.obs_dtype != "float32" else torch.float32
        shape = (B, 3, H, W) if self.obs_layout == "chw" else (B, H, W, 3)
        self.obs_buffer = torch.zeros(shape, dtype=obs_t, device=dev)
        self.cameras = torch.zeros((B, 5), dtype=torch.float32, device=dev)
        # Camera basis of the prebasis variants (minecu/__init__.py:237-241).
        self.basis = (torch.zeros((B, 14), dtype=torch.float32, device=dev)
                      if self.obs_dtype in _BASIS_OBS else None)
        # Action decode buffers (minecu/__init__.py:243-251): the step's
        # decoded controls, written by the fused step kernel.
        self.forward_in = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.strafe_in = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.delta_yaw_in = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.delta_pitch_in = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.jump_in = torch.zeros((B,), dtype=torch.bool, device=dev)
        self.do_break = torch.zeros((B,), dtype=torch.bool, device=dev)
        self.do_place = torch.zeros((B,), dtype=torch.bool, device=dev)
        self.speed_mult = torch.ones((B,), dtype=torch.float32, device=dev)
        self.rewards = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.step_count = torch.zeros((B,), dtype=torch.int32, device=dev)
        self.do_reset = torch.zeros((B,), dtype=torch.bool, device=dev)

    def _make_descriptors(self):
        c = _lib.McConfig()
        c.batch_size = self.batch_size
        c.world_x, c.world_y, c.world_z = self.world_x, self.world_y, self.world_z
        c.width, c.height = self.render_width, self.render_height
        c.max_steps = self.max_steps
        c.view_distance = self.view_distance
        c.fov_degrees = self.fov_degrees
        c.dt, c.gravity, c.walk_speed = self.dt, self.gravity, self.walk_speed
        c.sprint_mult, c.jump_vel = self.sprint_mult, self.jump_vel
        c.target_block = int(self.target_block)
        c.reward_value = self.reward_value
        c.spawn_x, c.spawn_y, c.spawn_z = self.spawn_x, self.spawn_y, self.spawn_z
        c.ground_height = self.ground_height
        c.seed = self.seed & 0xFFFFFFFF
        c.tree_x, c.tree_z, c.house_x, c.house_z = self.tree_x, self.tree_z, self.house_x, self.house_z
        c.episode_length = -1 if self.episode_length is None else int(self.episode_length)
        if self.obs_dtype == "uint8_prebasis_full" and c.episode_length <= 0:
            # Its C++ episode check reads 0 -- what it passes for None -- as no
            # limit (minecu/__init__.py:349, src/kernels.cu:1954-1968).
            c.episode_length = -1
        c.look_delta = self.look_delta
        c.obs_mode = _OBS_MODES[self.obs_dtype]
        if self.obs_layout == "chw":
            c.obs_mode = _lib.MC_OBS_U8_CHW if self.obs_dtype == "uint8" else _lib.MC_OBS_F32_CHW
        c.terrain = _lib.MC_TERRAIN_NOISE if self.terrain == "noise" else _lib.MC_TERRAIN_FLAT
        c.terrain_amplitude = self.terrain_amplitude
        c.collision = 1 if self.collision else 0
        s = _lib.McState()
        for name, t in (("voxels", self.voxels), ("positions", self.positions),
                        ("velocities", self.velocities), ("yaws", self.yaws),
                        ("pitches", self.pitches), ("on_ground", self.on_ground),
                        ("step_count", self.step_count), ("do_reset", self.do_reset),
                        ("cameras", self.cameras), ("obs", self.obs_buffer),
                        ("rewards", self.rewards), ("forward_in", self.forward_in),
                        ("strafe_in", self.strafe_in), ("delta_yaw_in", self.delta_yaw_in),
                        ("delta_pitch_in", self.delta_pitch_in), ("jump_in", self.jump_in),
                        ("do_break", self.do_break), ("do_place", self.do_place),
                        ("speed_mult", self.speed_mult)):
            setattr(s, name, t.data_ptr())
        s.basis = self.basis.data_ptr() if self.basis is not None else None
        return c, s

    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def _generate_world(self):
        """Flat world + tree + house in one pass (minecu/__init__.py:258-268)."""
        _lib.check(_lib.load().mc_regen_worlds(ctypes.byref(self._cfg), self._st.voxels, None,
                                               self._stream()), "generate world")

    def _sync(self):
        """Recompute the device episode flag from step_count."""
        _lib.check(_lib.load().mc_env_sync(self._handle, self._stream()), "sync")

    # -- API -----------------------------------------------------------------
    def reset(self, mask: Optional[torch.Tensor] = None) -> torch.Tensor:
        """minecu/__init__.py:270-302, without the host sync on ``mask.any()``."""
        mptr = None
        if mask is not None:
            mask = self._as_device(mask, torch.bool, (self.batch_size,), "mask")
            mptr = ctypes.c_void_p(mask.data_ptr())
        _lib.check(_lib.load().mc_env_reset(self._handle, mptr, self._stream()), "reset")
        return self.obs_buffer

    def step(self, buttons: torch.Tensor, look: Optional[torch.Tensor] = None
             ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """One step of every env (minecu/__init__.py:304-444).

        ``step(buttons, look)``: int8 [B,8] multi-hot + float32 [B,2] radians.
        ``step(actions)``: integer [B] discrete ids.
        Returns (obs, rewards, dones); obs and rewards alias internal buffers
        that the next step overwrites, dones is a fresh tensor (as in the
        reference).
        """
        B = self.batch_size
        dones = torch.empty((B,), dtype=torch.bool, device=self.device)
        a, kind, lk = self._prepare_actions(buttons, look)
        st = _lib.load().mc_env_step(self._handle, ctypes.c_void_p(a.data_ptr()), kind,
                                     ctypes.c_void_p(lk.data_ptr()) if lk is not None else None,
                                     ctypes.c_void_p(dones.data_ptr()), self._stream())
        _lib.check(st, "step")
        return self.obs_buffer, self.rewards, dones

    def _prepare_actions(self, buttons, look):
        """Validate one step's actions; returns (actions, kind, look)."""
        B = self.batch_size
        if look is None:
            return self._as_device(buttons, torch.int64, (B,), "actions"), _lib.MC_ACT_DISCRETE, None
        bt = self._as_device(buttons, torch.int8, (B, 8), "buttons")
        lk = self._as_device(look, torch.float32, (B, 2), "look")
        return bt, _lib.MC_ACT_BUTTONS, lk

    @property
    def shape(self) -> Tuple[int, int, int]:
        """Launch shape (threads per workgroup, waves per env, upward-ray exit
        0/1) of this group."""
        t, w, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().mc_env_get_shape(self._handle, ctypes.byref(t), ctypes.byref(w),
                                                ctypes.byref(l)))
        return t.value, w.value, l.value

    def set_shape(self, threads: int = 0, waves_per_env: int = 0, lift: int = -1) -> None:
        """Override the launch shape ((0, 0) restores the default size; lift
        -1 keeps the current value).  Results do not depend on it."""
        _lib.check(_lib.load().mc_env_set_shape(self._handle, threads, waves_per_env, lift),
                   "set_shape")

    def set_graph_mode(self, on: bool = True) -> None:
        """Make step() replayable from a captured graph: the episode-ring slot
        moves to device memory and each step also launches a one-thread slot
        advance.  Off by default (one launch per step).  Groups in graph mode
        cannot be stepped by EnvBatch / StepServer."""
        _lib.check(_lib.load().mc_env_set_graph_mode(self._handle, int(bool(on)), self._stream()),
                   "set_graph_mode")
        self._graph_mode = bool(on)

    def count_ray_steps(self, counters: Optional[torch.Tensor]) -> None:
        """Measurement hook: with an int64 [3] device tensor, every later step
        runs a build of the step kernel that adds to it [0] the DDA voxel
        lookups of every pixel, [1] the rays marched, [2] the pixels rendered
        (flat world, interleaved layouts).  None restores the production
        kernel.  Observations and state are the same either way."""
        ptr = None
        if counters is not None:
            if (counters.dtype != torch.int64 or tuple(counters.shape) != (3,)
                    or counters.device != self.device or not counters.is_contiguous()):
                raise ValueError("counters must be a contiguous int64 [3] tensor on the env's device")
            ptr = ctypes.c_void_p(counters.data_ptr())
        self._ray_counters = counters
        _lib.check(_lib.load().mc_env_set_ray_counter(self._handle, ptr), "count_ray_steps")

    def capture_step(self, actions: torch.Tensor, look: Optional[torch.Tensor] = None):
        """Capture one step on static input buffers into a torch CUDA graph
        (the reference captures its kernel sequence into a CUDA graph,
        src/kernels.cu:2338-2382).  Returns (graph, obs, rewards, dones): fill
        `actions` (and `look`) in place, call graph.replay(), read the outputs.
        Every replay is one environment step, exactly as step() would be."""
        self.set_graph_mode(True)
        a, kind, lk = self._prepare_actions(actions, look)
        if a.data_ptr() != actions.data_ptr() or (lk is not None and lk.data_ptr() != look.data_ptr()):
            raise ValueError("capture_step needs contiguous device tensors of the action dtype "
                             "(int64 [B] ids, or int8 [B,8] buttons with float32 [B,2] look)")
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            obs, rewards, dones = self.step(actions, look)
        return g, obs, rewards, dones

    def _render(self) -> torch.Tensor:
        """minecu/__init__.py:446-466: cameras from state, then render."""
        _lib.check(_lib.load().mc_env_render(self._handle, self._stream()), "render")
        return self.obs_buffer

    def get_obs(self) -> torch.Tensor:
        return self._render()

    def close(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            torch.cuda.current_stream(self.device).synchronize()
            _lib.load().mc_env_destroy(h)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def observation_shape(self) -> Tuple[int, int, int]:
        if self.obs_layout == "chw":
            return (3, self.render_height, self.render_width)
        return (self.render_height, self.render_width, 3)

    # -- checkpoint / resume ---------------------------------------------------
    _STATE_KEYS = ("voxels", "positions", "velocities", "yaws", "pitches", "on_ground",
                   "step_count", "do_reset", "cameras", "rewards", "obs_buffer", "forward_in",
                   "strafe_in", "delta_yaw_in", "delta_pitch_in", "jump_in", "do_break",
                   "do_place", "speed_mult")

    def _state_keys(self):
        return self._STATE_KEYS + (("basis",) if self.basis is not None else ())

    def state_dict(self) -> Dict[str, torch.Tensor]:
        """Copies of every state tensor; a step is a pure function of this
        state and the actions, so load_state_dict() resumes bit-exactly."""
        return {k: getattr(self, k).clone() for k in self._state_keys()}

    def load_state_dict(self, sd: Dict[str, torch.Tensor]) -> None:
        for k in self._state_keys():
            dst = getattr(self, k)
            src = sd[k]
            if tuple(src.shape) != tuple(dst.shape) or src.dtype != dst.dtype:
                raise ValueError(f"state {k}: expected {tuple(dst.shape)} {dst.dtype}, "
                                 f"got {tuple(src.shape)} {src.dtype}")
            dst.copy_(src)
        self._sync()

    # -- helpers -----------------------------------------------------------------
    def _as_device(self, t, dtype, shape, name) -> torch.Tensor:
        if not isinstance(t, torch.Tensor):
            t = torch.as_tensor(t)
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)} (got {tuple(t.shape)})")
        if dtype == torch.int64:
            if t.dtype.is_floating_point or t.dtype == torch.bool:
                raise TypeError(f"{name} must be an integer tensor (got {t.dtype})")
        elif dtype == torch.bool:
            pass
        elif t.dtype != dtype:
            raise RuntimeError(f"{name} must be {dtype} (got {t.dtype})")
        return t.to(device=self.device, dtype=dtype).contiguous()


from .batch import EnvBatch  # noqa: E402
