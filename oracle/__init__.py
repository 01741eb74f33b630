"""Oracle for the batched environment step -- TEST INFRASTRUCTURE ONLY.

A scalar CPU restatement of Infatoshi/mine.cu's env step, used as the checker
for the gfx950 HIP kernels.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this package; the product
(``mine.cu_amd/minecu``) never does.

* ``mc_oracle.cpp`` (built by ``Makefile`` into ``_build/libmc_oracle.so``)
  restates the reference kernels one by one (src/kernels.cu, cited per
  function).
* :class:`OracleEnv` restates the ``MineEnv`` orchestration
  (minecu/__init__.py:141-466) over numpy arrays, calling those kernels in the
  reference's order.

Pinning: every function here is checked by CPU tests against closed forms and
invariants (tests/test_oracle_*.py): worldgen layer/tree/house counts, free-fall
landing steps, shaded colour table, eye-inside-solid and all-air renders, an
independent float64 Amanatides-Woo DDA, raycast known answers and the episode
period.  The reference ships no golden vectors (it has no tests), and its CUDA
source cannot be built here, so nothing is compared with the reference binary.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmc_oracle.so")
_lib = None

# Block ids (src/kernels.cu:11-24).
AIR, GRASS, DIRT, STONE, OAKLOG, LEAVES = -1, 0, 1, 2, 3, 4
BRICK, COBBLESTONE, PLANKS, BEDROCK = 8, 9, 10, 12


def build() -> str:
    """Compile the oracle library (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _bind(L):
    """Declare the oracle library's signatures."""
    P = ctypes.c_void_p
    i, f, u = ctypes.c_int, ctypes.c_float, ctypes.c_uint
    L.orc_sincosf.argtypes = [P, P, P, i]
    L.orc_half_fov_tan.argtypes = [f]
    L.orc_half_fov_tan.restype = f
    L.orc_generate_world.argtypes = [P, i, i, i, i, i, u]
    L.orc_place_tree.argtypes = [P, i, i, i, i, i, i, i]
    L.orc_place_house.argtypes = [P, i, i, i, i, i, i, i]
    L.orc_fmix32.argtypes = [u]
    L.orc_fmix32.restype = u
    L.orc_terrain_column.argtypes = [i, i, i, i, i, u]
    L.orc_terrain_heights.argtypes = [P, i, i, i, i, i, u]
    L.orc_terrain_world.argtypes = [P, i, i, i, i, i, i, u, i, i, i, i, P]
    L.orc_reset.argtypes = [P, P, P, P, P, P, i, f, f, f]
    L.orc_discrete_to_buttons.argtypes = [P, i, f, P, P]
    L.orc_decode_actions.argtypes = [P, P, P, P, P, P, P, P, P, P, i, f]
    L.orc_physics.argtypes = [P, P, P, P, P, P, P, P, P, P, P, i, i, i, i, f, f, f, f]
    L.orc_physics_collide.argtypes = [P, P, P, P, P, P, P, P, P, P, P, P, i, i, i, i, f, f,
                                      f, f]
    L.orc_raycast_interact.argtypes = [P, P, P, P, P, P, P, i, i, i, i, i, f]
    L.orc_update_cameras.argtypes = [P, P, P, P, i]
    L.orc_render.argtypes = [P, P, P, i, i, i, i, i, i, i, i, f, f, P, i]
    L.orc_half_fov_tan_flat.argtypes = [f]
    L.orc_half_fov_tan_flat.restype = f
    L.orc_camera_basis.argtypes = [P, P, i, i, i, f]
    L.orc_render_flat.argtypes = [P, P, P, i, i, i, i, i, i, i, i, f, P, i]
    L.orc_f16.argtypes = [P, P, P, P, P, i]
    return L


def lib():
    """Load (building first if absent) the oracle shared library."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


_mutants = {}


def mutant_lib(k: int):
    """The oracle compiled with planted slip number k in its shading (see
    ORC_MUTANT in mc_oracle.cpp) -- only for the test that shows the render
    pins catch such a slip."""
    if k not in _mutants:
        out = os.path.join(_HERE, "_build", f"libmc_oracle_mutant{k}.so")
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c++17",
                        "-pthread", f"-DORC_MUTANT={k}",
                        "-I" + os.path.join(_HERE, "..", "mine.cu_amd", "csrc"), "-shared", "-o", out,
                        os.path.join(_HERE, "mc_oracle.cpp")], check=True)
        _mutants[k] = _bind(ctypes.CDLL(out))
    return _mutants[k]


def _p(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------------------
# Kernel-level wrappers
# ---------------------------------------------------------------------------

def sincosf(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().orc_sincosf(_p(x), _p(s), _p(c), x.size)
    return s, c


def half_fov_tan(fov_degrees: float) -> float:
    return float(lib().orc_half_fov_tan(fov_degrees))


def fresh_world(B, X, Y, Z, ground_height, tree_xz, house_xz, seed=42,
                terrain_amplitude=None, env_ids=None) -> np.ndarray:
    """generate_world + place_tree + place_house, the reference's regen order
    (minecu/__init__.py:258-268).  terrain_amplitude: the seeded noise terrain
    of that amplitude instead (None: the reference's flat world), seeded by
    each world's env index (env_ids: an int32 array of B indices, or None for
    0 .. B-1)."""
    v = np.empty((B, Y, X, Z), dtype=np.int8)
    L = lib()
    if terrain_amplitude is not None:
        ids = None if env_ids is None else np.ascontiguousarray(env_ids, dtype=np.int32)
        assert ids is None or ids.shape == (B,)
        L.orc_terrain_world(_p(v), B, X, Y, Z, ground_height, terrain_amplitude,
                            seed & 0xFFFFFFFF, tree_xz[0], tree_xz[1], house_xz[0], house_xz[1],
                            None if ids is None else _p(ids))
        return v
    L.orc_generate_world(_p(v), B, X, Y, Z, ground_height, seed)
    L.orc_place_tree(_p(v), B, X, Y, Z, tree_xz[0], tree_xz[1], ground_height)
    L.orc_place_house(_p(v), B, X, Y, Z, house_xz[0], house_xz[1], ground_height)
    return v


def terrain_heights(b, X, Z, ground_height, amplitude, seed) -> np.ndarray:
    """[X, Z] solid heights of env b's seeded terrain (before tree and house)."""
    h = np.empty((X, Z), dtype=np.int32)
    lib().orc_terrain_heights(_p(h), b, X, Z, ground_height, amplitude, seed & 0xFFFFFFFF)
    return h


def terrain_column(b, x, z, ground_height, amplitude, seed) -> int:
    return int(lib().orc_terrain_column(b, x, z, ground_height, amplitude, seed & 0xFFFFFFFF))


def fmix32(h: int) -> int:
    return int(lib().orc_fmix32(h & 0xFFFFFFFF))


def render(voxels, cameras, width, height, max_steps=64, view_distance=32.0,
           fov_degrees=70.0, out_u8=True, want_hits=False, nthreads=None, L=None):
    B, Y, X, Z = voxels.shape
    out = np.empty((B, height, width, 3), dtype=np.uint8 if out_u8 else np.float32)
    hits = np.empty((B, height, width), dtype=np.int32) if want_hits else None
    (L or lib()).orc_render(_p(voxels), _p(cameras), _p(out), int(out_u8), B, X, Y, Z, width, height,
                     max_steps, view_distance, fov_degrees, _p(hits),
                     nthreads or default_threads())
    return (out, hits) if want_hits else out


def half_fov_tan_flat(fov_degrees: float) -> float:
    return float(lib().orc_half_fov_tan_flat(fov_degrees))


def camera_basis(cameras, width, height, fov_degrees=70.0):
    """[B,14] camera basis (precompute_camera_basis_kernel, src/kernels.cu:511-561);
    column 13 stays zero."""
    cameras = np.ascontiguousarray(cameras, dtype=np.float32)
    B = cameras.shape[0]
    basis = np.zeros((B, 14), np.float32)
    lib().orc_camera_basis(_p(cameras), _p(basis), B, width, height, fov_degrees)
    return basis


def render_flat(voxels, cams, width, height, mode, max_steps=64, fov_degrees=70.0,
                want_hits=False, nthreads=None):
    """Flat-colour uint8 render.  mode "minimal": cams are [B,5] cameras
    (src/kernels.cu:985-1119); mode "prebasis": cams are a [B,14] basis
    (src/kernels.cu:568-700); mode "fp16": cameras, binary16 DDA
    (src/kernels.cu:1127-1262)."""
    B, Y, X, Z = voxels.shape
    cams = np.ascontiguousarray(cams, dtype=np.float32)
    out = np.empty((B, height, width, 3), dtype=np.uint8)
    hits = np.empty((B, height, width), dtype=np.int32) if want_hits else None
    lib().orc_render_flat(_p(voxels), _p(cams), _p(out), _FLAT_MODE_ID[mode], B, X, Y,
                          Z, width, height, max_steps, fov_degrees, _p(hits),
                          nthreads or default_threads())
    return (out, hits) if want_hits else out


_FLAT_MODE_ID = {"minimal": 0, "prebasis": 1, "fp16": 2}


def f16_ops(a, b):
    """Software binary16 (round to nearest even): (bits of half(a), bits of
    half(a) + half(b), half(a) < half(b)) for float arrays a, b."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    ha = np.empty(a.shape, np.uint16)
    s = np.empty(a.shape, np.uint16)
    lt = np.empty(a.shape, np.uint8)
    lib().orc_f16(_p(a), _p(b), _p(ha), _p(s), _p(lt), a.size)
    return ha, s, lt


def discrete_to_buttons(actions: np.ndarray, look_delta: float):
    actions = np.ascontiguousarray(actions, dtype=np.int64)
    B = actions.shape[0]
    buttons = np.empty((B, 8), dtype=np.int8)
    look = np.empty((B, 2), dtype=np.float32)
    lib().orc_discrete_to_buttons(_p(actions), B, look_delta, _p(buttons), _p(look))
    return buttons, look


# ---------------------------------------------------------------------------
# MineEnv orchestration restated (minecu/__init__.py)
# ---------------------------------------------------------------------------

# obs_dtype -> flat render mode of the step (None: shaded).
_FLAT = {"uint8": None, "float32": None, "uint8_minimal": "minimal",
         "uint8_prebasis": "prebasis", "uint8_prebasis_full": "prebasis",
         "uint8_smem": "prebasis", "uint8_fp16": "fp16"}


class OracleEnv:
    """numpy restatement of ``MineEnv`` (minecu/__init__.py:141-466) for the
    shaded ``uint8``/``float32`` modes and the flat-colour uint8 variants:
    their steps render flat colours (the prebasis variants from the [B,14]
    ``basis`` they update), their reset/get_obs renders stay shaded uint8."""

    def __init__(self, batch_size=4096, world_size=32, world_x=None, world_y=None, world_z=None,
                 render_width=64, render_height=64, max_steps=64, view_distance=32.0,
                 fov_degrees=70.0, dt=0.05, gravity=-20.0, walk_speed=4.0, sprint_mult=1.5,
                 jump_vel=8.0, ground_height=8, target_block=OAKLOG, reward_value=1.0,
                 episode_length=None, seed=42, obs_dtype="uint8", look_delta=0.1,
                 nthreads=None, terrain="flat", terrain_amplitude=3, collision=False,
                 env_ids=None):
        """env_ids: the env index of each env (default 0 .. B-1).  The seeded
        terrain and its spawn heights depend on it, so a few envs of a large
        batch are stepped alone by passing their indices."""
        assert obs_dtype in _FLAT, obs_dtype
        self.batch_size = B = batch_size
        self.env_ids = (np.arange(B, dtype=np.int32) if env_ids is None
                        else np.ascontiguousarray(env_ids, dtype=np.int32))
        assert self.env_ids.shape == (B,)
        self.world_x = world_x if world_x is not None else world_size
        self.world_y = world_y if world_y is not None else world_size
        self.world_z = world_z if world_z is not None else world_size
        self.render_width, self.render_height = render_width, render_height
        self.max_steps, self.view_distance, self.fov_degrees = max_steps, view_distance, fov_degrees
        self.dt, self.gravity, self.walk_speed = dt, gravity, walk_speed
        self.sprint_mult, self.jump_vel = sprint_mult, jump_vel
        self.ground_height, self.target_block, self.reward_value = ground_height, target_block, reward_value
        self.episode_length, self.seed, self.obs_dtype = episode_length, seed, obs_dtype
        self.look_delta = look_delta
        self.terrain_amplitude = terrain_amplitude if terrain == "noise" else None
        self.collision = bool(collision)
        self.nthreads = nthreads or default_threads()
        # Spawn / tree / house (minecu/__init__.py:196-206).
        self.spawn_x = self.world_x / 2.0
        self.spawn_y = float(ground_height + 2)
        self.spawn_z = self.world_z / 2.0
        self.tree_x, self.tree_z = int(self.spawn_x), int(self.spawn_z) + 2
        self.house_x, self.house_z = int(self.spawn_x) + 6, int(self.spawn_z) - 2
        # Per-env respawn height: on the seeded terrain, spawn_y shifted by the
        # height of the env's spawn column (int(spawn_x), int(spawn_z)) above
        # or below ground_height (the library's definition; the reference's
        # world is flat).
        self.spawn_ys = np.full(B, self.spawn_y, np.float32)
        if self.terrain_amplitude is not None:
            for b in range(B):
                h = terrain_column(int(self.env_ids[b]), int(self.spawn_x), int(self.spawn_z),
                                   ground_height, self.terrain_amplitude, seed)
                self.spawn_ys[b] = np.float32(self.spawn_y) + np.float32(h - ground_height)
        # State (minecu/__init__.py:214-256).
        X, Y, Z = self.world_x, self.world_y, self.world_z
        self.positions = np.zeros((B, 3), np.float32)
        self.velocities = np.zeros((B, 3), np.float32)
        self.yaws = np.zeros(B, np.float32)
        self.pitches = np.zeros(B, np.float32)
        self.on_ground = np.ones(B, np.uint8)
        self.obs_buffer = np.zeros((B, render_height, render_width, 3),
                                   np.float32 if obs_dtype == "float32" else np.uint8)
        self.flat = _FLAT[obs_dtype]
        self.basis = np.zeros((B, 14), np.float32) if self.flat == "prebasis" else None
        self.cameras = np.zeros((B, 5), np.float32)
        self.rewards = np.zeros(B, np.float32)
        self.step_count = np.zeros(B, np.int32)
        self.do_reset = np.zeros(B, np.uint8)
        self.forward_in, self.strafe_in = np.zeros(B, np.float32), np.zeros(B, np.float32)
        self.delta_yaw_in, self.delta_pitch_in = np.zeros(B, np.float32), np.zeros(B, np.float32)
        self.jump_in, self.do_break = np.zeros(B, np.uint8), np.zeros(B, np.uint8)
        self.do_place, self.speed_mult = np.zeros(B, np.uint8), np.ones(B, np.float32)
        self.voxels = self._fresh()
        self.hits = None

    def _spawn_heights(self):
        """Respawned envs stand spawn_ys[b] high (the flat world: spawn_y)."""
        m = self.do_reset.astype(bool)
        self.positions[m, 1] = self.spawn_ys[m]

    def _fresh(self):
        return fresh_world(self.batch_size, self.world_x, self.world_y, self.world_z,
                           self.ground_height, (self.tree_x, self.tree_z),
                           (self.house_x, self.house_z), self.seed, self.terrain_amplitude,
                           self.env_ids)

    def _regen(self):
        self.voxels[...] = self._fresh()

    def _render(self, want_hits=False):
        L = lib()
        L.orc_update_cameras(_p(self.cameras), _p(self.positions), _p(self.yaws), _p(self.pitches),
                             self.batch_size)
        r = render(self.voxels, self.cameras, self.render_width, self.render_height,
                   self.max_steps, self.view_distance, self.fov_degrees,
                   out_u8=self.obs_dtype != "float32", want_hits=want_hits, nthreads=self.nthreads)
        if want_hits:
            self.obs_buffer[...], self.hits = r
        else:
            self.obs_buffer[...] = r
        return self.obs_buffer

    # minecu/__init__.py:270-302
    def reset(self, mask: Optional[np.ndarray] = None, want_hits=False):
        if mask is None:
            self.do_reset[:] = 1
        else:
            self.do_reset[:] = np.asarray(mask).astype(np.uint8)
        lib().orc_reset(_p(self.positions), _p(self.velocities), _p(self.yaws), _p(self.pitches),
                        _p(self.on_ground), _p(self.do_reset), self.batch_size, self.spawn_x,
                        self.spawn_y, self.spawn_z)
        self._spawn_heights()
        self.step_count[self.do_reset.astype(bool)] = 0
        if mask is None or np.asarray(mask).any():
            self._regen()
        return self._render(want_hits)

    # minecu/__init__.py:304-444 (shaded uint8 / float32 modes)
    def step(self, buttons, look=None, want_hits=False):
        B = self.batch_size
        if look is None:
            buttons, look = discrete_to_buttons(np.asarray(buttons), self.look_delta)
        buttons = np.ascontiguousarray(buttons, dtype=np.int8)
        look = np.ascontiguousarray(look, dtype=np.float32)
        if self.obs_dtype == "uint8_prebasis_full":
            # episode_check_kernel: 0 (the reference's stand-in for None) is no
            # limit (src/kernels.cu:1954-1968, minecu/__init__.py:349).
            L_ = self.episode_length or 0
            self.do_reset = ((self.step_count >= L_) & (L_ > 0)).astype(np.uint8)
        elif self.episode_length is not None:
            self.do_reset = (self.step_count >= self.episode_length).astype(np.uint8)
        else:
            self.do_reset[:] = 0
        # The decode buffers are kept as attributes, as MineEnv's are
        # (minecu/__init__.py:243-251).
        fwd, strafe, dyaw, dpitch = (np.empty(B, np.float32) for _ in range(4))
        jump, brk, place = (np.empty(B, np.uint8) for _ in range(3))
        speed = np.empty(B, np.float32)
        (self.forward_in, self.strafe_in, self.delta_yaw_in, self.delta_pitch_in, self.jump_in,
         self.do_break, self.do_place, self.speed_mult) = (fwd, strafe, dyaw, dpitch, jump, brk,
                                                           place, speed)
        L = lib()
        L.orc_decode_actions(_p(buttons), _p(look), _p(fwd), _p(strafe), _p(dyaw), _p(dpitch),
                             _p(jump), _p(brk), _p(place), _p(speed), B, self.sprint_mult)
        X, Y, Z = self.world_x, self.world_y, self.world_z
        if self.collision:
            L.orc_physics_collide(_p(self.positions), _p(self.velocities), _p(self.yaws),
                                  _p(self.pitches), _p(self.on_ground), _p(fwd), _p(strafe),
                                  _p(dyaw), _p(dpitch), _p(jump), _p(speed), _p(self.voxels), B,
                                  X, Y, Z, self.dt, self.gravity, self.walk_speed, self.jump_vel)
        else:
            L.orc_physics(_p(self.positions), _p(self.velocities), _p(self.yaws),
                          _p(self.pitches), _p(self.on_ground), _p(fwd), _p(strafe), _p(dyaw),
                          _p(dpitch), _p(jump), _p(speed), B, X, Y, Z, self.dt, self.gravity,
                          self.walk_speed, self.jump_vel)
        L.orc_raycast_interact(_p(self.voxels), _p(self.positions), _p(self.yaws),
                               _p(self.pitches), _p(brk), _p(place), _p(self.rewards), B, X, Y, Z,
                               self.target_block, self.reward_value)
        L.orc_reset(_p(self.positions), _p(self.velocities), _p(self.yaws), _p(self.pitches),
                    _p(self.on_ground), _p(self.do_reset), B, self.spawn_x, self.spawn_y,
                    self.spawn_z)
        self._spawn_heights()
        if self.flat is None:
            self._render(want_hits)
        else:
            L.orc_update_cameras(_p(self.cameras), _p(self.positions), _p(self.yaws),
                                 _p(self.pitches), B)
            if self.flat == "prebasis":
                L.orc_camera_basis(_p(self.cameras), _p(self.basis), B, self.render_width,
                                   self.render_height, self.fov_degrees)
                cams = self.basis
            else:
                cams = self.cameras
            r = render_flat(self.voxels, cams, self.render_width, self.render_height, self.flat,
                            self.max_steps, self.fov_degrees, want_hits, self.nthreads)
            if want_hits:
                self.obs_buffer[...], self.hits = r
            else:
                self.obs_buffer[...] = r
        self.step_count += 1
        self.step_count[self.do_reset.astype(bool)] = 1
        if self.do_reset.any():
            self._regen()
        return self.obs_buffer, self.rewards, self.do_reset.astype(bool).copy()

    def state(self):
        """Copy of every state array, keyed like MineEnv's attributes."""
        d = {k: getattr(self, k).copy() for k in
             ("voxels", "positions", "velocities", "yaws", "pitches", "on_ground",
              "step_count", "cameras", "rewards", "obs_buffer")}
        if self.basis is not None:
            d["basis"] = self.basis.copy()
        return d
