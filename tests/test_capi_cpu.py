"""The C ABI without a GPU: the library loads, exports every symbol the header
declares, and rejects bad arguments with a status and a message before any
device work.  Also the Python-level argument errors the reference raises."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "minecu_hip.h")


def _lib():
    from minecu import _lib
    return _lib, _lib.load()


def test_library_exports_every_declared_symbol():
    _l, L = _lib()
    text = open(HEADER).read()
    declared = sorted(set(re.findall(r"^(?:int|const char\*)\s+(mc_[a-z0-9_]+)\(", text, re.M)))
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers exactly the declared API
    assert sorted(_l.SIGNATURES) == declared


def test_struct_layouts_match_header():
    _l, _ = _lib()
    # 31 four-byte fields (terrain, terrain_amplitude, collision last), 20 pointers (the
    # decode buffers last)
    assert ctypes.sizeof(_l.McConfig) == 31 * 4
    assert [f[0] for f in _l.McConfig._fields_][-3:] == ["terrain", "terrain_amplitude", "collision"]
    assert ctypes.sizeof(_l.McState) == 20 * 8
    assert [f[0] for f in _l.McState._fields_][-1] == "speed_mult"
    text = open(HEADER).read()
    state = text[text.index("typedef struct mc_state"):text.index("} mc_state;")]
    assert len(re.findall(r"\*\s*\w+;", state)) == 20


def _err(L):
    return L.mc_last_error().decode()


def test_null_and_range_errors():
    _l, L = _lib()
    assert L.mc_step(None, None, None, 0, None, None, None, None) == _l.MC_E_INVALID
    assert "config" in _err(L)
    assert L.mc_generate_world(None, 1, 8, 8, 8, 8, 0, None) == _l.MC_E_INVALID
    v = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    assert L.mc_generate_world(v, 0, 8, 8, 8, 8, 0, None) == _l.MC_E_INVALID
    # any world runs (from global memory when it does not fit LDS) up to 512 per axis
    assert L.mc_render(v, v, v, 0, 4, 600, 64, 64, 8, 8, 64, 32.0, 70.0, None) == _l.MC_E_UNSUPPORTED
    assert "at most 512" in _err(L)
    assert L.mc_render(v, v, v, 7, 4, 8, 8, 8, 8, 8, 64, 32.0, 70.0, None) == _l.MC_E_INVALID
    assert L.mc_render(v, v, v, 0, 4, 8, 8, 8, 0, 8, 64, 32.0, 70.0, None) == _l.MC_E_INVALID
    assert L.mc_env_create(None, None, None) == _l.MC_E_INVALID
    out = ctypes.c_void_p()
    cfg = _l.McConfig(batch_size=4, world_x=8, world_y=8, world_z=8, width=8, height=8,
                      obs_mode=0)
    st = _l.McState()
    assert L.mc_env_create(ctypes.byref(cfg), ctypes.byref(st), ctypes.byref(out)) == _l.MC_E_INVALID
    assert "non-NULL" in _err(L) and not out.value
    assert L.mc_env_step(None, v, 0, v, None, None) == _l.MC_E_INVALID
    # mc_step_static: a batch beyond its scratch limit is refused before any device call
    big = _l.McConfig(batch_size=_l.MC_STEP_STATIC_MAX_BATCH + 1, world_x=16, world_y=16,
                      world_z=16, width=8, height=8)
    assert L.mc_step_static(ctypes.byref(big), ctypes.byref(st), v, 1, None, None,
                            None) == _l.MC_E_INVALID
    assert "scratch limit" in _err(L)
    assert L.mc_debug_sincosf(None, None, None, 4, None) == _l.MC_E_INVALID
    # mc_render takes obs modes 0..4 (flat variants 2..4); the planar modes 5..6 render
    # from MineEnv state only, 7 is no mode; NULL basis is refused
    for mode in (5, 6, 7):
        assert L.mc_render(v, v, v, mode, 4, 8, 8, 8, 8, 8, 64, 32.0, 70.0, None) == _l.MC_E_INVALID
    assert L.mc_launch_precompute_camera_basis(None, v, 4, 8, 8, 70.0, None) == _l.MC_E_INVALID
    assert L.mc_launch_render_uint8_prebasis(v, None, v, 4, 8, 8, 8, 64, None) == _l.MC_E_INVALID
    args = [v] * 21
    args[18] = None  # basis
    assert L.mc_launch_step_uint8_prebasis(*args, 4, 8, 8, 8, 8, 8, 64, 32.0, 70.0, 0.05, -20.0,
                                           4.0, 1.5, 8.0, 3, 1.0, 4.0, 10.0, 4.0,
                                           None) == _l.MC_E_INVALID
    assert "basis" in _err(L)


def test_reference_step_rejects_partial_decode_buffers():
    _l, L = _lib()
    v = ctypes.c_void_p(0x1000)
    args = [v] * 20
    args[9] = None  # forward_in NULL while the other decode buffers are set
    st = L.mc_launch_step_uint8(*args, 4, 8, 8, 8, 64, 32.0, 70.0, 0.05, -20.0, 4.0, 1.5, 8.0,
                                3, 1.0, 4.0, 10.0, 4.0, None)
    assert st == _l.MC_E_INVALID and "decode buffers" in _err(L)


def test_check_raises_with_message():
    _l, L = _lib()
    L.mc_env_create(None, None, None)
    with pytest.raises(_l.MineCuError, match="invalid argument"):
        _l.check(_l.MC_E_INVALID, "x")


def test_mineenv_argument_errors():
    import minecu
    with pytest.raises(AssertionError):
        minecu.MineEnv(batch_size=4, device="cpu")
    with pytest.raises(AssertionError):
        minecu.MineEnv(batch_size=4, obs_dtype="int16")



def test_config_presets_match_reference():
    from minecu.config import DEFAULT_ENV, FAST_ENV, QUALITY_ENV, EnvConfig
    assert EnvConfig().episode_length == 60 and EnvConfig().world_size == 32
    assert (FAST_ENV.batch_size, FAST_ENV.render_width, FAST_ENV.max_steps) == (32768, 32, 32)
    assert (QUALITY_ENV.render_height, QUALITY_ENV.view_distance) == (128, 64.0)
    assert DEFAULT_ENV.memory_bytes() == EnvConfig().memory_bytes()


def test_terrain_and_collision_settings_are_validated_before_any_launch():
    _l, L = _lib()
    v = ctypes.c_void_p(16)  # never dereferenced: every case fails validation first

    def cfg(**kw):
        c = _l.McConfig()
        c.batch_size, c.world_x, c.world_y, c.world_z = 2, 16, 16, 16
        c.ground_height = 8
        for k, val in kw.items():
            setattr(c, k, val)
        return c

    for kw, code in [(dict(terrain=2), _l.MC_E_INVALID),
                     (dict(terrain=1, terrain_amplitude=9), _l.MC_E_INVALID),
                     (dict(terrain=1, terrain_amplitude=-1), _l.MC_E_INVALID),
                     (dict(terrain=1, terrain_amplitude=7), _l.MC_E_INVALID),  # 8 - 7 < 2
                     (dict(terrain=1, terrain_amplitude=3, obs_mode=5), _l.MC_E_UNSUPPORTED),
                     (dict(collision=2), _l.MC_E_INVALID),
                     (dict(collision=1, obs_mode=6), _l.MC_E_UNSUPPORTED)]:
        c = cfg(**kw)
        assert L.mc_regen_worlds(ctypes.byref(c), v, None, None) == code, kw
        assert _err(L), kw
