"""mc_step_static (the allocation-free step) against mc_step with a caller
workspace: identical state, observations, rewards and dones, bit for bit,
on consecutive calls -- the second group smaller than the first -- through
done steps, world regeneration and the full-height pass of deferred envs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

STATE = ("voxels", "positions", "velocities", "yaws", "pitches", "on_ground", "step_count",
         "do_reset", "cameras", "obs_buffer", "rewards", "forward_in", "strafe_in",
         "delta_yaw_in", "delta_pitch_in", "jump_in", "do_break", "do_place", "speed_mult")


def _twins(**kw):
    import minecu
    a = minecu.MineEnv(device="cuda", **kw)
    b = minecu.MineEnv(device="cuda", **kw)
    a.reset()
    b.reset()
    # Solids above the staged layers in a few envs, and agents placed high,
    # so the full-height pass runs from the deferred-env list.
    for e in (a, b):
        Y = e.world_y
        e.voxels[0:3, Y - 3, :, :] = 9
        e.positions[3:5, 1] = float(Y - 6)
    torch.cuda.synchronize()
    return a, b


def _run(a, b, steps, seed):
    from minecu import _lib
    L = _lib.load()
    B = a.batch_size
    ws = torch.zeros(_lib.MC_WORKSPACE_BYTES, dtype=torch.uint8, device="cuda")
    g = torch.Generator().manual_seed(seed)
    stream = torch.cuda.current_stream().cuda_stream
    for t in range(steps):
        acts = torch.randint(0, 11, (B,), generator=g).cuda()
        da = torch.empty(B, dtype=torch.bool, device="cuda")
        db = torch.empty(B, dtype=torch.bool, device="cuda")
        _lib.check(L.mc_step(ctypes.byref(a._cfg), ctypes.byref(a._st),
                             ctypes.c_void_p(acts.data_ptr()), _lib.MC_ACT_DISCRETE, None,
                             ctypes.c_void_p(da.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                             stream), "mc_step")
        _lib.check(L.mc_step_static(ctypes.byref(b._cfg), ctypes.byref(b._st),
                                    ctypes.c_void_p(acts.data_ptr()), _lib.MC_ACT_DISCRETE, None,
                                    ctypes.c_void_p(db.data_ptr()), stream), "mc_step_static")
        torch.cuda.synchronize()
        assert torch.equal(da, db), f"step {t}: dones"
        for k in STATE:
            x, y = getattr(a, k), getattr(b, k)
            if x.dtype.is_floating_point:
                x, y = x.view(torch.int32), y.view(torch.int32)
            assert torch.equal(x, y), f"step {t}: {k} differs"
    return int(a.do_reset.sum().item())


def test_static_step_equals_workspace_step():
    # First call sequence: 300 envs, 32^3 (12 of 18 staged layers + deferral),
    # uint8, done steps at 4 and 8.
    a, b = _twins(batch_size=300, world_size=32, render_width=32, render_height=24,
                  obs_dtype="uint8", episode_length=4)
    _run(a, b, 10, 1)
    # Second, smaller: 24 envs of 48^3 (the 512-thread full-height pass),
    # float32 observations.
    a, b = _twins(batch_size=24, world_size=48, render_width=24, render_height=16,
                  obs_dtype="float32", episode_length=3)
    _run(a, b, 7, 2)


def test_static_step_refuses_oversized_batch():
    from minecu import _lib
    L = _lib.load()
    cfg = _lib.McConfig(batch_size=_lib.MC_STEP_STATIC_MAX_BATCH + 1, world_x=16, world_y=16,
                        world_z=16, width=8, height=8)
    st = _lib.McState()
    assert L.mc_step_static(ctypes.byref(cfg), ctypes.byref(st), None, 1, None, None,
                            None) == _lib.MC_E_INVALID
    assert b"scratch limit" in L.mc_last_error()
