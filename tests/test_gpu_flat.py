"""GPU parity of the flat-colour observation variants (obs_dtype
uint8_minimal / uint8_prebasis / uint8_prebasis_full / uint8_smem /
uint8_fp16) and of the reference-named bindings that run them, against the
CPU oracle; the device binary16 arithmetic against the oracle's software
binary16."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from parity_util import Stats, compare_env, compare_exact, compare_obs, run_pair

pytestmark = pytest.mark.gpu

FLAT_MODES = ("uint8_minimal", "uint8_prebasis", "uint8_prebasis_full", "uint8_smem",
              "uint8_fp16")


def _envs(**kw):
    import minecu
    return minecu.MineEnv(device="cuda", **kw), oracle.OracleEnv(**kw)


def _break_heavy(steps, B, seed):
    g = torch.Generator().manual_seed(seed)
    u = torch.rand((steps, B), generator=g)
    move = torch.tensor([1, 2, 3, 4, 7, 8, 9, 10])[torch.randint(0, 8, (steps, B), generator=g)]
    anyid = torch.randint(0, 11, (steps, B), generator=g)
    return torch.where(u < 0.5, torch.full_like(anyid, 6), torch.where(u < 0.8, move, anyid)).numpy()


def _basis_equal(env, orc, where):
    if orc.basis is None:
        assert env.basis is None
        return
    compare_exact("basis", env.basis.cpu().numpy().view(np.int32), orc.basis.view(np.int32), where)


@pytest.mark.parametrize("mode", FLAT_MODES)
@pytest.mark.parametrize("ws,W,H", [(16, 32, 32), (32, 48, 40)])
def test_flat_env_matches_oracle(mode, ws, W, H):
    B, T = 96, 40
    env, orc = _envs(batch_size=B, world_size=ws, render_width=W, render_height=H,
                     obs_dtype=mode, episode_length=13)
    st = Stats()
    env.reset()
    orc.reset()
    compare_env(env, orc, st, f"{mode} reset")
    acts = _break_heavy(T, B, 21)
    for t in range(T):
        og, rg, dg = env.step(torch.as_tensor(acts[t]).cuda())
        oo, ro, do = orc.step(acts[t])
        where = f"{mode} ws{ws} step {t}"
        compare_exact("rewards", rg.cpu().numpy(), ro, where)
        compare_exact("dones", dg.cpu().numpy(), do, where)
        compare_env(env, orc, st, where)
        _basis_equal(env, orc, where)
    assert st.exact, st.summary()


@pytest.mark.parametrize("mode", ("uint8", "float32", "uint8_minimal", "uint8_prebasis"))
def test_images_wider_than_the_lds_tables(mode):
    """136 x 130 images: both sides above the 128-entry u and v tables, so u and
    v come from their formulas, the flat colours sit at the start of the item's
    LDS lead and the shaded renders keep no column terms; 17 tile columns
    (interleaved uint8 tile stores)."""
    B, T = 6, 14
    env, orc = _envs(batch_size=B, world_size=16, render_width=136, render_height=130,
                     obs_dtype=mode, episode_length=9)
    st = Stats()
    env.reset()
    orc.reset()
    compare_env(env, orc, st, f"{mode} reset")
    run_pair(env, orc, _break_heavy(T, B, 23), st, f"{mode} 136x130")
    _basis_equal(env, orc, mode)
    assert st.exact, st.summary()


@pytest.mark.parametrize("mode", ("uint8_minimal", "uint8_prebasis", "uint8_fp16"))
def test_flat_eye_outside_world_and_odd_image(mode):
    """9^3 world: the spawn eye (y 11.6) is above the world, so rays start
    outside it; width 17 takes the per-pixel store path."""
    env, orc = _envs(batch_size=24, world_size=9, render_width=17, render_height=11,
                     obs_dtype=mode, episode_length=20)
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, _break_heavy(30, 24, 4), st, f"{mode} 9^3")
    _basis_equal(env, orc, mode)
    assert st.exact, st.summary()


def test_prebasis_full_episode_convention():
    """episode_length None and 0 both mean no limit in uint8_prebasis_full."""
    import minecu
    for L in (None, 0):
        env = minecu.MineEnv(batch_size=4, world_size=8, render_width=8, render_height=8,
                             device="cuda", obs_dtype="uint8_prebasis_full", episode_length=L)
        env.reset()
        for _ in range(6):
            _, _, d = env.step(torch.zeros(4, dtype=torch.int64, device="cuda"))
            assert not d.any()
        assert env.step_count.cpu().tolist() == [6] * 4


def _ref_inputs(B, ws, seed):
    rng = np.random.default_rng(seed)
    vox = oracle.fresh_world(B, ws, ws, ws, 8, (ws // 2, ws // 2 + 2), (ws // 2 + 6, ws // 2 - 2))
    pos = np.stack([rng.uniform(1, ws - 1, B), rng.uniform(3, ws - 2, B),
                    rng.uniform(1, ws - 1, B)], 1).astype(np.float32)
    vel = np.zeros((B, 3), np.float32)
    vel[:, 1] = rng.normal(0, 2, B)
    yaw = rng.uniform(-4, 4, B).astype(np.float32)
    pitch = rng.uniform(-1.5, 1.5, B).astype(np.float32)
    og = (rng.random(B) < 0.5).astype(np.uint8)
    buttons = rng.integers(0, 2, (B, 8)).astype(np.int8)
    look = rng.normal(0, 0.3, (B, 2)).astype(np.float32)
    do_reset = (rng.random(B) < 0.3).astype(np.uint8)
    step_count = rng.integers(0, 9, B).astype(np.int32)
    return vox, pos, vel, yaw, pitch, og, buttons, look, do_reset, step_count


@pytest.mark.parametrize("fn", ["step_uint8_minimal", "step_uint8_prebasis",
                                "step_uint8_prebasis_full", "step_uint8_smem", "step_uint8_fp16"])
def test_reference_flat_step_bindings(fn):
    """The reference-named _C step bindings against the oracle's kernel
    sequence: decode -> physics -> raycast -> reset -> (episode update) ->
    cameras -> (basis) -> flat render."""
    from minecu import _C
    B, ws, W, H, EL = 40, 16, 32, 24, 5
    vox, pos, vel, yaw, pitch, og, buttons, look, do_reset, sc = _ref_inputs(B, ws, 8)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g = dict(voxels=t(vox), positions=t(pos), velocities=t(vel), yaws=t(yaw), pitches=t(pitch),
             on_ground=t(og.astype(bool)))
    dec = [torch.zeros(B, device="cuda") for _ in range(4)] + \
          [torch.zeros(B, dtype=torch.bool, device="cuda") for _ in range(3)] + \
          [torch.ones(B, device="cuda")]
    cams = torch.zeros((B, 5), device="cuda")
    basis = torch.zeros((B, 14), device="cuda")
    obs = torch.zeros((B, H, W, 3), dtype=torch.uint8, device="cuda")
    rew = torch.zeros(B, device="cuda")
    dr = t(do_reset.astype(bool))
    scd = t(sc)
    kw = dict(target_block=3, spawn_x=8.0, spawn_y=10.0, spawn_z=8.0)
    head = (g["voxels"], g["positions"], g["velocities"], g["yaws"], g["pitches"],
            g["on_ground"], t(buttons), t(look), dr)
    if fn in ("step_uint8_minimal", "step_uint8_fp16"):
        getattr(_C, fn)(*head, *dec, cams, obs, rew, W, H, **kw)
    elif fn == "step_uint8_prebasis_full":
        _C.step_uint8_prebasis_full(*head, scd, *dec, cams, basis, obs, rew, W, H,
                                    episode_length=EL, **kw)
    else:
        getattr(_C, fn)(*head, *dec, cams, basis, obs, rew, W, H, **kw)
    L = oracle.lib()
    P = oracle._p
    if fn == "step_uint8_prebasis_full":
        do_reset = (sc >= EL).astype(np.uint8)
    f = [np.zeros(B, np.float32) for _ in range(4)]
    bflags = [np.zeros(B, np.uint8) for _ in range(3)]
    sm = np.zeros(B, np.float32)
    L.orc_decode_actions(P(buttons), P(look), *(P(a) for a in f), *(P(a) for a in bflags), P(sm), B, 1.5)
    L.orc_physics(P(pos), P(vel), P(yaw), P(pitch), P(og), *(P(a) for a in f), P(bflags[0]), P(sm),
                  B, ws, ws, ws, 0.05, -20.0, 4.0, 8.0)
    r = np.zeros(B, np.float32)
    L.orc_raycast_interact(P(vox), P(pos), P(yaw), P(pitch), P(bflags[1]), P(bflags[2]), P(r), B,
                           ws, ws, ws, 3, 1.0)
    L.orc_reset(P(pos), P(vel), P(yaw), P(pitch), P(og), P(do_reset), B, 8.0, 10.0, 8.0)
    c = np.zeros((B, 5), np.float32)
    L.orc_update_cameras(P(c), P(pos), P(yaw), P(pitch), B)
    if fn in ("step_uint8_minimal", "step_uint8_fp16"):
        ref_obs = oracle.render_flat(vox, c, W, H, "minimal" if fn == "step_uint8_minimal" else "fp16")
    else:
        ref_basis = oracle.camera_basis(c, W, H)
        ref_obs = oracle.render_flat(vox, ref_basis, W, H, "prebasis")
    torch.cuda.synchronize()
    where = fn
    compare_exact("voxels", g["voxels"].cpu().numpy(), vox, where)
    for k, ref in (("positions", pos), ("velocities", vel), ("yaws", yaw), ("pitches", pitch)):
        compare_exact(k, g[k].cpu().numpy().view(np.int32), ref.view(np.int32), where)
    compare_exact("rewards", rew.cpu().numpy(), r, where)
    compare_exact("cameras", cams.cpu().numpy().view(np.int32), c.view(np.int32), where)
    if fn not in ("step_uint8_minimal", "step_uint8_fp16"):
        compare_exact("basis", basis.cpu().numpy().view(np.int32), ref_basis.view(np.int32), where)
    if fn == "step_uint8_prebasis_full":
        compare_exact("do_reset", dr.cpu().numpy().astype(np.uint8), do_reset, where)
        compare_exact("step_count", scd.cpu().numpy(), np.where(do_reset != 0, 1, sc + 1), where)
    st = Stats()
    compare_obs(obs.cpu().numpy(), ref_obs, st, f"{fn} obs")
    assert st.exact, st.summary()


def test_basis_and_flat_renders():
    """precompute_camera_basis, render_uint8_prebasis (from a stored basis,
    including one edited by hand) and the minimal camera render."""
    from minecu import _C, _lib
    B, ws, W, H = 12, 16, 40, 24
    rng = np.random.default_rng(3)
    vox = np.where(rng.random((B, ws, ws, ws)) < 0.06, rng.integers(0, 13, (B, ws, ws, ws)), -1).astype(np.int8)
    cams = np.zeros((B, 5), np.float32)
    cams[:, :3] = rng.uniform(-2, ws + 2, (B, 3))
    cams[:, 3] = rng.uniform(-3.2, 3.2, B)
    cams[:, 4] = rng.uniform(-1.5, 1.5, B)
    gv, gc = torch.from_numpy(vox).cuda(), torch.from_numpy(cams).cuda()
    gb = torch.zeros((B, 14), device="cuda")
    _C.precompute_camera_basis(gc, gb, W, H, 70.0)
    ref_b = oracle.camera_basis(cams, W, H, 70.0)
    compare_exact("basis", gb.cpu().numpy().view(np.int32), ref_b.view(np.int32), "basis")
    ref_b[:, 11] *= 0.75  # per-env scale_u read back from the basis
    gb.copy_(torch.from_numpy(ref_b))
    out = torch.zeros((B, H, W, 3), dtype=torch.uint8, device="cuda")
    _C.render_uint8_prebasis(gv, gb, out, 48)
    st = Stats()
    compare_obs(out.cpu().numpy(), oracle.render_flat(vox, ref_b, W, H, "prebasis", 48), st, "prebasis render")
    out.zero_()
    L = _lib.load()
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    _lib.check(L.mc_launch_render_uint8_minimal(p(gv), p(gc), p(out), B, ws, W, H, 64, 32.0, 70.0,
                                                torch.cuda.current_stream().cuda_stream))
    compare_obs(out.cpu().numpy(), oracle.render_flat(vox, cams, W, H, "minimal"), st, "minimal render")
    out.zero_()
    _lib.check(L.mc_launch_render_uint8_fp16(p(gv), p(gc), p(out), B, ws, W, H, 64, 32.0, 70.0,
                                             torch.cuda.current_stream().cuda_stream))
    compare_obs(out.cpu().numpy(), oracle.render_flat(vox, cams, W, H, "fp16"), st, "fp16 render")
    assert st.exact, st.summary()


def test_device_binary16_matches_software_binary16():
    """The GPU's native binary16 (conversion, sum, order) equals the
    oracle's software binary16 on 4M float pairs spanning the binary16
    range, including ties, subnormals and the overflow boundary."""
    from minecu import _lib
    from test_oracle_flat import _f16_cases
    a, b = _f16_cases(1 << 22, 11)
    b = np.abs(b)  # the DDA adds non-negative deltas
    a = np.abs(a)
    ha, s, lt = oracle.f16_ops(a, b)
    ga, gb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    gha = torch.empty(a.size, dtype=torch.int16, device="cuda")
    gs = torch.empty_like(gha)
    glt = torch.empty(a.size, dtype=torch.uint8, device="cuda")
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    _lib.check(_lib.load().mc_debug_f16(p(ga), p(gb), p(gha), p(gs), p(glt), a.size,
                                        torch.cuda.current_stream().cuda_stream))
    assert np.array_equal(gha.cpu().numpy().view(np.uint16), ha)
    assert np.array_equal(gs.cpu().numpy().view(np.uint16), s)
    assert np.array_equal(glt.cpu().numpy(), lt)


def test_flat_groups_in_a_batch_equal_alone():
    """A flat group, a prebasis group and a shaded group stepped together in
    one batched launch equal each stepped alone."""
    import minecu
    kws = [dict(batch_size=33, world_size=16, render_width=32, render_height=32,
                obs_dtype="uint8_minimal", episode_length=7),
           dict(batch_size=20, world_size=32, render_width=24, render_height=16,
                obs_dtype="uint8_prebasis", episode_length=9),
           dict(batch_size=50, world_size=16, render_width=16, render_height=16,
                obs_dtype="uint8", episode_length=None)]
    mk = lambda kw: minecu.MineEnv(device="cuda", **kw)
    alone, mixed = [mk(k) for k in kws], [mk(k) for k in kws]
    for e in alone + mixed:
        e.reset()
    acts = [torch.from_numpy(_break_heavy(15, k["batch_size"], 30 + i)).cuda() for i, k in enumerate(kws)]
    batch = minecu.EnvBatch()
    for t in range(15):
        batch.step(mixed, [a[t] for a in acts])
        for i, e in enumerate(alone):
            e.step(acts[i][t])
    for a, m in zip(alone, mixed):
        for k in a._state_keys():
            assert np.array_equal(getattr(a, k).cpu().numpy().view(np.uint8),
                                  getattr(m, k).cpu().numpy().view(np.uint8)), k
    batch.close()
