"""GPU parity: the gfx950 HIP step (through libminecu_hip.so) against the CPU
oracle on identical seeded inputs, every step, every state tensor."""
import numpy as np
import pytest
import torch

import oracle
from parity_util import Stats, compare_env, compare_exact, compare_obs, run_pair

pytestmark = pytest.mark.gpu


def _envs(**kw):
    import minecu
    look_delta = kw.pop("look_delta", 0.1)
    env = minecu.MineEnv(device="cuda", look_delta=look_delta, **kw)
    orc = oracle.OracleEnv(look_delta=look_delta, **kw)
    return env, orc


def _discrete(steps, B, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randint(0, 11, (steps, B), generator=g).numpy()


def _break_heavy(steps, B, seed):
    """Per env and step: break with p=.5, move/turn ids {1,2,3,4,7,8,9,10}
    with p=.3, uniform over 0..10 with p=.2."""
    g = torch.Generator().manual_seed(seed)
    u = torch.rand((steps, B), generator=g)
    move = torch.tensor([1, 2, 3, 4, 7, 8, 9, 10])[torch.randint(0, 8, (steps, B), generator=g)]
    anyid = torch.randint(0, 11, (steps, B), generator=g)
    a = torch.where(u < 0.5, torch.full_like(anyid, 6), torch.where(u < 0.8, move, anyid))
    return a.numpy()


def test_device_sincos_matches_host():
    from minecu import _lib
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    x = np.concatenate([x, rng.uniform(-300, 300, 1 << 18).astype(np.float32),
                        np.float32([0.0, -0.0, 1e-45, 3.14159265, 1.5707964, 1e30, -1e38])])
    xs, xc = oracle.sincosf(x)
    xg = torch.from_numpy(x).cuda()
    sg, cg = torch.empty_like(xg), torch.empty_like(xg)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().mc_debug_sincosf(xg.data_ptr(), sg.data_ptr(), cg.data_ptr(), x.size, st))
    torch.cuda.synchronize()
    assert np.array_equal(sg.cpu().numpy().view(np.uint32), xs.view(np.uint32))
    assert np.array_equal(cg.cpu().numpy().view(np.uint32), xc.view(np.uint32))


def test_world_generation_matches_oracle():
    env, orc = _envs(batch_size=3, world_size=16, render_width=8, render_height=8,
                     obs_dtype="uint8")
    torch.cuda.synchronize()
    compare_exact("voxels", env.voxels.cpu().numpy(), orc.voxels, "init")


@pytest.mark.parametrize("ws", [8, 16, 32])
def test_reset_matches_oracle(ws):
    env, orc = _envs(batch_size=16, world_size=ws, render_width=32, render_height=32,
                     obs_dtype="uint8")
    st = Stats()
    env.reset()
    orc.reset()
    compare_env(env, orc, st, f"reset ws={ws}")
    assert st.exact, st.summary()


def test_c1_small_plain_step():
    """B 256, 16^3, 64x64 uint8, 100 discrete steps, torch seed 0."""
    env, orc = _envs(batch_size=256, world_size=16, render_width=64, render_height=64,
                     obs_dtype="uint8", seed=0)
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, _discrete(100, 256, 0), st, "C1")
    print(st.summary())


def test_c2_shape_with_episode_regen():
    """16^3, 32x32 uint8, episode_length 60: two done/regeneration cycles."""
    env, orc = _envs(batch_size=512, world_size=16, render_width=32, render_height=32,
                     obs_dtype="uint8", episode_length=60)
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, _discrete(125, 512, 1), st, "C2-subset", check_every=5)
    print(st.summary())


def test_pockets_are_exact():
    """16^3, 32x32 uint8: agents move and look around while falling, land
    with the eye inside dirt, break the block at the eye (for most of them a
    one-voxel air pocket whose six neighbours are solid, so every ray ends on
    its first step), look around from inside it, then dig on (open pockets
    and tunnels).  Every image, and all state, equals the oracle bit for
    bit."""
    B = 128
    env, orc = _envs(batch_size=B, world_size=16, render_width=32, render_height=32,
                     obs_dtype="uint8", episode_length=None)
    g = torch.Generator().manual_seed(11)
    moves = torch.tensor([1, 2, 3, 4, 7, 8, 9, 10])[torch.randint(0, 8, (16, B), generator=g)]
    looks = torch.tensor([7, 8, 9, 10])[torch.randint(0, 4, (16, B), generator=g)]
    dig = torch.where(torch.rand((12, B), generator=g) < 0.5, torch.full((12, B), 6),
                      torch.tensor([7, 8, 9, 10])[torch.randint(0, 4, (12, B), generator=g)])
    acts = torch.cat([moves, torch.full((1, B), 6), looks, dig]).numpy()
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, acts, st, "sealed")
    assert st.exact, st.summary()


def test_c3_shape_break_heavy():
    """32^3, 64x64 uint8, 50% break actions, episode_length 60."""
    env, orc = _envs(batch_size=64, world_size=32, render_width=64, render_height=64,
                     obs_dtype="uint8", episode_length=60)
    st = Stats()
    env.reset()
    orc.reset()
    acts = _break_heavy(130, 64, 2)
    run_pair(env, orc, acts, st, "C3-subset", check_every=5)
    assert orc.rewards is not None
    print(st.summary())


def test_float32_observations():
    env, orc = _envs(batch_size=64, world_size=16, render_width=32, render_height=32,
                     obs_dtype="float32", episode_length=30)
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, _discrete(70, 64, 5), st, "f32")
    print(st.summary())


def test_float32_constant_images_of_odd_size():
    """17 x 11 float32 images (187 pixels, not a multiple of 4): constant images
    of an eye inside a block take the per-pixel path instead of the 16-byte
    fill; landed agents sit in dirt for most of the episode."""
    env, orc = _envs(batch_size=32, world_size=16, render_width=17, render_height=11,
                     obs_dtype="float32", episode_length=25)
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, _discrete(40, 32, 8), st, "f32 odd")
    assert st.exact, st.summary()


def test_buttons_and_look_continuous():
    """Multi-hot buttons (including non-0/1 int8 values) and continuous look
    with large yaw drift, non-square 48x40 image, 32^3 world, placement."""
    B, T = 96, 60
    rng = np.random.default_rng(11)
    buttons = rng.choice(np.array([0, 0, 0, 1, 1, 2, -1], np.int8), size=(T, B, 8))
    look = rng.normal(0, 0.6, size=(T, B, 2)).astype(np.float32)
    look[:, :8, 0] += 40.0  # drive a few yaws far from zero
    env, orc = _envs(batch_size=B, world_size=32, render_width=48, render_height=40,
                     obs_dtype="uint8", episode_length=45, target_block=-1)
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, buttons, st, "buttons", look=look)
    print(st.summary())


def test_buttons_from_unaligned_buffer():
    """Buttons whose [B, 8] buffer starts one byte past an aligned address
    (the step reads them byte by byte then) give the same states as the same
    buttons from an aligned buffer, and the oracle's."""
    import minecu
    B, T = 40, 24
    rng = np.random.default_rng(23)
    buttons = rng.choice(np.array([0, 0, 1, 1, 3, -1], np.int8), size=(T, B, 8))
    look = rng.normal(0, 0.5, size=(T, B, 2)).astype(np.float32)
    kw = dict(batch_size=B, world_size=16, render_width=32, render_height=24, obs_dtype="uint8",
              episode_length=10)
    a = minecu.MineEnv(device="cuda", **kw)
    u = minecu.MineEnv(device="cuda", **kw)
    orc = oracle.OracleEnv(**kw)
    a.reset()
    u.reset()
    orc.reset()
    raw = torch.zeros(B * 8 + 1, dtype=torch.int8, device="cuda")
    bt_u = raw[1:].view(B, 8)
    assert bt_u.data_ptr() % 4 != 0
    for t in range(T):
        bt_u.copy_(torch.from_numpy(buttons[t]))
        lk = torch.from_numpy(look[t]).cuda()
        oa, ra, da = a.step(torch.from_numpy(buttons[t]).cuda(), lk)
        ou, ru, du = u.step(bt_u, lk)
        oo, ro, do = orc.step(buttons[t], look[t])
        for name, x, y in (("obs", ou, oa), ("rewards", ru, ra), ("dones", du, da)):
            assert np.array_equal(x.cpu().numpy(), y.cpu().numpy()), f"t{t} {name}"
        assert np.array_equal(ou.cpu().numpy(), oo), f"t{t} obs vs oracle"
        assert np.array_equal(u.positions.cpu().numpy().view(np.int32),
                              orc.positions.view(np.int32)), f"t{t} positions vs oracle"
        assert np.array_equal(u.voxels.cpu().numpy(), orc.voxels), f"t{t} voxels vs oracle"


@pytest.mark.parametrize("max_steps", [8, 45, 46])
def test_step_limit_boundaries(max_steps):
    """16^3 (X + Y + Z - 3 = 45): with max_steps 46 no in-world ray can reach
    the step limit, at 45 the longest in-world path just fits, and at 8 the
    limit cuts the falling agents' long rays.  Bit-exact in every case."""
    env, orc = _envs(batch_size=32, world_size=16, render_width=32, render_height=32,
                     obs_dtype="uint8", episode_length=None, max_steps=max_steps)
    st = Stats()
    env.reset()
    orc.reset()
    compare_env(env, orc, st, f"reset max_steps={max_steps}")
    run_pair(env, orc, _discrete(20, 32, 13), st, f"max_steps={max_steps}")
    assert st.exact, st.summary()


@pytest.mark.parametrize("dims,img", [((24, 16, 20), (30, 20)), ((9, 9, 9), (17, 11)),
                                      ((16, 16, 16), (20, 12))])
def test_odd_shapes(dims, img):
    """Non-cubic world (Z not a multiple of 16), world size not a multiple of
    16 bytes, image width not a multiple of 4, image sides that end in half
    an 8x8 tile."""
    X, Y, Z = dims
    W, H = img
    env, orc = _envs(batch_size=24, world_x=X, world_y=Y, world_z=Z, render_width=W,
                     render_height=H, obs_dtype="uint8", episode_length=20)
    st = Stats()
    env.reset()
    orc.reset()
    compare_env(env, orc, st, "reset")
    run_pair(env, orc, _break_heavy(45, 24, 9), st, f"odd {dims} {img}")
    print(st.summary())


@pytest.mark.parametrize("dims,img", [((48, 48, 48), (32, 32)), ((56, 44, 60), (40, 24))])
def test_large_lds_worlds(dims, img):
    """Worlds above 64 KiB staged in one workgroup's LDS: 48^3 (the
    reference's largest benchmarked world, 110,592 B) and a non-cubic world
    near the 152 KiB budget, through falling, breaking and regeneration."""
    X, Y, Z = dims
    W, H = img
    env, orc = _envs(batch_size=6, world_x=X, world_y=Y, world_z=Z, render_width=W,
                     render_height=H, obs_dtype="uint8", episode_length=25)
    st = Stats()
    env.reset()
    orc.reset()
    compare_env(env, orc, st, "reset")
    run_pair(env, orc, _break_heavy(55, 6, 12), st, f"large {dims} {img}", check_every=3)
    print(st.summary())


@pytest.mark.parametrize("obs_dtype", ["uint8", "uint8_prebasis"])
def test_staged_height_deferral(obs_dtype):
    """Worlds above the LDS-staged layers: envs with solids up there (a ceiling
    slab over the view, one lone block) or with agents standing high take the
    full-height pass; every env equals the oracle on every step, through done
    and regeneration steps, with batched launches of the same group in between
    (they advance the episode slot without the staged path)."""
    import minecu
    B, T, Y = 16, 36, 32
    env, orc = _envs(batch_size=B, world_size=Y, render_width=32, render_height=24,
                     obs_dtype=obs_dtype, episode_length=12)
    st = Stats()
    env.reset()
    orc.reset()
    torch.cuda.synchronize()
    vox = env.voxels.cpu().numpy()
    pos = env.positions.cpu().numpy()
    vox[0:4, 24, :, :] = 9          # cobblestone ceiling over envs 0-3
    vox[4, 30, 16, 20] = 8          # one brick near the top of env 4
    pos[5:8, 1] = 26.0              # agents of envs 5-7 start high and fall
    env.voxels.copy_(torch.from_numpy(vox))
    env.positions.copy_(torch.from_numpy(pos))
    orc.voxels[...] = vox
    orc.positions[...] = pos
    acts = _break_heavy(T, B, 21)
    batch = minecu.EnvBatch()
    for t in range(T):
        a = acts[t]
        if t % 4 == 2:
            og, rg, dg = batch.step([env], [torch.as_tensor(a).cuda()])[0]
        else:
            og, rg, dg = env.step(torch.as_tensor(a).cuda())
        oo, ro, do = orc.step(np.asarray(a))
        compare_exact("rewards", rg.cpu().numpy(), ro, f"staged step {t}")
        compare_exact("dones", dg.cpu().numpy(), do, f"staged step {t}")
        compare_env(env, orc, st, f"staged step {t}")
    batch.close()
    print(st.summary())


def test_partial_reset_mask():
    env, orc = _envs(batch_size=32, world_size=16, render_width=32, render_height=32,
                     obs_dtype="uint8", episode_length=25)
    st = Stats()
    env.reset()
    orc.reset()
    run_pair(env, orc, _break_heavy(12, 32, 3), st, "pre-mask")
    mask = np.zeros(32, bool)
    mask[::3] = True
    env.reset(torch.from_numpy(mask).cuda())
    orc.reset(mask)
    compare_env(env, orc, st, "after partial reset")
    run_pair(env, orc, _break_heavy(30, 32, 4), st, "post-mask")
    empty = np.zeros(32, bool)
    env.reset(torch.from_numpy(empty).cuda())
    orc.reset(empty)
    compare_env(env, orc, st, "after empty-mask reset")
    print(st.summary())


def test_render_only_matches_oracle():
    """_C.render_uint8 / _C.render on arbitrary cameras (inside, outside and
    above the world, looking every way)."""
    from minecu import _C
    B, ws = 40, 16
    rng = np.random.default_rng(3)
    vox = oracle.fresh_world(B, ws, ws, ws, 8, (8, 10), (14, 6))
    # carve random holes so rays travel
    holes = rng.random(vox.shape) < 0.3
    vox[holes] = -1
    cams = np.zeros((B, 5), np.float32)
    cams[:, :3] = rng.uniform(-6, ws + 6, (B, 3))
    cams[:, 3] = rng.uniform(-7, 7, B)
    cams[:, 4] = rng.uniform(-1.56, 1.56, B)
    vg = torch.from_numpy(vox).cuda()
    cg = torch.from_numpy(cams).cuda()
    for u8 in (True, False):
        out = torch.empty((B, 24, 36, 3), dtype=torch.uint8 if u8 else torch.float32, device="cuda")
        (_C.render_uint8 if u8 else _C.render)(vg, cg, out, max_steps=40, view_distance=20.0,
                                               fov_degrees=80.0)
        ref = oracle.render(vox, cams, 36, 24, 40, 20.0, 80.0, out_u8=u8)
        st = Stats()
        compare_obs(out.cpu().numpy(), ref, st, f"render u8={u8}")
        assert st.exact, st.summary()


def test_reference_step_binding_with_mask():
    """_C.step_uint8 (launch_step_uint8 semantics: done mask given, no
    episode bookkeeping) against the oracle's kernel sequence."""
    from minecu import _C
    B, ws, W, H = 48, 16, 32, 32
    rng = np.random.default_rng(5)
    vox = oracle.fresh_world(B, ws, ws, ws, 8, (8, 10), (14, 6))
    pos = np.stack([rng.uniform(1, 15, B), rng.uniform(3, 12, B), rng.uniform(1, 15, B)], 1).astype(np.float32)
    vel = np.zeros((B, 3), np.float32)
    vel[:, 1] = rng.normal(0, 2, B)
    yaw = rng.uniform(-4, 4, B).astype(np.float32)
    pitch = rng.uniform(-1.5, 1.5, B).astype(np.float32)
    og = (rng.random(B) < 0.5).astype(np.uint8)
    buttons = rng.integers(0, 2, (B, 8)).astype(np.int8)
    look = rng.normal(0, 0.3, (B, 2)).astype(np.float32)
    do_reset = (rng.random(B) < 0.3).astype(np.uint8)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g = dict(voxels=t(vox), positions=t(pos), velocities=t(vel), yaws=t(yaw), pitches=t(pitch),
             on_ground=t(og.astype(bool)))
    dec = [torch.zeros(B, device="cuda") for _ in range(4)] + \
          [torch.zeros(B, dtype=torch.bool, device="cuda") for _ in range(3)] + \
          [torch.ones(B, device="cuda")]
    cams = torch.zeros((B, 5), device="cuda")
    obs = torch.zeros((B, H, W, 3), dtype=torch.uint8, device="cuda")
    rew = torch.zeros(B, device="cuda")
    _C.step_uint8(g["voxels"], g["positions"], g["velocities"], g["yaws"], g["pitches"],
                  g["on_ground"], t(buttons), t(look), t(do_reset.astype(bool)), *dec, cams, obs,
                  rew, W, H, target_block=3, spawn_x=8.0, spawn_y=10.0, spawn_z=8.0)
    # oracle: decode -> physics -> raycast -> reset -> cameras -> render
    L = oracle.lib()
    P = oracle._p
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
    ref_obs = oracle.render(vox, c, W, H)
    torch.cuda.synchronize()
    compare_exact("voxels", g["voxels"].cpu().numpy(), vox, "step_uint8")
    compare_exact("positions", g["positions"].cpu().numpy().view(np.int32), pos.view(np.int32), "step_uint8")
    compare_exact("velocities", g["velocities"].cpu().numpy().view(np.int32), vel.view(np.int32), "step_uint8")
    compare_exact("yaws", g["yaws"].cpu().numpy().view(np.int32), yaw.view(np.int32), "step_uint8")
    compare_exact("pitches", g["pitches"].cpu().numpy().view(np.int32), pitch.view(np.int32), "step_uint8")
    compare_exact("on_ground", g["on_ground"].cpu().numpy().astype(np.uint8), og, "step_uint8")
    compare_exact("rewards", rew.cpu().numpy(), r, "step_uint8")
    compare_exact("cameras", cams.cpu().numpy().view(np.int32), c.view(np.int32), "step_uint8")
    compare_exact("forward_in", dec[0].cpu().numpy(), f[0], "step_uint8")
    compare_exact("speed_mult", dec[7].cpu().numpy(), sm, "step_uint8")
    st = Stats()
    compare_obs(obs.cpu().numpy(), ref_obs, st, "step_uint8 obs")
    assert st.exact, st.summary()
