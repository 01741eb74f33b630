"""Pins for the oracle from closed forms, hand-derived known answers and an
independent float64 DDA.  The reference ships no tests or golden vectors, so
these are what fix the oracle's semantics (and through the GPU parity tests,
the HIP kernels')."""
import math

import numpy as np
import pytest

import oracle

AIR, GRASS, DIRT, OAKLOG, LEAVES = -1, 0, 1, 3, 4
BRICK, COBBLE, PLANKS, BEDROCK = 8, 9, 10, 12


def _counts(v):
    vals, cnt = np.unique(v, return_counts=True)
    return dict(zip(vals.tolist(), cnt.tolist()))


# Hand-derived block counts of one fresh world (layers by y: bedrock y=0, dirt
# 1..gh-2, grass gh-1; 3 logs + 1 leaf; 5x4 house with a 2-high door, corner
# logs, flat roof, clipped to the world -- src/kernels.cu:1793-1946).
EXPECTED_COUNTS = {
    16: {BEDROCK: 256, DIRT: 1536, GRASS: 256, OAKLOG: 3 + 6, LEAVES: 1, COBBLE: 8, BRICK: 8,
         PLANKS: 12},
    32: {BEDROCK: 1024, DIRT: 6144, GRASS: 1024, OAKLOG: 3 + 12, LEAVES: 1, COBBLE: 20,
         BRICK: 20, PLANKS: 28},
}


@pytest.mark.parametrize("ws", [16, 32])
def test_fresh_world_block_counts(ws):
    env = oracle.OracleEnv(batch_size=2, world_size=ws, render_width=4, render_height=4)
    exp = dict(EXPECTED_COUNTS[ws])
    exp[AIR] = ws ** 3 - sum(exp.values())
    for b in range(2):
        assert _counts(env.voxels[b]) == exp


def test_fresh_world_landmarks_32():
    v = oracle.OracleEnv(batch_size=1, world_size=32, render_width=4, render_height=4).voxels[0]
    # v[y, x, z]; tree at (16, 18), house corner at (22, 14), ground_height 8
    assert [v[y, 16, 18] for y in range(8, 13)] == [OAKLOG, OAKLOG, OAKLOG, LEAVES, AIR]
    # floor at y 8, door (x 24, z 14) at y 9-10, lintel plank at y 11
    assert v[8, 24, 14] == COBBLE
    assert v[9, 24, 14] == AIR and v[10, 24, 14] == AIR and v[11, 24, 14] == PLANKS
    assert v[9, 22, 14] == OAKLOG and v[12, 22, 14] == BRICK and v[9, 23, 15] == AIR
    assert v[9, 22, 15] == PLANKS and v[9, 26, 17] == OAKLOG and v[7, 22, 14] == GRASS


# Landing step / height from spawn y = gh + 2 = 10 under dt 0.05, g -20, with
# the physics ground at world_y/4 (src/kernels.cu:1552-1573), computed
# independently in float32.
@pytest.mark.parametrize("ws,step,y", [(8, 18, 2.0), (16, 15, 4.0), (32, 9, 8.0)])
def test_free_fall_landing(ws, step, y):
    env = oracle.OracleEnv(batch_size=1, world_size=ws, render_width=4, render_height=4)
    env.reset()
    for t in range(1, 40):
        env.step(np.zeros(1, np.int64))
        if env.on_ground[0]:
            break
    assert t == step and env.positions[0, 1] == np.float32(y)
    assert env.velocities[0, 1] == 0


def test_jump_arc_and_pitch_clamp_and_walls():
    env = oracle.OracleEnv(batch_size=1, world_size=32, render_width=4, render_height=4,
                           look_delta=0.5)
    env.reset()
    for _ in range(12):
        env.step(np.zeros(1, np.int64))
    assert env.on_ground[0] and env.positions[0, 1] == 8.0
    env.step(np.array([5]))  # jump: vy = 8 + g*dt = 7, y = 8 + 7*0.05
    assert env.velocities[0, 1] == np.float32(7.0)
    assert env.positions[0, 1] == np.float32(np.float32(8.0) + np.float32(7.0) * np.float32(0.05))
    for _ in range(10):
        env.step(np.array([9]))  # look up by 0.5 rad each
    assert env.pitches[0] == np.float32(1.56)
    for _ in range(10):
        env.step(np.array([10]))
    assert env.pitches[0] == np.float32(-1.56)
    for _ in range(200):
        env.step(np.array([1]))  # walk +z (yaw 0) into the wall
    assert env.positions[0, 2] == np.float32(31.0)


def test_sprint_only_with_forward():
    B = 3
    env = oracle.OracleEnv(batch_size=B, world_size=32, render_width=4, render_height=4)
    env.reset()
    bt = np.zeros((B, 8), np.int8)
    bt[0, [0, 7]] = 1      # forward + sprint
    bt[1, 0] = 1           # forward
    bt[2, [3, 7]] = 1      # strafe right + sprint: sprint ignored
    z0 = env.positions[:, 2].copy()
    x0 = env.positions[:, 0].copy()
    env.step(bt, np.zeros((B, 2), np.float32))
    dz = env.positions[:, 2] - z0
    # yaw 0: forward is +z; walk 4 b/s, sprint x1.5, dt 0.05
    assert math.isclose(dz[0], 0.3, rel_tol=1e-5)
    assert math.isclose(dz[1], 0.2, rel_tol=1e-5)
    assert dz[2] == 0 and math.isclose(env.positions[2, 0] - x0[2], 0.2, rel_tol=1e-5)


SHADE_TABLE = {  # (shade 1, 0.8, 0.9, 0.6) at fog 0, truncated to uint8
    GRASS: [(76, 165, 51), (61, 132, 40), (68, 149, 45), (45, 99, 30)],
    DIRT: [(140, 89, 51), (112, 71, 40), (126, 80, 45), (84, 53, 30)],
    OAKLOG: [(140, 102, 63), (112, 81, 51), (126, 91, 57), (84, 61, 38)],
    LEAVES: [(51, 140, 38), (40, 112, 30), (45, 126, 34), (30, 84, 22)],
    BRICK: [(178, 89, 76), (142, 71, 61), (160, 80, 68), (107, 53, 45)],
    COBBLE: [(102, 102, 102), (81, 81, 81), (91, 91, 91), (61, 61, 61)],
    PLANKS: [(191, 153, 102), (153, 122, 81), (172, 137, 91), (114, 91, 61)],
    BEDROCK: [(38, 38, 38), (30, 30, 30), (34, 34, 34), (22, 22, 22)],
}


def _render1(vox, cam, W=1, H=1, vd=1e30, max_steps=64, fov=70.0, L=None):
    return oracle.render(vox[None], np.asarray(cam, np.float32)[None], W, H, max_steps, vd, fov,
                         L=L)[0]


@pytest.mark.parametrize("block", sorted(SHADE_TABLE))
def test_shaded_colour_table(block):
    """Eye inside the block -> shade 1; a face one voxel away seen along -x
    -> 0.8, along +z -> 0.9, looking down onto it -> 0.6 (the reference darkens
    y-faces hit by downward rays, kernels.cu:451), looking up -> 1.0; with
    view_distance 1e30 so the fog underflows to exactly 0."""
    n = 8
    vox = np.full((n, n, n), AIR, np.int8)
    vox[4, 4, 4] = block
    # eye = cam + (0, 1.6, 0)
    inside = _render1(vox, [4.5, 4.5 - 1.6, 4.5, 0.0, 0.0], W=4, H=4)
    assert (inside.reshape(-1, 3) == SHADE_TABLE[block][0]).all()
    # 1x1 image -> the centre ray is exactly the forward vector.
    x_face = _render1(vox, [6.5, 4.5 - 1.6, 4.5, -math.pi / 2, 0.0])   # looking -x
    z_face = _render1(vox, [4.5, 4.5 - 1.6, 2.5, 0.0, 0.0])            # looking +z
    top = _render1(vox, [4.5, 6.5 - 1.6, 4.5, 0.0, -1.56])             # looking down
    bottom = _render1(vox, [4.5, 2.5 - 1.6, 4.5, 0.0, 1.56])           # looking up
    assert tuple(x_face[0, 0]) == SHADE_TABLE[block][1]
    assert tuple(z_face[0, 0]) == SHADE_TABLE[block][2]
    assert tuple(top[0, 0]) == SHADE_TABLE[block][3]
    assert tuple(bottom[0, 0]) == SHADE_TABLE[block][0]


def test_all_air_world_is_sky():
    vox = np.full((2, 8, 8, 8), AIR, np.int8)
    cams = np.float32([[4, 2, 4, 0.3, 0.2], [-5, 20, 3, 2.0, -1.0]])
    for u8, sky in ((True, (135, 206, 234)),):
        out = oracle.render(vox, cams, 16, 12, out_u8=u8)
        assert (out.reshape(-1, 3) == sky).all()
    outf = oracle.render(vox, cams, 16, 12, out_u8=False)
    assert np.allclose(outf.reshape(-1, 3), np.float32([0.53, 0.81, 0.92]), rtol=0, atol=0)


def _dda64(vox, eye, d, max_steps, vd):
    """Amanatides & Woo, "A Fast Voxel Traversal Algorithm for Ray Tracing"
    (Eurographics 1987), in float64, with the tie order x, y, z and the
    reference's termination rules (slab test against the world box, step and
    distance limits, leave once past the exit)."""
    Y, X, Z = vox.shape
    inv = [1.0 / c if c != 0 else 1e30 for c in d]
    lo = [(0.0 - e) * i for e, i in zip(eye, inv)]
    hi = [(s - e) * i for e, s, i in zip(eye, (X, Y, Z), inv)]
    t_entry = max(min(a, b) for a, b in zip(lo, hi))
    t_exit = min(max(a, b) for a, b in zip(lo, hi))
    if t_exit < 0 or t_entry > t_exit or t_entry > vd:
        return -1, None
    v = [math.floor(e) for e in eye]
    step = [1 if c >= 0 else -1 for c in d]
    tmax = [((v[k] + 1) - eye[k]) * inv[k] if d[k] >= 0 else (v[k] - eye[k]) * inv[k] for k in range(3)]
    tdel = [abs(i) for i in inv]
    t = 0.0
    for _ in range(max_steps):
        if t >= vd:
            break
        x, y, z = v
        if 0 <= x < X and 0 <= y < Y and 0 <= z < Z:
            if vox[y, x, z] >= 0:
                return y * X * Z + x * Z + z, tmax
        elif t > t_exit:
            break
        k = 0 if (tmax[0] < tmax[1] and tmax[0] < tmax[2]) else (1 if tmax[1] < tmax[2] else 2)
        t = tmax[k]
        tmax[k] += tdel[k]
        v[k] += step[k]
    return -1, tmax


def test_dda_against_independent_float64_traversal():
    """Hit voxels of the float32 oracle agree with a float64 traversal except
    on rays that pass within 1e-4 of a voxel edge/corner (where the two
    precisions may legitimately break the tie differently)."""
    rng = np.random.default_rng(1)
    B, n, W, H = 6, 10, 24, 18
    vox = np.where(rng.random((B, n, n, n)) < 0.08, rng.integers(0, 13, (B, n, n, n)), AIR).astype(np.int8)
    cams = np.zeros((B, 5), np.float32)
    cams[:, :3] = rng.uniform(-3, n + 3, (B, 3))
    cams[:, 3] = rng.uniform(-3.2, 3.2, B)
    cams[:, 4] = rng.uniform(-1.4, 1.4, B)
    _, hits = oracle.render(vox, cams, W, H, 48, 30.0, 75.0, want_hits=True)
    hf = oracle.half_fov_tan(75.0)
    agree = total = 0
    for b in range(B):
        x, y, z, yaw, pitch = [float(c) for c in cams[b]]
        eye = (x, y + 1.6, z)
        sy, cy, sp, cp = math.sin(yaw), math.cos(yaw), math.sin(pitch), math.cos(pitch)
        f = (sy * cp, sp, cy * cp)
        r = (cy, 0.0, -sy)
        u_ = (-sy * sp, cp, -cy * sp)
        for py in range(H):
            for px in range(W):
                u = (2 * (px + 0.5) / W - 1) * hf * (W / H)
                v = (1 - 2 * (py + 0.5) / H) * hf
                d = [f[k] + u * r[k] + v * u_[k] for k in range(3)]
                L = math.sqrt(sum(c * c for c in d))
                d = [c / L for c in d]
                h64, _ = _dda64(vox[b], eye, d, 48, 30.0)
                total += 1
                if h64 == hits[b, py, px]:
                    agree += 1
    assert total == B * W * H
    assert agree / total > 0.995, f"{agree}/{total}"


def test_raycast_break_known_answers_32():
    """Land at 32^3 facing +z: the break ray (32 samples, t = 0.1 + k/8)
    reaches the tree log at (y 9, x 16, z 18) -> reward 1, voxel -> AIR; the
    next break finds nothing within reach -> reward 0."""
    env = oracle.OracleEnv(batch_size=1, world_size=32, render_width=4, render_height=4)
    env.reset()
    for _ in range(9):
        env.step(np.zeros(1, np.int64))
    assert env.positions[0, 1] == 8.0
    assert env.voxels[0, 9, 16, 18] == OAKLOG
    _, r, _ = env.step(np.array([6]))
    assert r[0] == 1.0 and env.voxels[0, 9, 16, 18] == AIR
    _, r, _ = env.step(np.array([6]))
    assert r[0] == 0.0


def test_raycast_break_leaves_while_falling():
    """Right after reset the eye is at 11.6 - 0.05: the ray at y 11 hits the
    leaves (no reward: target is the oak log)."""
    env = oracle.OracleEnv(batch_size=1, world_size=32, render_width=4, render_height=4)
    env.reset()
    _, r, _ = env.step(np.array([6]))
    assert r[0] == 0.0 and env.voxels[0, 11, 16, 18] == AIR
    assert env.voxels[0, 10, 16, 18] == OAKLOG


def test_place_puts_dirt_in_front_of_hit():
    env = oracle.OracleEnv(batch_size=1, world_size=32, render_width=4, render_height=4)
    env.reset()
    for _ in range(9):
        env.step(np.zeros(1, np.int64))
    bt = np.zeros((1, 8), np.int8)
    bt[0, 6] = 1
    env.step(bt, np.zeros((1, 2), np.float32))
    assert env.voxels[0, 9, 16, 17] == DIRT and env.voxels[0, 9, 16, 18] == OAKLOG


def test_eye_inside_dirt_renders_dirt_at_16():
    """16^3: the agent lands at y 4 with its eye at 5.6 inside dirt, so every
    pixel is the shade-1 dirt colour."""
    env = oracle.OracleEnv(batch_size=2, world_size=16, render_width=16, render_height=16)
    env.reset()
    for _ in range(15):
        obs, _, _ = env.step(np.zeros(2, np.int64))
    assert (obs.reshape(-1, 3) == (140, 89, 51)).all()


def test_episode_period():
    """After reset, first done on step L+1, then every L steps
    (minecu/__init__.py:292, 363-366, 435-436)."""
    L = 5
    env = oracle.OracleEnv(batch_size=2, world_size=8, render_width=4, render_height=4,
                           episode_length=L)
    env.reset()
    done_steps = []
    for t in range(1, 18):
        _, _, d = env.step(np.zeros(2, np.int64))
        if d[0]:
            done_steps.append(t)
    assert done_steps == [6, 11, 16]


def test_batch_independence():
    """Each env evolves alone: a batch of 5 equals five batches of 1."""
    rng = np.random.default_rng(4)
    acts = rng.integers(0, 11, (25, 5))
    kw = dict(world_size=16, render_width=8, render_height=8, episode_length=9)
    big = oracle.OracleEnv(batch_size=5, **kw)
    big.reset()
    for t in range(25):
        big.step(acts[t])
    for b in range(5):
        one = oracle.OracleEnv(batch_size=1, **kw)
        one.reset()
        for t in range(25):
            one.step(acts[t, b:b + 1])
        assert np.array_equal(one.obs_buffer[0], big.obs_buffer[b])
        assert np.array_equal(one.positions[0], big.positions[b])


# --- fog blend (src/kernels.cu:455-464 uint8, 281-291 float32) ----------------
# colour = block * shade * (1 - fog) + sky * fog, fog = min(t / view_distance, 1)^2,
# then uint8 truncates colour * 255.  Expected values here come from that
# closed form in float64 at exactly known t; each case is chosen with every
# channel at least 1e-3 away from an integer boundary, so float32 rounding
# cannot move the truncated byte, while a fog that is not squared, a swapped
# (1 - fog) / fog, or a missing shade moves at least one channel.
PALETTE = {OAKLOG: (0.55, 0.40, 0.25), LEAVES: (0.20, 0.55, 0.15), GRASS: (0.30, 0.65, 0.20)}
SKY = (0.53, 0.81, 0.92)


def _fogged(block, shade, t, vd):
    fog = min(t / vd, 1.0) ** 2
    return [c * shade * (1.0 - fog) + s * fog for c, s in zip(PALETTE[block], SKY)]


def _u8(rgb):
    vals = [c * 255.0 for c in rgb]
    for v in vals:
        assert abs(v - round(v)) > 1e-3, "case too close to a rounding boundary"
    return tuple(int(math.floor(v)) for v in vals)


def _z_face_world(block):
    """32^3 air world with one block at (y 4, x 4, z 14); an eye at z = 4.25
    looking +z along the voxel row's centre hits its z face at t = 9.75."""
    vox = np.full((32, 32, 32), AIR, np.int8)
    vox[4, 4, 14] = block
    return vox, [4.5, 4.5 - 1.6, 4.25, 0.0, 0.0]


def test_fog_known_answer_oaklog_z_face():
    vox, cam = _z_face_world(OAKLOG)
    got = tuple(_render1(vox, cam, vd=32.0)[0, 0])
    assert got == (127, 102, 73) == _u8(_fogged(OAKLOG, 0.9, 9.75, 32.0))


@pytest.mark.parametrize("vd", [10.0, 12.5, 20.0, 64.0])
def test_fog_known_answers_vs_closed_form(vd):
    """The same z face at several view distances (fog from 0.023 to 0.95)."""
    vox, cam = _z_face_world(OAKLOG)
    assert tuple(_render1(vox, cam, vd=vd)[0, 0]) == _u8(_fogged(OAKLOG, 0.9, 9.75, vd))


def _check_fog_known_answers(L=None):
    """The z-face known answers at four view distances, uint8 and float32."""
    vox, cam = _z_face_world(OAKLOG)
    for vd in (32.0, 10.0, 12.5, 20.0):
        assert tuple(_render1(vox, cam, vd=vd, L=L)[0, 0]) == _u8(_fogged(OAKLOG, 0.9, 9.75, vd))
        out = oracle.render(vox[None], np.asarray(cam, np.float32)[None], 1, 1, 64, vd, 70.0,
                            out_u8=False, L=L)[0, 0, 0]
        assert np.max(np.abs(out.astype(np.float64) - _fogged(OAKLOG, 0.9, 9.75, vd))) < 1e-6


def test_fog_known_answers_all_view_distances():
    _check_fog_known_answers()


@pytest.mark.parametrize("k,slip", [(1, "fog not squared"), (2, "(1 - fog) and fog swapped"),
                                    (3, "face shade dropped")])
def test_fog_pins_catch_planted_slips(k, slip):
    """The oracle compiled with a planted slip in its shading fails the
    known answers."""
    with pytest.raises(AssertionError):
        _check_fog_known_answers(oracle.mutant_lib(k))


@pytest.mark.parametrize("vd", [9.75, 9.0, 2.0])
def test_fog_at_or_beyond_view_distance_is_sky(vd):
    """t >= view_distance ends the ray before the lookup (kernels.cu:413-417):
    the pixel is the shaded sky (135, 206, 234), not a fully fogged block."""
    vox, cam = _z_face_world(OAKLOG)
    assert tuple(_render1(vox, cam, vd=vd)[0, 0]) == (135, 206, 234)


@pytest.mark.parametrize("vd", [32.0, 12.5])
def test_fog_float32_obs_closed_form(vd):
    """float32 observations (render_kernel, kernels.cu:281-291) equal the
    closed form to float32 rounding (|error| < 1e-6 per channel)."""
    vox, cam = _z_face_world(OAKLOG)
    out = oracle.render(vox[None], np.asarray(cam, np.float32)[None], 1, 1, 64, vd, 70.0,
                        out_u8=False)[0, 0, 0]
    exp = _fogged(OAKLOG, 0.9, 9.75, vd)
    assert np.max(np.abs(out.astype(np.float64) - exp)) < 1e-6
    # beyond the view distance: the sky colour itself, exactly
    sky = oracle.render(vox[None], np.asarray(cam, np.float32)[None], 1, 1, 64, 9.0, 70.0,
                        out_u8=False)[0, 0, 0]
    assert np.array_equal(sky, np.float32(SKY))


def test_fog_x_and_top_faces():
    """Shades 0.8 (x face) and 0.6 (top face seen looking down) under fog."""
    vox = np.full((32, 32, 32), AIR, np.int8)
    vox[4, 20, 4] = GRASS
    # eye x = 4.5 looking +x (yaw pi/2: forward = (1, 0, cos(pi/2)))
    cam = [4.5, 4.5 - 1.6, 4.5, math.pi / 2, 0.0]
    t = 20.0 - 4.5
    assert tuple(_render1(vox, cam, vd=32.0)[0, 0]) == _u8(_fogged(GRASS, 0.8, t, 32.0))
    vox = np.full((32, 32, 32), AIR, np.int8)
    vox[4, 4, 4] = LEAVES
    # eye 7.25 above the top face y = 5, looking straight down is pitch
    # -pi/2, beyond the clamp; the kernel takes any pitch from the camera row.
    cam = [4.5, 12.25 - 1.6, 4.5, 0.0, -math.pi / 2]
    out = _render1(vox, cam, vd=16.0)[0, 0]
    assert tuple(out) == _u8(_fogged(LEAVES, 0.6, 12.25 - 5.0, 16.0))


# --- done-step order (minecu/__init__.py:435-442) ------------------
def _q3_env(obs_dtype):
    return oracle.OracleEnv(batch_size=1, world_size=32, render_width=1, render_height=1,
                            episode_length=12, obs_dtype=obs_dtype)


def _check_done_step_order(env):
    """Steps a fresh 32^3 one-env oracle (1x1 image) through the done step and
    checks what it must render; see test_done_step_renders_before_regeneration."""
    L = env.episode_length
    env.reset()
    assert env.voxels[0, 11, 16, 18] == LEAVES
    _, r, d = env.step(np.array([6]))
    assert r[0] == 0.0 and not d[0] and env.voxels[0, 11, 16, 18] == AIR
    for _ in range(1, L):
        _, _, d = env.step(np.array([0]))
        assert not d[0]
    assert env.positions[0, 1] == 8.0 and env.voxels[0, 9, 16, 18] == OAKLOG
    obs, r, d = env.step(np.array([6]))
    assert d[0] and r[0] == 1.0
    leaves = _fogged(LEAVES, 0.9, 2.0, 32.0)
    if env.obs_dtype == "uint8":
        assert tuple(obs[0, 0, 0]) == (135, 206, 234)
    else:
        assert np.array_equal(obs[0, 0, 0], np.float32(SKY))
    # spawn pose, regenerated world (leaves and log back)
    assert env.positions[0].tolist() == [16.0, 10.0, 16.0]
    assert np.array_equal(env.voxels, oracle.fresh_world(1, 32, 32, 32, 8, (16, 18), (22, 14)))
    obs, r, d = env.step(np.array([0]))
    assert not d[0]
    if env.obs_dtype == "uint8":
        assert tuple(obs[0, 0, 0]) == _u8(leaves) == (46, 126, 35)
    else:
        assert np.max(np.abs(obs[0, 0, 0].astype(np.float64) - leaves)) < 1e-6


@pytest.mark.parametrize("obs_dtype", ["uint8", "float32"])
def test_done_step_renders_before_regeneration(obs_dtype):
    """32^3, one env, 1x1 image (the centre ray is the forward vector).
    Step 0 breaks the leaves at (y 11, x 16, z 18) straight ahead of the
    falling agent's eye.  The done step (step index L after reset) respawns
    the agent facing +z at eye height 11.6; its image is rendered in the world
    as it was before regeneration, so the centre ray passes where the leaves
    were and leaves the world: sky.  The world is regenerated afterwards; the
    next step's centre ray hits the leaves' z face at t = 2.  The done step's
    own break (from the landed pose) still earns the log's reward."""
    _check_done_step_order(_q3_env(obs_dtype))


@pytest.mark.parametrize("obs_dtype", ["uint8", "float32"])
def test_done_step_order_pin_catches_regenerate_first(obs_dtype):
    """Planted on one oracle env: regenerate the worlds before the done
    step's render (the render sees do_reset already set) -- the check above
    fails."""
    env = _q3_env(obs_dtype)
    plain = env._render

    def regen_then_render(want_hits=False):
        if env.do_reset.any():
            env._regen()
        return plain(want_hits)

    env._render = regen_then_render
    with pytest.raises(AssertionError):
        _check_done_step_order(env)
