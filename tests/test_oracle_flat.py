"""Known answers for the oracle's flat-colour variants (obs_dtype
uint8_minimal / uint8_prebasis / uint8_prebasis_full / uint8_smem /
uint8_fp16, src/kernels.cu:479-1262): colour table, no shading, no fog and
no view-distance limit, the max_steps limit, the camera basis, an
independent float64 traversal, the software binary16 arithmetic against
numpy, and the MineEnv orchestration of these modes."""
import math

import numpy as np
import pytest

import oracle

AIR, DIRT, OAKLOG = oracle.AIR, oracle.DIRT, oracle.OAKLOG

# BLOCK_COLORS_U8 / SKY_COLOR_U8 as the reference lists them (kernels.cu:479-495).
FLAT = [(76, 166, 51), (140, 89, 51), (127, 127, 127), (140, 102, 64), (51, 140, 38),
        (229, 217, 153), (51, 102, 204), (217, 229, 242), (178, 89, 76), (102, 102, 102),
        (191, 153, 102), (242, 242, 250), (38, 38, 38)]
SKY = (135, 206, 235)
MODES = ("minimal", "prebasis")


def _flat(vox, cam, mode, W=1, H=1, max_steps=64, fov=70.0, want_hits=False):
    cam = np.asarray(cam, np.float32)[None]
    src = cam if mode == "minimal" else oracle.camera_basis(cam, W, H, fov)
    r = oracle.render_flat(vox[None], src, W, H, mode, max_steps, fov, want_hits)
    return (r[0][0], r[1][0]) if want_hits else r[0]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("block", range(13))
def test_flat_colour_table_and_no_shading(mode, block):
    """Eye inside the block -> its flat colour everywhere; the same block seen
    across an x face, a z face, from above and from below keeps that colour
    (the shaded modes would darken three of them)."""
    n = 8
    vox = np.full((n, n, n), AIR, np.int8)
    vox[4, 4, 4] = block
    inside = _flat(vox, [4.5, 4.5 - 1.6, 4.5, 0.0, 0.0], mode, W=4, H=4)
    assert (inside.reshape(-1, 3) == FLAT[block]).all()
    for cam in ([6.5, 4.5 - 1.6, 4.5, -math.pi / 2, 0.0], [4.5, 4.5 - 1.6, 2.5, 0.0, 0.0],
                [4.5, 6.5 - 1.6, 4.5, 0.0, -1.56], [4.5, 2.5 - 1.6, 4.5, 0.0, 1.56]):
        assert tuple(_flat(vox, cam, mode)[0, 0]) == FLAT[block]


@pytest.mark.parametrize("mode", MODES)
def test_all_air_world_is_flat_sky(mode):
    vox = np.full((8, 8, 8), AIR, np.int8)
    for cam in ([4, 2, 4, 0.3, 0.2], [-5, 20, 3, 2.0, -1.0]):
        out = _flat(vox, cam, mode, W=16, H=12)
        assert (out.reshape(-1, 3) == SKY).all()


@pytest.mark.parametrize("mode", MODES)
def test_no_view_distance_but_max_steps(mode):
    """A block 49.5 voxels ahead shows in the flat image (the shaded one,
    limited to view_distance 32, shows sky); max_steps counts lookups from the
    eye voxel: with the eye voxel at z=0 the block at z=50 is lookup 51."""
    vox = np.full((4, 4, 80), AIR, np.int8)
    vox[2, 2, 50] = OAKLOG
    cam = [2.5, 2.5 - 1.6, 0.5, 0.0, 0.0]
    assert tuple(_flat(vox, cam, mode, max_steps=64)[0, 0]) == FLAT[OAKLOG]
    assert tuple(_flat(vox, cam, mode, max_steps=51)[0, 0]) == FLAT[OAKLOG]
    assert tuple(_flat(vox, cam, mode, max_steps=50)[0, 0]) == SKY
    shaded = oracle.render(vox[None], np.float32([cam]), 1, 1, 64, 32.0)[0]
    assert tuple(shaded[0, 0]) == (135, 206, 234)


@pytest.mark.parametrize("mode", MODES)
def test_eye_outside_walks_in(mode):
    """No slab test: the DDA starts at the (out-of-bounds) eye voxel and walks
    into the world; looking straight down from above hits the top layer."""
    vox = np.full((8, 8, 8), AIR, np.int8)
    vox[0:3] = DIRT
    vox[2] = 0  # grass top at y = 2
    out = _flat(vox, [4.5, 20.0, 4.5, 0.0, -1.56], mode)
    assert tuple(out[0, 0]) == FLAT[0]
    # ...but only within max_steps lookups (eye voxel y = 21, grass at y = 2)
    assert tuple(_flat(vox, [4.5, 20.0, 4.5, 0.0, -1.56], mode, max_steps=19)[0, 0]) == SKY


def test_camera_basis_closed_form():
    """precompute_camera_basis_kernel (kernels.cu:511-561): yaw 0, pitch 0
    gives forward +z, right +x, up +y; scale_v = tan(35 deg) with the flat
    degree conversion (0.01745329), scale_u = scale_v * W/H; column 13 unset."""
    cams = np.float32([[1.0, 2.0, 3.0, 0.0, 0.0], [0.0, 0.0, 0.0, math.pi / 2, 0.3]])
    b = oracle.camera_basis(cams, 48, 32, 70.0)
    assert b[0, :3].tolist() == [1.0, np.float32(3.6), 3.0]
    assert b[0, 3:11].tolist() == [0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0]
    hf = oracle.half_fov_tan_flat(70.0)
    assert abs(hf - math.tan(35 * math.pi / 180)) < 2e-6
    assert b[0, 12] == np.float32(hf) and b[0, 11] == np.float32(np.float32(hf) * np.float32(1.5))
    assert b[0, 13] == 0.0 and b[1, 13] == 0.0
    # yaw pi/2 looks along +x; right is -z; pitch tilts forward up
    assert abs(b[1, 3] - math.cos(0.3)) < 1e-6 and abs(b[1, 4] - math.sin(0.3)) < 1e-6
    assert abs(b[1, 5]) < 1e-6 and abs(b[1, 7] + 1.0) < 1e-6


def test_minimal_equals_prebasis_on_square_images():
    """The two image planes differ only in how u is rounded, (a*hf)*aspect
    vs a*(hf*aspect): with aspect 1 they coincide bit for bit; otherwise the
    images still agree on almost every pixel."""
    rng = np.random.default_rng(5)
    B, n = 4, 12
    vox = np.where(rng.random((B, n, n, n)) < 0.1, rng.integers(0, 13, (B, n, n, n)), AIR).astype(np.int8)
    cams = np.zeros((B, 5), np.float32)
    cams[:, :3] = rng.uniform(1, n - 1, (B, 3))
    cams[:, 3] = rng.uniform(-3, 3, B)
    cams[:, 4] = rng.uniform(-1.4, 1.4, B)
    for W, H, exact in ((24, 24, True), (40, 24, False)):
        a = oracle.render_flat(vox, cams, W, H, "minimal")
        b = oracle.render_flat(vox, oracle.camera_basis(cams, W, H), W, H, "prebasis")
        if exact:
            assert np.array_equal(a, b)
        else:
            assert (a != b).any(axis=-1).mean() < 0.01


def _dda64_flat(vox, eye, d, max_steps):
    """Amanatides & Woo (Eurographics 1987) in float64 with the flat kernels'
    rules: start at the eye voxel, at most max_steps lookups, out-of-bounds
    voxels skipped, tie order x, y, z."""
    Y, X, Z = vox.shape
    inv = [1.0 / c if c != 0 else 1e30 for c in d]
    v = [math.floor(e) for e in eye]
    step = [1 if c >= 0 else -1 for c in d]
    tmax = [((v[k] + 1) - eye[k]) * inv[k] if d[k] >= 0 else (v[k] - eye[k]) * inv[k]
            for k in range(3)]
    tdel = [abs(i) for i in inv]
    for _ in range(max_steps):
        x, y, z = v
        if 0 <= x < X and 0 <= y < Y and 0 <= z < Z and vox[y, x, z] >= 0:
            return y * X * Z + x * Z + z
        k = 0 if (tmax[0] < tmax[1] and tmax[0] < tmax[2]) else (1 if tmax[1] < tmax[2] else 2)
        tmax[k] += tdel[k]
        v[k] += step[k]
    return -1


def test_flat_dda_against_independent_float64_traversal():
    rng = np.random.default_rng(6)
    B, n, W, H = 5, 10, 20, 16
    vox = np.where(rng.random((B, n, n, n)) < 0.07, rng.integers(0, 13, (B, n, n, n)), AIR).astype(np.int8)
    cams = np.zeros((B, 5), np.float32)
    cams[:, :3] = rng.uniform(-3, n + 3, (B, 3))
    cams[:, 3] = rng.uniform(-3.2, 3.2, B)
    cams[:, 4] = rng.uniform(-1.4, 1.4, B)
    _, hits = oracle.render_flat(vox, cams, W, H, "minimal", 40, 75.0, want_hits=True)
    hf = oracle.half_fov_tan_flat(75.0)
    agree = 0
    for b in range(B):
        x, y, z, yaw, pitch = [float(c) for c in cams[b]]
        sy, cy, sp, cp = math.sin(yaw), math.cos(yaw), math.sin(pitch), math.cos(pitch)
        f, r, up = (sy * cp, sp, cy * cp), (cy, 0.0, -sy), (-sy * sp, cp, -cy * sp)
        for py in range(H):
            for px in range(W):
                u = (2 * (px + 0.5) / W - 1) * hf * (W / H)
                v = (1 - 2 * (py + 0.5) / H) * hf
                d = [f[k] + u * r[k] + v * up[k] for k in range(3)]
                L = math.sqrt(sum(c * c for c in d))
                agree += _dda64_flat(vox[b], (x, y + 1.6, z), [c / L for c in d], 40) == hits[b, py, px]
    assert agree / (B * W * H) > 0.995


def test_env_flat_step_but_shaded_reset():
    """MineEnv renders the reset (and get_obs) with render_uint8 in every
    uint8 variant and the step with the variant's flat kernel
    (minecu/__init__.py:304-466): sky is (135,206,234) after reset and
    (135,206,235) after a step."""
    for dt in ("uint8_minimal", "uint8_prebasis", "uint8_prebasis_full", "uint8_smem",
               "uint8_fp16"):
        env = oracle.OracleEnv(batch_size=2, world_size=16, render_width=16, render_height=16,
                               obs_dtype=dt)
        o0 = env.reset().copy()
        assert ((o0.reshape(-1, 3) == (135, 206, 234)).all(axis=1)).any()
        assert not ((o0.reshape(-1, 3) == SKY).all(axis=1)).any()
        o1, _, _ = env.step(np.zeros(2, np.int64))
        assert ((o1.reshape(-1, 3) == SKY).all(axis=1)).any()
        if dt not in ("uint8_minimal", "uint8_fp16"):
            assert np.array_equal(env.basis, oracle.camera_basis(env.cameras, 16, 16))


def test_prebasis_full_episode_convention():
    """step_uint8_prebasis_full's C++ episode check: 0 (what MineEnv passes
    for None) means no limit (kernels.cu:1954-1968); otherwise the same period
    as the Python path."""
    for L, want in ((None, []), (0, []), (5, [6, 11, 16])):
        env = oracle.OracleEnv(batch_size=2, world_size=8, render_width=4, render_height=4,
                               episode_length=L, obs_dtype="uint8_prebasis_full")
        env.reset()
        got = [t for t in range(1, 18) if env.step(np.zeros(2, np.int64))[2][0]]
        assert got == want, (L, got)
    # the Python-path variants treat 0 literally: done on every step
    env = oracle.OracleEnv(batch_size=1, world_size=8, render_width=4, render_height=4,
                           episode_length=0, obs_dtype="uint8_prebasis")
    env.reset()
    assert all(env.step(np.zeros(1, np.int64))[2][0] for _ in range(4))


def _f16_cases(n, seed):
    """float32 pairs over the binary16 range: subnormals, ties, the overflow
    boundary 65504/65520, zeros, values far beyond the range."""
    rng = np.random.default_rng(seed)
    mags = np.exp2(rng.uniform(-27, 18, n)).astype(np.float32)
    a = mags * rng.choice([-1, 1], n).astype(np.float32)
    b = np.exp2(rng.uniform(-27, 18, n)).astype(np.float32)
    # exact binary16 midpoints (ties) and boundaries
    h = rng.integers(0, 0x7bff, n // 4).astype(np.uint16).view(np.float16).astype(np.float32)
    nxt = (rng.integers(0, 0x7bff, n // 4).astype(np.uint16) + 1).view(np.float16).astype(np.float32)
    a[: n // 4] = (h.astype(np.float64) + nxt.astype(np.float64)).astype(np.float32) / 2
    special = np.float32([0.0, -0.0, 65504.0, 65519.0, 65520.0, 65536.0, 1e6, 5.96e-8, 2.98e-8,
                          6.1035e-5, 1e30, 65504.0])
    a[-len(special):] = special
    b[-len(special):] = special[::-1]
    return a, b


def test_software_binary16_matches_numpy():
    """Conversion, sum (exact in binary64, rounded once) and order against
    numpy's binary16, whose float32/float64 -> float16 conversions round to
    nearest even (IEEE 754)."""
    a, b = _f16_cases(200000, 7)
    ha, s, lt = oracle.f16_ops(a, b)
    with np.errstate(over="ignore"):
        na, nb = a.astype(np.float16), b.astype(np.float16)
    assert np.array_equal(ha, na.view(np.uint16))
    with np.errstate(invalid="ignore", over="ignore"):
        exact = na.astype(np.float64) + nb.astype(np.float64)
        ref = exact.astype(np.float16).view(np.uint16)
    ok = ~np.isnan(exact)          # inf + -inf: any NaN (never formed by the DDA)
    assert ok.mean() > 0.99
    assert np.array_equal(s[ok], ref[ok])
    assert np.array_equal(lt.astype(bool), na < nb)
    # every finite binary16 value converts to itself
    allh = np.arange(0, 0x7c00, dtype=np.uint16)
    back, _, _ = oracle.f16_ops(allh.view(np.float16).astype(np.float32), np.zeros(allh.size, np.float32))
    assert np.array_equal(back, allh)


def test_fp16_variant_differs_only_by_precision():
    """The binary16 DDA sees the same blocks as the binary32 one on most
    pixels of a random scene, but not on all (its t values are coarser)."""
    rng = np.random.default_rng(9)
    B, n, W, H = 6, 16, 32, 32
    vox = np.where(rng.random((B, n, n, n)) < 0.06, rng.integers(0, 13, (B, n, n, n)), AIR).astype(np.int8)
    cams = np.zeros((B, 5), np.float32)
    cams[:, :3] = rng.uniform(1, n - 1, (B, 3))
    cams[:, 3] = rng.uniform(-3, 3, B)
    cams[:, 4] = rng.uniform(-1.4, 1.4, B)
    a = oracle.render_flat(vox, cams, W, H, "minimal")
    h = oracle.render_flat(vox, cams, W, H, "fp16")
    diff = (a != h).any(axis=-1).mean()
    assert 0 < diff < 0.05, diff
