"""Pins of the oracle's seeded terrain (MineEnv(terrain="noise")), which has no
reference algorithm (the reference's generator ignores its seed,
src/kernels.cu:1800-1828):
- the hash is MurmurHash3's fmix32: the finaliser of MurmurHash3_x86_32 on an
  empty input, whose published values are 0 (seed 0), 0x514E28B7 (seed 1) and
  0x81F16F39 (seed 0xFFFFFFFF);
- amplitude 0 reduces to the reference's flat world, bit for bit;
- heights stay in [ground - A, ground + A] and equal ground - A + corner at the
  lattice points (a plain-Python corner below);
- the lattice-first height map and the one-column evaluation agree, also for
  negative and out-of-world columns;
- worlds depend on seed and env, and the tree and house stand on the surface."""
import numpy as np
import pytest

import oracle

M = 0xFFFFFFFF


def test_fmix32_known_answers():
    assert oracle.fmix32(0) == 0
    assert oracle.fmix32(1) == 0x514E28B7
    assert oracle.fmix32(0xFFFFFFFF) == 0x81F16F39


def _py_fmix(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M
    return h ^ (h >> 16)


def _py_corner(seed, b, i, j, A):
    inner = _py_fmix((i * 0x632BE5AB + j) & M)
    return _py_fmix((seed ^ _py_fmix(((b * 0x9E3779B9) & M) ^ inner)) & M) % (2 * A + 1)


@pytest.mark.parametrize("seed,A,gh", [(42, 3, 8), (7, 1, 4), (123456789, 8, 10)])
def test_heights_range_and_lattice_points(seed, A, gh):
    for b in (0, 5):
        h = oracle.terrain_heights(b, 16, 16, gh, A, seed)
        assert h.min() >= gh - A and h.max() <= gh + A
        for i in range(4):
            for j in range(4):
                assert h[4 * i, 4 * j] == gh - A + _py_corner(seed, b, i, j, A)
        # between two lattice points of a row the height moves monotonically
        for j in range(4):
            for i in range(3):
                seg = h[4 * i:4 * i + 5, 4 * j]
                d = np.diff(seg)
                assert (d >= 0).all() or (d <= 0).all()


def test_column_form_matches_lattice_form():
    h = oracle.terrain_heights(3, 24, 20, 9, 4, 99)
    for x in range(24):
        for z in range(20):
            assert oracle.terrain_column(3, x, z, 9, 4, 99) == h[x, z]
    # any column (negative, beyond the world) is defined and in range
    for x, z in [(-1, -1), (-5, 7), (30, 2), (100, 100)]:
        assert 5 <= oracle.terrain_column(3, x, z, 9, 4, 99) <= 13


@pytest.mark.parametrize("ws", [8, 16, 32])
def test_amplitude_zero_is_the_flat_world(ws):
    kw = dict(B=3, X=ws, Y=ws, Z=ws, ground_height=8)
    t, h = (ws // 2, ws // 2 + 2), (ws // 2 + 6, ws // 2 - 2)
    flat = oracle.fresh_world(kw["B"], ws, ws, ws, 8, t, h, seed=5)
    for seed in (0, 5, 2 ** 31 + 3):
        noise = oracle.fresh_world(kw["B"], ws, ws, ws, 8, t, h, seed=seed, terrain_amplitude=0)
        assert np.array_equal(flat, noise)


def test_worlds_depend_on_seed_and_env_and_props_stand_on_the_surface():
    t, hs = (8, 10), (14, 6)
    a = oracle.fresh_world(4, 16, 16, 16, 8, t, hs, seed=1, terrain_amplitude=3)
    b = oracle.fresh_world(4, 16, 16, 16, 8, t, hs, seed=2, terrain_amplitude=3)
    a2 = oracle.fresh_world(4, 16, 16, 16, 8, t, hs, seed=1, terrain_amplitude=3)
    assert np.array_equal(a, a2) and not np.array_equal(a, b)
    assert not np.array_equal(a[0], a[1])
    for e in range(4):
        ht = oracle.terrain_column(e, t[0], t[1], 8, 3, 1)
        col = a[e, :, t[0], t[1]]
        assert col[ht - 1] == oracle.GRASS and list(col[ht:ht + 3]) == [oracle.OAKLOG] * 3
        assert col[ht + 3] == oracle.LEAVES
        hh = oracle.terrain_column(e, hs[0], hs[1], 8, 3, 1)
        assert a[e, hh, hs[0], hs[1]] == oracle.COBBLESTONE          # house floor corner
        assert a[e, hh + 4, hs[0] + 1, hs[1] + 1] == oracle.BRICK      # roof


@pytest.mark.parametrize("collision", [False, True])
def test_spawn_stands_on_the_spawn_column(collision):
    """On the seeded terrain every env respawns 2 above its own spawn column
    (ground_height + 2 on the flat world): at amplitude 8 no reset eye is inside
    a solid, and with voxel collision an agent falls onto its surface instead
    of climbing out of buried terrain."""
    B, gh, A = 64, 10, 8
    env = oracle.OracleEnv(batch_size=B, world_size=32, render_width=4, render_height=4,
                           terrain="noise", terrain_amplitude=A, ground_height=gh, seed=42,
                           collision=collision, episode_length=30)
    env.reset()
    sx, sz = int(env.spawn_x), int(env.spawn_z)
    heights = np.array([oracle.terrain_column(b, sx, sz, gh, A, 42) for b in range(B)])
    assert heights.min() < gh < heights.max()          # the columns differ
    assert np.array_equal(env.positions[:, 1], (heights + 2).astype(np.float32))
    for b in range(B):
        eye = env.positions[b] + np.float32([0, 1.6, 0])
        assert env.voxels[b, int(eye[1]), int(eye[0]), int(eye[2])] == -1
        assert env.voxels[b, heights[b] - 1, sx, sz] >= 0   # the surface block below
    for _ in range(4):
        env.step(np.zeros(B, np.int64))
    if collision:
        # falling, never rising: at most 2 above the column, landed at its top
        assert (env.positions[:, 1] <= heights + 2).all()
    # a done step respawns on the same column surface
    for _ in range(27):
        env.step(np.zeros(B, np.int64))
    assert env.do_reset.all()
    assert np.array_equal(env.positions[:, 1], (heights + 2).astype(np.float32))


def test_flat_world_spawn_is_spawn_y():
    env = oracle.OracleEnv(batch_size=3, world_size=16, render_width=4, render_height=4)
    env.reset()
    assert (env.positions[:, 1] == 10.0).all()


@pytest.mark.parametrize("collision", [False, True])
def test_env_ids_step_a_sample_as_the_full_batch(collision):
    """OracleEnv(env_ids=idx) steps envs idx of a larger batch alone: the
    seeded worlds and spawn heights follow the env index, so state and images
    equal those rows of the full batch through done steps and their
    regeneration."""
    B, T = 40, 26
    idx = np.array([0, 3, 17, 39], dtype=np.int32)
    kw = dict(world_size=16, render_width=16, render_height=12, obs_dtype="uint8",
              episode_length=10, terrain="noise", terrain_amplitude=4, collision=collision, seed=7)
    full = oracle.OracleEnv(batch_size=B, **kw)
    part = oracle.OracleEnv(batch_size=len(idx), env_ids=idx, **kw)
    plain = oracle.OracleEnv(batch_size=len(idx), **kw)
    for e in (full, part, plain):
        e.reset()
    rng = np.random.default_rng(5)
    for t in range(T):
        a = rng.integers(0, 11, B)
        full.step(a)
        part.step(a[idx])
        plain.step(a[idx])
        for k in ("voxels", "positions", "velocities", "obs_buffer", "rewards", "step_count"):
            assert np.array_equal(getattr(full, k)[idx], getattr(part, k)), (t, k)
    # without the indices, envs 1..3 of the sample get envs 1..3's worlds
    assert not np.array_equal(plain.voxels[1:], part.voxels[1:])
