"""Pins of the oracle's voxel collision (MineEnv(collision=True)), which the
reference does not have (its ground is the plane y = world_y / 4,
src/kernels.cu:1566).  Each case sets a state whose outcome the rule fixes:
landing on the grass top of the flat world (y = ground_height, not Y / 4),
a wall that stops a walk, a dug-out column that lets the agent fall to the
next solid, and a ceiling that stops a jump.  Buttons: forward, back, left,
right, jump, break, place, sprint."""
import numpy as np

import oracle


def _env(B=1, **kw):
    e = oracle.OracleEnv(batch_size=B, world_size=16, render_width=8, render_height=8,
                         collision=True, nthreads=1, **kw)
    e.reset()
    return e


def _press(e, steps, fwd=0, jump=0, yaw=None):
    b = np.zeros((e.batch_size, 8), np.int8)
    b[:, 0], b[:, 4] = fwd, jump
    look = np.zeros((e.batch_size, 2), np.float32)
    for _ in range(steps):
        e.step(b, look)


def test_lands_on_the_grass_top_not_the_plane():
    e = _env()
    ys = []
    for _ in range(40):
        _press(e, 1)
        ys.append(float(e.positions[0, 1]))
    assert min(ys) == 8.0 and ys[-1] == 8.0 and e.on_ground[0] == 1
    # the reference's plane would put it at Y / 4 = 4, inside the dirt
    ref = oracle.OracleEnv(batch_size=1, world_size=16, render_width=8, render_height=8,
                           nthreads=1)
    ref.reset()
    _press(ref, 40)
    assert float(ref.positions[0, 1]) == 4.0


def test_a_wall_stops_the_walk():
    e = _env()
    _press(e, 30)                       # land at y = 8
    # face +x (yaw pi/2: forward moves +x) just west of a 2-high wall at x = 10
    e.voxels[0, 8:10, 10, :] = oracle.STONE
    e.positions[0] = [9.5, 8.0, 4.5]
    e.yaws[0] = np.float32(np.pi / 2)
    _press(e, 10, fwd=1)
    assert float(e.positions[0, 0]) < 10.0
    e.voxels[0, 9, 10, :] = oracle.AIR  # one block high: still blocked at the feet layer
    _press(e, 10, fwd=1)
    assert float(e.positions[0, 0]) < 10.0
    e.voxels[0, 8, 10, :] = oracle.AIR  # opening: the walk continues
    _press(e, 10, fwd=1)
    assert float(e.positions[0, 0]) > 10.0


def test_falls_through_a_dug_column_to_the_next_solid():
    e = _env()
    _press(e, 30)
    x, z = (int(v) for v in e.positions[0, [0, 2]])
    e.voxels[0, 3:8, x, z] = oracle.AIR  # dig out y = 3..7 under the agent
    _press(e, 40)
    assert float(e.positions[0, 1]) == 3.0 and e.on_ground[0] == 1


def test_a_ceiling_stops_the_jump():
    e = _env()
    _press(e, 30)
    x, z = (int(v) for v in e.positions[0, [0, 2]])
    e.voxels[0, 10, x, z] = oracle.STONE  # the first rise puts the head (8.35 + 1.8) in layer 10
    ys = []
    for _ in range(12):
        _press(e, 1, jump=1)
        ys.append(float(e.positions[0, 1]))
    assert ys == [8.0] * 12
    free = _env()
    _press(free, 30)
    top = 8.0
    for _ in range(12):
        _press(free, 1, jump=1)
        top = max(top, float(free.positions[0, 1]))
    assert abs(top - 9.4) < 1e-5  # apex of the free jump: v0 = 8, g = -20, dt = 0.05
