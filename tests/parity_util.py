"""Helpers shared by the GPU parity tests: run MineEnv (HIP) and OracleEnv
(CPU restatement) side by side on identical actions and compare every output
and state tensor after every call.

Tolerances (stated by the parity contract for this path):
  * voxels, rewards, dones, step_count, on_ground: bit-exact;
  * positions, velocities, yaws, pitches, cameras: within 1 ulp of float32;
  * rendered RGB: bit-exact, or at most 0.1% of pixels differing, each by at
    most 1 LSB (uint8) -- float32 observations: at most 1 ulp.
Every comparison also records whether it was bit-exact, which is what the
build actually achieves (shared IEEE numerics, no FP contraction).
"""
from __future__ import annotations

import numpy as np

POS_ULP = 1
DECODE_BUFFERS = ("forward_in", "strafe_in", "delta_yaw_in", "delta_pitch_in", "jump_in",
                  "do_break", "do_place", "speed_mult")
OBS_DIFF_RATE = 1e-3
OBS_MAX_LSB = 1


def ulp_dist(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in units in the last place between float32 arrays."""
    ia = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


class Stats:
    def __init__(self):
        self.exact = True
        self.obs_diff_px = 0
        self.obs_px = 0
        self.max_state_ulp = 0

    def summary(self) -> str:
        rate = self.obs_diff_px / max(1, self.obs_px)
        return (f"bit-exact={self.exact} obs-diff-pixels={self.obs_diff_px}/{self.obs_px} "
                f"({rate:.2e}) max-state-ulp={self.max_state_ulp}")


def compare_obs(g: np.ndarray, o: np.ndarray, stats: Stats, where: str):
    assert g.shape == o.shape and g.dtype == o.dtype, where
    npx = int(np.prod(g.shape[:-1]))
    stats.obs_px += npx
    if g.dtype == np.uint8:
        d = np.abs(g.astype(np.int16) - o.astype(np.int16))
        bad = d.max(axis=-1) > 0
        nbad = int(bad.sum())
        stats.obs_diff_px += nbad
        if nbad:
            stats.exact = False
            assert d.max() <= OBS_MAX_LSB, f"{where}: obs differs by {d.max()} LSB"
            assert nbad <= OBS_DIFF_RATE * npx, f"{where}: {nbad}/{npx} pixels differ"
    else:
        u = ulp_dist(g, o)
        nbad = int((u.max(axis=-1) > 0).sum())
        stats.obs_diff_px += nbad
        if nbad:
            stats.exact = False
            assert u.max() <= 1, f"{where}: float obs differs by {u.max()} ulp"
            assert nbad <= OBS_DIFF_RATE * npx, f"{where}: {nbad}/{npx} pixels differ"


def compare_exact(name: str, g: np.ndarray, o: np.ndarray, where: str):
    assert g.shape == o.shape, f"{where}: {name} shape {g.shape} vs {o.shape}"
    if not np.array_equal(g, o):
        idx = np.argwhere(g != o)[:5]
        raise AssertionError(f"{where}: {name} differs at {idx.tolist()} "
                             f"gpu={g[tuple(idx[0])]} oracle={o[tuple(idx[0])]}")


def compare_float(name: str, g: np.ndarray, o: np.ndarray, stats: Stats, where: str):
    u = ulp_dist(g, o)
    m = int(u.max()) if u.size else 0
    stats.max_state_ulp = max(stats.max_state_ulp, m)
    if m:
        stats.exact = False
    assert m <= POS_ULP, f"{where}: {name} differs by {m} ulp at {np.argwhere(u == m)[:3].tolist()}"


def compare_env(env, orc, stats: Stats, where: str, check_obs=True):
    """Compare every state tensor of a MineEnv with an OracleEnv."""
    compare_exact("voxels", env.voxels.cpu().numpy(), orc.voxels, where)
    compare_exact("on_ground", env.on_ground.cpu().numpy().astype(np.uint8),
                  orc.on_ground.astype(np.uint8), where)
    compare_exact("step_count", env.step_count.cpu().numpy(), orc.step_count, where)
    for k in ("positions", "velocities", "yaws", "pitches", "cameras"):
        compare_float(k, getattr(env, k).cpu().numpy(), getattr(orc, k), stats, where)
    # The decode buffers (minecu/__init__.py:243-251): bit-exact.
    for k in DECODE_BUFFERS:
        g = getattr(env, k).cpu().numpy()
        o = getattr(orc, k)
        if g.dtype == np.bool_:
            compare_exact(k, g.astype(np.uint8), o.astype(np.uint8), where)
        else:
            compare_exact(k, g.view(np.int32), o.view(np.int32), where)
    if check_obs:
        compare_obs(env.obs_buffer.cpu().numpy(), orc.obs_buffer, stats, where)


def run_pair(env, orc, actions, stats: Stats, tag: str, look=None, check_every=1):
    """Step both envs over actions[t] (and look[t] when given); compare."""
    import torch
    for t in range(len(actions)):
        a = actions[t]
        if look is None:
            og, rg, dg = env.step(torch.as_tensor(a).to(env.device))
            oo, ro, do = orc.step(np.asarray(a))
        else:
            og, rg, dg = env.step(torch.as_tensor(a).to(env.device),
                                  torch.as_tensor(look[t]).to(env.device))
            oo, ro, do = orc.step(np.asarray(a), np.asarray(look[t]))
        where = f"{tag} step {t}"
        compare_exact("rewards", rg.cpu().numpy(), ro, where)
        compare_exact("dones", dg.cpu().numpy(), do, where)
        if (t % check_every) == 0 or t == len(actions) - 1:
            compare_env(env, orc, stats, where)
    return stats


SAMPLED_KEYS = ("voxels", "positions", "velocities", "yaws", "pitches", "on_ground", "step_count",
                "cameras", "rewards", "obs_buffer") + DECODE_BUFFERS


def compare_sampled(env, idx, orc, stats: Stats, where: str, chw: bool = False):
    """Envs idx of a (large) MineEnv against an OracleEnv that stepped exactly
    those envs (envs are independent, and episodes run in lockstep from one
    reset, so the group-wide regeneration happens on the same steps).  chw:
    the MineEnv writes planar [B, 3, H, W] images, compared transposed back to
    the oracle's [B, H, W, 3]."""
    import torch
    ti = torch.as_tensor(idx, device=env.device)
    for k in SAMPLED_KEYS:
        g = getattr(env, k).index_select(0, ti).cpu().numpy()
        o = getattr(orc, k)
        if k == "obs_buffer" and chw:      # compared as [B, H, W, 3] pixels
            g = np.ascontiguousarray(g.transpose(0, 2, 3, 1))
        if k == "obs_buffer":
            compare_obs(g, o, stats, where)
        elif g.dtype in (np.float32,) and k in ("positions", "velocities", "yaws", "pitches",
                                                 "cameras"):
            compare_float(k, g, o, stats, where)
        elif g.dtype == np.float32:
            compare_exact(k, g.view(np.int32), o.view(np.int32), where)
        else:
            compare_exact(k, g.astype(np.int64) if g.dtype == np.bool_ else g,
                          o.astype(np.int64) if g.dtype == np.bool_ else o, where)
