"""Continuous-batching step server over device-resident env groups.

``StepServer`` wraps the library's native server (mc_server_*): a C++ loop
thread that merges the pending step requests of every admitted ``MineEnv``
group into one persistent batched launch per free stream, admits groups FIFO
under an HBM budget, and timestamps every request (submit, launch,
completion).  The reference has no serving layer; this is the serving form of
its batched step.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib

WAITING, ADMITTED, REMOVED = 0, 1, 2


class Record(ctypes.Structure):
    _fields_ = [("req_id", ctypes.c_int64), ("group", ctypes.c_int64),
                ("t_submit_ns", ctypes.c_int64), ("t_launch_ns", ctypes.c_int64),
                ("t_done_ns", ctypes.c_int64)]


def env_bytes(env) -> int:
    """Device bytes a MineEnv group keeps resident (its state tensors)."""
    return sum(getattr(env, k).numel() * getattr(env, k).element_size() for k in env._STATE_KEYS)


class StepServer:
    def __init__(self, hbm_budget_bytes: int = 64 << 30, n_streams: int = 2):
        L = _lib.load()
        h = ctypes.c_void_p()
        st = L.mc_server_create(ctypes.c_int64(hbm_budget_bytes), n_streams, ctypes.byref(h))
        if st != 0:
            raise _lib.MineCuError(L.mc_server_last_error(None).decode())
        self._h = h
        self._envs: Dict[int, object] = {}

    def _check(self, st: int, what: str):
        if st != 0:
            raise _lib.MineCuError(f"{what}: {_lib.load().mc_server_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().mc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- groups -----------------------------------------------------------------
    def add_group(self, env, nbytes: Optional[int] = None) -> int:
        """Register a MineEnv group; admitted now if it fits the budget, else
        queued until earlier groups are removed."""
        gid = ctypes.c_int()
        self._check(_lib.load().mc_server_add_group(
            self._h, env._handle, ctypes.c_int64(env_bytes(env) if nbytes is None else nbytes),
            env.batch_size, ctypes.byref(gid)), "add_group")
        self._envs[gid.value] = env
        return gid.value

    def state(self, gid: int) -> int:
        s = ctypes.c_int()
        self._check(_lib.load().mc_server_group_state(self._h, gid, ctypes.byref(s)), "state")
        return s.value

    def remove_group(self, gid: int) -> None:
        self._check(_lib.load().mc_server_remove_group(self._h, gid), "remove_group")

    # -- requests -----------------------------------------------------------------
    def submit(self, gid: int, actions: torch.Tensor, look: Optional[torch.Tensor] = None,
               dones: Optional[torch.Tensor] = None) -> int:
        """Queue one step of group gid.  The tensors must stay alive until the
        request completes (see drain/records)."""
        env = self._envs[gid]
        a, kind, lk = env._prepare_actions(actions, look)
        if a.data_ptr() != actions.data_ptr():
            raise ValueError("actions must already be a device tensor of the right dtype")
        rid = ctypes.c_int64()
        self._check(_lib.load().mc_server_submit(
            self._h, gid, ctypes.c_void_p(a.data_ptr()), kind,
            ctypes.c_void_p(lk.data_ptr()) if lk is not None else None,
            ctypes.c_void_p(dones.data_ptr()) if dones is not None else None,
            ctypes.byref(rid)), "submit")
        return rid.value

    def replay(self, gid: int, actions: torch.Tensor) -> None:
        """Closed loop: step group gid through actions[0..T) (discrete ids,
        [T, B] int64 on the device), each step submitted when the previous
        completes."""
        env = self._envs[gid]
        if actions.dim() != 2 or actions.shape[1] != env.batch_size or actions.dtype != torch.int64 \
                or not actions.is_contiguous() or actions.device != env.device:
            raise ValueError("replay actions must be contiguous int64 [T, B] on the env's device")
        L = _lib.load()
        self._check(L.mc_server_set_replay(self._h, gid, ctypes.c_void_p(actions.data_ptr()),
                                           ctypes.c_int64(actions.stride(0) * 8),
                                           ctypes.c_int64(actions.shape[0]), _lib.MC_ACT_DISCRETE,
                                           None, ctypes.c_int64(0), None), "set_replay")
        self._check(L.mc_server_start_replay(self._h, gid), "start_replay")

    def schedule(self, gids, actions, t_offsets_s) -> None:
        """Open loop at planned times: request k steps group gids[k] with
        actions[k] (a [B] int64 device tensor, kept alive by the caller until
        drain) t_offsets_s[k] seconds from now (non-decreasing).  Each request
        is queued by the server's loop thread at its arrival time, which is
        also its recorded submit time."""
        n = len(gids)
        if len(actions) != n or len(t_offsets_s) != n:
            raise ValueError("gids, actions and t_offsets_s must have the same length")
        g = np.ascontiguousarray(np.asarray(gids, dtype=np.int32))
        ptr = np.empty(n, dtype=np.uint64)
        for k, (gid, a) in enumerate(zip(g.tolist(), actions)):
            env = self._envs[gid]
            if a.dtype != torch.int64 or a.device != env.device or a.numel() != env.batch_size \
                    or not a.is_contiguous():
                raise ValueError("scheduled actions must be contiguous int64 [B] on the env's device")
            ptr[k] = a.data_ptr()
        t = np.ascontiguousarray(np.round(np.asarray(t_offsets_s, dtype=np.float64) * 1e9).astype(np.int64))
        self._check(_lib.load().mc_server_schedule(
            self._h, g.ctypes.data_as(ctypes.c_void_p), ptr.ctypes.data_as(ctypes.c_void_p),
            _lib.MC_ACT_DISCRETE, t.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(n)), "schedule")

    def drain(self) -> None:
        self._check(_lib.load().mc_server_drain(self._h), "drain")

    def records(self) -> np.ndarray:
        """Completed requests since the last call: structured array with
        req_id, group, t_submit_ns, t_launch_ns, t_done_ns."""
        L = _lib.load()
        out: List[np.ndarray] = []
        buf = (Record * 65536)()
        while True:
            n = ctypes.c_int()
            self._check(L.mc_server_take_records(self._h, buf, 65536, ctypes.byref(n)), "records")
            if n.value == 0:
                break
            out.append(np.ctypeslib.as_array(buf)[: n.value].copy())
            if n.value < 65536:
                break
        dt = np.dtype([(f[0], np.int64) for f in Record._fields_])
        return np.concatenate(out).view(dt).reshape(-1) if out else np.zeros(0, dt)

    def stats(self) -> dict:
        v = [ctypes.c_int64() for _ in range(4)]
        self._check(_lib.load().mc_server_stats(self._h, *[ctypes.byref(x) for x in v]), "stats")
        return dict(zip(("launches", "env_steps", "requests", "used_bytes"), (x.value for x in v)))

    def autotune(self, steps: int = 12) -> None:
        dev = next(iter(self._envs.values())).device if self._envs else None
        stream = torch.cuda.current_stream(dev).cuda_stream if dev is not None else 0
        self._check(_lib.load().mc_server_autotune(self._h, steps, stream), "autotune")
