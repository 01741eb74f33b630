"""Batched env groups: several ``MineEnv`` groups -- uneven batch sizes, world
sizes, resolutions, obs modes -- stepped together in one persistent gfx950
launch (``mc_batch_step``), with launch shapes picked by the library's
autotuner (``mc_batch_autotune``).

The reference has no counterpart (one ``MineEnv`` per process-global CUDA
graph, src/kernels.cu:2338-2382).  Each group's results are bit-identical to
stepping it alone with ``MineEnv.step``.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib


class EnvBatch:
    """Steps a fixed or changing set of MineEnv groups in single launches."""

    def __init__(self, persistent: bool = True):
        h = ctypes.c_void_p()
        _lib.check(_lib.load().mc_batch_create(ctypes.byref(h)), "EnvBatch")
        self._h = h
        _lib.check(_lib.load().mc_batch_set_persistent(h, 1 if persistent else 0), "EnvBatch")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().mc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _handles(envs) -> ctypes.Array:
        arr = (ctypes.c_void_p * len(envs))()
        for i, e in enumerate(envs):
            arr[i] = e._handle.value
        return arr

    def autotune(self, envs: Sequence, steps: int = 12) -> List[Tuple[int, int]]:
        """Time candidate launch shapes per distinct group shape on scratch
        groups and apply the fastest to every env; returns (threads,
        waves_per_env) per env."""
        if not envs:
            return []
        dev = envs[0].device
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(_lib.load().mc_batch_autotune(self._h, self._handles(envs), len(envs), steps,
                                                 stream), "autotune")
        return [e.shape for e in envs]

    def step(self, envs: Sequence, actions: Sequence[torch.Tensor],
             looks: Optional[Sequence[Optional[torch.Tensor]]] = None):
        """One step of every env group in one launch.  actions[i] is an
        integer [B_i] tensor of discrete ids, or int8 [B_i, 8] buttons with
        looks[i] float32 [B_i, 2].  Returns [(obs, rewards, dones)] per group."""
        n = len(envs)
        if n == 0:
            return []
        if len(actions) != n or (looks is not None and len(looks) != n):
            raise ValueError("one actions (and look) entry per env group")
        dev = envs[0].device
        for e in envs:
            if e.device != dev:
                raise ValueError("all env groups of a batch must be on one device")
        keep = []
        act_p = (ctypes.c_void_p * n)()
        look_p = (ctypes.c_void_p * n)()
        dones_p = (ctypes.c_void_p * n)()
        kinds = (ctypes.c_int * n)()
        dones = []
        for i, e in enumerate(envs):
            lk = None if looks is None else looks[i]
            a, kind, l = e._prepare_actions(actions[i], lk)
            keep.extend([a, l])
            act_p[i] = a.data_ptr()
            kinds[i] = kind
            look_p[i] = l.data_ptr() if l is not None else None
            d = torch.empty((e.batch_size,), dtype=torch.bool, device=dev)
            dones.append(d)
            dones_p[i] = d.data_ptr()
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(_lib.load().mc_batch_step(self._h, self._handles(envs), act_p, kinds, look_p,
                                             dones_p, n, stream), "batch step")
        return [(e.obs_buffer, e.rewards, dones[i]) for i, e in enumerate(envs)]
