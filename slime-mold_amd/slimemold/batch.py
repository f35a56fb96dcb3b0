"""Batch: many independent simulations of uneven sizes advanced together.

No counterpart in the reference, which runs one simulation per process
(src/main.rs:147).  Each step is one agent launch and one trail launch over
all instances (an instance-offset table drives both), or with instance groups
one such pair per group, spread over four streams; instance k gives bit for
bit what a Simulation with the same size, agent count, seed and parameters
gives.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Union

import numpy as np

from . import _native as N
from .settings import Settings
from .simulation import _fptr, _params


@dataclass
class InstanceSpec:
    width: int
    height: int
    agent_count: int
    seed: int = 0
    params: Union[Settings, "N.Params", None] = None


class Batch:
    def __init__(self, instances: Sequence[InstanceSpec], storage: str = "f32",
                 resort_interval: int = 0, graph_chunk: int = 100, sort_cell_log2: int = 4,
                 device: int = 0, xcd_remap: bool = True, f64_positions: bool = False,
                 tiled_sensing: bool = False, stochastic_rounding: bool = False,
                 window_sensing: bool = False, persistent: bool = False,
                 gaussian: bool = False):
        cfg = N.default_batch_config()
        cfg.storage = N.STORAGE_NAMES[storage]
        cfg.device = device
        cfg.resort_interval = resort_interval
        cfg.graph_chunk = graph_chunk
        cfg.sort_cell_log2 = sort_cell_log2
        cfg.flags = ((0 if xcd_remap else N.FLAG_NO_XCD_REMAP)
                     | (N.FLAG_F64_POSITIONS if f64_positions else 0)
                     | (N.FLAG_TILED_SENSING if tiled_sensing else 0)
                     | (N.FLAG_WINDOW_SENSING if window_sensing else 0)
                     | (N.FLAG_PERSISTENT if persistent else 0)
                     | (N.FLAG_STOCHASTIC_ROUNDING if stochastic_rounding else 0)
                     | (N.FLAG_GAUSSIAN_DIFFUSE if gaussian else 0))
        descs = (N.InstanceDesc * len(instances))()
        for d, spec in zip(descs, instances):
            d.width, d.height = spec.width, spec.height
            d.agent_count, d.seed = spec.agent_count, spec.seed
            d.params = _params(spec.params)
        h = ctypes.c_void_p()
        N.check(N.lib().slm_batch_create(ctypes.byref(cfg), descs, len(instances), ctypes.byref(h)))
        self._h: Optional[ctypes.c_void_p] = h
        self.specs: List[InstanceSpec] = list(instances)
        self.storage = storage

    def close(self) -> None:
        if self._h is not None:
            N.check(N.lib().slm_batch_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.specs)

    @property
    def handle(self):
        if self._h is None:
            raise N.SlimeMoldError(N.SLM_ERR_INVALID_ARGUMENT, "batch is closed")
        return self._h

    def info(self) -> "N.BatchInfo":
        i = N.BatchInfo()
        N.check(N.lib().slm_batch_get_info(self.handle, ctypes.byref(i)))
        return i

    def step(self, n: int = 1) -> None:
        N.check(N.lib().slm_batch_step(self.handle, n))

    def step_timed(self, n: int) -> "N.Timing":
        t = N.Timing()
        N.check(N.lib().slm_batch_step_timed(self.handle, n, ctypes.byref(t)))
        return t

    def synchronize(self) -> None:
        N.check(N.lib().slm_batch_synchronize(self.handle))

    def render(self, k: int, lut, tex_w: int, tex_h: int) -> np.ndarray:
        """RGBA8 image (tex_h, tex_w, 4) of instance k's trail through a LUT."""
        from .simulation import _lut_bytes
        out = np.empty((tex_h, tex_w, 4), np.uint8)
        N.check(N.lib().slm_batch_render_rgba8(self.handle, k, _lut_bytes(lut), tex_w, tex_h,
                                               out.ctypes.data))
        return out

    # -- launch tuning ------------------------------------------------------
    def get_tuning(self) -> "N.Tuning":
        t = N.Tuning()
        N.check(N.lib().slm_batch_get_tuning(self.handle, ctypes.byref(t)))
        return t

    def set_tuning(self, tuning: "N.Tuning" = None, **fields) -> None:
        """Apply a tuning; keyword fields override those of `tuning` (or of
        the current one)."""
        t = N.Tuning.from_buffer_copy(tuning) if tuning is not None else self.get_tuning()
        for k, v in fields.items():
            setattr(t, k, v)
        N.check(N.lib().slm_batch_set_tuning(self.handle, ctypes.byref(t)))

    def autotune(self, trial_steps: int = 20) -> "N.Tuning":
        t = N.Tuning()
        N.check(N.lib().slm_batch_autotune(self.handle, trial_steps, ctypes.byref(t)))
        return t

    def set_params(self, k: int, params) -> None:
        p = _params(params)
        N.check(N.lib().slm_batch_set_params(self.handle, k, ctypes.byref(p)))

    def clear(self, k: int) -> None:
        N.check(N.lib().slm_batch_clear(self.handle, k))

    def read_agents(self, k: int) -> np.ndarray:
        n = self.specs[k].agent_count
        out = np.empty((n, 4), dtype=np.float32)
        N.check(N.lib().slm_batch_read_agents(self.handle, k, _fptr(out), n))
        return out

    def write_agents(self, k: int, agents: np.ndarray) -> None:
        a = np.ascontiguousarray(agents, dtype=np.float32).reshape(-1, 4)
        N.check(N.lib().slm_batch_write_agents(self.handle, k, _fptr(a), a.shape[0]))

    def read_trail(self, k: int) -> np.ndarray:
        s = self.specs[k]
        out = np.empty((s.height, s.width), dtype=np.float32)
        N.check(N.lib().slm_batch_read_trail(self.handle, k, _fptr(out), out.size))
        return out

    def write_trail(self, k: int, trail: np.ndarray) -> None:
        t = np.ascontiguousarray(trail, dtype=np.float32)
        N.check(N.lib().slm_batch_write_trail(self.handle, k, _fptr(t), t.size))

    def total_mass(self) -> np.ndarray:
        out = np.empty(len(self.specs), dtype=np.float64)
        N.check(N.lib().slm_batch_total_mass(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(self.specs)))
        return out
