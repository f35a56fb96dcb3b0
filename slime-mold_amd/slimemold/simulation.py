"""Simulation: the host-side handle over the C ABI.

It carries the operations the reference performs on its GPU buffers:
step (src/main.rs:1163-1235), clear (:908-913), resize (:921-1085), agent
count change (:682-791), speed reassignment (:101-145) and parameter upload
(:83-99, the preset switch :815-832).  Agents are exchanged as an (n, 4)
float32 array of (x, y, angle, speed) rows, the trail as an (height, width)
float32 array.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Union

import numpy as np

from . import _native as N
from .settings import Settings


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _lut_bytes(lut):
    """A LutData, or 768 bytes / uint8 values, as a ctypes buffer."""
    if hasattr(lut, "to_bytes"):
        lut = lut.to_bytes()
    if isinstance(lut, (bytes, bytearray)):
        data = bytes(lut)
    else:
        data = np.asarray(lut, dtype=np.uint8).tobytes()
    if len(data) != 768:
        raise ValueError("a LUT is 768 bytes: R[256] G[256] B[256]")
    return (ctypes.c_uint8 * 768).from_buffer_copy(data)


def _params(p: Union[Settings, "N.Params", None]) -> "N.Params":
    if p is None:
        return N.default_params()
    if isinstance(p, Settings):
        return p.to_params()
    return p


class Simulation:
    def __init__(self, width: int, height: int, agent_count: int, seed: int = 0,
                 params: Union[Settings, "N.Params", None] = None, storage: str = "f32",
                 resort_interval: int = 0, graph_chunk: int = 100, sort_cell_log2: int = 4,
                 device: int = 0, xcd_remap: bool = True, f64_positions: bool = False,
                 tiled_sensing: bool = False, stochastic_rounding: bool = False,
                 window_sensing: bool = False, persistent: bool = False,
                 gaussian: bool = False):
        cfg = N.default_config()
        cfg.width = width
        cfg.height = height
        cfg.agent_count = agent_count
        cfg.seed = seed
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
        self._params = _params(params)
        h = ctypes.c_void_p()
        N.check(N.lib().slm_create(ctypes.byref(cfg), ctypes.byref(self._params), ctypes.byref(h)))
        self._h: Optional[ctypes.c_void_p] = h
        self.storage = storage

    @classmethod
    def from_snapshot(cls, data: Union[bytes, bytearray, str], device: int = 0) -> "Simulation":
        """A new simulation holding a snapshot's state (bytes, or a file path)."""
        if isinstance(data, str):
            with open(data, "rb") as f:
                data = f.read()
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        h = ctypes.c_void_p()
        N.check(N.lib().slm_create_from_snapshot(buf, len(data), device, ctypes.byref(h)))
        sim = cls.__new__(cls)
        sim._h = h
        sim._params = N.Params()
        N.check(N.lib().slm_get_params(h, ctypes.byref(sim._params)))
        sim.storage = {v: k for k, v in N.STORAGE_NAMES.items()}[sim.info().storage]
        return sim

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if self._h is not None:
            N.check(N.lib().slm_destroy(self._h))
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

    @property
    def handle(self):
        if self._h is None:
            raise N.SlimeMoldError(N.SLM_ERR_INVALID_ARGUMENT, "simulation is closed")
        return self._h

    # -- state --------------------------------------------------------------
    def info(self) -> "N.Info":
        i = N.Info()
        N.check(N.lib().slm_get_info(self.handle, ctypes.byref(i)))
        return i

    @property
    def width(self) -> int:
        return self.info().width

    @property
    def height(self) -> int:
        return self.info().height

    @property
    def agent_count(self) -> int:
        return self.info().agent_count

    def set_params(self, params: Union[Settings, "N.Params"]) -> None:
        p = _params(params)
        N.check(N.lib().slm_set_params(self.handle, ctypes.byref(p)))
        self._params = p

    def get_params(self) -> "N.Params":
        p = N.Params()
        N.check(N.lib().slm_get_params(self.handle, ctypes.byref(p)))
        return p

    # -- stepping -----------------------------------------------------------
    def step(self, n: int = 1) -> None:
        N.check(N.lib().slm_step(self.handle, n))

    def step_timed(self, n: int) -> "N.Timing":
        t = N.Timing()
        N.check(N.lib().slm_step_timed(self.handle, n, ctypes.byref(t)))
        return t

    def synchronize(self) -> None:
        N.check(N.lib().slm_synchronize(self.handle))

    # -- display and snapshots --------------------------------------------
    def render(self, lut, tex_w: int, tex_h: int) -> np.ndarray:
        """RGBA8 image (tex_h, tex_w, 4) of the trail through a 768-byte LUT
        (R[256] G[256] B[256]; see slimemold.lut)."""
        lut = _lut_bytes(lut)
        out = np.empty((tex_h, tex_w, 4), np.uint8)
        N.check(N.lib().slm_render_rgba8(self.handle, lut, tex_w, tex_h, out.ctypes.data))
        return out

    def step_render(self, n_steps: int, lut, tex_w: int, tex_h: int) -> np.ndarray:
        """Advance n_steps (>= 1) steps and return the last one's image as the
        reference's frame shows it: after the agents' deposits and the decay,
        before the diffuse (src/main.rs:1163-1235)."""
        lut = _lut_bytes(lut)
        out = np.empty((tex_h, tex_w, 4), np.uint8)
        N.check(N.lib().slm_step_render_rgba8(self.handle, n_steps, lut, tex_w, tex_h,
                                              out.ctypes.data))
        return out

    def snapshot(self, path: Optional[str] = None) -> bytes:
        """The simulation's resumable state (written to `path` if given)."""
        n = ctypes.c_uint64()
        N.check(N.lib().slm_snapshot_size(self.handle, ctypes.byref(n)))
        buf = (ctypes.c_uint8 * n.value)()
        N.check(N.lib().slm_save_snapshot(self.handle, buf, n.value))
        data = bytes(buf)
        if path is not None:
            with open(path, "wb") as f:
                f.write(data)
        return data

    def load_snapshot(self, data: Union[bytes, bytearray, str]) -> None:
        if isinstance(data, str):
            with open(data, "rb") as f:
                data = f.read()
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        N.check(N.lib().slm_load_snapshot(self.handle, buf, len(data)))

    # -- launch tuning ------------------------------------------------------
    def get_tuning(self) -> "N.Tuning":
        t = N.Tuning()
        N.check(N.lib().slm_get_tuning(self.handle, ctypes.byref(t)))
        return t

    def set_tuning(self, tuning: "N.Tuning" = None, **fields) -> None:
        """Apply a tuning; keyword fields override those of `tuning` (or of
        the current one)."""
        t = N.Tuning.from_buffer_copy(tuning) if tuning is not None else self.get_tuning()
        for k, v in fields.items():
            setattr(t, k, v)
        N.check(N.lib().slm_set_tuning(self.handle, ctypes.byref(t)))

    def autotune(self, trial_steps: int = 20) -> "N.Tuning":
        t = N.Tuning()
        N.check(N.lib().slm_autotune(self.handle, trial_steps, ctypes.byref(t)))
        return t

    # -- edits --------------------------------------------------------------
    def clear(self) -> None:
        N.check(N.lib().slm_clear(self.handle))

    def resize(self, width: int, height: int) -> None:
        N.check(N.lib().slm_resize(self.handle, width, height))

    def set_agent_count(self, n: int) -> None:
        N.check(N.lib().slm_set_agent_count(self.handle, n))

    def reassign_speeds(self, speed_min: float, speed_max: float) -> None:
        N.check(N.lib().slm_reassign_speeds(self.handle, speed_min, speed_max))

    # -- transfers ------------------------------------------------------------
    def read_agents(self) -> np.ndarray:
        n = self.agent_count
        out = np.empty((n, 4), dtype=np.float32)
        N.check(N.lib().slm_read_agents(self.handle, _fptr(out), n))
        return out

    def write_agents(self, agents: np.ndarray) -> None:
        a = np.ascontiguousarray(agents, dtype=np.float32).reshape(-1, 4)
        N.check(N.lib().slm_write_agents(self.handle, _fptr(a), a.shape[0]))

    def read_trail(self) -> np.ndarray:
        i = self.info()
        out = np.empty((i.height, i.width), dtype=np.float32)
        N.check(N.lib().slm_read_trail(self.handle, _fptr(out), out.size))
        return out

    def write_trail(self, trail: np.ndarray) -> None:
        t = np.ascontiguousarray(trail, dtype=np.float32)
        N.check(N.lib().slm_write_trail(self.handle, _fptr(t), t.size))

    def total_mass(self) -> float:
        m = ctypes.c_double(0.0)
        N.check(N.lib().slm_total_mass(self.handle, ctypes.byref(m)))
        return m.value
