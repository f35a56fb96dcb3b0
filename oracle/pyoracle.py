"""ctypes binding of the C oracle (slm_oracle.h).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

STORAGE = {"f32": 0, "bf16": 1, "f16": 2, "bf16sr": 3}


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in (
        "decay_factor", "agent_jitter", "agent_speed_min", "agent_speed_max", "agent_turn_speed",
        "agent_sensor_angle", "agent_sensor_distance", "diffusion_rate",
        "pheromone_deposition_amount", "blur_radius", "blur_sigma")]


_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_sincos": (None, [ctypes.c_float, _F, _F]),
            "orc_set_libm_trig": (None, [ctypes.c_int]),
            "orc_wgsl_mod": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
            "orc_fract": (ctypes.c_float, [ctypes.c_float]),
            "orc_mix": (ctypes.c_float, [ctypes.c_float, ctypes.c_float, ctypes.c_float]),
            "orc_sign": (ctypes.c_float, [ctypes.c_float]),
            "orc_f32_to_i32": (ctypes.c_int32, [ctypes.c_float]),
            "orc_round_storage": (ctypes.c_float, [ctypes.c_float, ctypes.c_int]),
            "orc_sr_bits": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
            "orc_f32_to_bf16_sr": (ctypes.c_uint16, [ctypes.c_float, ctypes.c_uint32]),
            "orc_f32_to_bf16": (ctypes.c_uint16, [ctypes.c_float]),
            "orc_bf16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
            "orc_f32_to_f16": (ctypes.c_uint16, [ctypes.c_float]),
            "orc_f16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
            "orc_splitmix64_at": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
            "orc_uniform24": (ctypes.c_float, [ctypes.c_uint64, ctypes.c_uint64]),
            "orc_sample": (ctypes.c_float, [_F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                            ctypes.c_float]),
            "orc_init_agents": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint64, ctypes.c_float, ctypes.c_float, _F]),
            "orc_reassign_speeds": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                           ctypes.c_float, ctypes.c_float, _F]),
            "orc_agent_update": (ctypes.c_int64, [_F, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.POINTER(OrcParams), ctypes.c_uint32, _F]),
            "orc_trail_pass": (None, [_F, _U8, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(OrcParams), ctypes.c_int, _F, ctypes.c_int]),
            "orc_step": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _F, _F,
                                        ctypes.POINTER(OrcParams), ctypes.c_int, ctypes.c_uint32,
                                        ctypes.c_int]),
            "orc_step_f64": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _D,
                                            _D, ctypes.POINTER(OrcParams), ctypes.c_uint32,
                                            ctypes.c_int]),
            "orc_step_p64": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _D,
                                            _F, ctypes.POINTER(OrcParams), ctypes.c_int,
                                            ctypes.c_uint32, ctypes.c_int]),
            "orc_trail_pass_f64": (None, [_D, _U8, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.POINTER(OrcParams), _D, ctypes.c_int]),
            "orc_trail_pass_gauss": (None, [_F, _U8, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.POINTER(OrcParams), ctypes.c_int, _F,
                                            ctypes.c_int]),
            "orc_gauss_weights": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, _F]),
            "orc_set_gaussian": (None, [ctypes.c_int]),
            "orc_total_mass": (ctypes.c_double, [_F, ctypes.c_uint64]),
            "orc_total_mass_f64": (ctypes.c_double, [_D, ctypes.c_uint64]),
            "orc_render": (None, [_F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def params(p) -> OrcParams:
    """Accepts an object with the slm_params field names (Params, dict, Settings)."""
    if isinstance(p, OrcParams):
        return p
    if isinstance(p, dict):
        d = p
    elif hasattr(p, "to_params"):
        d = p.to_params().as_dict()
    else:
        d = {n: getattr(p, n) for n, _ in OrcParams._fields_}
    return OrcParams(**{n: float(d[n]) for n, _ in OrcParams._fields_})


def _f(a):
    return a.ctypes.data_as(_F)


def sincos(x: float):
    s, c = ctypes.c_float(), ctypes.c_float()
    lib().orc_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def set_libm_trig(on: bool) -> None:
    """Agent updates use the C library's sinf/cosf (True) or the shared trig."""
    lib().orc_set_libm_trig(1 if on else 0)


def init_agents(seed: int, w: int, h: int, n: int, smin: float, smax: float) -> np.ndarray:
    a = np.empty((n, 4), dtype=np.float32)
    lib().orc_init_agents(seed, w, h, n, smin, smax, _f(a))
    return a


def reassign_speeds(agents: np.ndarray, seed: int, stream: int, smin: float, smax: float) -> None:
    assert agents.dtype == np.float32 and agents.flags.c_contiguous
    lib().orc_reassign_speeds(seed, stream, agents.shape[0], smin, smax, _f(agents))


def sample(trail: np.ndarray, px: float, py: float) -> float:
    t = np.ascontiguousarray(trail, dtype=np.float32)
    return lib().orc_sample(_f(t), t.shape[1], t.shape[0], px, py)


def agent_update(trail: np.ndarray, p, index: int, agent) -> tuple:
    """One agent against a read-only trail -> (new agent (4,), deposit index or -1)."""
    t = np.ascontiguousarray(trail, dtype=np.float32)
    a = np.array(agent, dtype=np.float32).reshape(4)
    op = params(p)
    d = lib().orc_agent_update(_f(t), t.shape[1], t.shape[0], ctypes.byref(op), index, _f(a))
    return a, d


def trail_pass(trail: np.ndarray, p, mark=None, storage: str = "f32", nthreads: int = 0,
               gaussian: bool = False) -> np.ndarray:
    t = np.ascontiguousarray(trail, dtype=np.float32)
    out = np.empty_like(t)
    op = params(p)
    m = None
    if mark is not None:
        mk = np.ascontiguousarray(mark, dtype=np.uint8)
        m = mk.ctypes.data_as(_U8)
    fn = lib().orc_trail_pass_gauss if gaussian else lib().orc_trail_pass
    fn(_f(t), m, t.shape[1], t.shape[0], ctypes.byref(op), STORAGE[storage], _f(out), nthreads)
    return out


def gauss_weights(radius: float, sigma: float):
    """(R, float32 weights w[0..R]) of the Gaussian diffuse extension; R = 0 if invalid."""
    w = np.zeros(9, np.float32)
    r = lib().orc_gauss_weights(radius, sigma, _f(w))
    return r, w[:r + 1].copy()



def trail_pass_f64(trail: np.ndarray, p, mark=None, nthreads: int = 0) -> np.ndarray:
    """The trail pass in double precision (FP64 reference)."""
    t = np.ascontiguousarray(trail, dtype=np.float64)
    out = np.empty_like(t)
    m = None
    if mark is not None:
        mk = np.ascontiguousarray(mark, dtype=np.uint8)
        m = mk.ctypes.data_as(_U8)
    lib().orc_trail_pass_f64(t.ctypes.data_as(_D), m, t.shape[1], t.shape[0],
                             ctypes.byref(params(p)), out.ctypes.data_as(_D), nthreads)
    return out


def deposit_marks(agents: np.ndarray, w: int, h: int) -> np.ndarray:
    """Cells the agents (positions after their move) deposit into (compute.wgsl:136-141)."""
    x = agents[:, 0].astype(np.float64)
    y = agents[:, 1].astype(np.float64)
    ok = (x >= 0) & (x < w) & (y >= 0) & (y < h)
    mark = np.zeros((h, w), np.uint8)
    mark[np.floor(y[ok]).astype(np.int64), np.floor(x[ok]).astype(np.int64)] = 1
    return mark


def step(agents: np.ndarray, trail: np.ndarray, p, n_steps: int = 1, storage: str = "f32",
         nthreads: int = 0, gaussian: bool = False) -> None:
    """n_steps in place on (n,4) float32 agents and (h,w) float32 trail."""
    assert agents.dtype == np.float32 and agents.flags.c_contiguous
    assert trail.dtype == np.float32 and trail.flags.c_contiguous
    op = params(p)
    lib().orc_set_gaussian(1 if gaussian else 0)
    try:
        rc = lib().orc_step(trail.shape[1], trail.shape[0], agents.shape[0], _f(agents),
                            _f(trail), ctypes.byref(op), STORAGE[storage], n_steps, nthreads)
    finally:
        lib().orc_set_gaussian(0)
    if rc != 0:
        raise RuntimeError(f"orc_step failed: {rc}")


def step_p64(agents: np.ndarray, trail: np.ndarray, p, n_steps: int = 1, storage: str = "f32",
             nthreads: int = 0, gaussian: bool = False) -> None:
    """f64-position variant: (n,4) float64 agents, (h,w) float32 trail, in place."""
    assert agents.dtype == np.float64 and agents.flags.c_contiguous
    assert trail.dtype == np.float32 and trail.flags.c_contiguous
    op = params(p)
    lib().orc_set_gaussian(1 if gaussian else 0)
    try:
        rc = lib().orc_step_p64(trail.shape[1], trail.shape[0], agents.shape[0],
                                agents.ctypes.data_as(_D), _f(trail), ctypes.byref(op),
                                STORAGE[storage], n_steps, nthreads)
    finally:
        lib().orc_set_gaussian(0)
    if rc != 0:
        raise RuntimeError(f"orc_step_p64 failed: {rc}")


def step_f64(agents: np.ndarray, trail: np.ndarray, p, n_steps: int = 1, nthreads: int = 0) -> None:
    assert agents.dtype == np.float64 and trail.dtype == np.float64
    op = params(p)
    rc = lib().orc_step_f64(trail.shape[1], trail.shape[0], agents.shape[0],
                            agents.ctypes.data_as(_D), trail.ctypes.data_as(_D), ctypes.byref(op),
                            n_steps, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_step_f64 failed: {rc}")


def total_mass(trail: np.ndarray) -> float:
    t = np.ascontiguousarray(trail)
    if t.dtype == np.float64:
        return lib().orc_total_mass_f64(t.ctypes.data_as(_D), t.size)
    t = t.astype(np.float32, copy=False)
    return lib().orc_total_mass(_f(t), t.size)


def round_storage(values: np.ndarray, storage: str) -> np.ndarray:
    v = np.asarray(values, dtype=np.float32)
    out = np.empty_like(v)
    f = lib().orc_round_storage
    flat_in, flat_out = v.reshape(-1), out.reshape(-1)
    for i in range(flat_in.size):
        flat_out[i] = f(float(flat_in[i]), STORAGE[storage])
    return out


def render(trail: np.ndarray, lut, tex_w: int, tex_h: int) -> np.ndarray:
    """RGBA8 image (tex_h, tex_w, 4) of the trail through a 768-byte LUT."""
    t = np.ascontiguousarray(trail, dtype=np.float32)
    h, w = t.shape
    lut = np.ascontiguousarray(np.frombuffer(bytes(lut), dtype=np.uint8))
    assert lut.size == 768
    out = np.empty((tex_h, tex_w, 4), np.uint8)
    lib().orc_render(t.ctypes.data_as(_F), w, h, lut.ctypes.data, tex_w, tex_h, out.ctypes.data)
    return out
