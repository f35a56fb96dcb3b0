"""ctypes binding of libslimemold_hip.so (declarations: include/slimemold.h).

The library is loaded from this directory, where the build puts it.  There is
no fallback: if the library is missing every call raises NativeLibraryMissing.
"""
from __future__ import annotations

import ctypes
import os

LIB_NAME = "libslimemold_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

SLM_OK = 0
SLM_ERR_INVALID_ARGUMENT = -1
SLM_ERR_OUT_OF_MEMORY = -2
SLM_ERR_HIP = -3
SLM_ERR_NO_DEVICE = -4
SLM_ERR_UNSUPPORTED = -5

STORAGE_F32 = 0
STORAGE_BF16 = 1
STORAGE_F16 = 2
STORAGE_NAMES = {"f32": STORAGE_F32, "bf16": STORAGE_BF16, "f16": STORAGE_F16}

FLAG_NO_XCD_REMAP = 0x1
FLAG_F64_POSITIONS = 0x2
FLAG_TILED_SENSING = 0x4
FLAG_STOCHASTIC_ROUNDING = 0x8
FLAG_GAUSSIAN_DIFFUSE = 0x10
FLAG_WINDOW_SENSING = 0x20
FLAG_PERSISTENT = 0x40
SORT_HEADING_MASK = 0x7
SORT_AT_SENSOR = 0x10


class NativeLibraryMissing(RuntimeError):
    pass


class SlimeMoldError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"slimemold error {code}: {message}")
        self.code = code
        self.message = message


class Params(ctypes.Structure):
    """slm_params - SimSizeUniform after width/height (src/main.rs:29-46)."""

    _fields_ = [
        ("decay_factor", ctypes.c_float),
        ("agent_jitter", ctypes.c_float),
        ("agent_speed_min", ctypes.c_float),
        ("agent_speed_max", ctypes.c_float),
        ("agent_turn_speed", ctypes.c_float),
        ("agent_sensor_angle", ctypes.c_float),
        ("agent_sensor_distance", ctypes.c_float),
        ("diffusion_rate", ctypes.c_float),
        ("pheromone_deposition_amount", ctypes.c_float),
        ("blur_radius", ctypes.c_float),
        ("blur_sigma", ctypes.c_float),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Config(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("agent_count", ctypes.c_uint64),
        ("seed", ctypes.c_uint64),
        ("storage", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("resort_interval", ctypes.c_uint32),
        ("graph_chunk", ctypes.c_uint32),
        ("sort_cell_log2", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 6),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("agent_count", ctypes.c_uint64),
        ("storage", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("steps_done", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("resort_interval", ctypes.c_uint32),
        ("graph_chunk", ctypes.c_uint32),
    ]


class Timing(ctypes.Structure):
    _fields_ = [
        ("agents_ms", ctypes.c_double),
        ("trail_ms", ctypes.c_double),
        ("sort_ms", ctypes.c_double),
        ("agents_launches", ctypes.c_uint64),
        ("trail_launches", ctypes.c_uint64),
        ("sort_launches", ctypes.c_uint64),
    ]


class InstanceDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("agent_count", ctypes.c_uint64),
        ("seed", ctypes.c_uint64),
        ("params", Params),
    ]


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("storage", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("resort_interval", ctypes.c_uint32),
        ("graph_chunk", ctypes.c_uint32),
        ("sort_cell_log2", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 6),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint32),
        ("storage", ctypes.c_int32),
        ("total_agents", ctypes.c_uint64),
        ("total_cells", ctypes.c_uint64),
        ("steps_done", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
    ]


class Tuning(ctypes.Structure):
    """slm_tuning: launch layout of the step (never changes results)."""
    _fields_ = [
        ("resort_interval", ctypes.c_uint32),
        ("sort_cell_log2", ctypes.c_uint32),
        ("agents_per_group", ctypes.c_uint32),
        ("trail_rows", ctypes.c_uint32),
        ("tiled_sensing", ctypes.c_uint32),
        ("trials", ctypes.c_uint32),
        ("instance_groups", ctypes.c_uint32),
        ("streaming", ctypes.c_uint32),
        ("sort_key", ctypes.c_uint32),
        ("step_ms", ctypes.c_double),
        ("agents_ms", ctypes.c_double),
        ("trail_ms", ctypes.c_double),
        ("sort_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)

# name -> (argtypes) ; every function returns int
SIGNATURES = {
    "slm_device_count": [ctypes.POINTER(ctypes.c_int)],
    "slm_config_default": [ctypes.POINTER(Config)],
    "slm_params_default": [ctypes.POINTER(Params)],
    "slm_create": [ctypes.POINTER(Config), ctypes.POINTER(Params), ctypes.POINTER(_P)],
    "slm_destroy": [_P],
    "slm_set_params": [_P, ctypes.POINTER(Params)],
    "slm_get_params": [_P, ctypes.POINTER(Params)],
    "slm_step": [_P, ctypes.c_uint32],
    "slm_step_timed": [_P, ctypes.c_uint32, ctypes.POINTER(Timing)],
    "slm_synchronize": [_P],
    "slm_clear": [_P],
    "slm_resize": [_P, ctypes.c_uint32, ctypes.c_uint32],
    "slm_set_agent_count": [_P, ctypes.c_uint64],
    "slm_reassign_speeds": [_P, ctypes.c_float, ctypes.c_float],
    "slm_read_agents": [_P, _FP, ctypes.c_uint64],
    "slm_write_agents": [_P, _FP, ctypes.c_uint64],
    "slm_read_trail": [_P, _FP, ctypes.c_uint64],
    "slm_write_trail": [_P, _FP, ctypes.c_uint64],
    "slm_total_mass": [_P, ctypes.POINTER(ctypes.c_double)],
    "slm_get_info": [_P, ctypes.POINTER(Info)],
    "slm_batch_config_default": [ctypes.POINTER(BatchConfig)],
    "slm_batch_create": [ctypes.POINTER(BatchConfig), ctypes.POINTER(InstanceDesc), ctypes.c_uint32,
                         ctypes.POINTER(_P)],
    "slm_batch_destroy": [_P],
    "slm_batch_step": [_P, ctypes.c_uint32],
    "slm_batch_step_timed": [_P, ctypes.c_uint32, ctypes.POINTER(Timing)],
    "slm_batch_synchronize": [_P],
    "slm_batch_set_params": [_P, ctypes.c_uint32, ctypes.POINTER(Params)],
    "slm_batch_clear": [_P, ctypes.c_uint32],
    "slm_batch_read_agents": [_P, ctypes.c_uint32, _FP, ctypes.c_uint64],
    "slm_batch_write_agents": [_P, ctypes.c_uint32, _FP, ctypes.c_uint64],
    "slm_batch_read_trail": [_P, ctypes.c_uint32, _FP, ctypes.c_uint64],
    "slm_batch_write_trail": [_P, ctypes.c_uint32, _FP, ctypes.c_uint64],
    "slm_batch_total_mass": [_P, ctypes.POINTER(ctypes.c_double), ctypes.c_uint32],
    "slm_batch_get_info": [_P, ctypes.POINTER(BatchInfo)],
    "slm_render_rgba8": [_P, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p],
    "slm_render_rgba8_device": [_P, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_void_p],
    "slm_batch_render_rgba8": [_P, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                               ctypes.c_uint32, ctypes.c_void_p],
    "slm_render_rgba8_noalloc": [_P, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_void_p],
    "slm_step_render_rgba8": [_P, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                              ctypes.c_uint32, ctypes.c_void_p],
    "slm_step_render_rgba8_device": [_P, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_void_p],
    "slm_snapshot_size": [_P, ctypes.POINTER(ctypes.c_uint64)],
    "slm_save_snapshot": [_P, ctypes.c_void_p, ctypes.c_uint64],
    "slm_load_snapshot": [_P, ctypes.c_void_p, ctypes.c_uint64],
    "slm_create_from_snapshot": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32,
                                 ctypes.POINTER(_P)],
    "slm_get_tuning": [_P, ctypes.POINTER(Tuning)],
    "slm_set_tuning": [_P, ctypes.POINTER(Tuning)],
    "slm_autotune": [_P, ctypes.c_uint32, ctypes.POINTER(Tuning)],
    "slm_batch_get_tuning": [_P, ctypes.POINTER(Tuning)],
    "slm_batch_set_tuning": [_P, ctypes.POINTER(Tuning)],
    "slm_batch_autotune": [_P, ctypes.c_uint32, ctypes.POINTER(Tuning)],
}

_lib = None


def lib():
    """Load (once) and return the native library."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        L.slm_last_error.argtypes = []
        L.slm_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SLM_OK:
        raise SlimeMoldError(rc, lib().slm_last_error().decode(errors="replace"))


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = lib().slm_device_count(ctypes.byref(n))
    if rc == SLM_ERR_NO_DEVICE:
        return 0
    check(rc)
    return n.value


def default_params() -> Params:
    p = Params()
    check(lib().slm_params_default(ctypes.byref(p)))
    return p


def default_batch_config() -> BatchConfig:
    c = BatchConfig()
    check(lib().slm_batch_config_default(ctypes.byref(c)))
    return c


def default_config() -> Config:
    c = Config()
    check(lib().slm_config_default(ctypes.byref(c)))
    return c
