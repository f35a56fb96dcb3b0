"""slimemold - MI355X-native slime-mold (Physarum) step behind a C ABI.

Host-side mirror of the reference's library interface (src/lib.rs:1-3 exports
settings and presets) plus the Simulation handle that replaces its wgpu
compute dispatch.  The compute path is libslimemold_hip.so (HIP, gfx950).
"""
from ._native import (NativeLibraryMissing, Params, SlimeMoldError, Tuning, default_batch_config,
                      default_config, default_params, device_count, lib)
from .batch import Batch, InstanceSpec
from .lut import LutData, LutManager
from .presets import Preset, PresetManager, init_preset_manager
from .settings import Settings
from .simulation import Simulation

__all__ = [
    "Batch", "InstanceSpec", "LutData", "LutManager", "default_batch_config", "NativeLibraryMissing", "Params", "Preset", "PresetManager", "Settings", "SlimeMoldError", "Tuning",
    "Simulation", "default_config", "default_params", "device_count", "init_preset_manager", "lib",
]
