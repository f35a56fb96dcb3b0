"""The benchmark and parity workloads, as synthetic seeded inputs.

plain_small      1024x1024 f32, 2^18 agents, Default preset, seed 42
plain_at_scale   4096x4096 f32, 2^24 agents, Default -> Sponge -> Firecracker Trees, seed 1
precision        8192x8192 bf16 / f16 trail, 2^26 agents, f64 positions, seed 7
batch_mix        256 instances, grids from {256..2048}^2, 2^14..2^22 agents (log-uniform),
                 sensor angle / distance, turn speed, decay, diffusion drawn between the
                 minimum and maximum over the reference presets (src/presets.rs:45-155)
combined         1024 batch_mix instances plus one 16384x16384 instance with 2^28 agents

Agent initial state follows src/main.rs:273-280 with the library's seeded
SplitMix64 draws; the trail starts at zero.
"""
from __future__ import annotations

import dataclasses
from typing import List

import numpy as np

from .batch import InstanceSpec
from .presets import init_preset_manager
from .settings import Settings

PLAIN_PRESETS = ["Default", "Sponge", "Firecracker Trees"]


def preset(name: str) -> Settings:
    return init_preset_manager().get_preset(name).settings


def preset_ranges():
    """(min, max) over the reference presets of each swept parameter."""
    fields = ["agent_sensor_angle", "agent_sensor_distance", "agent_turn_speed",
              "pheromone_decay_factor", "pheromone_diffusion_rate", "pheromone_deposition_amount"]
    vals = {f: [getattr(p.settings, f) for p in init_preset_manager().presets] for f in fields}
    return {f: (min(v), max(v)) for f, v in vals.items()}


def batch_mix(count: int = 256, seed: int = 3) -> List[InstanceSpec]:
    """Uneven independent instances: grid side uniform in {256, 512, 1024, 2048},
    agents round(2^(14 + 8u)), swept parameters uniform in the preset range,
    jitter and speeds from the Default preset."""
    rng = np.random.default_rng(seed)
    r = preset_ranges()
    base = Settings()
    out = []
    for k in range(count):
        side = int(rng.choice([256, 512, 1024, 2048]))
        n = int(round(2.0 ** (14 + 8 * rng.uniform())))
        s = dataclasses.replace(
            base,
            agent_sensor_angle=float(rng.uniform(*r["agent_sensor_angle"])),
            agent_sensor_distance=float(rng.uniform(*r["agent_sensor_distance"])),
            agent_turn_speed=float(rng.uniform(*r["agent_turn_speed"])),
            pheromone_decay_factor=float(rng.uniform(*r["pheromone_decay_factor"])),
            pheromone_diffusion_rate=float(rng.uniform(*r["pheromone_diffusion_rate"])),
            pheromone_deposition_amount=float(rng.uniform(*r["pheromone_deposition_amount"])),
        )
        out.append(InstanceSpec(side, side, n, seed=seed * 1_000_003 + k, params=s))
    return out


def combined(count: int = 1024, seed: int = 4, big_side: int = 16384,
             big_agents: int = 1 << 28) -> List[InstanceSpec]:
    specs = batch_mix(count, seed)
    specs.append(InstanceSpec(big_side, big_side, big_agents, seed=seed, params=preset("Default")))
    return specs


def size_class(spec: InstanceSpec) -> str:
    return f"{spec.width}x{spec.height}"
