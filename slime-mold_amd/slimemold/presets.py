"""The reference's presets (src/presets.rs:6-155), values unchanged.

Preset / PresetManager mirror src/presets.rs:6-43; init_preset_manager mirrors
src/presets.rs:45-155 (8 presets, in the same order).  Fields a preset does
not set come from Settings() exactly as `..Settings::default()` does.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

from .settings import Settings


@dataclasses.dataclass
class Preset:
    name: str
    settings: Settings


class PresetManager:
    def __init__(self) -> None:
        self.presets: List[Preset] = []

    def add_preset(self, preset: Preset) -> None:
        self.presets.append(preset)

    def get_preset(self, name: str) -> Optional[Preset]:
        for p in self.presets:
            if p.name == name:
                return p
        return None

    def get_preset_names(self) -> List[str]:
        return [p.name for p in self.presets]


def _s(**kw) -> Settings:
    return dataclasses.replace(Settings(), **kw)


def init_preset_manager() -> PresetManager:
    m = PresetManager()
    m.add_preset(Preset("Default", Settings()))
    m.add_preset(Preset("Sponge", _s(
        agent_jitter=0.0, agent_speed_min=20.0, agent_speed_max=30.0, agent_turn_speed=0.43,
        agent_sensor_angle=0.3, agent_sensor_distance=20.0, pheromone_deposition_amount=1.0,
        pheromone_decay_factor=1.0, pheromone_diffusion_rate=1.0)))
    m.add_preset(Preset("Firecracker Trees", _s(
        agent_jitter=0.1, agent_speed_min=60.0, agent_speed_max=60.0, agent_turn_speed=1.47,
        agent_sensor_angle=0.3, agent_sensor_distance=20.0, pheromone_deposition_amount=1.0,
        pheromone_decay_factor=10.0, pheromone_diffusion_rate=1.0)))
    m.add_preset(Preset("Threads", _s(
        agent_jitter=0.0, agent_speed_min=70.0, agent_speed_max=80.0, agent_turn_speed=0.02,
        agent_sensor_angle=0.3, agent_sensor_distance=20.0, pheromone_deposition_amount=1.0,
        pheromone_decay_factor=10.0, pheromone_diffusion_rate=0.1)))
    m.add_preset(Preset("Curls", _s(
        agent_count=3_000_000, agent_jitter=5.0, agent_speed_min=70.0, agent_speed_max=80.0,
        agent_turn_speed=0.05, agent_sensor_angle=0.3, agent_sensor_distance=20.0,
        pheromone_deposition_amount=1.0, pheromone_decay_factor=75.0,
        pheromone_diffusion_rate=0.1)))
    m.add_preset(Preset("Waves", _s(
        agent_jitter=1.0, agent_speed_min=30.0, agent_speed_max=50.0, agent_turn_speed=6.0,
        agent_sensor_angle=0.3, agent_sensor_distance=20.0, pheromone_deposition_amount=1.0,
        pheromone_decay_factor=10.0, pheromone_diffusion_rate=0.1)))
    m.add_preset(Preset("Snake", _s(
        agent_jitter=3.0, agent_speed_min=100.0, agent_speed_max=120.0, agent_turn_speed=0.37,
        agent_sensor_angle=1.34, agent_sensor_distance=225.0, pheromone_deposition_amount=1.0,
        pheromone_decay_factor=10.0, pheromone_diffusion_rate=1.0)))
    m.add_preset(Preset("Mesh", _s(
        agent_jitter=3.0, agent_speed_min=100.0, agent_speed_max=120.0, agent_turn_speed=6.0,
        agent_sensor_angle=1.57, agent_sensor_distance=225.0, pheromone_deposition_amount=1.0,
        pheromone_decay_factor=10.0, pheromone_diffusion_rate=1.0)))
    return m
