"""Simulation settings, kept field for field from the reference.

Constants: src/settings.rs:3-27.  Settings struct and its Default:
src/settings.rs:29-70.  `to_params` builds the parameter block the way
SimSizeUniform::new does (src/main.rs:48-67): the raw decay factor is passed
through and scaled by 0.001 inside the step.
"""
from __future__ import annotations

import dataclasses

# General settings (src/settings.rs:4-6)
DEFAULT_WIDTH = 1600
DEFAULT_HEIGHT = 900
DEFAULT_IS_FULLSCREEN = False

# Agent settings (src/settings.rs:9-17)
AGENT_COUNT = 10_000_000
AGENT_SPEED_MIN = 30.0
AGENT_SPEED_MAX = 50.0
AGENT_TURN_SPEED = 0.43  # ~25 degrees
AGENT_POSSIBLE_STARTING_HEADINGS = (0.0, 360.0)  # Range 0.0..360.0, unused by the step
DEPOSITION_AMOUNT = 1.0
AGENT_JITTER = 0.0
AGENT_SENSOR_ANGLE = 0.3
AGENT_SENSOR_DISTANCE = 20.0

# Pheromone settings (src/settings.rs:19-27)
DECAY_FACTOR = 10.0
DIFFUSION_RATE = 1.0
BLUR_RADIUS = 2.0
BLUR_SIGMA = 1.0


@dataclasses.dataclass
class Settings:
    """src/settings.rs:29-47 (same field names; the Range is a (start, end) pair)."""

    agent_count: int = AGENT_COUNT
    agent_jitter: float = AGENT_JITTER
    agent_possible_starting_headings: tuple = AGENT_POSSIBLE_STARTING_HEADINGS
    agent_speed_max: float = AGENT_SPEED_MAX
    agent_speed_min: float = AGENT_SPEED_MIN
    agent_turn_speed: float = AGENT_TURN_SPEED
    pheromone_decay_factor: float = DECAY_FACTOR
    pheromone_diffusion_rate: float = DIFFUSION_RATE
    pheromone_deposition_amount: float = DEPOSITION_AMOUNT
    window_fullscreen: bool = DEFAULT_IS_FULLSCREEN
    window_height: int = DEFAULT_HEIGHT
    window_width: int = DEFAULT_WIDTH
    agent_sensor_angle: float = AGENT_SENSOR_ANGLE
    agent_sensor_distance: float = AGENT_SENSOR_DISTANCE
    blur_radius: float = BLUR_RADIUS
    blur_sigma: float = BLUR_SIGMA

    def clone(self) -> "Settings":
        return dataclasses.replace(self)

    def to_params(self):
        """The step parameter block (src/main.rs:48-67 without width/height)."""
        from ._native import Params

        return Params(
            decay_factor=self.pheromone_decay_factor,
            agent_jitter=self.agent_jitter,
            agent_speed_min=self.agent_speed_min,
            agent_speed_max=self.agent_speed_max,
            agent_turn_speed=self.agent_turn_speed,
            agent_sensor_angle=self.agent_sensor_angle,
            agent_sensor_distance=self.agent_sensor_distance,
            diffusion_rate=self.pheromone_diffusion_rate,
            pheromone_deposition_amount=self.pheromone_deposition_amount,
            blur_radius=self.blur_radius,
            blur_sigma=self.blur_sigma,
        )
