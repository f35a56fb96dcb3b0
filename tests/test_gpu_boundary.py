"""Round trips and invariants at the C ABI that the parity sweeps only imply:
the parameter block field by field (a1, src/main.rs:29-67), agents AoS -> SoA
-> AoS in the reference's record layout and order (a2, compute.wgsl:2,31-32),
and the speed clamp with its re-clamp after a preset switch (a6,
compute.wgsl:72)."""
import ctypes

import numpy as np
import pytest

from parity_util import preset

pytestmark = pytest.mark.gpu


def test_params_round_trip_field_by_field(slm):
    import slimemold._native as N
    fields = [name for name, _ in N.Params._fields_]
    assert len(fields) == 11
    p = N.Params()
    for i, name in enumerate(fields):  # distinct, exactly representable values
        setattr(p, name, 0.5 + 0.25 * i)
    with slm.Simulation(64, 48, 100, params=preset(slm, "Default")) as sim:
        sim.set_params(p)
        q = sim.get_params()
        sim.step(2)
        assert sim.get_params().agent_turn_speed == q.agent_turn_speed
    for name in fields:
        assert getattr(q, name) == getattr(p, name), name


@pytest.mark.parametrize("kw", [{}, {"resort_interval": 2}, {"f64_positions": True},
                                {"resort_interval": 1, "f64_positions": True}])
def test_agents_round_trip_in_reference_layout(slm, kw):
    w, h, n = 300, 200, 5000
    rng = np.random.default_rng(4)
    a = np.empty((n, 4), np.float32)
    a[:, 0] = rng.uniform(0, w, n)
    a[:, 1] = rng.uniform(0, h, n)
    a[:, 2] = rng.uniform(-10.0, 10.0, n)  # any finite heading
    a[:, 3] = rng.uniform(0.0, 80.0, n)
    a[:3, :2] = [[0.0, 0.0], [w, h], [w, 0.0]]  # the closed upper edge is accepted
    with slm.Simulation(w, h, n, params=preset(slm, "Default"), **kw) as sim:
        sim.write_agents(a)
        assert np.array_equal(sim.read_agents(), a)
        sim.step(3)  # re-sorts move agents internally
        b = sim.read_agents()
        sim.write_agents(b)
        assert np.array_equal(sim.read_agents(), b)
    bad = a.copy()
    bad[7, 0] = w + 0.5
    with slm.Simulation(w, h, n, params=preset(slm, "Default"), **kw) as sim:
        with pytest.raises(slm.SlimeMoldError):
            sim.write_agents(bad)


def test_speed_clamped_after_one_step_and_after_a_preset_switch(slm):
    import dataclasses
    s = preset(slm, "Default")
    with slm.Simulation(256, 256, 20000, seed=3, params=s, resort_interval=2) as sim:
        a = sim.read_agents()
        a[:, 3] = np.linspace(0.0, 100.0, a.shape[0], dtype=np.float32)  # outside [30, 50] too
        sim.write_agents(a)
        sim.step(1)
        v = sim.read_agents()[:, 3]
        assert v.min() == np.float32(s.agent_speed_min) and v.max() == np.float32(s.agent_speed_max)
        narrow = dataclasses.replace(s, agent_speed_min=38.0, agent_speed_max=41.0)
        sim.set_params(narrow)
        v = sim.read_agents()[:, 3]
        assert v.min() < 38.0  # parameters alone do not touch the agents
        sim.step(1)
        v = sim.read_agents()[:, 3]
        assert v.min() == np.float32(38.0) and v.max() == np.float32(41.0)
