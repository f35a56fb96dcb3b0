"""Seeded random parity sweep: the HIP path against the oracle, bit for bit.

Each case draws a grid shape (1 cell to 1100 per side, ragged widths
included), an agent count (0 to 2 per cell), a preset with every swept
parameter perturbed (diffusion rates outside [0, 1], zero or large sensor
distances, jitter on or off), a storage format, optionally a random starting
trail, and a launch tuning (re-sort interval, bin size and key, tiled or windowed
sensing, agents per workgroup, trail tile rows, streaming loads and stores,
graph chunk length;
for batches the instance groups).  A tuning never changes a result bit, so
every case must equal the oracle exactly.  The case is printed before it runs.
"""
import dataclasses

import numpy as np
import pytest

from parity_util import PRESET_NAMES, assert_bitwise, preset

pytestmark = pytest.mark.gpu

STORAGES = ["f32", "bf16", "f16", "bf16sr"]


def side(rng):
    return int(np.clip(np.round(np.exp(rng.uniform(0.0, np.log(1100.0)))), 1, 1100))


def random_settings(slm, rng):
    s = preset(slm, PRESET_NAMES[rng.integers(len(PRESET_NAMES))])
    smin = float(rng.uniform(0.0, 60.0))
    return dataclasses.replace(
        s,
        agent_sensor_angle=float(rng.uniform(0.0, 1.6)),
        agent_sensor_distance=float(rng.choice([0.0, rng.uniform(0.5, 60.0), 230.0])),
        agent_turn_speed=float(rng.uniform(0.0, 6.5)),
        agent_jitter=float(rng.choice([0.0, rng.uniform(0.0, 0.5)])),
        agent_speed_min=smin,
        agent_speed_max=smin + float(rng.uniform(0.0, 60.0)),
        pheromone_decay_factor=float(rng.uniform(0.0, 100.0)),
        pheromone_diffusion_rate=float(rng.choice([0.0, 1.0, rng.uniform(-0.2, 1.2)])),
        pheromone_deposition_amount=float(rng.choice([1.0, rng.uniform(0.2, 3.0)])),
    )


def random_tuning(rng, batch, extra=None):
    """A random launch tuning.  `extra` (its own generator, so the draws of
    `rng` and hence the shapes and parameters of every case stay put) adds
    the newer options: refined re-sort keys and windowed sensing."""
    resort = int(rng.choice([0, 1, 3, 7]))
    t = dict(resort_interval=resort,
             sort_cell_log2=int(rng.integers(2, 7)),
             agents_per_group=int(rng.choice([256, 512, 1024, 2048, 4096])),
             trail_rows=int(rng.choice([0, 4, 8, 16, 32, 64])),
             streaming=int(rng.integers(0, 4)),
             tiled_sensing=int(resort > 0 and rng.random() < 0.3))
    if batch and not t["tiled_sensing"]:
        t["instance_groups"] = int(rng.choice([1, 2, 3, 8]))
    if extra is not None:
        if not t["tiled_sensing"] and extra.random() < 0.2:
            t["tiled_sensing"] = 2  # windowed sensing
        if resort > 0 and t["tiled_sensing"] != 1 and extra.random() < 0.5:
            t["sort_key"] = int(extra.choice([3, 5, 0x10, 0x13, 0x15]))
    return t


def extra_rng(kind, case):
    return np.random.default_rng([kind, case, 7])


def storage_kw(storage):
    sr = storage == "bf16sr"
    return dict(storage="bf16" if sr else storage, stochastic_rounding=sr)


@pytest.mark.parametrize("case", range(32))
def test_random_single_simulation(slm, orc, case):
    rng = np.random.default_rng(1000 + case)
    w, h = side(rng), side(rng)
    n = int(rng.integers(0, min(2 * w * h, 30000) + 1))
    s = random_settings(slm, rng)
    storage = STORAGES[rng.integers(len(STORAGES))]
    tun = random_tuning(rng, batch=False, extra=extra_rng(1, case))
    chunk = int(rng.choice([0, 1, 5, 16]))
    steps = int(rng.integers(1, 13))
    seed = int(rng.integers(0, 2**31))
    trail0 = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32) if rng.random() < 0.3 else None
    print(f"[random {case}] {w}x{h} n={n} {storage} steps={steps} chunk={chunk} {tun} "
          f"params={s}", flush=True)
    with slm.Simulation(w, h, n, seed=seed, params=s, graph_chunk=chunk,
                        **storage_kw(storage)) as sim:
        sim.set_tuning(**tun)
        if trail0 is not None:
            sim.write_trail(trail0)
        a, t = sim.read_agents(), sim.read_trail()
        sim.step(steps)
        ga, gt = sim.read_agents(), sim.read_trail()
        mass = sim.total_mass()
    orc.step(a, t, s, steps, storage=storage)
    assert_bitwise(ga, a, f"case {case} agents")
    assert_bitwise(gt, t, f"case {case} trail")
    assert mass == pytest.approx(t.astype(np.float64).sum(), rel=1e-12, abs=1e-9)


@pytest.mark.parametrize("case", range(8))
def test_random_batch(slm, orc, case):
    rng = np.random.default_rng(2000 + case)
    count = int(rng.integers(2, 9))
    specs = []
    for k in range(count):
        w, h = side(rng), side(rng)
        n = int(rng.integers(0, min(2 * w * h, 12000) + 1))
        specs.append(slm.InstanceSpec(w, h, n, seed=int(rng.integers(0, 2**31)),
                                      params=random_settings(slm, rng)))
    storage = STORAGES[rng.integers(len(STORAGES))]
    tun = random_tuning(rng, batch=True, extra=extra_rng(2, case))
    chunk = int(rng.choice([0, 3, 10]))
    steps = int(rng.integers(1, 11))
    print(f"[random batch {case}] {[(sp.width, sp.height, sp.agent_count) for sp in specs]} "
          f"{storage} steps={steps} chunk={chunk} {tun}", flush=True)
    with slm.Batch(specs, graph_chunk=chunk, **storage_kw(storage)) as b:
        b.set_tuning(**tun)
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(count)]
        b.step(steps)
        got = [(b.read_agents(k), b.read_trail(k)) for k in range(count)]
    for k, sp in enumerate(specs):
        a, t = starts[k]
        orc.step(a, t, sp.params, steps, storage=storage)
        assert_bitwise(got[k][0], a, f"batch {case} instance {k} agents")
        assert_bitwise(got[k][1], t, f"batch {case} instance {k} trail")


def gaussian_settings(slm, rng):
    s = random_settings(slm, rng)
    return dataclasses.replace(s, blur_radius=float(rng.integers(1, 9)) + float(rng.uniform(-0.45, 0.45)),
                               blur_sigma=float(rng.uniform(0.3, 5.0)))


@pytest.mark.parametrize("case", range(12))
def test_random_gaussian_single(slm, orc, case):
    """The Gaussian diffuse extension under random shapes, parameters, storage
    and tunings: equal to the oracle's restatement bit for bit."""
    rng = np.random.default_rng(3000 + case)
    w, h = side(rng), side(rng)
    if rng.random() < 0.3:  # whole 16-byte-chunk tiles
        w = int(rng.choice([256, 512, 768, 1024]))
    n = int(rng.integers(0, min(2 * w * h, 30000) + 1))
    s = gaussian_settings(slm, rng)
    storage = STORAGES[rng.integers(len(STORAGES))]
    tun = random_tuning(rng, batch=False, extra=extra_rng(3, case))
    chunk = int(rng.choice([0, 1, 5, 16]))
    steps = int(rng.integers(1, 9))
    seed = int(rng.integers(0, 2**31))
    trail0 = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32) if rng.random() < 0.5 else None
    print(f"[random gaussian {case}] {w}x{h} n={n} {storage} steps={steps} chunk={chunk} {tun} "
          f"r={s.blur_radius:.3f} sigma={s.blur_sigma:.3f}", flush=True)
    with slm.Simulation(w, h, n, seed=seed, params=s, graph_chunk=chunk, gaussian=True,
                        **storage_kw(storage)) as sim:
        sim.set_tuning(**tun)
        if trail0 is not None:
            sim.write_trail(trail0)
        a, t = sim.read_agents(), sim.read_trail()
        sim.step(steps)
        ga, gt = sim.read_agents(), sim.read_trail()
    orc.step(a, t, s, steps, storage=storage, gaussian=True)
    assert_bitwise(ga, a, f"case {case} agents")
    assert_bitwise(gt, t, f"case {case} trail")


@pytest.mark.parametrize("case", range(4))
def test_random_gaussian_batch(slm, orc, case):
    rng = np.random.default_rng(4000 + case)
    count = int(rng.integers(2, 9))
    specs = []
    for k in range(count):
        w, h = side(rng), side(rng)
        n = int(rng.integers(0, min(2 * w * h, 12000) + 1))
        specs.append(slm.InstanceSpec(w, h, n, seed=int(rng.integers(0, 2**31)),
                                      params=gaussian_settings(slm, rng)))
    storage = STORAGES[rng.integers(len(STORAGES))]
    tun = random_tuning(rng, batch=True, extra=extra_rng(4, case))
    chunk = int(rng.choice([0, 3, 10]))
    steps = int(rng.integers(1, 9))
    print(f"[random gaussian batch {case}] {[(sp.width, sp.height, sp.agent_count) for sp in specs]} "
          f"{storage} steps={steps} chunk={chunk} {tun}", flush=True)
    with slm.Batch(specs, graph_chunk=chunk, gaussian=True, **storage_kw(storage)) as b:
        b.set_tuning(**tun)
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(count)]
        b.step(steps)
        got = [(b.read_agents(k), b.read_trail(k)) for k in range(count)]
    for k, sp in enumerate(specs):
        a, t = starts[k]
        orc.step(a, t, sp.params, steps, storage=storage, gaussian=True)
        assert_bitwise(got[k][0], a, f"batch {case} instance {k} agents")
        assert_bitwise(got[k][1], t, f"batch {case} instance {k} trail")
