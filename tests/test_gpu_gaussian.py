"""The Gaussian diffuse extension (SLM_FLAG_GAUSSIAN_DIFFUSE) on the GPU against
the oracle, bit for bit: grids smaller than the kernel and than one tile,
ragged tile edges, radii 1..8, every storage format, re-sorting, graph chunks,
batches with a radius per instance and instance groups, a radius changed
under captured graphs; and the argument checks."""
import dataclasses

import numpy as np
import pytest

from parity_util import assert_bitwise, preset

pytestmark = pytest.mark.gpu


def storage_kw(storage):
    sr = storage == "bf16sr"
    return dict(storage="bf16" if sr else storage, stochastic_rounding=sr)


CASES = [
    # w, h, agents, radius, sigma, rate, storage, resort, chunk, steps
    (1, 1, 3, 2.0, 1.0, 1.0, "f32", 0, 0, 3),
    (3, 5, 20, 4.0, 1.5, 0.8, "f32", 0, 0, 4),
    (257, 17, 2000, 2.0, 1.0, 1.0, "f32", 5, 4, 9),
    (300, 70, 5000, 3.0, 2.0, 0.35, "bf16", 0, 0, 6),
    (600, 40, 9000, 8.0, 3.0, 1.0, "f16", 3, 0, 5),
    (512, 512, 40000, 2.0, 1.0, 1.0, "bf16sr", 10, 16, 20),
    (1000, 33, 6000, 1.0, 0.6, 0.9, "f32", 0, 5, 7),
    (64, 300, 4000, 5.0, 2.2, 1.0, "bf16", 2, 3, 8),
    # whole 256-column tiles (16-byte chunk loads and stores), short and tall maps
    (1024, 5, 3000, 2.0, 1.0, 1.0, "f32", 0, 4, 5),
    (2048, 40, 30000, 3.0, 1.3, 0.7, "f16", 5, 0, 6),
    (4096, 20, 20000, 8.0, 3.5, 1.0, "bf16sr", 0, 0, 4),
    (768, 700, 50000, 2.0, 1.0, 1.0, "bf16", 10, 8, 12),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_gaussian_single_matches_oracle(slm, orc, case):
    w, h, n, radius, sigma, rate, storage, resort, chunk, steps = CASES[case]
    s = dataclasses.replace(preset(slm, "Default"), blur_radius=radius, blur_sigma=sigma,
                            pheromone_diffusion_rate=rate)
    print(f"[gauss {case}] {w}x{h} n={n} r={radius} sigma={sigma} rate={rate} {storage} "
          f"resort={resort} chunk={chunk} steps={steps}", flush=True)
    rng = np.random.default_rng(case)
    with slm.Simulation(w, h, n, seed=case + 11, params=s, resort_interval=resort,
                        graph_chunk=chunk, gaussian=True, **storage_kw(storage)) as sim:
        sim.write_trail(rng.uniform(0.0, 1.0, (h, w)).astype(np.float32))
        a, t = sim.read_agents(), sim.read_trail()
        sim.step(steps)
        ga, gt = sim.read_agents(), sim.read_trail()
    orc.step(a, t, s, steps, storage=storage, gaussian=True)
    assert_bitwise(ga, a, f"case {case} agents")
    assert_bitwise(gt, t, f"case {case} trail")


@pytest.mark.parametrize("deposit,storage,w,h", [
    (0.3, "bf16sr", 512, 64), (0.3, "f32", 300, 50), (2.5, "bf16sr", 512, 64),
    (1.0, "f16", 256, 40), (0.75, "bf16", 260, 33)])
def test_gaussian_deposit_amounts(slm, orc, deposit, storage, w, h):
    """Deposits below 1 take the general deposit + clamp staging; 1 and above
    the one-select form (a marked cell becomes exactly 1 before the decay);
    whole-tile and ragged maps, in every storage."""
    s = dataclasses.replace(preset(slm, "Default"), pheromone_deposition_amount=deposit,
                            blur_radius=2.0, blur_sigma=1.0, pheromone_diffusion_rate=0.85)
    print(f"[gauss deposit] {deposit} {storage} {w}x{h}", flush=True)
    rng = np.random.default_rng(int(deposit * 10) + w)
    with slm.Simulation(w, h, w * h // 2, seed=9, params=s, graph_chunk=3, gaussian=True,
                        **storage_kw(storage)) as sim:
        sim.write_trail(rng.uniform(0.0, 1.0, (h, w)).astype(np.float32))
        a, t = sim.read_agents(), sim.read_trail()
        sim.step(7)
        ga, gt = sim.read_agents(), sim.read_trail()
    orc.step(a, t, s, 7, storage=storage, gaussian=True)
    assert_bitwise(ga, a, "agents")
    assert_bitwise(gt, t, "trail")


@pytest.mark.parametrize("groups", [1, 3])
def test_gaussian_batch_radius_per_instance(slm, orc, groups):
    rng = np.random.default_rng(40 + groups)
    specs = []
    for k in range(7):
        w, h = int(rng.integers(1, 700)), int(rng.integers(1, 300))
        s = dataclasses.replace(preset(slm, ["Default", "Sponge", "Threads"][k % 3]),
                                blur_radius=float(1 + k % 8), blur_sigma=float(rng.uniform(0.4, 3.0)),
                                pheromone_diffusion_rate=float(rng.uniform(0.2, 1.0)))
        specs.append(slm.InstanceSpec(w, h, int(rng.integers(0, 3 * w * h // 2 + 2)),
                                      seed=100 + k, params=s))
    print(f"[gauss batch] groups={groups} {[(sp.width, sp.height, sp.agent_count) for sp in specs]}",
          flush=True)
    with slm.Batch(specs, graph_chunk=4, resort_interval=3, gaussian=True, storage="bf16") as b:
        b.set_tuning(instance_groups=groups)
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
        b.step(9)
        got = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
    for k, sp in enumerate(specs):
        a, t = starts[k]
        orc.step(a, t, sp.params, 9, storage="bf16", gaussian=True)
        assert_bitwise(got[k][0], a, f"instance {k} agents")
        assert_bitwise(got[k][1], t, f"instance {k} trail")


@pytest.mark.parametrize("storage", ["f32", "bf16sr"])
def test_gaussian_radius_change_under_captured_graphs(slm, orc, storage):
    """Radius changes between captured chunks, within a radius class (the
    chunks replay) and across classes 2 / 4 / 8 (the pass's LDS tile changes
    and the chunks are recaptured; in 16-bit storage the tile height and the
    tile plan change too), back and forth."""
    s2 = dataclasses.replace(preset(slm, "Default"), blur_radius=2.0, blur_sigma=1.0)
    seq = [(s2, 10),
           (dataclasses.replace(s2, blur_radius=1.0, blur_sigma=0.7), 5),   # same class
           (dataclasses.replace(s2, blur_radius=8.0, blur_sigma=4.0), 10),  # class 8
           (dataclasses.replace(s2, blur_radius=3.0, blur_sigma=1.5), 5),   # class 4
           (s2, 10)]                                                          # class 2
    print(f"[gauss radius change] 400x200 n=20000 chunk=5 radii 2,1,8,3,2 {storage}", flush=True)
    sr = storage == "bf16sr"
    with slm.Simulation(400, 200, 20000, seed=5, params=s2, graph_chunk=5, gaussian=True,
                        storage="bf16" if sr else storage, stochastic_rounding=sr) as sim:
        a, t = sim.read_agents(), sim.read_trail()
        for s, n in seq:
            sim.set_params(s)
            sim.step(n)
        ga, gt = sim.read_agents(), sim.read_trail()
    for s, n in seq:
        orc.step(a, t, s, n, storage=storage, gaussian=True)
    assert_bitwise(ga, a, "agents")
    assert_bitwise(gt, t, "trail")


@pytest.mark.parametrize("groups", [1, 2])
def test_gaussian_batch_radius_class_change(slm, orc, groups):
    """One instance of a batch moves its radius to another class while the
    others keep theirs: the pass's tile follows the largest radius (also with
    the instances in two groups on their own streams)."""
    base = dataclasses.replace(preset(slm, "Default"), blur_radius=2.0, blur_sigma=1.0)
    specs = [slm.InstanceSpec(w, h, n, seed=11 + k, params=base)
             for k, (w, h, n) in enumerate([(256, 64, 3000), (130, 70, 2000), (300, 40, 4000)])]
    print(f"[gauss batch radius class] 3 instances, chunk 4, instance 1 radius 2 -> 6 -> 2, "
          f"{groups} groups", flush=True)
    with slm.Batch(specs, graph_chunk=4, gaussian=True, storage="bf16",
                   stochastic_rounding=True) as b:
        b.set_tuning(instance_groups=groups)
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
        plan = [(None, 8), (dataclasses.replace(base, blur_radius=6.0, blur_sigma=2.5), 8),
                (base, 8)]
        for s, n in plan:
            if s is not None:
                b.set_params(1, s)
            b.step(n)
        got = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
    for k in range(len(specs)):
        a, t = starts[k]
        for s, n in plan:
            orc.step(a, t, s if (s is not None and k == 1) else base, n, storage="bf16sr",
                     gaussian=True)
        assert_bitwise(got[k][0], a, f"instance {k} agents")
        assert_bitwise(got[k][1], t, f"instance {k} trail")


def test_gaussian_params_are_checked(slm):
    base = preset(slm, "Default")
    for radius, sigma in [(0.0, 1.0), (0.4, 1.0), (8.6, 1.0), (2.0, 0.0), (2.0, -3.0)]:
        bad = dataclasses.replace(base, blur_radius=radius, blur_sigma=sigma)
        with pytest.raises(slm.SlimeMoldError):
            slm.Simulation(64, 64, 10, params=bad, gaussian=True)
        with slm.Simulation(64, 64, 10, params=bad) as sim:  # the box diffuse ignores them
            sim.step(1)
        with slm.Simulation(64, 64, 10, params=base, gaussian=True) as sim:
            with pytest.raises(slm.SlimeMoldError):
                sim.set_params(bad)
            sim.step(1)  # still usable
        with pytest.raises(slm.SlimeMoldError):
            slm.Batch([slm.InstanceSpec(32, 32, 5, params=base),
                       slm.InstanceSpec(32, 32, 5, params=bad)], gaussian=True)
    # a snapshot whose parameters do not describe a Gaussian cannot be loaded into one
    bad = dataclasses.replace(base, blur_radius=0.0)
    with slm.Simulation(64, 64, 10, params=bad) as src:
        snap = src.snapshot()
    with slm.Simulation(64, 64, 10, params=base, gaussian=True) as sim:
        with pytest.raises(slm.SlimeMoldError):
            sim.load_snapshot(snap)
