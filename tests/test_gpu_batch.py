"""Batches and the f64-position mode on the GPU, bit for bit.

A batch instance must equal the same simulation run alone (and the oracle);
the f64-position mode must equal the oracle's f64-position restatement.
"""
import numpy as np
import pytest

from parity_util import PRESET_NAMES, assert_bitwise, preset, run_parity

pytestmark = pytest.mark.gpu


def mixed_specs(slm, seed=0):
    """Uneven instances: powers of two, ragged widths (one-cell-per-lane path),
    tiny grids, an empty instance and every preset."""
    rng = np.random.default_rng(seed)
    sizes = [(256, 256), (512, 128), (37, 23), (1030, 40), (64, 64), (1, 1), (260, 2),
             (128, 512), (96, 80), (1032, 50)]
    specs = []
    for k, (w, h) in enumerate(sizes):
        n = int(min(w * h * 2, rng.integers(0, 20000)))
        if k == 4:
            n = 0
        specs.append(slm.InstanceSpec(w, h, n, seed=100 + k,
                                      params=preset(slm, PRESET_NAMES[k % len(PRESET_NAMES)])))
    return specs


@pytest.mark.parametrize("storage", ["f32", "bf16", "bf16sr"])
def test_batch_instances_equal_single_runs_and_oracle(slm, orc, storage):
    specs = mixed_specs(slm)
    steps = 12
    print(f"[batch] {len(specs)} instances storage={storage} steps={steps}", flush=True)
    sr = storage == "bf16sr"  # bf16 storage, stochastic rounding in the trail pass
    kind = "bf16" if sr else storage
    with slm.Batch(specs, storage=kind, graph_chunk=5, stochastic_rounding=sr) as b:
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
        b.step(steps)
        got = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
        masses = b.total_mass()
    for k, s in enumerate(specs):
        with slm.Simulation(s.width, s.height, s.agent_count, seed=s.seed, params=s.params,
                            storage=kind, graph_chunk=0, stochastic_rounding=sr) as sim:
            assert_bitwise(sim.read_agents(), starts[k][0], f"instance {k} init")
            sim.step(steps)
            assert_bitwise(got[k][0], sim.read_agents(), f"instance {k} agents vs single")
            assert_bitwise(got[k][1], sim.read_trail(), f"instance {k} trail vs single")
        a, t = starts[k]
        orc.step(a, t, s.params, steps, storage=storage)
        assert_bitwise(got[k][0], a, f"instance {k} agents vs oracle")
        assert_bitwise(got[k][1], t, f"instance {k} trail vs oracle")
        assert masses[k] == pytest.approx(got[k][1].astype(np.float64).sum(), rel=1e-12, abs=1e-9)


def test_batch_resort_and_graphs_do_not_change_results(slm):
    specs = mixed_specs(slm, seed=1)
    outs = []
    for chunk, resort, cl, tiled in [(0, 0, 4, False), (10, 3, 4, False), (7, 2, 2, True),
                                     (5, 5, 5, True)]:
        with slm.Batch(specs, graph_chunk=chunk, resort_interval=resort, sort_cell_log2=cl,
                       tiled_sensing=tiled) as b:
            b.step(21)
            outs.append([(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))])
    for other in outs[1:]:
        for k in range(len(specs)):
            assert_bitwise(other[k][0], outs[0][k][0], f"instance {k} agents")
            assert_bitwise(other[k][1], outs[0][k][1], f"instance {k} trail")


def test_batch_per_instance_params_and_clear(slm, orc):
    specs = [slm.InstanceSpec(128, 128, 3000, seed=s, params=preset(slm, "Default"))
             for s in range(3)]
    with slm.Batch(specs, graph_chunk=4) as b:
        states = [(b.read_agents(k), b.read_trail(k)) for k in range(3)]
        b.set_params(1, preset(slm, "Mesh"))
        b.step(8)
        b.clear(2)
        b.step(3)
        for k, nm in enumerate(["Default", "Mesh", "Default"]):
            a, t = states[k]
            orc.step(a, t, preset(slm, nm), 8)
            if k == 2:
                t[:] = 0.0
            orc.step(a, t, preset(slm, nm), 3)
            assert_bitwise(b.read_agents(k), a, f"instance {k} agents")
            assert_bitwise(b.read_trail(k), t, f"instance {k} trail")


def test_batch_rejects_bad_instances(slm):
    with pytest.raises(slm.SlimeMoldError):
        slm.Batch([slm.InstanceSpec(0, 10, 5)])
    with pytest.raises(slm.SlimeMoldError):
        slm.Batch([])
    with slm.Batch([slm.InstanceSpec(16, 16, 5)]) as b:
        import ctypes
        buf = np.zeros(256, np.float32)
        ptr = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        assert slm.lib().slm_batch_read_trail(b.handle, 1, ptr, 256) == -1
        assert "out of range" in slm.lib().slm_last_error().decode()
        assert slm.lib().slm_batch_read_trail(b.handle, 0, ptr, 255) == -1


@pytest.mark.parametrize("name", ["Default", "Waves", "Snake"])
def test_f64_positions_match_oracle(slm, orc, name):
    w, h, n, steps = 700, 500, 1 << 14, 16
    s = preset(slm, name)
    with slm.Simulation(w, h, n, seed=8, params=s, f64_positions=True, graph_chunk=4,
                        resort_interval=5) as sim:
        a0 = sim.read_agents()
        t = sim.read_trail()
        sim.step(steps)
        ga, gt = sim.read_agents(), sim.read_trail()
    a = a0.astype(np.float64)
    orc.step_p64(a, t, s, steps)
    assert_bitwise(ga, a.astype(np.float32), "f64-position agents (rounded to f32)")
    assert_bitwise(gt, t, "trail")


def test_f64_positions_in_a_batch(slm):
    specs = mixed_specs(slm, seed=2)[:6]
    with slm.Batch(specs, f64_positions=True, storage="f16", resort_interval=3,
                   tiled_sensing=True) as b:
        b.step(9)
        got = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
    for k, sp in enumerate(specs):
        with slm.Simulation(sp.width, sp.height, sp.agent_count, seed=sp.seed, params=sp.params,
                            storage="f16", f64_positions=True) as sim:
            sim.step(9)
            assert_bitwise(got[k][0], sim.read_agents(), f"instance {k} agents")
            assert_bitwise(got[k][1], sim.read_trail(), f"instance {k} trail")


@pytest.mark.parametrize("groups,chunk,resort,storage", [
    (2, 0, 0, "f32"), (3, 7, 3, "bf16sr"), (4, 10, 4, "f16"), (16, 5, 0, "f32"), (8, 6, 1, "bf16")])
def test_instance_groups_equal_one_group(slm, orc, groups, chunk, resort, storage):
    """Instance groups step concurrently on their own streams, forked and joined
    around every re-sort: eager and captured, with odd stretches between
    re-sorts (the trail buffers swap parity per step), more groups than
    instances, and f64 positions in one case."""
    specs = mixed_specs(slm, seed=3)
    sr = storage == "bf16sr"
    kind = "bf16" if sr else storage
    f64 = groups == 8
    outs = []
    for g in (1, groups):
        print(f"[groups] {g} chunk={chunk} resort={resort} {storage}", flush=True)
        with slm.Batch(specs, storage=kind, graph_chunk=chunk, resort_interval=resort,
                       stochastic_rounding=sr, f64_positions=f64) as b:
            b.set_tuning(instance_groups=g)
            assert b.get_tuning().instance_groups == g
            if g == 1:
                starts = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
            b.step(13)
            b.step(chunk + 3 if chunk else 4)
            outs.append([(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))])
            masses = b.total_mass()
    steps = 13 + (chunk + 3 if chunk else 4)
    for k in range(len(specs)):
        assert_bitwise(outs[1][k][0], outs[0][k][0], f"instance {k} agents, {groups} groups")
        assert_bitwise(outs[1][k][1], outs[0][k][1], f"instance {k} trail, {groups} groups")
        assert masses[k] == pytest.approx(outs[1][k][1].astype(np.float64).sum(), rel=1e-12,
                                          abs=1e-9)
    if not f64:
        for k, s in enumerate(specs):
            a, t = starts[k]
            orc.step(a, t, s.params, steps, storage=storage)
            assert_bitwise(outs[1][k][0], a, f"instance {k} agents vs oracle")
            assert_bitwise(outs[1][k][1], t, f"instance {k} trail vs oracle")


@pytest.mark.parametrize("sort_key,groups", [(3, 1), (0x13, 3), (0x15, 1), (0x10, 2)])
def test_refined_sort_keys_track_the_oracle(slm, orc, sort_key, groups):
    """Re-sorts keyed by heading class and by the centre sensor's bin (each
    instance with its own sensor distance, up to 225 cells off small grids):
    only the agents' slots change, so every instance still equals the oracle."""
    specs = mixed_specs(slm, seed=6)
    print(f"[sort key] {sort_key:#x} groups={groups}", flush=True)
    with slm.Batch(specs, graph_chunk=5, resort_interval=3, sort_cell_log2=3) as b:
        b.set_tuning(sort_key=sort_key, instance_groups=groups)
        assert b.get_tuning().sort_key == sort_key
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
        b.step(4)
        b.step(13)
        got = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
    for k, s in enumerate(specs):
        a, t = starts[k]
        orc.step(a, t, s.params, 17)
        assert_bitwise(got[k][0], a, f"instance {k} agents")
        assert_bitwise(got[k][1], t, f"instance {k} trail")


def test_instance_groups_switched_mid_run(slm):
    specs = mixed_specs(slm, seed=4)
    with slm.Batch(specs, graph_chunk=4, resort_interval=3) as ref:
        ref.step(40)
        want = [(ref.read_agents(k), ref.read_trail(k)) for k in range(len(specs))]
    with slm.Batch(specs, graph_chunk=4, resort_interval=3) as b:
        for g, n in [(4, 9), (1, 5), (7, 11), (2, 8), (4, 7)]:
            b.set_tuning(instance_groups=g)
            b.step(n)
        for k in range(len(specs)):
            assert_bitwise(b.read_agents(k), want[k][0], f"instance {k} agents")
            assert_bitwise(b.read_trail(k), want[k][1], f"instance {k} trail")
        # groups and tiled sensing do not combine
        import ctypes
        t = b.get_tuning()
        t.tiled_sensing = 1
        assert slm.lib().slm_batch_set_tuning(b.handle, ctypes.byref(t)) == -5
        b.set_tuning(instance_groups=1, tiled_sensing=1)
        b.step(2)


def test_instance_groups_with_params_clear_render_and_mass(slm):
    """Calls between grouped steps (parameter changes, clear, render, total
    mass, reads) see every group's work finished, and change nothing."""
    specs = mixed_specs(slm, seed=6)
    lut = np.random.default_rng(2).integers(0, 256, 768, dtype=np.uint8).tobytes()
    outs = []
    for g in (1, 5):
        with slm.Batch(specs, storage="bf16", stochastic_rounding=True, graph_chunk=6,
                       resort_interval=4) as b:
            b.set_tuning(instance_groups=g, streaming=3)
            b.step(9)
            b.set_params(1, preset(slm, "Mesh"))
            b.step(7)
            b.clear(3)
            masses = b.total_mass()
            images = [b.render(k, lut, 96, 64) for k in range(len(specs))]
            b.step(13)
            outs.append((masses, images,
                         [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]))
    (m1, i1, s1), (m5, i5, s5) = outs
    assert np.array_equal(m1, m5)
    for k in range(len(specs)):
        assert_bitwise(i5[k], i1[k], f"instance {k} render")
        assert_bitwise(s5[k][0], s1[k][0], f"instance {k} agents")
        assert_bitwise(s5[k][1], s1[k][1], f"instance {k} trail")


@pytest.mark.parametrize("storage,window", [("f32", False), ("bf16sr", True)])
def test_persistent_passes_equal_oracle_over_graph_replays(slm, orc, storage, window):
    """Persistent agent and trail passes (resident workgroups claiming the
    pass's work from counters the kernels reset themselves): every instance
    equals the oracle after eager steps and several replays of the captured
    chunks, and instance groups are refused."""
    specs = mixed_specs(slm, seed=4)
    sr = storage == "bf16sr"
    print(f"[persistent] {len(specs)} instances {storage} window={window}", flush=True)
    with slm.Batch(specs, storage="bf16" if sr else storage, stochastic_rounding=sr,
                   resort_interval=3, graph_chunk=4, sort_cell_log2=3, persistent=True,
                   window_sensing=window) as b:
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
        b.step(3)   # eager
        b.step(12)  # three replays of captured chunks
        b.synchronize()
        b.step(5)   # one more replay and an eager step
        for k, sp in enumerate(specs):
            a, t = starts[k]
            orc.step(a, t, sp.params, 20, storage=storage)
            assert_bitwise(b.read_agents(k), a, f"instance {k} agents")
            assert_bitwise(b.read_trail(k), t, f"instance {k} trail")
        with pytest.raises(slm.SlimeMoldError) as e:
            b.set_tuning(instance_groups=2)
        assert e.value.code == -5
    s = preset(slm, "Default")
    ga, gt, oa, ot = run_parity(slm, orc, 1024, 512, 1 << 17, 2, s, 11, resort_interval=5,
                                sort_cell_log2=3, graph_chunk=5, persistent=True)
    assert_bitwise(ga, oa, "single persistent agents")
    assert_bitwise(gt, ot, "single persistent trail")


@pytest.mark.parametrize("storage", ["f32", "bf16sr"])
def test_c3_mix_sampled_instances_16_steps(slm, orc, storage):
    """The batched workload: the 256-instance mix with an f32 trail (the mix
    as configured) and in bf16 with stochastic rounding (the batched leg),
    128 instance groups, re-sort every 20 steps, 16 steps in graph chunks.
    Instances are independent, so a sample of them (every grid size, small and
    large agent counts) is checked against the oracle bit for bit."""
    import slimemold.workloads as wl
    specs = wl.batch_mix(256)
    sample = sorted({0, 1, 2, 3, 64, 127, 128, 200, 255} |
                    {min(range(256), key=lambda k: specs[k].agent_count),
                     max(range(256), key=lambda k: specs[k].agent_count)})
    print(f"[c3 parity] 256 instances {storage}, sampled {sample}", flush=True)
    sr = storage == "bf16sr"
    with slm.Batch(specs, storage="bf16" if sr else storage, stochastic_rounding=sr,
                   resort_interval=20, graph_chunk=4, sort_cell_log2=3) as b:
        b.set_tuning(trail_rows=8, instance_groups=128)
        starts = {k: (b.read_agents(k), b.read_trail(k)) for k in sample}
        b.step(16)
        got = {k: (b.read_agents(k), b.read_trail(k)) for k in sample}
    for k in sample:
        a, t = starts[k]
        orc.step(a, t, specs[k].params, 16, storage=storage)
        assert_bitwise(got[k][0], a, f"instance {k} agents")
        assert_bitwise(got[k][1], t, f"instance {k} trail")


def test_c4_big_instance_two_steps(slm, orc):
    """The combined workload (1024 mixed instances plus one 16384^2 instance
    with 2^28 agents, bf16 with stochastic rounding, 512 instance groups):
    after 30 steps, two more steps of the big instance and two small ones
    equal the oracle's bit for bit."""
    import slimemold.workloads as wl
    specs = wl.combined()
    big = len(specs) - 1
    sample = [0, 511, big]
    print(f"[c4 parity] {len(specs)} instances, big {specs[big].width}^2", flush=True)
    with slm.Batch(specs, storage="bf16", stochastic_rounding=True, resort_interval=20,
                   graph_chunk=4, sort_cell_log2=3) as b:
        b.set_tuning(trail_rows=8, instance_groups=512)
        b.step(30)
        starts = {k: (b.read_agents(k), b.read_trail(k)) for k in sample}
        b.step(2)
        got = {k: (b.read_agents(k), b.read_trail(k)) for k in sample}
    for k in sample:
        a, t = starts[k]
        orc.step(a, t, specs[k].params, 2, storage="bf16sr")
        assert_bitwise(got[k][0], a, f"instance {k} agents")
        assert_bitwise(got[k][1], t, f"instance {k} trail")


@pytest.mark.parametrize("storage,positions", [("bf16", "f64"), ("f16", "f64"),
                                               ("bf16", "f32"), ("f16", "f32")])
def test_c2_shape_4_steps(slm, orc, storage, positions):
    """The precision configuration itself: 8192^2 trail in bf16 or f16, 2^26
    agents, seed 7, Default preset, re-sorting on, with f32 positions (as the
    configuration states) and with f64 positions (the mode meant for grids of
    8192 and larger); 4 steps bit for bit against the oracle's f32 or
    f64-position restatement."""
    s = preset(slm, "Default")
    p64 = positions == "f64"
    print(f"[c2 parity] 8192^2 n=2^26 {storage} {positions} positions", flush=True)
    with slm.Simulation(8192, 8192, 1 << 26, seed=7, params=s, storage=storage,
                        f64_positions=p64, resort_interval=2, graph_chunk=2) as sim:
        a0 = sim.read_agents()
        t = sim.read_trail()
        sim.step(4)
        ga, gt = sim.read_agents(), sim.read_trail()
    if p64:
        a = a0.astype(np.float64)
        orc.step_p64(a, t, s, 4, storage=storage)
        assert_bitwise(ga, a.astype(np.float32), "f64-position agents (rounded to f32)")
    else:
        a = a0.copy()
        orc.step(a, t, s, 4, storage=storage)
        assert_bitwise(ga, a, "f32-position agents")
    assert_bitwise(gt, t, "trail")
