"""HIP path vs oracle, bit for bit, through the C ABI (ctypes).

The parity bar is per-agent position/heading within 1e-6 relative and the
trail within 1e-5 absolute for the first 16 steps; both sides evaluate the same
f32 operations in the same order (including sin/cos from slm_trig.h), so the
tests demand exact equality, which is stricter.
"""
import numpy as np
import pytest

from parity_util import PRESET_NAMES, assert_bitwise, preset, run_parity

pytestmark = pytest.mark.gpu


def test_init_matches_oracle(slm, orc):
    s = preset(slm, "Default")
    for (w, h, n, seed) in [(1024, 1024, 1 << 18, 42), (37, 23, 1000, 7), (1600, 900, 5000, 0)]:
        with slm.Simulation(w, h, n, seed=seed, params=s) as sim:
            got = sim.read_agents()
            assert np.all(sim.read_trail() == 0.0)
        want = orc.init_agents(seed, w, h, n, s.agent_speed_min, s.agent_speed_max)
        assert_bitwise(got, want, f"init {w}x{h} n={n}")


def test_default_preset_c0_shape_16_steps(slm, orc):
    """1024x1024, 2^18 agents, Default preset, seed 42: the small headless case."""
    ga, gt, oa, ot = run_parity(slm, orc, 1024, 1024, 1 << 18, 42, preset(slm, "Default"), 16)
    assert_bitwise(ga, oa, "agents")
    assert_bitwise(gt, ot, "trail")


@pytest.mark.parametrize("name", PRESET_NAMES)
def test_every_preset_16_steps(slm, orc, name):
    ga, gt, oa, ot = run_parity(slm, orc, 256, 256, 1 << 14, 3, preset(slm, name), 16)
    assert_bitwise(ga, oa, f"{name} agents")
    assert_bitwise(gt, ot, f"{name} trail")


@pytest.mark.parametrize("storage", ["bf16", "f16", "bf16sr"])
@pytest.mark.parametrize("name", ["Default", "Curls"])
def test_16bit_storage_16_steps(slm, orc, storage, name):
    ga, gt, oa, ot = run_parity(slm, orc, 512, 384, 1 << 15, 5, preset(slm, name), 16,
                                storage=storage)
    assert_bitwise(ga, oa, f"{storage} agents")
    assert_bitwise(gt, ot, f"{storage} trail")


@pytest.mark.parametrize("w,h", [(37, 23), (1, 1), (1, 17), (17, 1), (2, 3), (1028, 60),
                                 (1030, 40), (1032, 50), (1600, 900), (260, 2)])
@pytest.mark.parametrize("storage", ["f32", "bf16", "bf16sr"])
def test_ragged_grids(slm, orc, w, h, storage):
    s = preset(slm, "Waves")
    n = min(4000, w * h * 2)
    ga, gt, oa, ot = run_parity(slm, orc, w, h, n, 11, s, 12, storage=storage)
    assert_bitwise(ga, oa, f"{w}x{h} agents")
    assert_bitwise(gt, ot, f"{w}x{h} trail")


def test_random_state_start(slm, orc):
    rng = np.random.default_rng(12)
    w, h, n = 300, 200, 20000
    agents = np.stack([rng.uniform(0, w, n), rng.uniform(0, h, n), rng.uniform(-10, 20, n),
                       rng.uniform(0, 90, n)], axis=1).astype(np.float32)
    agents[:5, 0] = w  # on the far edges
    agents[5:10, 1] = h
    trail = rng.uniform(0, 1.2, (h, w)).astype(np.float32)
    s = preset(slm, "Mesh")
    ga, gt, oa, ot = run_parity(slm, orc, w, h, n, 0, s, 5, trail0=trail, agents0=agents)
    assert_bitwise(ga, oa, "agents")
    assert_bitwise(gt, ot, "trail")


def test_zero_and_one_agent(slm, orc):
    rng = np.random.default_rng(13)
    trail = rng.uniform(0, 1, (64, 64)).astype(np.float32)
    for n in (0, 1):
        ga, gt, oa, ot = run_parity(slm, orc, 64, 64, n, 1, preset(slm, "Default"), 7,
                                    trail0=trail)
        assert_bitwise(ga, oa, f"n={n} agents")
        assert_bitwise(gt, ot, f"n={n} trail")


@pytest.mark.parametrize("interval,tiled", [(1, False), (3, False), (1, True), (4, True)])
def test_resort_gives_identical_results(slm, orc, interval, tiled):
    """Morton re-sort permutes agents internally and tiled sensing stages trail
    windows in LDS; outputs are in original order and the result must not change."""
    s = preset(slm, "Firecracker Trees")
    ga, gt, oa, ot = run_parity(slm, orc, 512, 512, 1 << 16, 9, s, 20, resort_interval=interval,
                                graph_chunk=0, tiled_sensing=tiled)
    assert_bitwise(ga, oa, "agents")
    assert_bitwise(gt, ot, "trail")


WINDOW_CASES = [
    # (w, h, n, storage, preset, resort, bin log2, steps): ragged and tiny grids,
    # rows of whole 16-byte chunks only (others sense globally), every storage,
    # sensor distances from 20 to 225 (windows too large fall back), no re-sort
    dict(w=512, h=512, n=1 << 16, storage="f32", name="Firecracker Trees", resort=1, cl=3, steps=20),
    dict(w=1024, h=256, n=60000, storage="f32", name="Default", resort=7, cl=3, steps=24),
    dict(w=520, h=300, n=40000, storage="bf16", name="Waves", resort=2, cl=4, steps=16),
    dict(w=256, h=200, n=30000, storage="f16", name="Sponge", resort=5, cl=3, steps=16),
    dict(w=600, h=600, n=50000, storage="f32", name="Snake", resort=3, cl=3, steps=12),
    dict(w=96, h=17, n=3000, storage="bf16", name="Mesh", resort=1, cl=2, steps=12),
    dict(w=8, h=8, n=500, storage="f32", name="Curls", resort=1, cl=2, steps=12),
    dict(w=333, h=129, n=20000, storage="f32", name="Threads", resort=4, cl=3, steps=12),
    dict(w=400, h=300, n=25000, storage="f32", name="Default", resort=0, cl=4, steps=12),
]


@pytest.mark.parametrize("case", WINDOW_CASES,
                         ids=lambda c: f"{c['w']}x{c['h']}-{c['storage']}-{c['name'].replace(' ', '')}")
def test_window_sensing_gives_identical_results(slm, orc, case):
    """Windowed sensing stages each workgroup's bounding-box window of the
    trail in LDS; taps outside it and workgroups whose window does not fit
    read global memory, so the result must be the oracle's bit for bit."""
    s = preset(slm, case["name"])
    ga, gt, oa, ot = run_parity(slm, orc, case["w"], case["h"], case["n"], 31, s, case["steps"],
                                storage=case["storage"], resort_interval=case["resort"],
                                sort_cell_log2=case["cl"], graph_chunk=4, window_sensing=True)
    assert_bitwise(ga, oa, "agents")
    assert_bitwise(gt, ot, "trail")


def test_window_sensing_with_f64_positions_and_groups(slm, orc):
    """The window's bounding box from f64 positions, and windowed sensing in a
    batch with instance groups (each group's launch offsets its tables)."""
    s = preset(slm, "Default")
    ga, gt, oa, ot = run_parity(slm, orc, 640, 480, 40000, 5, s, 10, resort_interval=3,
                                sort_cell_log2=3, graph_chunk=0, window_sensing=True,
                                f64_positions=True)
    # f64 positions: the oracle's f32 path equals the f64 one where positions
    # stay exact; compare against the f64-position restatement instead
    with slm.Simulation(640, 480, 40000, seed=5, params=s, resort_interval=3,
                        sort_cell_log2=3, graph_chunk=0, f64_positions=True) as ref:
        ref.step(10)
        assert_bitwise(ga, ref.read_agents(), "f64 agents vs global sensing")
        assert_bitwise(gt, ref.read_trail(), "f64 trail vs global sensing")
    specs = [slm.InstanceSpec(w, h, n, seed=k, params=preset(slm, nm))
             for k, (w, h, n, nm) in enumerate([(256, 256, 20000, "Default"),
                                                (512, 128, 30000, "Waves"),
                                                (128, 128, 5000, "Snake"),
                                                (64, 64, 700, "Sponge")])]
    with slm.Batch(specs, resort_interval=2, sort_cell_log2=3, graph_chunk=3) as b:
        b.set_tuning(instance_groups=3, tiled_sensing=2)
        assert b.get_tuning().tiled_sensing == 2
        starts = [(b.read_agents(k), b.read_trail(k)) for k in range(len(specs))]
        b.step(9)
        for k, sp in enumerate(specs):
            a, t = starts[k]
            orc.step(a, t, sp.params, 9)
            assert_bitwise(b.read_agents(k), a, f"instance {k} agents")
            assert_bitwise(b.read_trail(k), t, f"instance {k} trail")


def test_graphs_match_eager(slm):
    s = preset(slm, "Snake")
    outs = []
    for chunk, resort, tiled in [(0, 0, False), (10, 0, False), (10, 4, False), (7, 3, True)]:
        with slm.Simulation(384, 256, 30000, seed=4, params=s, graph_chunk=chunk,
                            resort_interval=resort, tiled_sensing=tiled) as sim:
            sim.step(33)
            sim.step(10)
            outs.append((sim.read_agents(), sim.read_trail()))
    for a, t in outs[1:]:
        assert_bitwise(a, outs[0][0], "agents")
        assert_bitwise(t, outs[0][1], "trail")


def test_graph_replays_track_parameter_changes(slm, orc):
    """The preset switch (src/main.rs:815-832) lands in captured graphs."""
    names = ["Default", "Sponge", "Firecracker Trees"]
    w, h, n = 256, 256, 1 << 14
    with slm.Simulation(w, h, n, seed=1, params=preset(slm, names[0]), graph_chunk=5) as sim:
        a = sim.read_agents()
        t = sim.read_trail()
        for nm in names:
            sim.set_params(preset(slm, nm))
            sim.step(10)
            orc.step(a, t, preset(slm, nm), 10)
        assert_bitwise(sim.read_agents(), a, "agents")
        assert_bitwise(sim.read_trail(), t, "trail")


def test_clear_resize_count_speeds(slm, orc):
    s = preset(slm, "Default")
    with slm.Simulation(200, 100, 5000, seed=21, params=s) as sim:
        sim.step(5)
        assert sim.read_trail().max() > 0
        sim.clear()
        assert np.all(sim.read_trail() == 0.0)
        # resize: positions scaled by the f32 ratio, trail zero (src/main.rs:993-1015)
        before = sim.read_agents()
        sim.resize(300, 150)
        after = sim.read_agents()
        sx = np.float32(np.float32(300) / np.float32(200))
        sy = np.float32(np.float32(150) / np.float32(100))
        assert np.array_equal(after[:, 0], (before[:, 0] * sx).astype(np.float32))
        assert np.array_equal(after[:, 1], (before[:, 1] * sy).astype(np.float32))
        assert np.array_equal(after[:, 2:], before[:, 2:])
        t = sim.read_trail()
        assert t.shape == (150, 300) and np.all(t == 0.0)
        sim.step(4)
        orc.step(after, t, s, 4)
        assert_bitwise(sim.read_agents(), after, "after resize")
        # reassign speeds (src/main.rs:101-145): one fresh stream per call
        a = sim.read_agents()
        sim.reassign_speeds(10.0, 20.0)
        orc.reassign_speeds(a, 21, 1, 10.0, 20.0)
        assert_bitwise(sim.read_agents(), a, "reassign")
        # agent count: everything re-initialised from the seed (src/main.rs:682-791)
        sim.set_agent_count(7000)
        want = orc.init_agents(21, 300, 150, 7000, s.agent_speed_min, s.agent_speed_max)
        assert_bitwise(sim.read_agents(), want, "set_agent_count")


def test_total_mass_f64(slm):
    rng = np.random.default_rng(14)
    t = rng.uniform(0, 1, (333, 517)).astype(np.float32)
    for storage in ("f32", "bf16", "f16"):
        with slm.Simulation(517, 333, 0, params=preset(slm, "Default"), storage=storage) as sim:
            sim.write_trail(t)
            stored = sim.read_trail().astype(np.float64)
            assert sim.total_mass() == pytest.approx(stored.sum(), rel=1e-13)


def test_rejects_bad_input_without_launching(slm):
    with slm.Simulation(64, 64, 10, params=preset(slm, "Default")) as sim:
        a = sim.read_agents()
        bad = a.copy()
        bad[3, 0] = 65.0  # outside [0, width]
        with pytest.raises(slm.SlimeMoldError):
            sim.write_agents(bad)
        bad = a.copy()
        bad[0, 2] = np.nan
        with pytest.raises(slm.SlimeMoldError):
            sim.write_agents(bad)
        t = np.zeros((64, 64), np.float32)
        t[1, 1] = -0.5
        with pytest.raises(slm.SlimeMoldError):
            sim.write_trail(t)
        with pytest.raises(slm.SlimeMoldError):
            sim.write_trail(np.zeros((10, 10), np.float32))
        # -0 is accepted and stored as +0 (the sign bit is the deposit mark)
        t = np.zeros((64, 64), np.float32)
        t[2, 2] = -0.0
        sim.write_trail(t)
        assert not np.signbit(sim.read_trail()).any()
        assert np.array_equal(sim.read_agents(), a)


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_16bit_write_rounding_matches_oracle(slm, orc, storage):
    """write_trail rounds to the storage format to nearest-even on the device;
    the oracle's integer rounding (pinned against numpy/torch) must agree,
    including subnormal results and exact ties."""
    rng = np.random.default_rng(15)
    h, w = 64, 96
    vals = np.concatenate([rng.uniform(0, 1, 3000), rng.uniform(0, 1e-4, 1000),
                           rng.uniform(0, 1e-37, 500), rng.uniform(0, 2, 500)])
    # exact ties halfway between neighbours in both formats
    ties = np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -11, 0.5 + 2 ** -12],
                    dtype=np.float32)
    vals = np.concatenate([vals.astype(np.float32), ties])
    t = np.zeros(h * w, np.float32)
    t[:vals.size] = vals
    t = t.reshape(h, w)
    with slm.Simulation(w, h, 0, params=preset(slm, "Default"), storage=storage) as sim:
        sim.write_trail(t)
        got = sim.read_trail()
    assert_bitwise(got, orc.round_storage(t, storage), f"{storage} rounding")


def test_stochastic_rounding_needs_bf16(slm):
    for storage in ("f32", "f16"):
        with pytest.raises(slm.SlimeMoldError) as e:
            slm.Simulation(16, 16, 10, storage=storage, stochastic_rounding=True)
        assert e.value.code == -5


@pytest.mark.parametrize("resort", [0, 3])
def test_every_agent_moves_exactly_once_above_the_reference_dispatch_limit(slm, resort):
    """2^24 + 3 agents, far above the 4,194,240 the reference's 2-D dispatch
    covers once per step (compute.wgsl:58-63).  Sensors placed off the grid all
    read 0, so nobody turns, and every speed is clamped to one value: after K
    steps each agent has moved exactly K * speed * 0.016 cells along its
    heading, toroidally.  An agent updated twice or skipped in a step is off
    by a whole move (0.8 cells)."""
    import dataclasses
    n, W, H, K = (1 << 24) + 3, 4096, 4096, 8
    s = dataclasses.replace(preset(slm, "Default"), agent_sensor_distance=100000.0,
                            agent_speed_min=50.0, agent_speed_max=50.0, agent_jitter=0.0)
    print(f"[exactly once] n={n} resort={resort}", flush=True)
    with slm.Simulation(W, H, n, seed=9, params=s, resort_interval=resort) as sim:
        a0 = sim.read_agents()
        sim.step(K)
        a1 = sim.read_agents()
    assert np.all(a1[:, 3] == np.float32(50.0))
    assert np.array_equal(a1[:, 2], a0[:, 2])  # headings in [0, 2pi) are left as they are
    step = K * float(np.float32(50.0) * np.float32(0.016))
    th = a0[:, 2].astype(np.float64)
    for axis, size, trig in ((0, W, np.cos), (1, H, np.sin)):
        want = a0[:, axis].astype(np.float64) + step * trig(th)
        d = np.abs(((a1[:, axis] - want) + size / 2) % size - size / 2)
        assert d.max() < 0.02, (axis, float(d.max()), int(np.argmax(d)))


def test_c1_headline_shape_16_steps(slm, orc):
    """The plain leg's own configuration, bit for bit against the oracle for the
    first 16 steps: 4096x4096 f32, 2^24 agents, seed 1, presets Default ->
    Sponge -> Firecracker Trees switched between graph chunks, re-sort every
    15 steps into 8x8 bins, streaming trail pass."""
    names = ["Default", "Sponge", "Firecracker Trees"]
    print("[c1 parity] 4096x4096 n=2^24 16 steps", flush=True)
    with slm.Simulation(4096, 4096, 1 << 24, seed=1, params=preset(slm, names[0]),
                        resort_interval=15, graph_chunk=4, sort_cell_log2=3) as sim:
        sim.set_tuning(trail_rows=16, streaming=3)
        a, t = sim.read_agents(), sim.read_trail()
        for name, n in zip(names, (4, 4, 8)):
            sim.set_params(preset(slm, name))
            sim.step(n)
        ga, gt = sim.read_agents(), sim.read_trail()
    for name, n in zip(names, (4, 4, 8)):
        orc.step(a, t, preset(slm, name), n)
    assert_bitwise(ga, a, "agents")
    assert_bitwise(gt, t, "trail")
