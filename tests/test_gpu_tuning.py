"""Launch tuning and the autotuner on the GPU.

A tuning only changes how the step is laid out (agents per workgroup, trail
tile rows, re-sort interval, bin size and key, tiled sensing), never a result bit:
every tuning, switched mid-run, must track the oracle exactly, and autotune
must leave the state it found, so a run that autotunes halfway equals one that
does not.
"""
import ctypes

import numpy as np
import pytest

from parity_util import assert_bitwise, preset

pytestmark = pytest.mark.gpu

TUNINGS = [
    dict(agents_per_group=512, streaming=1),
    dict(agents_per_group=4096, trail_rows=4),
    dict(trail_rows=64, streaming=3),
    dict(trail_rows=8, streaming=2),
    dict(resort_interval=3, sort_cell_log2=3),
    dict(resort_interval=3, sort_cell_log2=3, sort_key=3),
    dict(resort_interval=2, sort_cell_log2=4, sort_key=0x15, tiled_sensing=2),
    dict(resort_interval=5, sort_cell_log2=2, sort_key=0x10, tiled_sensing=0),
    dict(resort_interval=2, sort_cell_log2=5, tiled_sensing=1, agents_per_group=2048,
         sort_key=0),
    dict(resort_interval=0, tiled_sensing=1, trail_rows=8),
]


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_every_tuning_switched_mid_run_tracks_the_oracle(slm, orc, storage):
    s = preset(slm, "Firecracker Trees")  # jitter on: the original index matters
    w, h, n = 520, 300, 40000
    with slm.Simulation(w, h, n, seed=21, params=s, storage=storage, graph_chunk=4) as sim:
        a, t = sim.read_agents(), sim.read_trail()
        for k, fields in enumerate(TUNINGS):
            print(f"[tuning] {storage} {fields}", flush=True)
            sim.set_tuning(**fields)
            got = sim.get_tuning()
            for name, v in fields.items():
                assert getattr(got, name) == v, name
            sim.step(5 + k)
            orc.step(a, t, s, 5 + k, storage=storage)
            assert_bitwise(sim.read_agents(), a, f"agents after {fields}")
            assert_bitwise(sim.read_trail(), t, f"trail after {fields}")
        assert sim.info().steps_done == sum(5 + k for k in range(len(TUNINGS)))


def test_set_tuning_rejects_bad_values(slm):
    with slm.Simulation(64, 64, 100, seed=1) as sim:
        before = sim.get_tuning().as_dict()
        for fields in [dict(agents_per_group=300), dict(agents_per_group=8192),
                       dict(agents_per_group=128), dict(trail_rows=3), dict(trail_rows=128),
                       dict(sort_cell_log2=13), dict(tiled_sensing=3),
                       dict(instance_groups=4097), dict(streaming=4), dict(sort_key=6),
                       dict(sort_key=0x20), dict(sort_key=0x8)]:
            with pytest.raises(slm.SlimeMoldError) as e:
                sim.set_tuning(**fields)
            assert e.value.code == -1, fields
        # tiled sensing's tile work lists need the plain re-sort key
        with pytest.raises(slm.SlimeMoldError) as e:
            sim.set_tuning(resort_interval=4, tiled_sensing=1, sort_key=0x13)
        assert e.value.code == -5
        assert sim.get_tuning().as_dict() == before
        sim.step(3)  # still usable


@pytest.mark.parametrize("resort", [0, 4])
def test_autotune_restores_state(slm, orc, resort):
    s = preset(slm, "Waves")
    w, h, n = 640, 480, 60000
    with slm.Simulation(w, h, n, seed=5, params=s, resort_interval=resort, graph_chunk=10) as sim:
        sim.step(7)
        a, t = sim.read_agents(), sim.read_trail()
        res = sim.autotune(trial_steps=6)
        print(f"[autotune] {res.as_dict()}", flush=True)
        assert res.trials >= 10
        assert res.step_ms > 0 and res.trail_ms > 0 and res.agents_ms > 0
        assert sim.info().steps_done == 7
        assert_bitwise(sim.read_agents(), a, "agents after autotune")
        assert_bitwise(sim.read_trail(), t, "trail after autotune")
        chosen = sim.get_tuning()
        for k in ("resort_interval", "sort_cell_log2", "agents_per_group", "trail_rows",
                  "tiled_sensing", "sort_key"):
            assert getattr(chosen, k) == getattr(res, k), k
        sim.step(23)
        orc.step(a, t, s, 23)
        assert_bitwise(sim.read_agents(), a, "agents 23 steps after autotune")
        assert_bitwise(sim.read_trail(), t, "trail 23 steps after autotune")


def test_batch_autotune_and_tunings_keep_results(slm):
    specs = [slm.InstanceSpec(w, h, n, seed=40 + k, params=preset(slm, name))
             for k, (w, h, n, name) in enumerate([(256, 256, 9000, "Default"),
                                                  (1030, 40, 3000, "Snake"),
                                                  (512, 128, 0, "Curls"),
                                                  (64, 96, 5000, "Mesh"),
                                                  (300, 200, 20000, "Firecracker Trees")])]
    with slm.Batch(specs, graph_chunk=5) as ref:
        ref.step(30)
        want = [(ref.read_agents(k), ref.read_trail(k)) for k in range(len(specs))]
    with slm.Batch(specs, graph_chunk=5) as b:
        b.step(10)
        res = b.autotune(trial_steps=4)
        print(f"[batch autotune] {res.as_dict()}", flush=True)
        assert res.trials >= 10
        b.step(10)
        b.set_tuning(resort_interval=3, agents_per_group=1024, trail_rows=16)
        b.step(10)
        assert b.info().steps_done == 30
        for k in range(len(specs)):
            assert_bitwise(b.read_agents(k), want[k][0], f"instance {k} agents")
            assert_bitwise(b.read_trail(k), want[k][1], f"instance {k} trail")


def test_f64_positions_survive_retuning(slm):
    s = preset(slm, "Default")
    with slm.Simulation(300, 200, 5000, seed=2, params=s, f64_positions=True) as ref:
        ref.step(12)
        wa, wt = ref.read_agents(), ref.read_trail()
    with slm.Simulation(300, 200, 5000, seed=2, params=s, f64_positions=True) as sim:
        sim.step(4)
        sim.set_tuning(resort_interval=2)
        sim.step(4)
        sim.set_tuning(resort_interval=0, agents_per_group=512)
        sim.step(4)
        assert_bitwise(sim.read_agents(), wa, "agents")
        assert_bitwise(sim.read_trail(), wt, "trail")


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_step_timed_matches_step(slm, storage):
    """slm_step_timed replays graph chunks with event-record nodes (and eager
    steps for the remainder): same bits as slm_step, every kernel timed."""
    s = preset(slm, "Firecracker Trees")
    kw = dict(seed=8, params=s, storage=storage, graph_chunk=10, resort_interval=3)
    with slm.Simulation(300, 200, 20000, **kw) as a, slm.Simulation(300, 200, 20000, **kw) as b:
        a.step(35)
        t = b.step_timed(35)
        t2 = b.step_timed(20)  # replays the cached timed graphs
        a.step(20)
        assert_bitwise(b.read_agents(), a.read_agents(), "agents")
        assert_bitwise(b.read_trail(), a.read_trail(), "trail")
        for tm in (t, t2):
            print(f"[timed] {tm.agents_ms:.3f} {tm.trail_ms:.3f} {tm.sort_ms:.3f} "
                  f"{tm.agents_launches} {tm.trail_launches} {tm.sort_launches}", flush=True)
        assert t.agents_launches == 35 and t.trail_launches == 35 and t2.trail_launches == 20
        # re-sorts on absolute steps 0, 3, 6, ...: 12 in steps 0..34, 7 in 35..54
        assert t.sort_launches == 12 and t2.sort_launches == 7
        assert t.agents_ms > 0 and t.trail_ms > 0 and t.sort_ms > 0
        assert t2.agents_ms > 0 and t2.trail_ms > 0
        # per launch, times are those of real kernels (not stale or zero events)
        assert 0.0005 < t.trail_ms / t.trail_launches < 5.0


def test_resort_schedule_is_absolute_across_call_sizes(slm):
    """Re-sorts fall on steps 0, k, 2k, ... of the simulation whatever the
    calls' sizes and whether steps run eagerly or from captured chunks (a
    chunk that k does not divide is captured per starting phase)."""
    s = preset(slm, "Default")
    kw = dict(seed=9, params=s, graph_chunk=20, resort_interval=15)
    with slm.Simulation(256, 128, 9000, **kw) as a, slm.Simulation(256, 128, 9000, **kw) as b:
        total = 0
        for n in (7, 100, 13, 41, 60):  # eager remainders and chunks at shifted phases
            tm = b.step_timed(n)
            want = sum(1 for i in range(total, total + n) if i % 15 == 0)
            print(f"[schedule] steps {total}..{total + n - 1}: {tm.sort_launches} re-sorts "
                  f"(want {want})", flush=True)
            assert tm.sort_launches == want and tm.agents_launches == n
            total += n
        for n in (7, 100, 13, 41, 60):
            a.step(n)  # the captured-graph path of slm_step, same schedule
        a.synchronize()
        assert_bitwise(b.read_agents(), a.read_agents(), "agents")
        assert_bitwise(b.read_trail(), a.read_trail(), "trail")


def test_too_many_resort_bins_is_an_error(slm):
    """A wide grid with small bins needs 2^40 bins here: creation (and a tuning
    asking for it) fails with an error instead of wrapping the count."""
    with pytest.raises(slm.SlimeMoldError) as e:
        slm.Simulation(1 << 20, 64, 1000, seed=1, resort_interval=5, sort_cell_log2=0)
    assert e.value.code == -1
    with slm.Simulation(1 << 20, 64, 1000, seed=1, resort_interval=5, sort_cell_log2=10) as sim:
        for fields in (dict(sort_cell_log2=0), dict(sort_cell_log2=4, sort_key=5)):
            with pytest.raises(slm.SlimeMoldError) as e:
                sim.set_tuning(**fields)
            assert e.value.code == -1, fields
        sim.step(6)  # still usable after the refusals
