"""The display colour map and snapshots on the GPU.

Render (src/display.wgsl:29-87): the HIP pass equals the oracle's restatement
byte for byte, for grids and textures of every aspect and storage format, into
host or device memory; rendered at the reference's own sampling point (after
the deposits and the decay, before the diffuse: src/main.rs:1183-1217) it
equals the oracle's render of that intermediate trail.  Snapshots: a run saved and resumed (in the same
simulation, a fresh one, or one created from the snapshot) equals the run that
was never interrupted, bit for bit; a snapshot built by an independent encoder
of the documented layout loads as the state it describes.
"""
import ctypes

import numpy as np
import pytest

import snapshot_util as su
from parity_util import assert_bitwise, display_point, preset

pytestmark = pytest.mark.gpu


def random_lut(seed):
    return np.random.default_rng(seed).integers(0, 256, 768, dtype=np.uint8).tobytes()


@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
def test_render_matches_oracle(slm, orc, storage):
    lut = random_lut(3)
    for (w, h, n) in [(1600, 900, 30000), (257, 1031, 5000), (64, 64, 2000)]:
        with slm.Simulation(w, h, n, seed=2, params=preset(slm, "Default"), storage=storage) as sim:
            sim.step(9)
            t = sim.read_trail()
            for tw, th in [(w, h), (1920, 1080), (1000, 1000), (7, 3), (3, 7), (4096, 100),
                           (w * 2, h * 2)]:
                print(f"[render] {w}x{h} {storage} -> {tw}x{th}", flush=True)
                got = sim.render(lut, tw, th)
                want = orc.render(t, lut, tw, th)
                assert_bitwise(got, want, f"render {w}x{h} -> {tw}x{th}")


def hip_runtime():
    """The HIP runtime the library links against (device memory for the test)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_render_into_device_memory_and_batches(slm, orc):
    lut = slm.lut.gray()
    hip = hip_runtime()
    with slm.Simulation(640, 360, 8000, seed=4, params=preset(slm, "Snake")) as sim:
        sim.step(5)
        nbytes = 1280 * 720 * 4
        dev = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(dev), nbytes) == 0
        try:
            N = slm._native
            N.check(N.lib().slm_render_rgba8_device(sim.handle, slm.simulation._lut_bytes(lut),
                                                    1280, 720, dev))
            sim.synchronize()
            got = np.empty((720, 1280, 4), np.uint8)
            assert hip.hipMemcpy(got.ctypes.data, dev, nbytes, 2) == 0  # device to host
        finally:
            hip.hipFree(dev)
        assert_bitwise(got, orc.render(sim.read_trail(), lut.to_bytes(), 1280, 720),
                       "device render")
    specs = [slm.InstanceSpec(w, h, 3000, seed=k, params=preset(slm, "Waves"))
             for k, (w, h) in enumerate([(300, 200), (128, 512), (37, 23)])]
    with slm.Batch(specs) as b:
        b.step(6)
        for k in range(len(specs)):
            lut2 = random_lut(10 + k)
            assert_bitwise(b.render(k, lut2, 333, 222),
                           orc.render(b.read_trail(k), lut2, 333, 222), f"batch render {k}")


STEP_RENDER_CASES = [
    dict(storage="f32", name="Default", resort=0, gaussian=False, shape=(1600, 900, 30000)),
    dict(storage="bf16", name="Firecracker Trees", resort=5, gaussian=False,
         shape=(257, 1031, 6000)),
    dict(storage="f16", name="Mesh", resort=3, gaussian=False, shape=(300, 200, 9000)),
    dict(storage="f32", name="Snake", resort=4, gaussian=True, shape=(128, 96, 3000)),
]


@pytest.mark.parametrize("case", STEP_RENDER_CASES,
                         ids=lambda c: f"{c['storage']}-{c['name'].replace(' ', '')}")
def test_step_render_shows_the_reference_sampling_point(slm, orc, case):
    w, h, n = case["shape"]
    s = preset(slm, case["name"])
    lut = random_lut(21)
    kw = dict(seed=6, params=s, storage=case["storage"], resort_interval=case["resort"],
              graph_chunk=4, gaussian=case["gaussian"])
    with slm.Simulation(w, h, n, **kw) as sim:
        sim.step(7)
        a, t = sim.read_agents(), sim.read_trail()
        for k, (steps, tw, th) in enumerate([(1, w, h), (6, 1920, 1080), (3, 7, 3)]):
            print(f"[step_render] {case} steps={steps} tex={tw}x{th}", flush=True)
            got = sim.step_render(steps, lut, tw, th)
            orc.step(a, t, s, steps - 1, storage=case["storage"], gaussian=case["gaussian"])
            shown = display_point(orc, a, t, s, case["storage"], gaussian=case["gaussian"])
            assert_bitwise(got, orc.render(shown, lut, tw, th), f"step_render {k}")
            # the state afterwards is slm_step's
            assert_bitwise(sim.read_agents(), a, f"agents after step_render {k}")
            assert_bitwise(sim.read_trail(), t, f"trail after step_render {k}")
            if s.pheromone_diffusion_rate > 0 and k == 0:
                # the post-diffuse image (slm_render_rgba8) is a different one
                assert not np.array_equal(sim.render(lut, tw, th), got)


def test_step_render_rejects_bad_arguments(slm):
    with slm.Simulation(32, 32, 10) as sim:
        L = slm.lib()
        lut = slm.simulation._lut_bytes(slm.lut.gray())
        out = (ctypes.c_uint8 * 64)()
        assert L.slm_step_render_rgba8(sim.handle, 0, lut, 4, 4, out) == -1
        assert L.slm_step_render_rgba8(sim.handle, 1, None, 4, 4, out) == -1
        assert L.slm_step_render_rgba8_device(sim.handle, 1, lut, 0, 4, out) == -1
        assert sim.info().steps_done == 0


def test_render_rejects_bad_arguments(slm):
    with slm.Simulation(32, 32, 10) as sim:
        L = slm.lib()
        lut = slm.simulation._lut_bytes(slm.lut.gray())
        out = (ctypes.c_uint8 * 16)()
        assert L.slm_render_rgba8(sim.handle, lut, 0, 4, out) == -1
        assert L.slm_render_rgba8(sim.handle, None, 2, 2, out) == -1
        assert L.slm_render_rgba8(sim.handle, lut, 1 << 16, 1 << 15, out) == -1
        with pytest.raises(ValueError):
            sim.render(b"\0" * 100, 4, 4)


CASES = [
    dict(storage="f32", resort=0, f64=False, name="Firecracker Trees"),
    dict(storage="bf16", resort=5, f64=False, name="Default"),
    dict(storage="f16", resort=3, f64=False, name="Mesh"),
    dict(storage="f32", resort=4, f64=True, name="Curls"),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['storage']}-r{c['resort']}-f64{c['f64']}")
def test_resume_equals_uninterrupted(slm, case):
    s = preset(slm, case["name"])
    kw = dict(seed=5, params=s, storage=case["storage"], resort_interval=case["resort"],
              graph_chunk=4, f64_positions=case["f64"])
    w, h, n = 400, 300, 25000
    with slm.Simulation(w, h, n, **kw) as a:
        a.step(13)
        a.reassign_speeds(20.0, 90.0)
        snap = a.snapshot()
        a.step(11)
        a.reassign_speeds(10.0, 60.0)
        a.step(6)
        want = (a.read_agents(), a.read_trail(), a.info().steps_done)

    def finish(sim, label):
        sim.step(11)
        sim.reassign_speeds(10.0, 60.0)
        sim.step(6)
        assert_bitwise(sim.read_agents(), want[0], f"{label} agents")
        assert_bitwise(sim.read_trail(), want[1], f"{label} trail")
        assert sim.info().steps_done == want[2]

    with slm.Simulation.from_snapshot(snap) as b:
        assert b.info().storage == {"f32": 0, "bf16": 1, "f16": 2}[case["storage"]]
        finish(b, "created from snapshot")
    with slm.Simulation(w, h, n, **dict(kw, seed=77, params=preset(slm, "Sponge"))) as c:
        c.step(3)
        c.load_snapshot(snap)
        assert c.get_params().agent_turn_speed == pytest.approx(s.agent_turn_speed, rel=1e-7)
        finish(c, "loaded into another simulation")


def test_independently_encoded_snapshot_loads_as_described(slm):
    rng = np.random.default_rng(8)
    w, h, n = 96, 64, 500
    agents = np.stack([rng.uniform(0, w, n), rng.uniform(0, h, n), rng.uniform(0, 6.28, n),
                       rng.uniform(30, 50, n)], axis=1).astype(np.float32)
    trail = rng.uniform(0, 1, (h, w)).astype(np.float32)
    data = su.build(w, h, agents, trail, steps=41, seed=3)
    with slm.Simulation.from_snapshot(data) as sim:
        assert sim.info().steps_done == 41
        assert_bitwise(sim.read_agents(), agents, "agents")
        assert_bitwise(sim.read_trail(), trail, "trail")
        assert sim.snapshot() == data  # saving again reproduces the same bytes
    bf = (trail.view(np.uint32) >> 16).astype(np.uint16)  # bf16 bits (truncated: any bits load)
    data16 = su.build(w, h, agents, bf, storage=1)
    with slm.Simulation.from_snapshot(data16) as sim:
        want = (bf.astype(np.uint32) << 16).view(np.float32)
        assert_bitwise(sim.read_trail(), want, "bf16 trail")


def test_snapshot_errors(slm):
    with slm.Simulation(50, 40, 100, seed=1) as a, slm.Simulation(50, 41, 100, seed=1) as b:
        snap = a.snapshot()
        with pytest.raises(slm.SlimeMoldError) as e:
            b.load_snapshot(snap)
        assert e.value.code == -1
        bad = bytearray(snap)
        bad[-20] ^= 0x40
        with pytest.raises(slm.SlimeMoldError) as e:
            a.load_snapshot(bytes(bad))
        assert "checksum" in str(e.value)
        # a trail cell marked (sign bit) or out of range is refused even with a valid checksum
        ag = a.read_agents()
        t = a.read_trail().copy()
        t[3, 4] = -0.5
        with pytest.raises(slm.SlimeMoldError):
            a.load_snapshot(su.build(50, 40, ag, t, seed=1))
        ag2 = ag.copy()
        ag2[0, 0] = 51.0
        with pytest.raises(slm.SlimeMoldError):
            a.load_snapshot(su.build(50, 40, ag2, a.read_trail(), seed=1))
        a.step(2)  # still usable
