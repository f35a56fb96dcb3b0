"""slm_render_rgba8_noalloc through the C ABI: the same bytes as
slm_render_rgba8 (which stages through a buffer it allocates per call), for
two calls in a row on the fixed per-GPU staging array - the second with a
smaller texture, so stale bytes of the first would show - and for a second
handle and storage format on the same GPU; textures beyond the array's size
are refused before any launch."""
import ctypes

import numpy as np
import pytest

from parity_util import assert_bitwise, preset

pytestmark = pytest.mark.gpu

MAX_BYTES = 16 << 20  # SLM_RENDER_NOALLOC_MAX_BYTES


def lut_bytes(seed):
    return np.random.default_rng(seed).integers(0, 256, 768, dtype=np.uint8).tobytes()


def render_both(slm, sim, lut, tw, th):
    L = slm.lib()
    N = slm._native
    buf = (ctypes.c_uint8 * 768).from_buffer_copy(lut)
    got = np.full((th, tw, 4), 7, np.uint8)
    N.check(L.slm_render_rgba8_noalloc(sim.handle, buf, tw, th, got.ctypes.data))
    want = np.empty((th, tw, 4), np.uint8)
    N.check(L.slm_render_rgba8(sim.handle, buf, tw, th, want.ctypes.data))
    return got, want


def test_noalloc_render_equals_render_on_two_calls(slm):
    with slm.Simulation(640, 360, 20000, seed=3, params=preset(slm, "Default")) as sim:
        sim.step(6)
        print("[noalloc] f32 640x360 -> 2048x2048 (the whole array)", flush=True)
        got, want = render_both(slm, sim, lut_bytes(1), 2048, 2048)
        assert_bitwise(got, want, "first call, 2048x2048")
        sim.step(3)
        print("[noalloc] f32 640x360 -> 333x222", flush=True)
        got, want = render_both(slm, sim, lut_bytes(2), 333, 222)
        assert_bitwise(got, want, "second call, 333x222")
    with slm.Simulation(300, 500, 5000, seed=4, params=preset(slm, "Mesh"), storage="bf16") as b:
        b.step(5)
        print("[noalloc] bf16 300x500 -> 1920x1080", flush=True)
        got, want = render_both(slm, b, lut_bytes(3), 1920, 1080)
        assert_bitwise(got, want, "bf16 handle, 1920x1080")
        got, want = render_both(slm, b, lut_bytes(4), 1, 1)
        assert_bitwise(got, want, "bf16 handle, 1x1")


def test_noalloc_render_refuses_textures_beyond_its_array(slm):
    L = slm.lib()
    with slm.Simulation(64, 64, 100, seed=1) as sim:
        buf = (ctypes.c_uint8 * 768).from_buffer_copy(lut_bytes(5))
        out = np.empty(MAX_BYTES + 4, np.uint8)
        assert L.slm_render_rgba8_noalloc(sim.handle, buf, 2048, 2049, out.ctypes.data) == -1
        assert "SLM_RENDER_NOALLOC_MAX_BYTES" in L.slm_last_error().decode()
        assert L.slm_render_rgba8_noalloc(sim.handle, None, 4, 4, out.ctypes.data) == -1
        assert L.slm_render_rgba8_noalloc(None, buf, 4, 4, out.ctypes.data) == -1
        # the limit itself is accepted
        assert L.slm_render_rgba8_noalloc(sim.handle, buf, 4096, 1024, out.ctypes.data) == 0
