"""Pins for the oracle's display colour map (src/display.wgsl:29-87) against
closed forms: identity and integer-scale fits, letter- and pillar-boxing, the
LUT channel layout (R[256] G[256] B[256], src/lut_manager.rs:174-176) and the
intensity -> entry map u32(clamp(T, 0, 1) * 255)."""
import numpy as np


def gray_lut():
    r = np.arange(256, dtype=np.uint8)
    return np.concatenate([r, r, r]).tobytes()


def distinct_lut():
    i = np.arange(256)
    return np.concatenate([i, 255 - i, (7 * i) & 255]).astype(np.uint8).tobytes()


def test_identity_fit_maps_pixel_to_cell(orc):
    rng = np.random.default_rng(1)
    t = rng.uniform(0, 1, (9, 16)).astype(np.float32)
    img = orc.render(t, gray_lut(), 16, 9)
    k = (t * np.float32(255.0)).astype(np.uint32)
    assert np.array_equal(img[..., 0], k) and np.array_equal(img[..., 1], k)
    assert np.array_equal(img[..., 2], k) and np.all(img[..., 3] == 255)


def test_integer_upscale_repeats_cells(orc):
    rng = np.random.default_rng(2)
    t = rng.uniform(0, 1, (5, 7)).astype(np.float32)
    img = orc.render(t, distinct_lut(), 21, 15)  # scale 3 exactly
    small = orc.render(t, distinct_lut(), 7, 5)
    assert np.array_equal(img, np.repeat(np.repeat(small, 3, axis=0), 3, axis=1))


def test_letterbox_and_pillarbox_are_black(orc):
    t = np.full((4, 4), 0.5, np.float32)
    wide = orc.render(t, gray_lut(), 12, 4)  # texture wider: fit height, bars left/right
    assert np.all(wide[:, :4, :3] == 0) and np.all(wide[:, 8:, :3] == 0)
    assert np.all(wide[:, 4:8, 0] == 127) and np.all(wide[..., 3] == 255)
    tall = orc.render(t, gray_lut(), 4, 10)  # texture taller: fit width, bars top/bottom
    assert np.all(tall[:3, :, :3] == 0) and np.all(tall[7:, :, :3] == 0)
    assert np.all(tall[3:7, :, 0] == 127)


def test_lut_channels_and_intensity_map(orc):
    vals = np.array([[0.0, 0.5, 1.0, 7.0, 0.999, 1.0 / 255.0]], np.float32)
    img = orc.render(vals, distinct_lut(), 6, 1)
    k = np.array([0, 127, 255, 255, 254, 1])
    assert np.array_equal(img[0, :, 0], k)
    assert np.array_equal(img[0, :, 1], 255 - k)
    assert np.array_equal(img[0, :, 2], (7 * k) & 255)


def test_non_integer_fit_samples_nearest_lower_cell(orc):
    t = np.arange(10, dtype=np.float32).reshape(1, 10) / 10.0
    img = orc.render(t, gray_lut(), 25, 3)  # scale 2.5 on x; the fit is width-limited
    # pixel px samples cell floor(px / 2.5) on the single grid row, which sits at
    # rows [off_y, off_y + 2.5) with off_y = (3 - 2.5) / 2 = 0.25
    row = img[1, :, 0]
    want = (t[0, (np.arange(25) / 2.5).astype(int)] * np.float32(255)).astype(np.uint8)
    assert np.array_equal(row, want)


def test_display_point_is_the_step_without_its_diffuse(orc, slm):
    """The reference displays its trail between the decay and the diffuse
    (src/main.rs:1183-1217).  parity_util.display_point restates that point
    as the trail pass at diffusion rate 0 over the moved agents' cells; with
    the rate already 0 it must equal the whole oracle step's trail, and with
    decay 0 too it is the old trail with exactly the agents' truncated cells
    raised by the deposit."""
    from parity_util import display_point, preset

    rng = np.random.default_rng(11)
    w, h, n = 97, 61, 700
    trail = rng.uniform(0, 1, (h, w)).astype(np.float32)
    for name in ("Default", "Firecracker Trees", "Snake"):
        s = preset(slm, name).clone()
        s.pheromone_diffusion_rate = 0.0
        a = orc.init_agents(5, w, h, n, s.agent_speed_min, s.agent_speed_max)
        a1, t1 = a.copy(), trail.copy()
        shown = display_point(orc, a1, t1, s)
        a2, t2 = a.copy(), trail.copy()
        orc.step(a2, t2, s, 1)
        assert np.array_equal(shown, t2) and np.array_equal(a1, a2), name
        s.pheromone_decay_factor = 0.0
        a3, t3 = a.copy(), trail.copy()
        shown = display_point(orc, a3, t3, s)
        want = trail.copy()
        xs = np.floor(a3[:, 0]).astype(np.int64)
        ys = np.floor(a3[:, 1]).astype(np.int64)
        ok = (xs < w) & (ys < h)
        want[ys[ok], xs[ok]] = np.minimum(want[ys[ok], xs[ok]] + np.float32(1.0), np.float32(1.0))
        assert np.array_equal(shown, want), name
