"""Pins for the oracle's Gaussian diffuse (the library's opt-in extension of
the 3x3 box; the reference carries blur_radius / blur_sigma, src/main.rs:43-44,
but never reads them), against math that does not share its code: the
weights' closed form, the impulse response (an outer product of the weights),
identity at rate 0, a constant field, mass conservation, and a float64 np.roll
convolution with the separable kernel, on grids smaller than the kernel too.
"""
import math

import numpy as np
import pytest

DEFAULT = dict(decay_factor=0.0, agent_jitter=0.0, agent_speed_min=30.0, agent_speed_max=50.0,
               agent_turn_speed=0.43, agent_sensor_angle=0.3, agent_sensor_distance=20.0,
               diffusion_rate=1.0, pheromone_deposition_amount=1.0, blur_radius=2.0,
               blur_sigma=1.0)


def params(**kw):
    d = dict(DEFAULT)
    d.update(kw)
    return d


@pytest.mark.parametrize("radius,sigma", [(1.0, 0.5), (2.0, 1.0), (2.4, 3.0), (5.0, 2.0),
                                          (8.0, 4.0), (7.6, 0.7)])
def test_weights_closed_form(orc, radius, sigma):
    r, w = orc.gauss_weights(radius, sigma)
    assert r == int(math.floor(radius + 0.5))
    sg = float(np.float32(sigma))  # the parameter is an f32 field
    e = [math.exp(-i * i / (2.0 * sg * sg)) for i in range(r + 1)]
    s = e[0] + 2.0 * sum(e[1:])
    assert list(w) == [np.float32(x / s) for x in e]
    assert abs(float(w[0]) + 2.0 * float(w[1:].astype(np.float64).sum()) - 1.0) < 1e-6


@pytest.mark.parametrize("radius,sigma", [(0.0, 1.0), (0.49, 1.0), (8.5, 1.0), (2.0, 0.0),
                                          (2.0, -1.0), (float("nan"), 1.0), (2.0, float("inf"))])
def test_weights_reject_invalid(orc, radius, sigma):
    assert orc.gauss_weights(radius, sigma)[0] == 0


def test_impulse_response_is_the_weight_outer_product(orc):
    h, w = 21, 23
    t = np.zeros((h, w), np.float32)
    t[10, 11] = 1.0
    for radius, sigma in [(1.0, 0.8), (2.0, 1.0), (4.0, 2.5)]:
        r, wt = orc.gauss_weights(radius, sigma)
        out = orc.trail_pass(t, params(blur_radius=radius, blur_sigma=sigma), gaussian=True)
        want = np.zeros((h, w), np.float32)
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                want[10 + dy, 11 + dx] = np.float32(wt[abs(dy)]) * np.float32(wt[abs(dx)])
        assert np.array_equal(out, want)


def test_rate_zero_is_decayed_identity(orc):
    rng = np.random.default_rng(5)
    t = rng.uniform(0.0, 1.0, (17, 19)).astype(np.float32)
    out = orc.trail_pass(t, params(diffusion_rate=0.0, decay_factor=30.0), gaussian=True)
    assert np.array_equal(out, np.maximum(t - np.float32(30.0) * np.float32(0.001), 0))


def test_constant_field_and_mass(orc):
    t = np.full((40, 36), 0.625, np.float32)
    out = orc.trail_pass(t, params(blur_radius=3.0, blur_sigma=1.7), gaussian=True)
    assert np.abs(out - 0.625).max() < 1e-6
    rng = np.random.default_rng(6)
    t = rng.uniform(0.0, 1.0, (64, 48)).astype(np.float32)
    out = orc.trail_pass(t, params(blur_radius=5.0, blur_sigma=2.0), gaussian=True)
    assert out.astype(np.float64).sum() == pytest.approx(t.astype(np.float64).sum(), rel=2e-6)


@pytest.mark.parametrize("h,w,radius,sigma,rate", [(30, 26, 2.0, 1.0, 1.0), (9, 40, 3.0, 2.0, 0.35),
                                                   (3, 5, 4.0, 1.5, 0.8), (1, 1, 2.0, 1.0, 1.0),
                                                   (2, 17, 8.0, 3.0, 1.0)])
def test_matches_float64_roll_convolution(orc, h, w, radius, sigma, rate):
    rng = np.random.default_rng(h * 100 + w)
    t = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32)
    mark = (rng.random((h, w)) < 0.1).astype(np.uint8)
    p = params(blur_radius=radius, blur_sigma=sigma, diffusion_rate=rate, decay_factor=12.0)
    out = orc.trail_pass(t, p, mark=mark, gaussian=True)
    r, wt = orc.gauss_weights(radius, sigma)
    t2 = np.where(mark == 1, np.clip(t.astype(np.float64) + 1.0, 0, 1), t.astype(np.float64))
    t2 = np.maximum(t2 - 0.012, 0.0)
    hs = sum(float(wt[abs(i)]) * np.roll(t2, -i, axis=1) for i in range(-r, r + 1))
    g = sum(float(wt[abs(j)]) * np.roll(hs, -j, axis=0) for j in range(-r, r + 1))
    want = t2 * (1.0 - rate) + g * rate
    assert np.abs(out - want).max() < 2e-6


def test_steps_use_the_gaussian_only_when_asked(orc):
    rng = np.random.default_rng(8)
    h, w = 24, 28
    a0 = orc.init_agents(3, w, h, 200, 30.0, 50.0)
    t0 = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32)
    p = params(decay_factor=10.0, blur_radius=3.0, blur_sigma=1.2)
    a1, t1 = a0.copy(), t0.copy()
    orc.step(a1, t1, p, 1, gaussian=True)
    a2, t2 = a0.copy(), t0.copy()
    orc.step(a2, t2, p, 1)
    assert np.array_equal(a1, a2)  # same agent pass
    assert not np.array_equal(t1, t2)
    mark = orc.deposit_marks(a1, w, h)
    assert np.array_equal(t1, orc.trail_pass(t0, p, mark=mark, gaussian=True))
    assert np.array_equal(t2, orc.trail_pass(t0, p, mark=mark))
