"""Pins for the oracle's three passes, against closed forms, invariants and
brute force that do not share its code.

sample (compute.wgsl:7-29): constant fields, exact planes, the non-toroidal edge.
agent (compute.wgsl:57-145): the four turn cases, speed clamp, exactly-once
kinematics in a static field, the jitter's range, the deposit cell.
trail (compute.wgsl:140,147-195): decay closed form, identity at rate 0, the
1/9 impulse response, constant fixed point, mass conservation, and a numpy
np.roll convolution in float64.
"""
import math

import numpy as np
import pytest

DEFAULT = dict(decay_factor=10.0, agent_jitter=0.0, agent_speed_min=30.0, agent_speed_max=50.0,
               agent_turn_speed=0.43, agent_sensor_angle=0.3, agent_sensor_distance=20.0,
               diffusion_rate=1.0, pheromone_deposition_amount=1.0, blur_radius=2.0,
               blur_sigma=1.0)


def params(**kw):
    d = dict(DEFAULT)
    d.update(kw)
    return d


# ---- bilinear sample ------------------------------------------------------------

def test_sample_constant_field(orc):
    t = np.full((32, 48), 0.375, np.float32)
    for px, py in [(0.0, 0.0), (10.3, 7.9), (46.99, 30.5), (3.5, 30.999)]:
        assert orc.sample(t, px, py) == pytest.approx(0.375, abs=1e-7)


def test_sample_plane_is_exact_at_dyadic_points(orc):
    h, w = 16, 20
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    t = (0.03125 * xx + 0.0625 * yy + 0.125).astype(np.float32)  # exactly representable
    for px, py in [(2.5, 3.25), (7.75, 1.125), (0.0, 0.0), (18.5, 14.5)]:
        assert orc.sample(t, px, py) == 0.03125 * px + 0.0625 * py + 0.125


def test_sample_edges_return_zero(orc):
    t = np.ones((8, 8), np.float32)
    # x1 = x0 + 1 must be < width: the last column and row are never interpolated
    for px, py in [(7.0, 3.0), (7.5, 3.0), (3.0, 7.2), (-0.01, 3.0), (3.0, -0.5), (8.0, 8.0),
                   (-300.0, 2.0)]:
        assert orc.sample(t, px, py) == 0.0
    assert orc.sample(t, 6.999, 6.999) == pytest.approx(1.0)
    one = np.ones((1, 1), np.float32)
    assert orc.sample(one, 0.0, 0.0) == 0.0


# ---- agent update -------------------------------------------------------------------

def _field_with(values, w=64, h=64):
    t = np.zeros((h, w), np.float32)
    for (x, y), v in values.items():
        t[y, x] = v
    return t


def test_turn_cases(orc):
    w = h = 64
    x, y = 30.0, 30.0
    angle = 1.0
    p = params(agent_sensor_angle=0.3, agent_sensor_distance=10.0, agent_turn_speed=0.43,
               agent_speed_min=0.0, agent_speed_max=0.0)
    sl, cl = orc.sincos(np.float32(angle) - np.float32(0.3))
    sr, cr = orc.sincos(np.float32(angle) + np.float32(0.3))
    sc, cc = orc.sincos(angle)
    def cell(s, c):
        return (int(math.floor(x + 10.0 * c)), int(math.floor(y + 10.0 * s)))
    L, R, C = cell(sl, cl), cell(sr, cr), cell(sc, cc)
    big = lambda cells: {k: 1.0 for k in cells}  # noqa: E731 (2x2 blocks around a tap)
    def around(c):
        return [(c[0], c[1]), (c[0] + 1, c[1]), (c[0], c[1] + 1), (c[0] + 1, c[1] + 1)]
    tau = np.float32(6.28318530718)
    a32 = np.float32(angle)
    left_turn = np.float32(a32 + np.float32(-min(np.float32(0.43), abs(np.float32((a32 - tau) - a32)))))
    right_turn = np.float32(a32 + np.float32(min(np.float32(0.43), abs(np.float32((a32 + tau) - a32)))))
    # left strongest -> turn by -turn_speed
    a, _ = orc.agent_update(_field_with(big(around(L))), p, 0, [x, y, angle, 0.0])
    assert a[2] == left_turn
    # right strongest -> +turn_speed
    a, _ = orc.agent_update(_field_with(big(around(R))), p, 0, [x, y, angle, 0.0])
    assert a[2] == right_turn
    # centre strongest -> straight
    a, _ = orc.agent_update(_field_with(big(around(C))), p, 0, [x, y, angle, 0.0])
    assert a[2] == np.float32(angle)
    # all equal (empty field) -> straight, no random turn in the tie case
    a, _ = orc.agent_update(np.zeros((h, w), np.float32), p, 0, [x, y, angle, 0.0])
    assert a[2] == np.float32(angle)


def test_turn_speed_above_tau_is_limited_by_diff(orc):
    # |d| = |(a - TAU) - a| ~ TAU bounds the step: a turn of 10 rad only moves ~TAU
    p = params(agent_turn_speed=10.0, agent_speed_min=0.0, agent_speed_max=0.0,
               agent_sensor_distance=5.0)
    t = np.zeros((64, 64), np.float32)
    t[:, :] = 0.0
    # left sensor region bright: the sensors at angle 0 +- 0.3 point right, left one is y-negative
    s, c = orc.sincos(np.float32(0.0) - np.float32(0.3))
    lx, ly = int(math.floor(30 + 5 * c)), int(math.floor(30 + 5 * s))
    t[ly:ly + 2, lx:lx + 2] = 1.0
    a, _ = orc.agent_update(t, p, 0, [30.0, 30.0, 0.0, 0.0])
    tau = np.float32(6.28318530718)
    d = np.float32(np.float32(np.float32(0.0) - tau) - np.float32(0.0))
    turned = np.float32(np.float32(0.0) + np.float32(-abs(d)))
    expect = np.float32(turned - np.float32(tau * np.trunc(np.float32(turned / tau))))
    if expect < 0:
        expect = np.float32(expect + tau)
    assert a[2] == expect


def test_speed_clamped_and_written(orc):
    p = params(agent_speed_min=30.0, agent_speed_max=50.0)
    t = np.zeros((64, 64), np.float32)
    for v, want in [(10.0, 30.0), (70.0, 50.0), (40.0, 40.0)]:
        a, _ = orc.agent_update(t, p, 0, [10.0, 10.0, 0.0, v])
        assert a[3] == want
        # the move uses the clamped speed: angle 0 -> x advances speed*0.016
        assert a[0] == np.float32(np.float32(10.0) + np.float32(np.float32(want) * np.float32(0.016)))


def test_static_field_kinematics_exactly_once(orc):
    """Zero trail, no deposit, no jitter: heading fixed, x_{t+1} = wrap(x_t + m*cos),
    which numpy float32 reproduces from the pinned sin/cos alone."""
    w, h, n = 37, 23, 500
    p = params(pheromone_deposition_amount=0.0, agent_jitter=0.0)
    ag = orc.init_agents(7, w, h, n, 30.0, 50.0)
    ref = ag.copy()
    trail = np.zeros((h, w), np.float32)
    steps = 25
    orc.step(ag, trail, p, steps)
    assert np.all(trail == 0.0)
    for i in range(n):
        x, y, a, v = (np.float32(q) for q in ref[i])
        s, c = orc.sincos(float(a))
        m = np.float32(v * np.float32(0.016))
        for _ in range(steps):
            x = np.float32(x + np.float32(m * np.float32(c)))
            y = np.float32(y + np.float32(m * np.float32(s)))
            x = np.float32(x - np.float32(np.float32(w) * np.trunc(np.float32(x / np.float32(w)))))
            if x < 0:
                x = np.float32(x + np.float32(w))
            y = np.float32(y - np.float32(np.float32(h) * np.trunc(np.float32(y / np.float32(h)))))
            if y < 0:
                y = np.float32(y + np.float32(h))
        assert (ag[i, 0], ag[i, 1], ag[i, 2]) == (x, y, a), i


def test_jitter_range_and_zero(orc):
    t = np.zeros((64, 64), np.float32)
    deltas = []
    for i in range(3000):
        a, _ = orc.agent_update(t, params(agent_jitter=0.5, agent_speed_min=0.0,
                                          agent_speed_max=0.0), i, [20.0, 20.0, 3.0, 0.0])
        deltas.append(float(a[2]) - 3.0)
        b, _ = orc.agent_update(t, params(agent_jitter=0.0, agent_speed_min=0.0,
                                          agent_speed_max=0.0), i, [20.0, 20.0, 3.0, 0.0])
        assert b[2] == np.float32(3.0)
    deltas = np.array(deltas)
    assert deltas.min() >= -0.5 - 1e-6 and deltas.max() <= 0.5 + 1e-6
    assert abs(deltas.mean()) < 0.03 and deltas.std() == pytest.approx(0.5 / math.sqrt(3), rel=0.1)


def test_jitter_hash_matches_double_precision_closed_form(orc):
    """compute.wgsl:115-118: r = fract(sin(f32(i)*12.9898 + x*78.233 + y*37.719) * 43758.5453)
    with the pre-move position and the agent's index.  The argument is formed in
    f32 (left to right, each product rounded) and sin is taken in double here;
    the oracle's f32 sin is within 2 ulp, which the x43758 scale turns into at
    most ~5e-3, plus the f32 product's own ulp (<= 2^-8 near 4e4): so
    |r - r_ref| <= 0.02 (circular).  A wrong constant, operand, index or
    post-move position gives unrelated values and fails."""
    f = np.float32
    t = np.zeros((4096, 4096), np.float32)
    rng = np.random.default_rng(12)
    J = f(0.5)
    worst = 0.0
    for _ in range(400):
        i = int(rng.integers(0, 1 << 26))
        x, y = f(rng.uniform(0, 4096)), f(rng.uniform(0, 4096))
        a, _ = orc.agent_update(t, params(agent_jitter=float(J), agent_speed_min=0.0,
                                          agent_speed_max=0.0), i, [x, y, 3.0, 0.0])
        r_got = (float(a[2]) - 3.0) / float(J) * 0.5 + 0.5
        arg = f(f(f(i) * f(12.9898)) + f(x * f(78.233))) + f(y * f(37.719))
        p = f(f(math.sin(float(arg))) * f(43758.5453))
        r_ref = float(p - np.floor(p))
        d = abs(r_got - r_ref)
        worst = max(worst, min(d, 1.0 - d))
    assert worst <= 0.02, worst


def test_deposit_cell_is_truncated_position(orc):
    t = np.zeros((16, 16), np.float32)
    a, d = orc.agent_update(t, params(agent_speed_min=0.0, agent_speed_max=0.0), 0,
                            [5.7, 9.2, 0.0, 0.0])
    assert d == 9 * 16 + 5
    a, d = orc.agent_update(t, params(agent_speed_min=0.0, agent_speed_max=0.0), 0,
                            [15.99, 0.0, 0.0, 0.0])
    assert d == 15


# ---- trail pass ------------------------------------------------------------------------

def test_decay_closed_form_at_rate_zero(orc):
    rng = np.random.default_rng(5)
    t = rng.uniform(0, 1, (19, 23)).astype(np.float32)
    out = orc.trail_pass(t, params(diffusion_rate=0.0, decay_factor=10.0))
    k = np.float32(np.float32(10.0) * np.float32(0.001))
    assert np.array_equal(out, np.maximum(t - k, np.float32(0)).astype(np.float32))


def test_rate_is_clamped(orc):
    rng = np.random.default_rng(6)
    t = rng.uniform(0, 1, (9, 11)).astype(np.float32)
    assert np.array_equal(orc.trail_pass(t, params(diffusion_rate=2.0)),
                          orc.trail_pass(t, params(diffusion_rate=1.0)))
    assert np.array_equal(orc.trail_pass(t, params(diffusion_rate=-1.0)),
                          orc.trail_pass(t, params(diffusion_rate=0.0)))


def test_impulse_response_is_one_ninth(orc):
    t = np.zeros((7, 9), np.float32)
    t[3, 4] = 1.0
    out = orc.trail_pass(t, params(decay_factor=0.0, diffusion_rate=1.0))
    want = np.zeros_like(t)
    want[2:5, 3:6] = np.float32(1.0) / np.float32(9.0)
    assert np.array_equal(out, want)
    # toroidal: an impulse in the corner spreads to the opposite edges
    t = np.zeros((5, 6), np.float32)
    t[0, 0] = 1.0
    out = orc.trail_pass(t, params(decay_factor=0.0))
    nz = {(int(y), int(x)) for y, x in zip(*np.nonzero(out))}
    assert nz == {(y, x) for y in (4, 0, 1) for x in (5, 0, 1)}


def test_constant_field_fixed_point(orc):
    t = np.full((12, 10), 0.5, np.float32)
    out = orc.trail_pass(t, params(decay_factor=0.0, diffusion_rate=0.37))
    assert np.allclose(out, 0.5, atol=2e-7)


def test_matches_float64_roll_convolution(orc):
    rng = np.random.default_rng(8)
    for (h, w) in [(16, 16), (5, 13), (1, 7), (3, 1), (2, 2)]:
        t = rng.uniform(0, 1, (h, w)).astype(np.float32)
        rate = 0.6
        out = orc.trail_pass(t, params(decay_factor=0.0, diffusion_rate=rate))
        t64 = t.astype(np.float64)
        box = sum(np.roll(np.roll(t64, dy, 0), dx, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1))
        want = t64 * (1 - rate) + box / 9.0 * rate
        assert np.max(np.abs(out - want)) <= 4 * np.spacing(np.float32(1.0)), (h, w)


def test_mass_conserved_without_decay_or_deposit(orc):
    rng = np.random.default_rng(9)
    t = rng.uniform(0, 1, (64, 48)).astype(np.float32)
    out = orc.trail_pass(t, params(decay_factor=0.0, diffusion_rate=0.8))
    assert orc.total_mass(out) == pytest.approx(orc.total_mass(t), rel=1e-6)
    # f64 mode: conserved to rounding of doubles
    ag = np.zeros((0, 4), np.float64)
    t64 = t.astype(np.float64)
    m0 = t64.sum()
    orc.step_f64(ag, t64, params(decay_factor=0.0, diffusion_rate=0.8), 10)
    assert abs(t64.sum() - m0) <= 1e-12 * m0


def test_deposit_saturates_and_collisions_are_idempotent(orc):
    t = np.full((6, 6), 0.25, np.float32)
    mark = np.zeros((6, 6), np.uint8)
    mark[2, 3] = 1
    out = orc.trail_pass(t, params(decay_factor=0.0, diffusion_rate=0.0), mark=mark)
    assert out[2, 3] == 1.0 and out[0, 0] == np.float32(0.25)
    p = params(pheromone_deposition_amount=0.5, decay_factor=0.0, diffusion_rate=0.0)
    out = orc.trail_pass(t, p, mark=mark)
    assert out[2, 3] == np.float32(0.75)
    # k agents landing in one cell deposit once (S3): a full step with 50 agents
    # parked on cell (2,3) gives the same trail as one agent there
    p = params(agent_speed_min=0.0, agent_speed_max=0.0, decay_factor=0.0, diffusion_rate=0.0,
               pheromone_deposition_amount=0.5)
    one = np.array([[3.5, 2.5, 0.0, 0.0]], np.float32)
    many = np.repeat(one, 50, axis=0)
    t1, t2 = t.copy(), t.copy()
    orc.step(one, t1, p, 1)
    orc.step(many, t2, p, 1)
    assert np.array_equal(t1, t2) and t1[2, 3] == np.float32(0.75)


def test_f32_and_f64_oracles_agree_for_a_few_steps(orc):
    """Same algorithm in two precisions: catches a wrong index or term in either."""
    w, h, n = 96, 80, 2000
    p = params()
    a32 = orc.init_agents(11, w, h, n, 30.0, 50.0)
    a64 = a32.astype(np.float64)
    t32 = np.zeros((h, w), np.float32)
    t64 = np.zeros((h, w), np.float64)
    orc.step(a32, t32, p, 4)
    orc.step_f64(a64, t64, p, 4)
    dpos = np.abs(a32[:, :2] - a64[:, :2]).max(axis=1)
    assert np.mean(dpos < 1e-3) > 0.99
    assert np.abs(t32 - t64).max() < 1e-3 or np.mean(np.abs(t32 - t64) < 1e-4) > 0.99
    assert orc.total_mass(t32) == pytest.approx(orc.total_mass(t64), rel=1e-3)


def test_storage_rounding_modes(orc):
    rng = np.random.default_rng(10)
    t = rng.uniform(0, 1, (8, 8)).astype(np.float32)
    for st in ("bf16", "f16"):
        tt = orc.round_storage(t, st)
        out = orc.trail_pass(tt, params(), storage=st)
        f32out = orc.trail_pass(tt, params(), storage="f32")
        assert np.array_equal(out, orc.round_storage(f32out, st))


def test_f64_trail_pass_matches_roll_convolution(orc):
    rng = np.random.default_rng(16)
    t = rng.uniform(0, 1, (24, 17))
    mark = (rng.uniform(0, 1, t.shape) < 0.1).astype(np.uint8)
    p = params(decay_factor=2.0, diffusion_rate=0.75, pheromone_deposition_amount=0.5)
    got = orc.trail_pass_f64(t, p, mark)
    k = float(np.float32(2.0)) * 0.001
    t1 = np.where(mark == 1, np.clip(t + 0.5, 0.0, 1.0), t)
    t2 = np.maximum(t1 - k, 0.0)
    box = sum(np.roll(np.roll(t2, dy, 0), dx, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    assert np.abs(got - (t2 * 0.25 + box / 9.0 * 0.75)).max() <= 1e-15
