"""Pins for the oracle's primitives, each against something independent of it.

sin/cos: libm's double sin/cos and mpmath (error bound, symmetry, specials).
Storage rounding: numpy's float16 cast and torch's bfloat16 cast (both RNE).
SplitMix64: the generator's published first outputs for seed 0.
WGSL '%', mix, fract, sign, i32(): special cases whose answers the definitions fix.
"""
import math
import struct

import numpy as np
import pytest


def f32(x):
    return struct.unpack("<f", struct.pack("<f", x))[0]


def ulp(x):
    x = abs(f32(x))
    return float(np.spacing(np.float32(x))) if x > 0 else 1.4e-45


def test_sincos_accuracy_dense_small_range(orc):
    # Dense sample of the range the sensors and moves use; the exhaustive run over
    # every float measured max 1.43 ulp on [0, 8) and 1.80 ulp overall.
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.linspace(-8.0, 8.0, 20001, dtype=np.float32),
                         rng.uniform(-20, 20, 20000).astype(np.float32)])
    worst = 0.0
    for x in xs:
        s, c = orc.sincos(float(x))
        rs, rc = math.sin(float(x)), math.cos(float(x))
        worst = max(worst, abs(s - rs) / ulp(rs), abs(c - rc) / ulp(rc))
    assert worst <= 1.5, worst


def test_sincos_large_arguments_vs_mpmath(orc):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 60
    rng = np.random.default_rng(2)
    xs = [4096.0, 4096.5, 1e5, 3.4906585e9, 2.0 ** 31, 2.0 ** 32 - 256, 1e20, 3.4e38]
    xs += list(rng.uniform(4096, 4.3e9, 300).astype(np.float32))
    for x in xs:
        x = f32(float(x))
        s, c = orc.sincos(x)
        rs, rc = float(mp.sin(mp.mpf(x))), float(mp.cos(mp.mpf(x)))
        assert abs(s - rs) <= 2.0 * ulp(rs), (x, s, rs)
        assert abs(c - rc) <= 2.0 * ulp(rc), (x, c, rc)


def test_sincos_symmetry_and_specials(orc):
    for x in [0.1, 1.0, 2.5, 7.9, 1234.5, 5e6, 3e9]:
        s, c = orc.sincos(x)
        ns, nc = orc.sincos(-x)
        assert ns == -s and nc == c
    s, c = orc.sincos(0.0)
    assert s == 0.0 and c == 1.0
    s, _ = orc.sincos(-0.0)
    assert math.copysign(1.0, s) == -1.0
    for x in (math.inf, -math.inf, math.nan):
        s, c = orc.sincos(x)
        assert math.isnan(s) and math.isnan(c)
    # quadrant boundaries: sin(pi/2) = cos(0) = 1, cos(pi) = -1
    assert orc.sincos(f32(math.pi / 2))[0] == 1.0
    assert orc.sincos(f32(math.pi))[1] == -1.0


def test_wgsl_mod_special_cases(orc):
    L = orc.lib()
    tau = f32(6.28318530718)
    assert L.orc_wgsl_mod(1.0, tau) == 1.0
    assert L.orc_wgsl_mod(tau + 1.0, tau) == f32(f32(tau + 1.0) - tau)  # exact by Sterbenz
    assert L.orc_wgsl_mod(-1.0, tau) == -1.0                           # truncated, keeps sign
    assert L.orc_wgsl_mod(7.5, 2.0) == 1.5
    assert L.orc_wgsl_mod(-7.5, 2.0) == -1.5
    # just below b: a/b stays below 1 (ulp(a)/b > 2^-25 for every float a < b),
    # so trunc gives 0 and the result is a itself
    for b in (tau, 1024.0, 1600.0, 900.0, 3.0):
        below = float(np.nextafter(np.float32(b), np.float32(0)))
        assert L.orc_wgsl_mod(below, b) == below
    # wrap of a position exactly at the width gives 0
    assert L.orc_wgsl_mod(1024.0, 1024.0) == 0.0


def test_mix_fract_sign_i32(orc):
    L = orc.lib()
    assert L.orc_mix(2.0, 6.0, 0.25) == 3.0
    assert L.orc_mix(0.7, 0.2, 0.0) == f32(0.7)
    assert L.orc_mix(0.7, 0.2, 1.0) == f32(0.2)
    assert L.orc_fract(2.75) == 0.75
    assert L.orc_fract(-0.25) == 0.75
    assert L.orc_sign(3.0) == 1.0 and L.orc_sign(-2.0) == -1.0 and L.orc_sign(0.0) == 0.0
    assert L.orc_f32_to_i32(3.9) == 3 and L.orc_f32_to_i32(-3.9) == -3
    assert L.orc_f32_to_i32(3e10) == 2 ** 31 - 1 and L.orc_f32_to_i32(-3e10) == -2 ** 31
    assert L.orc_f32_to_i32(math.nan) == 0


def test_f16_rounding_matches_numpy(orc):
    L = orc.lib()
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.uniform(0, 1, 20000), rng.uniform(0, 1e-4, 5000),
                           rng.uniform(0, 7e4, 2000), [0.0, 6.1e-5, 5.96e-8, 2.98e-8, 65504.0,
                                                       65519.99, 65520.0, 1.0, 0.5]])
    vals = vals.astype(np.float32)
    for v in vals:
        want = np.float16(v).view(np.uint16)
        assert L.orc_f32_to_f16(float(v)) == int(want), v
        assert L.orc_f16_to_f32(int(want)) == float(np.float16(v).astype(np.float32))


def test_bf16_rounding_matches_torch(orc):
    torch = pytest.importorskip("torch")
    L = orc.lib()
    rng = np.random.default_rng(4)
    vals = np.concatenate([rng.uniform(0, 1, 20000), rng.uniform(0, 1e-30, 2000),
                           [0.0, 1.0, 0.00390625, 1.00390625, 1.01171875, 3e38]]).astype(np.float32)
    t = torch.from_numpy(vals).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    for v, want in zip(vals, t):
        assert L.orc_f32_to_bf16(float(v)) == int(want), v
        assert L.orc_bf16_to_f32(int(want)) == float(np.float32(struct.unpack("<f", struct.pack("<I", int(want) << 16))[0]))


def test_splitmix64_published_outputs(orc):
    L = orc.lib()
    assert [L.orc_splitmix64_at(0, n) for n in range(3)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform24_resolution_and_range(orc):
    L = orc.lib()
    us = np.array([L.orc_uniform24(42, n) for n in range(20000)])
    assert us.min() >= 0.0 and us.max() < 1.0
    assert np.all((us * 2 ** 24) == np.floor(us * 2 ** 24))  # multiples of 2^-24
    assert abs(us.mean() - 0.5) < 0.01


def test_stochastic_bf16_rounding(orc):
    """Deterministic stochastic rounding: results are the two neighbouring bf16
    values; over many cells the fraction rounded up equals the discarded
    fraction of the value (so the mean is unbiased, where round-to-nearest is
    biased by the whole discarded part), for several fractions; the random
    bits spread over [0, 2^16); exact bf16 values are kept."""
    L = orc.lib()
    cells = np.arange(1 << 16)
    for frac in (0.05, 0.3, 0.5, 0.77, 0.99):
        base = np.float32(1.0) + np.float32(frac * 2 ** -7)
        bits = int(base.view(np.uint32))
        lo, hi = bits >> 16, (bits >> 16) + 1
        got = np.array([L.orc_f32_to_bf16_sr(float(base), int(c)) for c in cells])
        assert set(np.unique(got)) <= {lo, hi}
        want_up = (bits & 0xFFFF) / 65536.0
        assert abs(float(np.mean(got == hi)) - want_up) < 0.01, frac
        vals = (got.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        assert abs(vals.mean() - float(base)) < 2e-5 * float(base), frac
    r = np.array([L.orc_sr_bits(int(c), 0x3F800000) for c in cells[:4096]])
    assert r.min() >= 0 and r.max() < 65536
    hist = np.histogram(r, bins=16, range=(0, 65536))[0]
    assert hist.min() > 4096 / 16 * 0.7 and hist.max() < 4096 / 16 * 1.3
    exact = np.float32(0.8125)  # a bf16 value
    assert all(L.orc_f32_to_bf16_sr(float(exact), c) == int(exact.view(np.uint32)) >> 16
               for c in range(1000))
    base = np.float32(1.0) + np.float32(0.3 * 2 ** -7)
    assert L.orc_f32_to_bf16(float(base)) == int(base.view(np.uint32)) >> 16  # nearest: down


def test_stochastic_rounding_neighbours_draw_independently(orc):
    """The random bits of horizontally and vertically adjacent cells (rows of
    4096 and 16384 cells) are uncorrelated and jointly uniform, for a fixed
    value and for the value a smooth field would hold in both cells; a
    plausible slip (a dropped cell term, a multiply by an even constant, the
    wrong half of the product) makes neighbours draw alike and fails here."""
    L = orc.lib()
    rng = np.random.default_rng(3)
    cells = rng.integers(0, (1 << 28) - (1 << 15), 50000)
    for bits in (0x3F800000, 0x3E4CCCCD, 0x3F3504F3):
        r0 = np.array([L.orc_sr_bits(int(c), bits) for c in cells], np.float64)
        for step in (1, 4096, 16384):
            r1 = np.array([L.orc_sr_bits(int(c) + step, bits) for c in cells], np.float64)
            rho = np.corrcoef(r0, r1)[0, 1]
            assert abs(rho) < 0.02, (hex(bits), step, rho)
            h = np.histogram2d(r0 // 4096, r1 // 4096, bins=16, range=[[0, 16], [0, 16]])[0]
            e = len(cells) / 256.0
            chi2 = float(((h - e) ** 2 / e).sum())
            assert chi2 < 255 + 5 * np.sqrt(2 * 255), (hex(bits), step, chi2)


def test_stochastic_rounding_is_unbiased_over_random_values(orc):
    """Random values in [0.25, 1), one per cell: the stochastically rounded
    values average to the values within four standard errors of the mean.
    On values whose discarded fraction lies in [0, 1/2) round-to-nearest-even
    always rounds down (mean error about -1/4 ulp, many standard errors from
    zero); stochastic rounding of the same values stays unbiased."""
    L = orc.lib()
    rng = np.random.default_rng(11)
    n = 40000

    def sr_error(vals):
        got = np.array([L.orc_f32_to_bf16_sr(float(x), k) for k, x in enumerate(vals)], np.uint32)
        return (got << 16).view(np.float32).astype(np.float64) - vals.astype(np.float64)

    def se(e):
        return float(e.std()) / np.sqrt(len(e))

    v = rng.uniform(0.25, 1.0, n).astype(np.float32)
    err = sr_error(v)
    assert abs(err.mean()) < 4 * se(err), (err.mean(), se(err))
    low = ((v.view(np.uint32) & 0xFFFF0000) |
           rng.integers(0, 0x8000, n).astype(np.uint32)).view(np.float32)
    rne = np.array([L.orc_f32_to_bf16(float(x)) for x in low], np.uint32)
    rne_err = (rne << 16).view(np.float32).astype(np.float64) - low.astype(np.float64)
    sr_low = sr_error(low)
    assert abs(sr_low.mean()) < 4 * se(sr_low), (sr_low.mean(), se(sr_low))
    assert rne_err.mean() < -50 * se(sr_low), (rne_err.mean(), se(sr_low))
