"""Derive the constants used by slm_wgsl.h from mathematics (mpmath, 300 digits).

Run: python gen_wgsl_constants.py  -> prints the constant block pasted into slm_wgsl.h.
Nothing here is taken from another math library: the bits of 2/pi and pi/2 are
computed by mpmath, the polynomial coefficients are the Taylor coefficients
(-1)^k/(2k+1)! and (-1)^k/(2k)! rounded to the nearest float.
"""
import struct
import mpmath as mp

mp.mp.dps = 300


def f32_bits(v):
    return struct.unpack("<I", struct.pack("<f", float(v)))[0]


def f32(v):
    return struct.unpack("<f", struct.pack("<f", float(v)))[0]


def f64_bits(v):
    return struct.unpack("<Q", struct.pack("<d", float(v)))[0]


def main():
    two_over_pi = 2 / mp.pi
    # 2/pi < 1: emit 8 words of 32 fractional bits, bits 1..256 after the binary point.
    frac = two_over_pi
    words = []
    for _ in range(8):
        frac *= 2 ** 32
        w = int(mp.floor(frac))
        words.append(w)
        frac -= w
    print("/* bits 1..256 of 2/pi after the binary point, 32 per word, most significant first */")
    print("static const uint32_t SLM_TWO_OVER_PI_BITS[8] = {")
    print("    " + ", ".join("0x%08xu" % w for w in words) + "};")

    pio2 = mp.pi / 2
    c1 = f32(pio2)
    c2 = f32(pio2 - c1)
    c3 = f32(pio2 - c1 - c2)
    print("/* pi/2 = C1 + C2 + C3 (each the nearest float to the remainder) */")
    for n, v in (("SLM_PIO2_C1", c1), ("SLM_PIO2_C2", c2), ("SLM_PIO2_C3", c3)):
        print("#define %s slm_f32_from_bits(0x%08xu) /* %.9e */" % (n, f32_bits(v), v))
    print("#define SLM_TWO_OVER_PI_F slm_f32_from_bits(0x%08xu) /* %.9e */"
          % (f32_bits(two_over_pi), f32(two_over_pi)))
    scaled = float(pio2 * mp.mpf(2) ** -53)
    print("#define SLM_PIO2_X2M53_D slm_f64_from_bits(0x%016xull) /* pi/2 * 2^-53 */" % f64_bits(scaled))

    print("/* Taylor coefficients of sin: (-1)^k/(2k+1)!, k = 1..4 */")
    for k in range(1, 5):
        v = mp.mpf(-1) ** k / mp.factorial(2 * k + 1)
        print("#define SLM_SIN_T%d slm_f32_from_bits(0x%08xu) /* %.9e */" % (k, f32_bits(v), f32(v)))
    print("/* Taylor coefficients of cos: (-1)^k/(2k)!, k = 1..5 */")
    for k in range(1, 6):
        v = mp.mpf(-1) ** k / mp.factorial(2 * k)
        print("#define SLM_COS_T%d slm_f32_from_bits(0x%08xu) /* %.9e */" % (k, f32_bits(v), f32(v)))


if __name__ == "__main__":
    main()
