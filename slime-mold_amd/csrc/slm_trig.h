/*
 * slm_trig.h - the single-precision sin/cos used by every slime-mold kernel.
 *
 * WGSL leaves the accuracy of sin/cos to the executor (compute.wgsl:80-90 for
 * the sensors, :117 for the jitter hash, :126-127 for the move), so the bits the
 * reference produces depend on the driver.  This library fixes one definition and
 * compiles it for both the GPU kernels and the CPU test oracle, so that the two
 * agree bit for bit; its accuracy against the true sin/cos is tested separately.
 *
 * Method (textbook, no library constants):
 *   - |x| <  SLM_TRIG_SMALL_MAX: Cody-Waite reduction, k = rint(x*2/pi),
 *     r = x - k*(C1+C2+C3) with three fused multiply-adds (Cody & Waite,
 *     "Software Manual for the Elementary Functions", 1980).
 *   - larger |x| (the jitter hash reaches ~3.5e9 at 2^28 agents): Payne-Hanek
 *     reduction (Payne & Hanek, SIGNUM Newsletter 18(1), 1983) - the 24-bit
 *     mantissa times a 96-bit window of the bits of 2/pi, in integers, giving
 *     x*2/pi mod 4 with 62 fractional bits.
 *   - sin and cos of r in [-pi/4, pi/4] by their Taylor polynomials of degree 9
 *     and 10, evaluated by Horner's rule with explicit fmaf (truncation error
 *     below 0.05 ulp on that interval).
 * All constants are produced by gen_wgsl_constants.py from mpmath.
 *
 * Only IEEE-754 single/double add, multiply, fma, rint and integer operations
 * are used, so gcc on x86-64 (SSE, -ffp-contract=off) and hipcc for gfx950 give
 * identical results.  The header is plain C so the oracle can include it.
 */
#ifndef SLM_TRIG_H
#define SLM_TRIG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define SLM_TRIG_FN static inline __host__ __device__
#else
#define SLM_TRIG_FN static inline
#endif

SLM_TRIG_FN float slm_f32_from_bits(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, sizeof f);
  return f;
}
SLM_TRIG_FN uint32_t slm_f32_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, sizeof u);
  return u;
}
SLM_TRIG_FN double slm_f64_from_bits(uint64_t u) {
  double d;
  __builtin_memcpy(&d, &u, sizeof d);
  return d;
}

/* ---- constants: output of gen_wgsl_constants.py ------------------------- */
/* pi/2 = C1 + C2 + C3 (each the nearest float to the remainder) */
#define SLM_PIO2_C1 slm_f32_from_bits(0x3fc90fdbu)       /* 1.570796371e+00 */
#define SLM_PIO2_C2 slm_f32_from_bits(0xb33bbd2eu)       /* -4.371138829e-08 */
#define SLM_PIO2_C3 slm_f32_from_bits(0xa6f72cedu)       /* -1.715124510e-15 */
#define SLM_TWO_OVER_PI_F slm_f32_from_bits(0x3f22f983u) /* 6.366197467e-01 */
#define SLM_PIO2_X2M53_D slm_f64_from_bits(0x3ca921fb54442d18ull) /* pi/2 * 2^-53 */
/* Taylor coefficients of sin: (-1)^k/(2k+1)!, k = 1..4 */
#define SLM_SIN_T1 slm_f32_from_bits(0xbe2aaaabu) /* -1.666666716e-01 */
#define SLM_SIN_T2 slm_f32_from_bits(0x3c088889u) /* 8.333333768e-03 */
#define SLM_SIN_T3 slm_f32_from_bits(0xb9500d01u) /* -1.984127011e-04 */
#define SLM_SIN_T4 slm_f32_from_bits(0x3638ef1du) /* 2.755731884e-06 */
/* Taylor coefficients of cos: (-1)^k/(2k)!, k = 1..5 */
#define SLM_COS_T1 slm_f32_from_bits(0xbf000000u) /* -5.000000000e-01 */
#define SLM_COS_T2 slm_f32_from_bits(0x3d2aaaabu) /* 4.166666791e-02 */
#define SLM_COS_T3 slm_f32_from_bits(0xbab60b61u) /* -1.388888923e-03 */
#define SLM_COS_T4 slm_f32_from_bits(0x37d00d01u) /* 2.480158764e-05 */
#define SLM_COS_T5 slm_f32_from_bits(0xb493f27eu) /* -2.755731998e-07 */

/* Cody-Waite is used below this magnitude, Payne-Hanek at and above it. */
#define SLM_TRIG_SMALL_MAX 4096.0f

/*
 * Bits of 2/pi after the binary point, as a 288-bit string preceded by one zero
 * word: word 0 stands for the (zero) bits -31..0, words 1..8 hold bits 1..256.
 */
#define SLM_2OPI_W1 0xa2f9836eu
#define SLM_2OPI_W2 0x4e441529u
#define SLM_2OPI_W3 0xfc2757d1u
#define SLM_2OPI_W4 0xf534ddc0u
#define SLM_2OPI_W5 0xdb629599u
#define SLM_2OPI_W6 0x3c439041u
#define SLM_2OPI_W7 0xfe5163abu
#define SLM_2OPI_W8 0xdebbc561u

/* 32 bits of the string starting s bits into word a, b the next word
 * (0 <= s < 32): the top half of (a:b) << s. */
SLM_TRIG_FN uint32_t slm_funnel32(uint32_t a, uint32_t b, int s) {
  uint64_t pair = ((uint64_t)a << 32) | b;
  return (uint32_t)((pair << s) >> 32);
}

/*
 * Reduce a finite ax >= SLM_TRIG_SMALL_MAX: returns r in [-pi/4, pi/4] and the
 * quadrant q with ax = (4n + q)*pi/2 + r.
 */
SLM_TRIG_FN float slm_reduce_payne_hanek(float ax, int* q_out) {
  uint32_t ux = slm_f32_bits(ax);
  int e = (int)(ux >> 23) - 150;            /* ax = m * 2^e */
  uint64_t m = (ux & 0x7fffffu) | 0x800000u;
  /* Bits k <= e-2 of 2/pi contribute multiples of 4 to ax*2/pi: start the
   * 96-bit window at bit k0 = e-1, i.e. extended position p = k0 + 31
   * (19 <= p <= 134 for finite ax >= 4096, so words w .. w+3 with w <= 4). */
  int p = e + 30;
  int w = p >> 5, s = p & 31;
  /* words w .. w+3, picked by selects on w (no table in memory, no branches) */
  uint32_t a0 = 0u, a1 = SLM_2OPI_W1, a2 = SLM_2OPI_W2, a3 = SLM_2OPI_W3;
  a0 = w == 1 ? SLM_2OPI_W1 : a0;
  a1 = w == 1 ? SLM_2OPI_W2 : a1;
  a2 = w == 1 ? SLM_2OPI_W3 : a2;
  a3 = w == 1 ? SLM_2OPI_W4 : a3;
  a0 = w == 2 ? SLM_2OPI_W2 : a0;
  a1 = w == 2 ? SLM_2OPI_W3 : a1;
  a2 = w == 2 ? SLM_2OPI_W4 : a2;
  a3 = w == 2 ? SLM_2OPI_W5 : a3;
  a0 = w == 3 ? SLM_2OPI_W3 : a0;
  a1 = w == 3 ? SLM_2OPI_W4 : a1;
  a2 = w == 3 ? SLM_2OPI_W5 : a2;
  a3 = w == 3 ? SLM_2OPI_W6 : a3;
  a0 = w == 4 ? SLM_2OPI_W4 : a0;
  a1 = w == 4 ? SLM_2OPI_W5 : a1;
  a2 = w == 4 ? SLM_2OPI_W6 : a2;
  a3 = w == 4 ? SLM_2OPI_W7 : a3;
  uint64_t hi = slm_funnel32(a0, a1, s);
  uint64_t mid = slm_funnel32(a1, a2, s);
  uint64_t lo = slm_funnel32(a2, a3, s);
  /* (m * window) mod 2^96 = (ax*2/pi mod 4) in fixed point with 94 fraction bits */
  uint64_t acc = m * lo;
  acc = (acc >> 32) + m * mid;
  uint32_t w1 = (uint32_t)acc;
  acc = (acc >> 32) + m * hi;
  uint32_t w2 = (uint32_t)acc;
  uint64_t top = ((uint64_t)w2 << 32) | w1; /* 2 integer bits, 62 fraction bits */
  int q = (int)(top >> 62);
  int64_t frac = (int64_t)(top & 0x3fffffffffffffffull);
  if (frac >= ((int64_t)1 << 61)) { /* round to the nearest multiple of pi/2 */
    frac -= (int64_t)1 << 62;
    q = (q + 1) & 3;
  }
  /* frac * 2^-62 in [-1/2, 1/2); keep 53 bits so the conversion is exact */
  int64_t f53 = frac / 512;
  *q_out = q;
  return (float)((double)f53 * SLM_PIO2_X2M53_D);
}

/* sin and cos of a finite float argument (see the header comment): callers
 * whose arguments are finite by construction (the agent update) skip the
 * NaN/infinity test of slm_sincosf below; the results are the same. */
SLM_TRIG_FN void slm_sincosf_finite(float x, float* s_out, float* c_out) {
  float ax = __builtin_fabsf(x);
  float r;
  int q;
  if (ax < SLM_TRIG_SMALL_MAX) {
    float k = __builtin_rintf(ax * SLM_TWO_OVER_PI_F);
    r = __builtin_fmaf(-k, SLM_PIO2_C1, ax);
    r = __builtin_fmaf(-k, SLM_PIO2_C2, r);
    r = __builtin_fmaf(-k, SLM_PIO2_C3, r);
    q = (int)k & 3;
  } else {
    r = slm_reduce_payne_hanek(ax, &q);
  }
  float r2 = r * r;
  float ps = __builtin_fmaf(r2, SLM_SIN_T4, SLM_SIN_T3);
  ps = __builtin_fmaf(r2, ps, SLM_SIN_T2);
  ps = __builtin_fmaf(r2, ps, SLM_SIN_T1);
  float r3 = r * r2;
  float s = __builtin_fmaf(r3, ps, r);
  float pc = __builtin_fmaf(r2, SLM_COS_T5, SLM_COS_T4);
  pc = __builtin_fmaf(r2, pc, SLM_COS_T3);
  pc = __builtin_fmaf(r2, pc, SLM_COS_T2);
  pc = __builtin_fmaf(r2, pc, SLM_COS_T1);
  float c = __builtin_fmaf(r2, pc, 1.0f);
  float sn = (q & 1) ? c : s;
  float cs = (q & 1) ? s : c;
  if (q & 2) sn = -sn;
  if ((q + 1) & 2) cs = -cs;
  if (slm_f32_bits(x) >> 31) sn = -sn;
  *s_out = sn;
  *c_out = cs;
}

/* sin and cos of one float argument; NaN for NaN or infinity. */
SLM_TRIG_FN void slm_sincosf(float x, float* s_out, float* c_out) {
  if (!(__builtin_fabsf(x) <= 3.4028234663852886e38f)) { /* NaN or infinity */
    float nan = slm_f32_from_bits(0x7fc00000u);
    *s_out = nan;
    *c_out = nan;
    return;
  }
  slm_sincosf_finite(x, s_out, c_out);
}

SLM_TRIG_FN float slm_sinf(float x) {
  float s, c;
  slm_sincosf(x, &s, &c);
  return s;
}

SLM_TRIG_FN float slm_cosf(float x) {
  float s, c;
  slm_sincosf(x, &s, &c);
  return c;
}

#endif /* SLM_TRIG_H */
