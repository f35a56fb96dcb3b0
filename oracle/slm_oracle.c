/*
 * slm_oracle.c - CPU oracle for the slime-mold step.  TEST INFRASTRUCTURE ONLY
 * (see slm_oracle.h for who may use it and for the semantics S1-S6).
 *
 * Built with gcc -O2 -ffp-contract=off -fopenmp (oracle/Makefile): SSE
 * single-precision arithmetic with no contraction is WGSL's f32 arithmetic.
 * Each function cites the part of the reference it restates.
 */
#include "slm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <omp.h>
#include <string.h>

#include "../slime-mold_amd/csrc/slm_trig.h"

/* compute.wgsl:4 TAU and :121 2.0*3.14159265359 both round to 0x40c90fdb */
#define ORC_TAU 6.28318530718f
/* compute.wgsl:55 */
#define ORC_TIME_STEP 0.016f

/* ---- primitives ---------------------------------------------------------- */

void orc_sincos(float x, float* s, float* c) { slm_sincosf(x, s, c); }

/* The agent update's sin/cos: the shared implementation, or (for tests that
 * model a reference whose transcendental bits differ) the C library's. */
static int g_libm_trig = 0;
void orc_set_libm_trig(int on) { g_libm_trig = on != 0; }
static inline void agent_sincos(float x, float* s, float* c) {
  if (g_libm_trig) {
    *s = sinf(x);
    *c = cosf(x);
  } else {
    slm_sincosf(x, s, c);
  }
}
static inline float agent_sin(float x) { return g_libm_trig ? sinf(x) : slm_sinf(x); }

/* WGSL f32 '%': e1 - e2 * trunc(e1 / e2), each operation rounded (S5). */
float orc_wgsl_mod(float a, float b) {
  float q = a / b;
  float t = truncf(q);
  float p = b * t;
  return a - p;
}

/* WGSL fract(e) = e - floor(e) (compute.wgsl:117). */
float orc_fract(float v) { return v - floorf(v); }

/* WGSL mix(e1, e2, e3) = e1 * (1 - e3) + e2 * e3 (compute.wgsl:26-28, :194). */
float orc_mix(float a, float b, float t) {
  float u = 1.0f - t;
  float l = a * u;
  float r = b * t;
  return l + r;
}

/* WGSL sign (compute.wgsl:104,109). */
float orc_sign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }

/* WGSL i32(f32): round toward zero, then clamp to the i32 range; NaN -> 0. */
int32_t orc_f32_to_i32(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return 2147483647;
  if (v <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)v;
}

static float wgsl_clamp(float e, float lo, float hi) {
  float m = e > lo ? e : lo;
  return m < hi ? m : hi;
}

static float wgsl_min(float a, float b) { return a < b ? a : b; }
static float wgsl_max(float a, float b) { return a > b ? a : b; }

static uint32_t f32_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float f32_from(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* Round-to-nearest-even f32 -> bfloat16 (IEEE-style RNE on the top 16 bits). */
uint16_t orc_f32_to_bf16(float v) {
  uint32_t x = f32_bits(v);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x0040u);
  uint32_t lsb = (x >> 16) & 1u;
  return (uint16_t)((x + 0x7fffu + lsb) >> 16);
}

float orc_bf16_to_f32(uint16_t h) { return f32_from((uint32_t)h << 16); }

/* Round-to-nearest-even f32 -> IEEE binary16, with subnormals and overflow. */
uint16_t orc_f32_to_f16(float v) {
  uint32_t x = f32_bits(v);
  uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
  uint32_t ax = x & 0x7fffffffu;
  if (ax > 0x7f800000u) return sign | 0x7e00u;       /* NaN */
  if (ax >= 0x477ff000u) return sign | 0x7c00u;      /* >= 65520 rounds to inf */
  if (ax >= 0x38800000u) {                           /* normal binary16 */
    uint32_t m = ax - 0x38000000u;                   /* rebias 127 -> 15 */
    m = m + 0x0fffu + ((m >> 13) & 1u);
    return (uint16_t)(sign | (m >> 13));
  }
  /* below 2^-14: multiples of 2^-24, scaling by 2^24 is exact */
  float n = rintf(f32_from(ax) * 16777216.0f);
  return (uint16_t)(sign | (uint16_t)n);
}

float orc_f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0) {
    float f = (float)m * (1.0f / 16777216.0f);
    return sign ? -f : f;
  }
  if (e == 31) return f32_from(sign | 0x7f800000u | (m << 13));
  return f32_from(sign | ((e + 112u) << 23) | (m << 13));
}

float orc_round_storage(float v, int storage) {
  if (storage == ORC_STORAGE_BF16 || storage == ORC_STORAGE_BF16_SR)
    return orc_bf16_to_f32(orc_f32_to_bf16(v));
  if (storage == ORC_STORAGE_F16) return orc_f16_to_f32(orc_f32_to_f16(v));
  return v;
}

/* Stochastic rounding to bfloat16, deterministic: the 16 discarded bits of a
 * non-negative finite f32 are added to 16 pseudo-random bits and the sum
 * truncated.  The bits come from the cell and the value: x = (cell *
 * 0x9e3779b1) ^ bits, then two rounds of x ^= x >> 16 and a multiply of x's
 * low 24 bits by an odd 24-bit constant (0xebca6b, then 0x9e3779), all mod
 * 2^32; the draw is bits 16..31 of the result.  The xor-shifts break the
 * lattice a single multiplicative hash of the linear cell sequence leaves
 * between neighbouring cells; the 24-bit products are full-rate
 * v_mul_u32_u24 on the GPU.  The result rounds up with probability
 * (discarded bits) / 2^16, so it is unbiased. */
uint32_t orc_sr_bits(uint32_t cell, uint32_t bits) {
  uint32_t x = (cell * 0x9e3779b1u) ^ bits;
  x ^= x >> 16;
  x = (x & 0xffffffu) * 0xebca6bu;
  x ^= x >> 16;
  x = (x & 0xffffffu) * 0x9e3779u;
  return x >> 16;
}

uint16_t orc_f32_to_bf16_sr(float v, uint32_t cell) {
  uint32_t b = f32_bits(v);
  return (uint16_t)((b + orc_sr_bits(cell, b)) >> 16);
}

float orc_round_storage_at(float v, int storage, uint32_t cell) {
  if (storage == ORC_STORAGE_BF16_SR) return orc_bf16_to_f32(orc_f32_to_bf16_sr(v, cell));
  return orc_round_storage(v, storage);
}

/* SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): output n (0-based) of the
 * generator seeded with `seed`. */
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t n) {
  uint64_t z = seed + (n + 1u) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* Top 24 bits as a multiple of 2^-24 in [0, 1), the resolution of rand 0.9's
 * standard f32 (src/main.rs:277-280 draws four of these per agent). */
float orc_uniform24(uint64_t seed, uint64_t n) {
  return (float)(orc_splitmix64_at(seed, n) >> 40) * (1.0f / 16777216.0f);
}

/* ---- state initialisation ---------------------------------------------- */

/* src/main.rs:273-280: x = u*W, y = u*H, angle = u*2*PI, speed = min + u*range */
void orc_init_agents(uint64_t seed, uint32_t w, uint32_t h, uint64_t n, float speed_min,
                     float speed_max, float* a) {
  const float pi = 3.14159265358979323846f; /* std::f32::consts::PI */
  float range = speed_max - speed_min;
  for (uint64_t i = 0; i < n; i++) {
    float u0 = orc_uniform24(seed, 4 * i), u1 = orc_uniform24(seed, 4 * i + 1);
    float u2 = orc_uniform24(seed, 4 * i + 2), u3 = orc_uniform24(seed, 4 * i + 3);
    a[4 * i + 0] = u0 * (float)w;
    a[4 * i + 1] = u1 * (float)h;
    a[4 * i + 2] = (u2 * 2.0f) * pi;
    a[4 * i + 3] = speed_min + u3 * range;
  }
}

/* src/main.rs:132-137: speed = min + u*(max - min) for every agent. */
void orc_reassign_speeds(uint64_t seed, uint64_t stream, uint64_t n, float speed_min,
                         float speed_max, float* a) {
  uint64_t s = seed + stream * 0xd1b54a32d192ed03ull;
  float range = speed_max - speed_min;
  for (uint64_t i = 0; i < n; i++) a[4 * i + 3] = speed_min + orc_uniform24(s, i) * range;
}

/* ---- agent pass ---------------------------------------------------------- */

/* compute.wgsl:7-29.  Not toroidal: any tap outside the grid gives 0. */
float orc_sample(const float* t, uint32_t w, uint32_t h, float px, float py) {
  int32_t x0 = orc_f32_to_i32(floorf(px));
  int32_t y0 = orc_f32_to_i32(floorf(py));
  int64_t x1 = (int64_t)x0 + 1, y1 = (int64_t)y0 + 1;
  if (x0 < 0 || x1 >= (int64_t)w || y0 < 0 || y1 >= (int64_t)h) return 0.0f;
  float dx = px - (float)x0;
  float dy = py - (float)y0;
  float v00 = t[(int64_t)y0 * w + x0];
  float v10 = t[(int64_t)y0 * w + x1];
  float v01 = t[y1 * w + x0];
  float v11 = t[y1 * w + x1];
  float v0 = orc_mix(v00, v10, dx);
  float v1 = orc_mix(v01, v11, dx);
  return orc_mix(v0, v1, dy);
}

/* compute.wgsl:57-145 for one agent; the deposit is returned, not applied (S2/S3). */
int64_t orc_agent_update(const float* t, uint32_t w, uint32_t h, const orc_params* p,
                         uint32_t index, float* ag) {
  float x = ag[0], y = ag[1], angle = ag[2], speed = ag[3];
  speed = wgsl_clamp(speed, p->agent_speed_min, p->agent_speed_max); /* :72 */

  float sa = p->agent_sensor_angle, sd = p->agent_sensor_distance;
  float al = angle - sa, ar = angle + sa, ac = angle; /* :75-77 */
  float sl, cl, sr, cr, sc, cc;
  agent_sincos(al, &sl, &cl);
  agent_sincos(ar, &sr, &cr);
  agent_sincos(ac, &sc, &cc);
  float lx = x + sd * cl, ly = y + sd * sl; /* :79-90 */
  float rx = x + sd * cr, ry = y + sd * sr;
  float cx = x + sd * cc, cy = y + sd * sc;
  float lv = orc_sample(t, w, h, lx, ly); /* :93-95, read order L, R, C */
  float rv = orc_sample(t, w, h, rx, ry);
  float cv = orc_sample(t, w, h, cx, cy);

  if (cv > lv && cv > rv) { /* :98 straight */
  } else if (lv > rv) {     /* :100-104 */
    float target = angle - ORC_TAU;
    float diff = target - angle;
    angle += wgsl_min(p->agent_turn_speed, fabsf(diff)) * orc_sign(diff);
  } else if (rv > lv) { /* :105-109 */
    float target = angle + ORC_TAU;
    float diff = target - angle;
    angle += wgsl_min(p->agent_turn_speed, fabsf(diff)) * orc_sign(diff);
  }

  /* :115-118 - left-to-right evaluation of the hash argument */
  float hx = (float)(int32_t)index * 12.9898f;
  hx = hx + x * 78.233f;
  hx = hx + y * 37.719f;
  float rnd = orc_fract(agent_sin(hx) * 43758.5453f);
  angle += (rnd * 2.0f - 1.0f) * p->agent_jitter;

  angle = orc_wgsl_mod(angle, ORC_TAU); /* :121-122 */
  if (angle < 0.0f) angle = angle + ORC_TAU;

  float move = speed * ORC_TIME_STEP; /* :125-127 */
  float ms, mc;
  agent_sincos(angle, &ms, &mc);
  x = x + move * mc;
  y = y + move * ms;

  float fw = (float)w, fh = (float)h; /* :130-133 */
  x = orc_wgsl_mod(x, fw);
  if (x < 0.0f) x = x + fw;
  y = orc_wgsl_mod(y, fh);
  if (y < 0.0f) y = y + fh;

  int64_t dep = -1; /* :136-141 */
  int32_t dx = orc_f32_to_i32(x), dy = orc_f32_to_i32(y);
  if (dx >= 0 && (int64_t)dx < (int64_t)w && dy >= 0 && (int64_t)dy < (int64_t)h)
    dep = (int64_t)dy * w + dx;

  ag[0] = x; /* :144 */
  ag[1] = y;
  ag[2] = angle;
  ag[3] = speed;
  return dep;
}

/* f64-position variant (the library's SLM_FLAG_F64_POSITIONS): x, y in double;
 * sensor offsets and the move step are the f32 products of the f32 path, added
 * to the double position; floor, wrap and deposit cell in double; bilinear
 * weights and the jitter hash use the f32 roundings.  agent4 = (x, y, angle,
 * speed) with angle and speed holding f32 values. */
static float sample_p64(const float* t, uint32_t w, uint32_t h, double px, double py) {
  double fx = floor(px), fy = floor(py);
  float dx = (float)(px - fx), dy = (float)(py - fy);
  int32_t x0 = (fx != fx) ? 0 : (fx >= 2147483648.0 ? 2147483647 : (fx <= -2147483648.0 ? (int32_t)0x80000000u : (int32_t)fx));
  int32_t y0 = (fy != fy) ? 0 : (fy >= 2147483648.0 ? 2147483647 : (fy <= -2147483648.0 ? (int32_t)0x80000000u : (int32_t)fy));
  int64_t x1 = (int64_t)x0 + 1, y1 = (int64_t)y0 + 1;
  if (x0 < 0 || x1 >= (int64_t)w || y0 < 0 || y1 >= (int64_t)h) return 0.0f;
  float v00 = t[(int64_t)y0 * w + x0], v10 = t[(int64_t)y0 * w + x1];
  float v01 = t[y1 * w + x0], v11 = t[y1 * w + x1];
  return orc_mix(orc_mix(v00, v10, dx), orc_mix(v01, v11, dx), dy);
}

static double mod_d(double a, double b) {
  double q = a / b;
  double tr = trunc(q);
  double pr = b * tr;
  return a - pr;
}

int64_t orc_agent_update_p64(const float* t, uint32_t w, uint32_t h, const orc_params* p,
                             uint32_t index, double* ag) {
  double xd = ag[0], yd = ag[1];
  float angle = (float)ag[2], speed = (float)ag[3];
  float x = (float)xd, y = (float)yd;
  speed = wgsl_clamp(speed, p->agent_speed_min, p->agent_speed_max);
  float sa = p->agent_sensor_angle, sd = p->agent_sensor_distance;
  float sl, cl, sr, cr, sc, cc;
  agent_sincos(angle - sa, &sl, &cl);
  agent_sincos(angle + sa, &sr, &cr);
  agent_sincos(angle, &sc, &cc);
  float lv = sample_p64(t, w, h, xd + (double)(sd * cl), yd + (double)(sd * sl));
  float rv = sample_p64(t, w, h, xd + (double)(sd * cr), yd + (double)(sd * sr));
  float cv = sample_p64(t, w, h, xd + (double)(sd * cc), yd + (double)(sd * sc));
  if (cv > lv && cv > rv) {
  } else if (lv > rv) {
    float target = angle - ORC_TAU;
    float diff = target - angle;
    angle += wgsl_min(p->agent_turn_speed, fabsf(diff)) * orc_sign(diff);
  } else if (rv > lv) {
    float target = angle + ORC_TAU;
    float diff = target - angle;
    angle += wgsl_min(p->agent_turn_speed, fabsf(diff)) * orc_sign(diff);
  }
  float hx = (float)(int32_t)index * 12.9898f;
  hx = hx + x * 78.233f;
  hx = hx + y * 37.719f;
  float rnd = orc_fract(agent_sin(hx) * 43758.5453f);
  angle += (rnd * 2.0f - 1.0f) * p->agent_jitter;
  angle = orc_wgsl_mod(angle, ORC_TAU);
  if (angle < 0.0f) angle = angle + ORC_TAU;
  float move = speed * ORC_TIME_STEP;
  float ms, mc;
  agent_sincos(angle, &ms, &mc);
  double fw = (double)w, fh = (double)h;
  xd = xd + (double)(move * mc);
  yd = yd + (double)(move * ms);
  xd = mod_d(xd, fw);
  if (xd < 0.0) xd = xd + fw;
  yd = mod_d(yd, fh);
  if (yd < 0.0) yd = yd + fh;
  int64_t dep = -1;
  if (xd >= 0.0 && xd < fw && yd >= 0.0 && yd < fh) dep = (int64_t)yd * w + (int64_t)xd;
  ag[0] = xd;
  ag[1] = yd;
  ag[2] = angle;
  ag[3] = speed;
  return dep;
}

/* ---- trail pass ---------------------------------------------------------- */

void orc_trail_pass(const float* tin, const uint8_t* mark, uint32_t w, uint32_t h,
                    const orc_params* p, int storage, float* tout, int nthreads) {
  uint64_t cells = (uint64_t)w * h;
  float* t2 = (float*)malloc(cells * sizeof(float));
  float k = p->decay_factor * 0.001f; /* compute.wgsl:159 */
  float dep = p->pheromone_deposition_amount;
  float rate = wgsl_clamp(p->diffusion_rate, 0.0f, 1.0f); /* :173 */
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t c = 0; c < (int64_t)cells; c++) {
    float v = tin[c];
    if (mark && mark[c]) v = wgsl_clamp(v + dep, 0.0f, 1.0f); /* :140 (S3) */
    t2[c] = wgsl_max(v - k, 0.0f);                            /* :160 */
  }
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t y = 0; y < (int64_t)h; y++) {
    for (int64_t x = 0; x < (int64_t)w; x++) {
      float sum = 0.0f; /* :176-189, dy outer, dx inner */
      float count = 0.0f;
      for (int dy = -1; dy <= 1; dy++) {
        for (int dx = -1; dx <= 1; dx++) {
          int64_t nx = (x + dx + (int64_t)w) % (int64_t)w;
          int64_t ny = (y + dy + (int64_t)h) % (int64_t)h;
          sum += t2[ny * w + nx];
          count += 1.0f;
        }
      }
      float avg = sum / count; /* :193 */
      float out = orc_mix(t2[y * w + x], avg, rate); /* :194 */
      tout[y * w + x] = orc_round_storage_at(out, storage, (uint32_t)(y * w + x));
    }
  }
  free(t2);
}

/* ---- Gaussian diffuse (the library's SLM_FLAG_GAUSSIAN_DIFFUSE extension) ----
 * Not in the reference: it carries blur_radius and blur_sigma (src/main.rs:43-44,
 * src/settings.rs:25-27,45-46) but its shader only runs the 3x3 box.  Radius
 * R = round(blur_radius) in 1..8, sigma > 0; weights w[i] = exp(-i^2/(2 sigma^2))
 * normalised over i = -R..R in double and rounded to float.  Returns R, or 0
 * (all weights 0) when the fields do not describe such a Gaussian. */
int orc_gauss_weights(float radius, float sigma, float* w) {
  for (int i = 0; i <= 8; i++) w[i] = 0.0f;
  if (!(radius >= 0.5f && radius < 8.5f) || !(sigma > 0.0f) || !isfinite(sigma)) return 0;
  int r = (int)floor((double)radius + 0.5);
  double s2 = 2.0 * (double)sigma * sigma, e[9], sum = 0.0;
  for (int i = 0; i <= r; i++) {
    e[i] = exp(-(double)(i * i) / s2);
    sum += i == 0 ? e[i] : 2.0 * e[i];
  }
  for (int i = 0; i <= r; i++) w[i] = (float)(e[i] / sum);
  return r;
}

/* Deposit and decay as orc_trail_pass, then H = horizontal taps x+i, i = -R..R
 * ascending, G = vertical taps y+j of H likewise (toroidal): the first tap's
 * product, then one fmaf per further tap; and out = mix(T2, G, clamp(rate, 0, 1)). */
void orc_trail_pass_gauss(const float* tin, const uint8_t* mark, uint32_t w, uint32_t h,
                          const orc_params* p, int storage, float* tout, int nthreads) {
  uint64_t cells = (uint64_t)w * h;
  float* t2 = (float*)malloc(cells * sizeof(float));
  float* hs = (float*)malloc(cells * sizeof(float));
  float wt[9];
  int r = orc_gauss_weights(p->blur_radius, p->blur_sigma, wt);
  float k = p->decay_factor * 0.001f;
  float dep = p->pheromone_deposition_amount;
  float rate = wgsl_clamp(p->diffusion_rate, 0.0f, 1.0f);
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t c = 0; c < (int64_t)cells; c++) {
    float v = tin[c];
    if (mark && mark[c]) v = wgsl_clamp(v + dep, 0.0f, 1.0f);
    t2[c] = wgsl_max(v - k, 0.0f);
  }
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t y = 0; y < (int64_t)h; y++)
    for (int64_t x = 0; x < (int64_t)w; x++) {
      float s = 0.0f;
      for (int i = -r; i <= r; i++) {
        int64_t nx = ((x + i) % (int64_t)w + (int64_t)w) % (int64_t)w;
        float wi = wt[i < 0 ? -i : i], a = t2[y * w + nx];
        s = i == -r ? wi * a : fmaf(wi, a, s);
      }
      hs[y * w + x] = s;
    }
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t y = 0; y < (int64_t)h; y++)
    for (int64_t x = 0; x < (int64_t)w; x++) {
      float g = 0.0f;
      for (int j = -r; j <= r; j++) {
        int64_t ny = ((y + j) % (int64_t)h + (int64_t)h) % (int64_t)h;
        float wj = wt[j < 0 ? -j : j], a = hs[ny * w + x];
        g = j == -r ? wj * a : fmaf(wj, a, g);
      }
      float out = orc_mix(t2[y * w + x], g, rate);
      tout[y * w + x] = orc_round_storage_at(out, storage, (uint32_t)(y * w + x));
    }
  free(t2);
  free(hs);
}

static int g_gaussian = 0;
void orc_set_gaussian(int on) { g_gaussian = on != 0; }

static void trail_pass_mode(const float* tin, const uint8_t* mark, uint32_t w, uint32_t h,
                            const orc_params* p, int storage, float* tout, int nthreads) {
  if (g_gaussian)
    orc_trail_pass_gauss(tin, mark, w, h, p, storage, tout, nthreads);
  else
    orc_trail_pass(tin, mark, w, h, p, storage, tout, nthreads);
}

/* The trail pass in double precision: the FP64 reference for the 16-bit
 * storage modes (same deposit marks, decay, 3x3 toroidal box and mix). */
void orc_trail_pass_f64(const double* tin, const uint8_t* mark, uint32_t w, uint32_t h,
                        const orc_params* p, double* tout, int nthreads) {
  uint64_t cells = (uint64_t)w * h;
  double* t2 = (double*)malloc(cells * sizeof(double));
  double k = (double)p->decay_factor * 0.001, dep = p->pheromone_deposition_amount;
  double rate = p->diffusion_rate < 0.0f ? 0.0 : (p->diffusion_rate > 1.0f ? 1.0 : p->diffusion_rate);
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t c = 0; c < (int64_t)cells; c++) {
    double v = tin[c];
    if (mark && mark[c]) v = v + dep < 0.0 ? 0.0 : (v + dep > 1.0 ? 1.0 : v + dep);
    v -= k;
    t2[c] = v > 0.0 ? v : 0.0;
  }
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t y = 0; y < (int64_t)h; y++) {
    for (int64_t x = 0; x < (int64_t)w; x++) {
      double sum = 0.0;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
          sum += t2[((y + dy + h) % h) * w + (x + dx + w) % w];
      double c = t2[y * w + x];
      tout[y * w + x] = c * (1.0 - rate) + (sum / 9.0) * rate;
    }
  }
  free(t2);
}

int orc_step(uint32_t w, uint32_t h, uint64_t n, float* ag, float* trail, const orc_params* p,
             int storage, uint32_t n_steps, int nthreads) {
  uint64_t cells = (uint64_t)w * h;
  if (w == 0 || h == 0) return -1;
  uint8_t* mark = (uint8_t*)malloc(cells);
  float* next = (float*)malloc(cells * sizeof(float));
  if (!mark || !next) {
    free(mark);
    free(next);
    return -2;
  }
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
  for (uint32_t s = 0; s < n_steps; s++) {
    memset(mark, 0, cells);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)n; i++) {
      int64_t d = orc_agent_update(trail, w, h, p, (uint32_t)i, ag + 4 * i);
      if (d >= 0) {
#pragma omp atomic write
        mark[d] = 1;
      }
    }
    trail_pass_mode(trail, mark, w, h, p, storage, next, nt);
    memcpy(trail, next, cells * sizeof(float));
  }
  free(mark);
  free(next);
  return 0;
}

int orc_step_p64(uint32_t w, uint32_t h, uint64_t n, double* ag, float* trail,
                 const orc_params* p, int storage, uint32_t n_steps, int nthreads) {
  uint64_t cells = (uint64_t)w * h;
  if (w == 0 || h == 0) return -1;
  uint8_t* mark = (uint8_t*)malloc(cells);
  float* next = (float*)malloc(cells * sizeof(float));
  if (!mark || !next) {
    free(mark);
    free(next);
    return -2;
  }
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
  for (uint32_t s = 0; s < n_steps; s++) {
    memset(mark, 0, cells);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)n; i++) {
      int64_t d = orc_agent_update_p64(trail, w, h, p, (uint32_t)i, ag + 4 * i);
      if (d >= 0) {
#pragma omp atomic write
        mark[d] = 1;
      }
    }
    trail_pass_mode(trail, mark, w, h, p, storage, next, nt);
    memcpy(trail, next, cells * sizeof(float));
  }
  free(mark);
  free(next);
  return 0;
}

/* ---- double-precision step ---------------------------------------------- */

static double sample_f64(const double* t, uint32_t w, uint32_t h, double px, double py) {
  double fx = floor(px), fy = floor(py);
  if (!(fx >= 0.0) || fx + 1.0 >= (double)w || !(fy >= 0.0) || fy + 1.0 >= (double)h) return 0.0;
  int64_t x0 = (int64_t)fx, y0 = (int64_t)fy;
  double dx = px - fx, dy = py - fy;
  double v00 = t[y0 * w + x0], v10 = t[y0 * w + x0 + 1];
  double v01 = t[(y0 + 1) * w + x0], v11 = t[(y0 + 1) * w + x0 + 1];
  double v0 = v00 * (1.0 - dx) + v10 * dx;
  double v1 = v01 * (1.0 - dx) + v11 * dx;
  return v0 * (1.0 - dy) + v1 * dy;
}

static double mod_f64(double a, double b) { return a - b * trunc(a / b); }

int orc_step_f64(uint32_t w, uint32_t h, uint64_t n, double* ag, double* trail,
                 const orc_params* p, uint32_t n_steps, int nthreads) {
  const double tau = 6.28318530718, ts = 0.016;
  uint64_t cells = (uint64_t)w * h;
  if (w == 0 || h == 0) return -1;
  uint8_t* mark = (uint8_t*)malloc(cells);
  double* t2 = (double*)malloc(cells * sizeof(double));
  if (!mark || !t2) {
    free(mark);
    free(t2);
    return -2;
  }
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
  double smin = p->agent_speed_min, smax = p->agent_speed_max;
  double sa = p->agent_sensor_angle, sd = p->agent_sensor_distance;
  double turn = p->agent_turn_speed, jit = p->agent_jitter;
  double k = (double)p->decay_factor * 0.001, dep = p->pheromone_deposition_amount;
  double rate = p->diffusion_rate < 0.0f ? 0.0 : (p->diffusion_rate > 1.0f ? 1.0 : p->diffusion_rate);
  for (uint32_t s = 0; s < n_steps; s++) {
    memset(mark, 0, cells);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)n; i++) {
      double* a = ag + 4 * i;
      double x = a[0], y = a[1], angle = a[2], speed = a[3];
      speed = speed < smin ? smin : (speed > smax ? smax : speed);
      double lv = sample_f64(trail, w, h, x + sd * cos(angle - sa), y + sd * sin(angle - sa));
      double rv = sample_f64(trail, w, h, x + sd * cos(angle + sa), y + sd * sin(angle + sa));
      double cv = sample_f64(trail, w, h, x + sd * cos(angle), y + sd * sin(angle));
      if (cv > lv && cv > rv) {
      } else if (lv > rv) {
        angle -= turn < tau ? turn : tau;
      } else if (rv > lv) {
        angle += turn < tau ? turn : tau;
      }
      double hx = (double)(int32_t)(uint32_t)i * 12.9898 + x * 78.233 + y * 37.719;
      double v = sin(hx) * 43758.5453;
      angle += ((v - floor(v)) * 2.0 - 1.0) * jit;
      angle = mod_f64(angle, tau);
      if (angle < 0.0) angle += tau;
      double mv = speed * ts;
      x += mv * cos(angle);
      y += mv * sin(angle);
      x = mod_f64(x, (double)w);
      if (x < 0.0) x += w;
      y = mod_f64(y, (double)h);
      if (y < 0.0) y += h;
      if (x >= 0.0 && x < (double)w && y >= 0.0 && y < (double)h) {
#pragma omp atomic write
        mark[(int64_t)y * w + (int64_t)x] = 1;
      }
      a[0] = x;
      a[1] = y;
      a[2] = angle;
      a[3] = speed;
    }
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t c = 0; c < (int64_t)cells; c++) {
      double v = trail[c];
      if (mark[c]) v = v + dep < 0.0 ? 0.0 : (v + dep > 1.0 ? 1.0 : v + dep);
      v -= k;
      t2[c] = v > 0.0 ? v : 0.0;
    }
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t y = 0; y < (int64_t)h; y++) {
      for (int64_t x = 0; x < (int64_t)w; x++) {
        double sum = 0.0;
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++)
            sum += t2[((y + dy + h) % h) * w + (x + dx + w) % w];
        double c = t2[y * w + x];
        trail[y * w + x] = c * (1.0 - rate) + (sum / 9.0) * rate;
      }
    }
  }
  free(mark);
  free(t2);
  return 0;
}

double orc_total_mass(const float* t, uint64_t cells) {
  double s = 0.0;
  for (uint64_t c = 0; c < cells; c++) s += (double)t[c];
  return s;
}

double orc_total_mass_f64(const double* t, uint64_t cells) {
  double s = 0.0;
  for (uint64_t c = 0; c < cells; c++) s += t[c];
  return s;
}

/* src/display.wgsl:29-87: fit the grid into a tex_w x tex_h texture keeping
 * its aspect (f32 scale and offsets), nearest cell by truncation, LUT entry
 * u32(clamp(T, 0, 1) * 255) of the 768-byte R[256] G[256] B[256] table, alpha
 * 255; outside the grid black.  rgba: tex_w * tex_h * 4 bytes, row-major. */
void orc_render(const float* trail, uint32_t w, uint32_t h, const uint8_t* lut, uint32_t tex_w,
                uint32_t tex_h, uint8_t* rgba) {
  float sim_w = (float)w, sim_h = (float)h, tw = (float)tex_w, th = (float)tex_h;
  float sim_aspect = sim_w / sim_h;                   /* :55 */
  float tex_aspect = tw / th;                         /* :56 */
  float scale, off_x = 0.0f, off_y = 0.0f;
  if (tex_aspect > sim_aspect) {                      /* :62-66 */
    scale = th / sim_h;
    off_x = (tw - sim_w * scale) * 0.5f;
  } else {                                            /* :67-71 */
    scale = tw / sim_w;
    off_y = (th - sim_h * scale) * 0.5f;
  }
  for (uint32_t py = 0; py < tex_h; py++) {
    for (uint32_t px = 0; px < tex_w; px++) {
      float fx = ((float)px - off_x) / scale;         /* :74-75 */
      float fy = ((float)py - off_y) / scale;
      uint8_t* o = rgba + 4 * ((size_t)py * tex_w + px);
      o[0] = o[1] = o[2] = 0;
      o[3] = 255;
      if (fx >= 0.0f && fx < sim_w && fy >= 0.0f && fy < sim_h) { /* :78 */
        int32_t x = orc_f32_to_i32(fx), y = orc_f32_to_i32(fy);
        float t = trail[(size_t)y * w + (size_t)x];
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);  /* :83, :31 */
        uint32_t k = (uint32_t)(t * 255.0f);            /* :34 */
        o[0] = lut[k];                                  /* :37-39 */
        o[1] = lut[k + 256];
        o[2] = lut[k + 512];
      }
    }
  }
}
