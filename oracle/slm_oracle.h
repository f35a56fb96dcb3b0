/*
 * slm_oracle.h - CPU oracle for the slime-mold step.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load this library, and only to check results (or as the timed CPU baseline).  The
 * product (libslimemold_hip.so) never links, loads or calls it.
 *
 * It restates src/compute.wgsl of the reference (agents :57-145, decay
 * :147-161, diffuse :163-195) and the host-side init of src/main.rs:262-282 in
 * plain C, under these race-free semantics (the reference's own in-place passes
 * have data races, so it has no unique answer):
 *   S1 each agent is updated exactly once per step (no 65535-stride dispatch);
 *   S2 sensing reads the trail as it was at the start of the agent pass;
 *   S3 a cell that at least one agent lands in becomes clamp(T + DEP, 0, 1),
 *      which equals the reference's racy read-modify-write whenever DEP >= 1
 *      (every preset in src/presets.rs);
 *   S4 decay, then a Jacobi (double-buffered) toroidal 3x3 box diffuse;
 *   S5 WGSL op semantics: no fp contraction, IEEE division, '%' as
 *      a - b*trunc(a/b), mix(a,b,t) = a*(1-t) + b*t, fract(v) = v - floor(v);
 *      sin/cos from slm_trig.h (the build's single definition of WGSL sin/cos);
 *   S6 the jitter hash uses the agent's original index.
 * Storage modes round the stored trail to bf16/fp16 after each step while the
 * deposit, decay and diffuse arithmetic stays in f32.  The f64 mode evaluates
 * the same step in double precision with libm sin/cos.
 */
#ifndef SLM_ORACLE_H
#define SLM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields and order as the reference's SimSizeUniform after width/height
 * (src/main.rs:29-46); decay_factor is raw and scaled by 0.001 per step. */
typedef struct orc_params {
  float decay_factor;
  float agent_jitter;
  float agent_speed_min;
  float agent_speed_max;
  float agent_turn_speed;
  float agent_sensor_angle;
  float agent_sensor_distance;
  float diffusion_rate;
  float pheromone_deposition_amount;
  float blur_radius;
  float blur_sigma;
} orc_params;

/* BF16_SR: bfloat16 storage with deterministic stochastic rounding in the
 * trail pass (orc_f32_to_bf16_sr); other conversions round to nearest. */
enum { ORC_STORAGE_F32 = 0, ORC_STORAGE_BF16 = 1, ORC_STORAGE_F16 = 2, ORC_STORAGE_BF16_SR = 3 };

/* ---- primitives (exposed for the known-answer tests) -------------------- */
void orc_sincos(float x, float* s, float* c);
/* Switch the agent update between the shared sin/cos (default, 0) and the C
 * library's sinf/cosf (1): a stand-in for a reference whose transcendental
 * bits differ, for statistical (not bitwise) comparisons. */
void orc_set_libm_trig(int on);
/* The Gaussian diffuse extension: weights (returns the radius, 0 if invalid),
 * one pass, and a switch making orc_step / orc_step_p64 use it (default 0). */
int orc_gauss_weights(float radius, float sigma, float* w);
void orc_set_gaussian(int on);
float orc_wgsl_mod(float a, float b);
float orc_fract(float v);
float orc_mix(float a, float b, float t);
float orc_sign(float v);
int32_t orc_f32_to_i32(float v);
float orc_round_storage(float v, int storage);
uint32_t orc_sr_bits(uint32_t cell, uint32_t bits);
uint16_t orc_f32_to_bf16_sr(float v, uint32_t cell);
float orc_round_storage_at(float v, int storage, uint32_t cell);
uint16_t orc_f32_to_bf16(float v);
float orc_bf16_to_f32(uint16_t h);
uint16_t orc_f32_to_f16(float v);
float orc_f16_to_f32(uint16_t h);
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t n);
float orc_uniform24(uint64_t seed, uint64_t n);

/* Bilinear sample of the trail (compute.wgsl:7-29). */
float orc_sample(const float* trail, uint32_t w, uint32_t h, float px, float py);

/* ---- state initialisation (src/main.rs:262-282, :101-145) -------------- */
/* agents_aos holds n records (x, y, angle, speed). Agent i uses generator
 * outputs 4i..4i+3 of SplitMix64 seeded with `seed`. */
void orc_init_agents(uint64_t seed, uint32_t w, uint32_t h, uint64_t n, float speed_min,
                     float speed_max, float* agents_aos);
/* speed_i = min + u*(max-min), u from stream `stream` (seed + stream*0xd1b54a32d192ed03). */
void orc_reassign_speeds(uint64_t seed, uint64_t stream, uint64_t n, float speed_min,
                         float speed_max, float* agents_aos);

/* ---- passes ------------------------------------------------------------- */
/* One agent update against a read-only trail; returns the deposit cell index
 * or -1 (compute.wgsl:57-145). `index` is the agent's original index. */
int64_t orc_agent_update(const float* trail, uint32_t w, uint32_t h, const orc_params* p,
                         uint32_t index, float* agent4);
/* Deposit marks + decay + diffuse: trail_in -> trail_out (compute.wgsl:140,
 * :147-161, :163-195).  mark may be NULL (no deposits). */
void orc_trail_pass(const float* trail_in, const uint8_t* mark, uint32_t w, uint32_t h,
                    const orc_params* p, int storage, float* trail_out, int nthreads);
/* The same pass with the Gaussian diffuse extension in place of the 3x3 box. */
void orc_trail_pass_gauss(const float* trail_in, const uint8_t* mark, uint32_t w, uint32_t h,
                          const orc_params* p, int storage, float* trail_out, int nthreads);

/* The trail pass in double precision (FP64 reference for the storage modes). */
void orc_trail_pass_f64(const double* trail_in, const uint8_t* mark, uint32_t w, uint32_t h,
                        const orc_params* p, double* trail_out, int nthreads);

/* n_steps full steps in place.  nthreads <= 0 uses all OpenMP threads. */
int orc_step(uint32_t w, uint32_t h, uint64_t n, float* agents_aos, float* trail,
             const orc_params* p, int storage, uint32_t n_steps, int nthreads);

/* f64-position variant of orc_agent_update / orc_step: agents as (x, y,
 * angle, speed) doubles, angle and speed holding f32 values. */
int64_t orc_agent_update_p64(const float* trail, uint32_t w, uint32_t h, const orc_params* p,
                             uint32_t index, double* agent4);
int orc_step_p64(uint32_t w, uint32_t h, uint64_t n, double* agents_aos, float* trail,
                 const orc_params* p, int storage, uint32_t n_steps, int nthreads);

/* The same step in double precision (agents as (x,y,angle,speed) doubles). */
int orc_step_f64(uint32_t w, uint32_t h, uint64_t n, double* agents_aos, double* trail,
                 const orc_params* p, uint32_t n_steps, int nthreads);

double orc_total_mass(const float* trail, uint64_t cells);
double orc_total_mass_f64(const double* trail, uint64_t cells);

/* Display colour map (src/display.wgsl:29-87) into tex_w * tex_h RGBA8. */
void orc_render(const float* trail, uint32_t w, uint32_t h, const uint8_t* lut, uint32_t tex_w,
                uint32_t tex_h, uint8_t* rgba);

#ifdef __cplusplus
}
#endif

#endif /* SLM_ORACLE_H */
