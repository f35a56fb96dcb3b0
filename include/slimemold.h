/*
 * slimemold.h - C ABI of the MI355X slime-mold step (libslimemold_hip.so).
 *
 * This is the seam that replaces the reference's wgpu compute dispatch: the
 * three compute pipelines built from src/compute.wgsl
 * (src/pipeline_manager.rs:37-54), their bind group (agents rw, trail rw,
 * uniform; src/pipeline_manager.rs:77-113, src/bind_group_manager.rs:50-65,
 * :101-126) and the host code in src/main.rs that uploads parameters and
 * dispatches, clears, resizes and re-initialises the simulation.
 *
 * Conventions
 *   - Every function returns SLM_OK (0) or a negative SLM_ERR_* code and never
 *     aborts or throws.  slm_last_error() gives a message for the calling
 *     thread's most recent failure.
 *   - A failed call leaves the handle usable, with one exception: when
 *     slm_resize, slm_set_agent_count or a tuning change (slm_set_tuning,
 *     slm_autotune and their batch forms) fails part-way through
 *     re-laying its buffers, every later call on that handle but destroy
 *     returns the same status.  (Inside slm_autotune, a trial that fails
 *     this way is undone when the final re-layout and restore succeed.)
 *   - A simulation handle owns all of its device memory and one HIP stream
 *     (a batch with instance groups also three helper streams, forked from
 *     and joined back into its own within every call).  It may be used by
 *     one host thread at a time; distinct handles are independent.  slm_step
 *     only enqueues work; the read_* and total_mass calls wait for it.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - Agents are exchanged as the reference lays them out: n records of four
 *     f32 (x, y, angle, speed) (src/compute.wgsl:2,31-32), in the original
 *     agent order whatever re-sorting the library does internally.
 *   - The trail is exchanged as row-major f32, index y*width + x
 *     (src/compute.wgsl:34-35), whatever the storage format.
 *
 * Semantics: the per-step update is the reference's (src/compute.wgsl:57-195)
 * made race-free: every agent moves exactly once per step; sensing reads the
 * trail as it was when the step began; a cell that agents land in becomes
 * clamp(T + deposit, 0, 1) (the reference's result for deposit >= 1, i.e. for
 * every preset); decay then a double-buffered toroidal 3x3 box diffuse.
 */
#ifndef SLIMEMOLD_H
#define SLIMEMOLD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLM_ABI_VERSION 1

/* ---- status codes --------------------------------------------------------- */
#define SLM_OK 0
#define SLM_ERR_INVALID_ARGUMENT (-1)
#define SLM_ERR_OUT_OF_MEMORY (-2)
#define SLM_ERR_HIP (-3)
#define SLM_ERR_NO_DEVICE (-4)
#define SLM_ERR_UNSUPPORTED (-5)

/* ---- parameters ----------------------------------------------------------
 * The fields of the reference's SimSizeUniform after width and height, in the
 * same order (src/main.rs:29-46, built from Settings by src/main.rs:48-67).
 * decay_factor is the raw setting (e.g. 10.0); the step subtracts
 * decay_factor * 0.001 (src/compute.wgsl:159).  blur_radius and blur_sigma
 * are carried for layout compatibility; the reference shader never reads them
 * (src/compute.wgsl:49) and neither does the 3x3 box diffuse. */
typedef struct slm_params {
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
} slm_params;

/* Trail storage in HBM.  Deposit, decay and diffuse arithmetic is f32 in all
 * three; the 16-bit formats round the stored result to nearest-even. */
typedef enum slm_storage {
  SLM_STORAGE_F32 = 0,
  SLM_STORAGE_BF16 = 1,
  SLM_STORAGE_F16 = 2
} slm_storage;

/* Creation-time configuration.  Start from slm_config_default(). */
typedef struct slm_config {
  uint32_t width;           /* grid width in cells  (1 .. 2^24 - 1) */
  uint32_t height;          /* grid height in cells (1 .. 2^24 - 1); width*height <= 2^31 */
  uint64_t agent_count;     /* >= 0, < 2^32 */
  uint64_t seed;            /* agent initialisation seed (SplitMix64) */
  int32_t storage;          /* slm_storage */
  int32_t device;           /* HIP device ordinal */
  uint32_t resort_interval; /* Morton re-sort of agents every k steps; 0 = never.
                               The re-sorts fall on the simulation's steps 0, k,
                               2k, ... however the steps are split into calls
                               and whether they run eagerly or from captured
                               graphs (a captured chunk is kept per starting
                               point in that schedule).  Results do not depend
                               on when re-sorts happen. */
  uint32_t graph_chunk;     /* steps per captured hipGraph; 0 = launch eagerly */
  uint32_t sort_cell_log2;  /* re-sort bin edge = 2^k cells (default 4) */
  uint32_t flags;           /* SLM_FLAG_* */
  uint32_t reserved[6];     /* must be zero */
} slm_config;

/* Map consecutive workgroups of a launch onto one XCD's L2 (default on). */
#define SLM_FLAG_NO_XCD_REMAP 0x1u
/* Keep agent positions in f64 (for grids of 8192 and larger, where an f32
 * position's ulp approaches 1e-3 cell).  Sensor offsets and the move step stay
 * the shader's f32 products; they are added to the f64 position, and the wrap,
 * floor and deposit cell are taken in f64.  Agents are still exchanged as f32. */
#define SLM_FLAG_F64_POSITIONS 0x2u
/* With re-sorting on, stage each 64x64-cell tile's trail window (tile plus
 * sensor reach plus drift since the re-sort) in LDS and sense from there; taps
 * outside the window read global memory, so results are unchanged.  Pays off
 * for short sensor distances and frequent re-sorts. */
#define SLM_FLAG_TILED_SENSING 0x4u
/* With BF16 storage: the trail pass stores each cell rounded up with
 * probability equal to the discarded fraction of its f32 value, instead of to
 * nearest-even.  The decision adds the 16 discarded bits to 16 pseudo-random
 * bits: x = (cell * 0x9e3779b1) ^ value bits, then twice x ^= x >> 16 and
 * x = (x mod 2^24) * K (K = 0xebca6b, then 0x9e3779), mod 2^32, and bits
 * 16..31 of x; so runs stay deterministic and reproducible.  Unbiased: removes the drift of a
 * long round-to-nearest BF16 run's mass.  Other storage: SLM_ERR_UNSUPPORTED.
 * Values given to slm_write_trail are still rounded to nearest. */
#define SLM_FLAG_STOCHASTIC_ROUNDING 0x8u
/* Diffuse with a separable Gaussian instead of the reference's 3x3 box
 * (an extension: the reference carries blur_radius and blur_sigma but never
 * reads them).  Per instance, R = round(blur_radius) must be 1..8 and
 * blur_sigma > 0, else SLM_ERR_INVALID_ARGUMENT.  With T2 the deposited and
 * decayed trail and w[i] = exp(-i^2 / (2 sigma^2)) / sum_{|j|<=R} exp(...)
 * (double, rounded to f32), the pass computes in f32:
 *   H[y][x] = sum_{i=-R..R} w[|i|] * T2[y][(x+i) mod W]   (i ascending)
 *   G[y][x] = sum_{j=-R..R} w[|j|] * H[(y+j) mod H][x]    (j ascending)
 *   out     = mix(T2[y][x], G[y][x], clamp(diffusion_rate, 0, 1)),
 * each sum starting from its first product and adding every further tap with
 * one fused multiply-add (a single rounding); mix as the 3x3 pass. */
#define SLM_FLAG_GAUSSIAN_DIFFUSE 0x10u
/* Windowed sensing: each 256-agent workgroup of the agent pass stages the
 * trail around its agents' bounding box, widened by the sensor reach, in LDS
 * (at most 31 KiB) and senses from there; taps outside the window, and
 * workgroups whose window would not fit, read global memory, so results are
 * unchanged.  Pays off after Morton re-sorts, which keep a workgroup's agents
 * together.  Exclusive with SLM_FLAG_TILED_SENSING. */
#define SLM_FLAG_WINDOW_SENSING 0x20u
/* Persistent passes: the agent and trail passes of a step each launch only
 * the workgroups the GPU holds at once, which claim the pass's work from
 * device counters (reset by the kernels themselves, so captured graphs replay
 * them) until none is left.  One agent launch and one trail launch per step
 * over every instance, as without the flag; not with instance groups
 * (SLM_ERR_UNSUPPORTED from the tuning calls).  Results are unchanged. */
#define SLM_FLAG_PERSISTENT 0x40u

typedef struct slm_info {
  uint32_t width, height;
  uint64_t agent_count;
  int32_t storage;
  int32_t device;
  uint64_t steps_done;
  uint64_t device_bytes; /* device memory held by the handle */
  uint32_t resort_interval, graph_chunk;
} slm_info;

/* Per-kernel device time (HIP events on the handle's stream). */
typedef struct slm_timing {
  double agents_ms;   /* total over the call */
  double trail_ms;    /* fused deposit + decay + diffuse */
  double sort_ms;     /* Morton re-sorts */
  uint64_t agents_launches, trail_launches, sort_launches;
} slm_timing;

typedef struct slm_sim slm_sim;

/* Thread-local message for the most recent failing call (never NULL). */
const char* slm_last_error(void);

/* Number of HIP devices visible (0 when none). */
int slm_device_count(int* out_count);

/* Fill *cfg with defaults: 1600x900 (src/settings.rs:4-5), 0 agents, seed 0,
 * f32 storage, device 0, no re-sort, 100-step graphs, 16-cell sort bins. */
int slm_config_default(slm_config* cfg);

/* Fill *params with the reference's defaults (src/settings.rs:9-27). */
int slm_params_default(slm_params* params);

/* Allocate a simulation: agents initialised on the device from cfg->seed as
 * src/main.rs:262-282 does on the host; trail zero (src/main.rs:286-293,
 * wgpu zero-initialises new buffers).  Replaces the buffer and pipeline
 * creation of src/main.rs:263-323 and PipelineManager::new
 * (src/pipeline_manager.rs:20-75). */
int slm_create(const slm_config* cfg, const slm_params* params, slm_sim** out);

/* Release everything the handle owns (waits for its stream). */
int slm_destroy(slm_sim* sim);

/* Replace the parameters for subsequent steps; agents and trail persist.
 * Replaces update_settings / the preset switch (src/main.rs:83-99, :815-832).
 * Captured graphs stay valid: kernels read parameters from device memory.
 * One exception, with SLM_FLAG_GAUSSIAN_DIFFUSE: a 16-bit pool's Gaussian pass
 * sizes its LDS tile (and its tile height) for the largest radius class
 * (R <= 2, <= 4, <= 8) of its instances, so a change of class waits for the
 * handle's stream, drops the captured graphs (the next slm_step recaptures
 * them) and, when the tile height changes, rebuilds the tile plan. */
int slm_set_params(slm_sim* sim, const slm_params* params);
int slm_get_params(const slm_sim* sim, slm_params* out);

/* Enqueue n steps: agents (src/compute.wgsl:57-145) then the fused deposit +
 * decay + diffuse pass (:147-195).  Replaces the dispatch block
 * src/main.rs:1163-1235 (the display pass there only reads the trail). */
int slm_step(slm_sim* sim, uint32_t n_steps);

/* As slm_step, with HIP events around every kernel; waits and adds the
 * per-kernel device times to *out (zeroed first).  Whole graph chunks replay
 * captured graphs holding the event records (the launch path of slm_step);
 * the remaining steps are launched eagerly. */
int slm_step_timed(slm_sim* sim, uint32_t n_steps, slm_timing* out);

/* Wait for all work enqueued on the handle's stream. */
int slm_synchronize(slm_sim* sim);

/* Zero the trail (src/main.rs:908-913). */
int slm_clear(slm_sim* sim);

/* New grid size: x *= new_w/old_w and y *= new_h/old_h in f32, trail
 * reallocated and zeroed (src/main.rs:921-1085).  Unlike the reference the
 * current parameters are kept (it re-uploads the start-up decay factor). */
int slm_resize(slm_sim* sim, uint32_t width, uint32_t height);

/* New agent count: every agent re-initialised from the seed, as the N key
 * does (src/main.rs:682-791), using the current speed range. */
int slm_set_agent_count(slm_sim* sim, uint64_t agent_count);

/* speed = min + u*(max-min) for every agent, u from a fresh generator stream
 * per call (src/main.rs:101-145). */
int slm_reassign_speeds(slm_sim* sim, float speed_min, float speed_max);

/* Copy agents out/in as n (x, y, angle, speed) f32 records; n must equal the
 * agent count.  Written positions must be finite and inside [0,width] x
 * [0,height]; angle and speed must be finite. */
int slm_read_agents(slm_sim* sim, float* agents_xyzw, uint64_t n);
int slm_write_agents(slm_sim* sim, const float* agents_xyzw, uint64_t n);

/* Copy the trail out/in as width*height f32.  Written values must lie in
 * [0, 2^100] (-0 is stored as +0); 16-bit storage rounds them to nearest-even. */
int slm_read_trail(slm_sim* sim, float* trail, uint64_t cells);
int slm_write_trail(slm_sim* sim, const float* trail, uint64_t cells);

/* Sum of the trail, accumulated in f64 on the device in a fixed order. */
int slm_total_mass(slm_sim* sim, double* out);

int slm_get_info(const slm_sim* sim, slm_info* out);

/* ---- batches ---------------------------------------------------------------
 * Many independent simulations of uneven sizes advanced together: every step
 * is one agent launch and one trail launch over all instances, driven by an
 * instance-offset table (no per-instance launches).  Instance k of a batch
 * gives bit for bit the result of a single simulation created with the same
 * width, height, agent count, seed and parameters.  No counterpart in the
 * reference, which runs one simulation per process (src/main.rs:147). */
typedef struct slm_instance_desc {
  uint32_t width, height; /* as slm_config */
  uint64_t agent_count;
  uint64_t seed;
  slm_params params;
} slm_instance_desc;

typedef struct slm_batch_config {
  int32_t storage;          /* slm_storage, shared by all instances */
  int32_t device;
  uint32_t resort_interval; /* as slm_config */
  uint32_t graph_chunk;
  uint32_t sort_cell_log2;
  uint32_t flags;           /* SLM_FLAG_* */
  uint32_t reserved[6];     /* must be zero */
} slm_batch_config;

typedef struct slm_batch_info {
  uint32_t count;           /* instances */
  int32_t storage;
  uint64_t total_agents;
  uint64_t total_cells;
  uint64_t steps_done;
  uint64_t device_bytes;
} slm_batch_info;

typedef struct slm_batch slm_batch;

int slm_batch_config_default(slm_batch_config* cfg);
/* 1 <= count <= 65535 instances; total agents < 2^32, total cells <= 2^40. */
int slm_batch_create(const slm_batch_config* cfg, const slm_instance_desc* instances,
                     uint32_t count, slm_batch** out);
int slm_batch_destroy(slm_batch* batch);
int slm_batch_step(slm_batch* batch, uint32_t n_steps);
int slm_batch_step_timed(slm_batch* batch, uint32_t n_steps, slm_timing* out);
int slm_batch_synchronize(slm_batch* batch);
int slm_batch_set_params(slm_batch* batch, uint32_t index, const slm_params* params);
int slm_batch_clear(slm_batch* batch, uint32_t index);
int slm_batch_read_agents(slm_batch* batch, uint32_t index, float* agents_xyzw, uint64_t n);
int slm_batch_write_agents(slm_batch* batch, uint32_t index, const float* agents_xyzw,
                           uint64_t n);
int slm_batch_read_trail(slm_batch* batch, uint32_t index, float* trail, uint64_t cells);
int slm_batch_write_trail(slm_batch* batch, uint32_t index, const float* trail, uint64_t cells);
/* f64 trail sums of all instances: out[k] for instance k, count == instances. */
int slm_batch_total_mass(slm_batch* batch, double* out, uint32_t count);
int slm_batch_get_info(const slm_batch* batch, slm_batch_info* out);

/* ---- launch tuning and the autotuner ---------------------------------------
 * How the step is laid out on the GPU.  None of these settings changes a
 * result bit; they only move time between kernels.  No counterpart in the
 * reference, whose dispatch shape is fixed (agents: 64-wide workgroups,
 * src/compute.wgsl:57 and src/main.rs:1168-1179; decay and diffuse: 16x16
 * workgroups, src/compute.wgsl:148,164 and src/main.rs:1186-1233). */
typedef struct slm_tuning {
  uint32_t resort_interval;  /* Morton re-sort of agents every k steps; 0 = never */
  uint32_t sort_cell_log2;   /* re-sort bin edge = 2^k cells, k <= 12 */
  uint32_t agents_per_group; /* agents one workgroup of the agent pass updates:
                                256, 512, 1024, 2048 or 4096 */
  uint32_t trail_rows;       /* rows of one trail-pass tile (64*vector-width
                                columns): 4, 8, 16, 32 or 64; 0 = chosen from the
                                grid so the launch has >= 8192 waves */
  uint32_t tiled_sensing;    /* 1 = SLM_FLAG_TILED_SENSING (needs re-sorting),
                                2 = SLM_FLAG_WINDOW_SENSING, 0 = neither */
  uint32_t trials;           /* autotune: configurations timed; else 0 */
  uint32_t instance_groups;  /* 1..4096 (0 reads as 1): between re-sorts the
                                instances step as this many contiguous groups
                                of about equal cost (agents + cells/8), dealt
                                to 4 streams (parallel branches of the
                                captured graph), so kernels of different
                                groups overlap and each covers a few
                                instances.  More groups than instances act as
                                one group per instance.  Not with
                                tiled_sensing (SLM_ERR_UNSUPPORTED). */
  uint32_t streaming;        /* trail-pass row accesses streamed past the
                                caches: bit 0 non-temporal stores, bit 1
                                non-temporal loads (0..3).  Pays for maps
                                beyond the 256 MB Infinity Cache in f32
                                (8192^2: 3 gives 75% of HBM peak against 72%
                                with 0); costs on 16384^2 bf16 */
  uint32_t sort_key;         /* re-sort key refinement: the agents of a bin
                                are further split into 2^k heading classes
                                (k = sort_key & SLM_SORT_HEADING_MASK, 0..5),
                                and with SLM_SORT_AT_SENSOR the bin is the one
                                holding the point the centre sensor reads
                                instead of the agent's own cell.  Either
                                gathers each wave's sensor taps onto fewer
                                cache lines.  Not with tiled_sensing = 1
                                (SLM_ERR_UNSUPPORTED). */
  /* autotune, for the chosen tuning: step_ms = device time per step from the
     first launch to the end of the last (eager launches, gaps included);
     agents/trail/sort_ms = per-step sums of those kernels' own times (sorts
     amortised over their interval); else 0 */
  double step_ms, agents_ms, trail_ms, sort_ms;
} slm_tuning;

#define SLM_SORT_HEADING_MASK 0x7u
#define SLM_SORT_AT_SENSOR 0x10u

int slm_get_tuning(const slm_sim* sim, slm_tuning* out);
/* Applies a tuning.  Agent and trail state, the step count and the parameters
 * are kept; captured graphs are rebuilt on the next slm_step. */
int slm_set_tuning(slm_sim* sim, const slm_tuning* tuning);
/* Times candidate tunings on the simulation's own state and keeps the
 * fastest: the trail tile rows by trail-pass time, streaming (0..3), then the
 * re-sort interval
 * (0, 5, 10, 20 or 40), the bin size, the sensing mode (0, 1, 2), agents per workgroup and
 * (batches) instance groups (powers of 4 up to the instance count, then the
 * best one's neighbours x2 and /2) by total step time (coordinate descent).
 * Step times are spans over steps launched as slm_step launches them; kernel
 * times come from further steps with events around each whole-pool kernel.  Every trial starts from a copy of
 * the current state, warms up 40 steps and is timed over the next
 * `trial_steps` steps rounded up to a multiple of 40, so all candidates see
 * the same stretch of the simulation.  The state the call
 * found is restored afterwards, so stepping on gives the same bits as without
 * the call.  out (may be NULL) receives the chosen tuning and its timing.
 * Needs device memory for one extra copy of the agents and the trail, plus the
 * re-sort buffers when re-sorting was off; candidates that do not fit are
 * skipped. */
int slm_autotune(slm_sim* sim, uint32_t trial_steps, slm_tuning* out);

int slm_batch_get_tuning(const slm_batch* batch, slm_tuning* out);
int slm_batch_set_tuning(slm_batch* batch, const slm_tuning* tuning);
int slm_batch_autotune(slm_batch* batch, uint32_t trial_steps, slm_tuning* out);

/* ---- display colour map ------------------------------------------------------
 * Replaces the display compute pass (src/display.wgsl:29-87, dispatched at
 * src/main.rs:1199-1215) and the LUT buffer it reads (src/main.rs:330-342).
 * lut: 768 bytes, R[256] then G[256] then B[256], the layout of the reference's
 * .lut files and LutData (src/lut_manager.rs:4-18,162-186).  Renders the
 * current trail (as it stands after the last step, i.e. after the diffuse;
 * slm_step_render_rgba8 below renders at the reference's own sampling point)
 * into a tex_w x tex_h RGBA8 image, row-major, 4 bytes per pixel
 * (r, g, b, 255): the grid is fitted into the texture keeping its aspect ratio
 * and centred (f32 scale and offsets as the shader computes them), each pixel
 * takes its cell by truncation, the intensity clamp(T, 0, 1) selects table
 * entry u32(intensity * 255); pixels outside the grid are (0, 0, 0, 255).
 * 1 <= tex_w, tex_h <= 2^24 and tex_w * tex_h <= 2^30. */
int slm_render_rgba8(slm_sim* sim, const uint8_t* lut, uint32_t tex_w, uint32_t tex_h,
                     uint8_t* rgba);
/* The same into device memory (tex_w * tex_h * 4 bytes on the handle's
 * device), enqueued on the handle's stream: slm_synchronize waits for it. */
int slm_render_rgba8_device(slm_sim* sim, const uint8_t* lut, uint32_t tex_w, uint32_t tex_h,
                            void* rgba_device);
int slm_batch_render_rgba8(slm_batch* batch, uint32_t index, const uint8_t* lut, uint32_t tex_w,
                           uint32_t tex_h, uint8_t* rgba);
/* slm_render_rgba8 without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (a display loop that
 * renders every step, code under a harness that forbids hipMalloc).  It does
 * what slm_render_rgba8 does, with the same bytes, but stages the image in a
 * fixed per-GPU device array of SLM_RENDER_NOALLOC_MAX_BYTES (16 MiB, so
 * tex_w * tex_h <= 4194304 pixels; larger textures return
 * SLM_ERR_INVALID_ARGUMENT) that the library holds until the process ends,
 * instead of a buffer allocated and freed per call.  Every call on one GPU
 * shares that array and keeps the last image in it (each call writes every
 * pixel it returns, so nothing needs clearing): calls that overlap on one GPU
 * - from different threads, whatever the handle - race, and the caller must
 * serialise them.  Waits for the image. */
#define SLM_RENDER_NOALLOC_MAX_BYTES (16u << 20)
int slm_render_rgba8_noalloc(slm_sim* sim, const uint8_t* lut, uint32_t tex_w, uint32_t tex_h,
                             uint8_t* rgba);
/* The calls above show the trail after the last step's diffuse.  The
 * reference's frame displays it earlier: agents, decay, display, diffuse
 * (src/main.rs:1163-1235; the display pass at :1199-1215 samples the trail
 * between the decay and diffuse submits).  slm_step_render_rgba8 enqueues
 * n_steps >= 1 steps, as slm_step would, and renders the last one at that
 * point: each pixel's cell shows clamp(T + deposit, 0, 1) where an agent of
 * that step landed, else T, minus the decay and floored at 0 - the trail
 * after deposit and decay, before the diffuse (rendered in f32, not rounded
 * to the storage format).  The image is the one the reference would show
 * for that frame; the state afterwards equals slm_step(n_steps)'s.  The last
 * step is launched eagerly (not from a captured graph).  The host form waits
 * for the image; the device form only enqueues. */
int slm_step_render_rgba8(slm_sim* sim, uint32_t n_steps, const uint8_t* lut, uint32_t tex_w,
                          uint32_t tex_h, uint8_t* rgba);
int slm_step_render_rgba8_device(slm_sim* sim, uint32_t n_steps, const uint8_t* lut,
                                 uint32_t tex_w, uint32_t tex_h, void* rgba_device);

/* ---- snapshots ------------------------------------------------------------------
 * A resumable copy of a single simulation.  Loading it into a simulation of
 * the same shape (grid, agent count, storage, position precision), or creating
 * one from it, and stepping on gives bit for bit the result of the run that
 * was never interrupted.  No counterpart in the reference, which keeps no
 * state across runs.
 * Layout, little-endian: a 128-byte header
 *   char magic[8] = "SLMSNAP1"; uint32 version = 1; uint32 header_bytes = 128;
 *   uint32 width, height; uint64 agent_count; uint64 seed; int32 storage;
 *   uint32 flags (SLM_FLAG_F64_POSITIONS or 0); uint64 steps_done;
 *   uint64 speed_reassignments; slm_params params (44 bytes); uint32 0;
 *   uint64 agent_bytes; uint64 trail_bytes;
 * then agent_count f32 records (x, y, angle, speed) in original order; then,
 * with f64 positions, agent_count f64 (x, y) pairs; then width*height trail
 * elements in the storage format, row-major; then a uint64 checksum, FNV-1a
 * over all preceding bytes taken as little-endian 64-bit words (the last one
 * zero-padded). */
int slm_snapshot_size(const slm_sim* sim, uint64_t* bytes);
int slm_save_snapshot(slm_sim* sim, void* buf, uint64_t bytes);
int slm_load_snapshot(slm_sim* sim, const void* buf, uint64_t bytes);
/* A new simulation (default launch tuning, 100-step graphs) holding the
 * snapshot's state. */
int slm_create_from_snapshot(const void* buf, uint64_t bytes, int32_t device, slm_sim** out);


#ifdef __cplusplus
}
#endif

#endif /* SLIMEMOLD_H */
