// slm_internal.h - device-side data layout and kernel launchers of
// libslimemold_hip.so.  Not part of the public ABI (see include/slimemold.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/slimemold.h"

namespace slm {

// Per-simulation parameters as the kernels read them, from device memory so a
// captured graph picks up slm_set_params without re-capture.  Derived values
// are computed on the host with the same f32 operations the reference shader
// does (compute.wgsl:159 decay scale, :173 rate clamp, mix's 1 - t).
struct DevParams {
  float decay_k;         // decay_factor * 0.001f
  float jitter;          // agent_jitter
  float speed_min;       // agent_speed_min
  float speed_max;       // agent_speed_max
  float turn;            // agent_turn_speed
  float sensor_angle;    // agent_sensor_angle
  float sensor_dist;     // agent_sensor_distance
  float rate;            // clamp(diffusion_rate, 0, 1)
  float one_minus_rate;  // 1.0f - rate
  float deposit;         // pheromone_deposition_amount
  float blur_radius;
  float blur_sigma;
  // Gaussian diffuse mode: integer radius round(blur_radius) in 1..kGaussMaxR
  // (0 when blur_radius / blur_sigma do not describe one) and the normalised
  // weights w[i] = exp(-i^2 / (2 sigma^2)) / sum, i = 0..radius, computed in
  // double and rounded to f32.
  int gauss_r;
  float gauss_w[9];
};
constexpr int kGaussMaxR = 8;
// The Gaussian pass's LDS tile is sized for the largest radius of a pool,
// rounded up to 2, 4 or kGaussMaxR (gauss_class): the presets' radius 2 then
// needs 37 KiB per bf16 tile instead of 52, and four workgroups fit a CU.
constexpr int gauss_class(int rmax) { return rmax <= 2 ? 2 : (rmax <= 4 ? 4 : kGaussMaxR); }

// Gaussian weights for blur_radius / blur_sigma; returns the radius, or 0 if
// they do not describe a Gaussian of radius 1..kGaussMaxR.
int gauss_weights(float blur_radius, float blur_sigma, float* w);

DevParams make_dev_params(const slm_params& p);

// Agents in structure-of-arrays form.  Positions are f32 (x, y) or, in the
// f64-position mode, f64 (x64, y64).  id[j] is the original (instance-local)
// index of the agent stored at slot j; the jitter hash and slm_read_agents
// use it.  It is only kept when re-sorting is enabled; otherwise id == nullptr
// and the slot minus the instance's first slot is the original index.
struct AgentSoA {
  float* x;
  float* y;
  double* x64;
  double* y64;
  float* th;
  float* v;
  uint32_t* id;
};

// One simulation inside a pool of several (a single simulation is a pool of
// one).  Offsets index the pooled agent arrays (slots) and trail buffers
// (elements).
struct Instance {
  uint32_t width, height;
  uint32_t agent_off, agent_count;  // slots [agent_off, agent_off + agent_count)
  uint64_t trail_off;               // first element of this instance's trail
  uint32_t unit_off, units;         // trail-pass wave units [unit_off, unit_off + units)
  uint32_t tile_rows, ntx;          // strip tiles (ntx > 0) or 64-cell units (ntx == 0)
  uint32_t bin_off;                 // re-sort keys [bin_off, bin_off + nbins): bins x heading classes
  uint32_t nbins;
  uint32_t nbx, nby;                // bins across / down
  uint32_t tile_log2;               // sensing tiles of 2^k x 2^k cells (aligned bin squares)
  uint32_t tile_off, tiles;         // tiles [tile_off, tile_off + tiles), Morton order
  uint32_t gtile_off, gtx;          // Gaussian-pass tiles [gtile_off, ...), gtx across
};

// One workgroup's share of the tiled agent pass: agent slots [begin, end) of
// tile `tile` (instance-local Morton index) of instance `inst`.
struct WorkItem {
  uint32_t inst, tile, begin, end;
};
constexpr uint32_t kItemAgents = 4096;  // agents per work item at most

// A launchable description of a pool, with device copies of its tables.
struct Pool {
  uint32_t count;            // instances
  const Instance* inst;      // device, count entries
  const DevParams* params;   // device, count entries
  uint32_t total_agents;     // slots
  uint32_t total_units;      // trail-pass wave units
  uint32_t total_bins;
  const uint32_t* agent_block_inst;  // device: first instance of each agent block
  const uint32_t* unit_block_inst;   // device: first instance of each trail block
  int storage;    // slm_storage
  int pos64;      // f64 positions
  int xcd_remap;  // group consecutive logical blocks on one XCD
  uint32_t sort_cell_log2;
  uint32_t sort_key;           // heading classes (low bits) and binning at the centre sensor
  // tiled sensing (valid after a re-sort has built the work list)
  uint32_t total_tiles;
  const uint32_t* tile_inst;   // device: instance of each tile
  WorkItem* items;             // device: work list, at most max_items entries
  uint32_t* item_total;        // device: number of valid work items
  uint32_t max_items;
  uint32_t resort_interval;    // steps agents may drift from their tile between re-sorts
  uint32_t agents_per_group;   // agents per workgroup of the global agent pass (k * kAgentBlock)
  uint32_t preload_rows;       // 4, 8 or 16 when every strip instance has tiles of that height, else 0
  int stochastic_rounding;     // bf16 trail-pass stores round stochastically (deterministic hash)
  int streaming;               // trail-pass rows: 1 non-temporal stores, 2 non-temporal loads
  int gaussian;                // trail pass diffuses with the Gaussian, not the 3x3 box
  int gauss_rclass;            // Gaussian pass LDS sized for radii up to 2, 4 or 8
  int window_sensing;          // one-agent-per-thread workgroups sense through an LDS window
  uint32_t* claim_agents;      // persistent agent / trail passes: kClaimWords counters each
  uint32_t* claim_trail;       // (nullptr: one workgroup per logical workgroup)
  uint32_t total_gtiles;       // Gaussian-pass tiles of the pool
};

// Gaussian-pass tiles: kGaussTW columns x gauss_th(storage, radius class)
// rows per workgroup.  Taller tiles re-read and re-filter fewer halo rows but
// hold more LDS: 16 rows for f32 (HBM-bound; measured 72.6% of peak against
// 70.5% with 32); for the 16-bit formats (VALU-bound) 16 rows up to radius 4
// (five to seven workgroups per CU: 230.8 against 241.1 us at radius 2) and
// 32 rows up to radius 8 (597.8 against 686.6 us at radius 8, 16384^2 bf16).
constexpr uint32_t kGaussTW = 256;
constexpr uint32_t gauss_th(int storage, int rclass) {
  return storage == SLM_STORAGE_F32 || rclass <= 4 ? 16u : 32u;
}

constexpr uint32_t kTileLog2 = 6;  // sensing tiles of 64 x 64 cells

constexpr int kAgentBlock = 256;
constexpr int kTrailBlock = 256;
constexpr int kTrailWavesPerBlock = kTrailBlock / 64;

// Workgroups of one launch are dealt round-robin to the 8 XCDs; this gives
// XCD k a contiguous range of logical blocks (a bijection on [0, nb)), so
// neighbouring blocks' data share that XCD's L2.
__device__ inline uint32_t xcd_logical_block(uint32_t b, uint32_t nb) {
  uint32_t xcd = b & 7u, idx = b >> 3;
  uint32_t per = nb >> 3, rem = nb & 7u;
  return xcd < rem ? xcd * (per + 1u) + idx : rem * (per + 1u) + (xcd - rem) * per + idx;
}

// ---- persistent launches -------------------------------------------------------
// A persistent pass launches only the workgroups that are resident at once
// (persistent_grid); each claims logical workgroups of the pass one at a time
// from device counters until none are left, so every workgroup exits when its
// claims run out.  Workgroup b claims from the contiguous range of logical
// workgroups that xcd_logical_block gives XCD b % 8 in a one-workgroup-per-
// block launch, so neighbouring work still shares an L2.  claim[0..7] count
// the claims per range, claim[8] the workgroups that have finished claiming;
// the last of those resets all nine to zero, ready for the next launch (they
// are zeroed once when allocated), so no memset is needed between launches.
constexpr int kClaimWords = 9;

template <class Body>
__device__ inline void persistent_blocks(uint32_t* claim, uint32_t nb, uint32_t* s_claim,
                                         Body&& body) {
  const uint32_t xcd = blockIdx.x & 7u;
  const uint32_t per = nb >> 3, rem = nb & 7u;
  const uint32_t lo = xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per;
  const uint32_t cnt = per + (xcd < rem ? 1u : 0u);
  for (;;) {
    if (threadIdx.x == 0) *s_claim = atomicAdd(&claim[xcd], 1u);
    __syncthreads();
    const uint32_t k = *s_claim;
    __syncthreads();
    if (k >= cnt) break;  // uniform
    body(lo + k);
  }
  if (threadIdx.x == 0 && atomicAdd(&claim[8], 1u) == gridDim.x - 1u)
    for (int x = 0; x < kClaimWords; x++) atomicExch(&claim[x], 0u);
}

// Workgroups of `kernel` resident on the whole device at once (at least 8, a
// multiple of 8, and no more than the pass's nb logical workgroups rounded up
// to 8), for a persistent launch.
// Resident workgroups per kernel (0 until measured), keyed by the kernel.
inline int& resident_of(const void* kernel) {
  static std::mutex mu;
  static const void* keys[128];
  static int vals[128];
  static int n = 0;
  static int overflow = 0;
  std::lock_guard<std::mutex> lock(mu);
  for (int i = 0; i < n; i++)
    if (keys[i] == kernel) return vals[i];
  if (n == 128) return overflow = 0;
  keys[n] = kernel;
  vals[n] = 0;
  return vals[n++];
}

template <class K>
uint32_t persistent_grid(K* kernel, uint32_t nb, int block) {
  int& resident = resident_of(reinterpret_cast<const void*>(kernel));
  if (resident <= 0) {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess)
      per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 8;
    (void)hipGetLastError();
    resident = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 8);
  }
  uint64_t g = static_cast<uint64_t>(resident) < nb ? static_cast<uint64_t>(resident) : nb;
  g = (g + 7u) & ~uint64_t(7u);
  return static_cast<uint32_t>(g < 8u ? 8u : g);
}

// Fill persistent_grid's per-kernel caches outside any stream capture (the
// engine calls these when it is created with persistent passes).
void prime_persistent_agents();
void prime_persistent_trail();

inline uint64_t storage_bytes(int storage) { return storage == SLM_STORAGE_F32 ? 4 : 2; }
inline uint32_t blocks_for(uint64_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

// Host-side layout of one instance's trail pass: fills tile_rows, ntx, units.
// rows > 0 fixes the tile height; 0 takes the tallest of 64..4 rows that still
// gives target_waves units.
void plan_trail(Instance& in, int storage, uint32_t target_waves, uint32_t rows);
// Host-side layout of one instance's re-sort bins: fills nbins (clamped to
// 2^31), nbx, nby, and returns the exact bin count.
uint64_t plan_bins(Instance& in, uint32_t cell_log2);
// Tiles of 2^tile_log2 cells (>= the bin edge) for tiled sensing: fills
// tile_log2 and tiles.
void plan_tiles(Instance& in, uint32_t cell_log2, uint32_t tile_log2);

// ---- launchers (all asynchronous on `stream`) ----------------------------
void launch_agents(const Pool& pool, const AgentSoA& a, void* trail, hipStream_t stream);
// The same update with each tile's trail window staged in LDS (needs the work
// list of the latest re-sort); bit-identical to launch_agents.
void launch_agents_tiled(const Pool& pool, const AgentSoA& a, void* trail, hipStream_t stream);
// LDS bytes the tiled pass stages per workgroup for this storage format.
constexpr uint32_t kWindowBytes = 65536;
uint32_t window_side(int storage);
void launch_trail(const Pool& pool, const void* trail_in, void* trail_out, hipStream_t stream);
void launch_set_params(DevParams* dst, const DevParams& src, hipStream_t stream);
void launch_init_agents(const Instance& in, const AgentSoA& a, int pos64, uint64_t seed,
                        float speed_min, float speed_max, hipStream_t stream);
void launch_reassign_speeds(const Instance& in, const AgentSoA& a, uint64_t stream_seed,
                            float speed_min, float speed_max, hipStream_t stream);
void launch_scale_positions(const Instance& in, const AgentSoA& a, int pos64, float sx, float sy,
                            hipStream_t stream);
void launch_pack_agents(const Instance& in, const AgentSoA& a, int pos64, float* aos,
                        hipStream_t stream);
void launch_unpack_agents(const Instance& in, const AgentSoA& a, int pos64, const float* aos,
                          hipStream_t stream);
void launch_decode_trail(const void* trail, int storage, uint64_t cells, float* out,
                         hipStream_t stream);
void launch_encode_trail(const float* in, int storage, uint64_t cells, void* trail,
                         hipStream_t stream);
// Deterministic two-pass f64 sum per instance; `partials` holds
// count * kMassBlocks doubles and `out` count doubles.
constexpr int kMassBlocks = 256;
void launch_total_mass(const Pool& pool, const void* trail, double* partials, double* out,
                       hipStream_t stream);

// ---- display colour map --------------------------------------------------
struct LutBytes {
  uint32_t word[192];  // the 768 table bytes R[256] G[256] B[256], little-endian words
};
struct RenderArgs {
  const void* trail;  // one instance's trail
  // nullptr: the stored values (after the diffuse).  Else the instance's
  // parameters, and each sampled cell shows its deposit and decay applied to
  // the raw element - the trail as the reference displays it, between its
  // decay and diffuse passes (src/main.rs:1183-1217) - when `trail` is the
  // buffer the agent pass has just marked.
  const DevParams* decode;
  uint32_t width;     // cells per trail row
  float sim_w, sim_h;
  uint32_t tex_w, tex_h;
  float scale, off_x, off_y;
  uint32_t* out;      // tex_w * tex_h RGBA8 pixels, row-major
};
// Fills the fit fields of `a` as src/display.wgsl:48-72 computes them.
void render_fit(uint32_t width, uint32_t height, uint32_t tex_w, uint32_t tex_h, RenderArgs& a);
void launch_render(const RenderArgs& a, const LutBytes& lut, int storage, hipStream_t stream);

// ---- snapshots: f64 positions in original order ----------------------------
void launch_pack_positions64(const Instance& in, const AgentSoA& a, double* xy,
                             hipStream_t stream);
void launch_unpack_positions64(const Instance& in, const AgentSoA& a, const double* xy,
                               hipStream_t stream);

// ---- Morton re-sort ------------------------------------------------------
struct SortScratch {
  uint32_t* rank;      // per slot: rank inside its bin (the bin is recomputed from x, y)
  uint32_t* count;     // per bin
  uint32_t* offset;    // per bin, exclusive prefix sum; offset[total_bins] = total agents
  uint32_t* tile_items;  // per tile: work items, then their exclusive prefix sum
  uint32_t* level[3];  // tile sums of the multi-level scan
  uint32_t level_len[3];
  int levels;
};
// Sizes the scan levels for `nbins` bins; returns the number of u32 the
// levels need in total (they are carved out of one allocation).
uint32_t plan_scan(SortScratch& s, uint32_t nbins);
void launch_morton_sort(const Pool& pool, const AgentSoA& src, const AgentSoA& dst,
                        const SortScratch& s, hipStream_t stream);
// Copies every agent of the pool from src into dst in original order (slot
// agent_off + id, or the same slot when src has no id table); dst.id, if
// present, receives the identity.  src and dst must not overlap.
void launch_canonical_copy(const Pool& pool, const AgentSoA& src, const AgentSoA& dst,
                           hipStream_t stream);

}  // namespace slm
