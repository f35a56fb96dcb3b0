// slm_internal.h - device-side data layout and kernel launchers of
// libslimemold_hip.so.  Not part of the public ABI (see include/slimemold.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
};

DevParams make_dev_params(const slm_params& p);

// Agents in structure-of-arrays form.  id[j] is the original index of the
// agent stored at slot j (the jitter hash and slm_read_agents use it); it is
// only maintained when re-sorting is enabled, otherwise id == nullptr and the
// slot is the original index.
struct AgentSoA {
  float* x;
  float* y;
  float* th;
  float* v;
  uint32_t* id;
};

struct StepGeom {
  uint32_t width;
  uint32_t height;
  uint32_t agents;
  int storage;   // slm_storage
  int xcd_remap;  // 1: group consecutive logical blocks on one XCD
};

// ---- launchers (all asynchronous on `stream`) ----------------------------
void launch_agents(const StepGeom& g, const AgentSoA& a, void* trail, const DevParams* p,
                   hipStream_t stream);
void launch_trail(const StepGeom& g, const void* trail_in, void* trail_out, const DevParams* p,
                  hipStream_t stream);
void launch_set_params(DevParams* dst, const DevParams& src, hipStream_t stream);
void launch_init_agents(const AgentSoA& a, uint32_t n, uint64_t seed, uint32_t width,
                        uint32_t height, float speed_min, float speed_max, hipStream_t stream);
void launch_reassign_speeds(const AgentSoA& a, uint32_t n, uint64_t stream_seed,
                            float speed_min, float speed_max, hipStream_t stream);
void launch_scale_positions(const AgentSoA& a, uint32_t n, float sx, float sy,
                            hipStream_t stream);
void launch_pack_agents(const AgentSoA& a, uint32_t n, float* aos, hipStream_t stream);
void launch_unpack_agents(const AgentSoA& a, uint32_t n, const float* aos, hipStream_t stream);
void launch_decode_trail(const void* trail, int storage, uint64_t cells, float* out,
                         hipStream_t stream);
void launch_encode_trail(const float* in, int storage, uint64_t cells, void* trail,
                         hipStream_t stream);
// Deterministic two-pass f64 sum; `partials` holds kMassBlocks doubles.
constexpr int kMassBlocks = 1024;
void launch_total_mass(const void* trail, int storage, uint64_t cells, double* partials,
                       double* out, hipStream_t stream);

// ---- Morton re-sort ------------------------------------------------------
struct SortScratch {
  uint32_t* key;      // per agent: bin (Morton code)
  uint32_t* rank;     // per agent: rank inside its bin
  uint32_t* count;    // per bin
  uint32_t* offset;   // per bin, exclusive prefix sum
  uint32_t* block_sums;
  uint32_t nbins;     // power of four
  uint32_t cell_log2; // bin edge = 2^cell_log2 cells
};
uint32_t sort_bins_for(uint32_t width, uint32_t height, uint32_t cell_log2);
uint32_t sort_block_sums_for(uint32_t nbins);
void launch_morton_sort(const StepGeom& g, const AgentSoA& src, const AgentSoA& dst,
                        const SortScratch& s, hipStream_t stream);

inline uint64_t storage_bytes(int storage) { return storage == SLM_STORAGE_F32 ? 4 : 2; }

}  // namespace slm
