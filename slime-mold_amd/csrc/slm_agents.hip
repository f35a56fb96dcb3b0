// slm_agents.hip - the agent update (reference src/compute.wgsl:57-145) and the
// device-side agent initialisation / speed reassignment / resize scaling that
// replace the reference's host loops (src/main.rs:101-145, :262-282, :955-997).
//
// One lane per agent over structure-of-arrays state.  Sensing reads the trail
// as it was when the step began: a deposit does not write a value, it sets the
// sign bit of the target cell (trail values are never negative, so the bit is
// free); sensing reads |T|, and the trail pass that follows turns marked cells
// into clamp(|T| + deposit, 0, 1).  The mark is a plain load and store of that
// one cell, not an atomic: every agent landing in a cell stores the same bits
// (-|T|), and a sensor reading the cell meanwhile sees T or -T, whose |.| is the
// same, so the race is benign.  The result is therefore independent of the
// order agents run in.  (A global atomic OR per agent cost ~0.45 ms per 2^24
// agents and serialised badly once agents are spatially sorted.)
//
// Arithmetic follows the reference shader operation by operation: this file is
// compiled with -ffp-contract=off, division is IEEE, '%' is a - b*trunc(a/b),
// and sin/cos come from slm_trig.h.
#include "slm_internal.h"
#include "slm_trig.h"

namespace slm {
namespace {

constexpr float kTau = 6.28318530718f;    // compute.wgsl:4 and :121 (0x40c90fdb)
constexpr float kTimeStep = 0.016f;       // compute.wgsl:55
constexpr int kAgentBlock = 256;

// ---- trail access by storage format -------------------------------------
template <int S>
struct TrailRead;

// load_pair(t, i) returns |T[i]| and |T[i+1]| from one (4- or 2-byte aligned)
// load; gfx950 allows unaligned global loads, and the compiler only fuses the
// two cells into one instruction where the target permits it.
template <>
struct TrailRead<SLM_STORAGE_F32> {
  __device__ static void load_pair(const void* t, uint32_t i, float& a, float& b) {
    uint2 v;
    __builtin_memcpy(&v, static_cast<const uint32_t*>(t) + i, sizeof v);
    a = __uint_as_float(v.x & 0x7fffffffu);
    b = __uint_as_float(v.y & 0x7fffffffu);
  }
  __device__ static void mark(void* t, uint32_t i) {
    uint32_t* p = static_cast<uint32_t*>(t) + i;
    *p = *p | 0x80000000u;
  }
};

// 16-bit cells are marked with a 2-byte store, which leaves the neighbour in
// the same 32-bit word untouched.
__device__ inline void mark16(void* t, uint32_t i) {
  uint16_t* p = static_cast<uint16_t*>(t) + i;
  *p = static_cast<uint16_t>(*p | 0x8000u);
}

__device__ inline uint32_t load_u16_pair(const void* t, uint32_t i) {
  uint32_t w;
  __builtin_memcpy(&w, static_cast<const uint16_t*>(t) + i, sizeof w);
  return w;
}

template <>
struct TrailRead<SLM_STORAGE_BF16> {
  __device__ static void load_pair(const void* t, uint32_t i, float& a, float& b) {
    uint32_t w = load_u16_pair(t, i);
    a = __uint_as_float((w & 0x7fffu) << 16);
    b = __uint_as_float(((w >> 16) & 0x7fffu) << 16);
  }
  __device__ static void mark(void* t, uint32_t i) { mark16(t, i); }
};

template <>
struct TrailRead<SLM_STORAGE_F16> {
  __device__ static void load_pair(const void* t, uint32_t i, float& a, float& b) {
    uint32_t w = load_u16_pair(t, i);
    a = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0x7fffu)));
    b = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>((w >> 16) & 0x7fffu)));
  }
  __device__ static void mark(void* t, uint32_t i) { mark16(t, i); }
};

// ---- WGSL operations (S5) -------------------------------------------------
__device__ inline float wgsl_mix(float a, float b, float t) {
  float u = 1.0f - t;
  float l = a * u;
  float r = b * t;
  return l + r;
}
__device__ inline float wgsl_mod(float a, float b) {
  float q = a / b;
  float t = __builtin_truncf(q);
  float p = b * t;
  return a - p;
}
__device__ inline float wgsl_sign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
__device__ inline float wgsl_min(float a, float b) { return a < b ? a : b; }
__device__ inline float wgsl_clamp(float e, float lo, float hi) {
  float m = e > lo ? e : lo;
  return m < hi ? m : hi;
}

// Bilinear trail sample, not toroidal (compute.wgsl:7-29).  Positions are kept
// finite and within [-2^24, 2^24] by the API, so the i32 conversions are exact.
// Branch-free: an out-of-range sample loads cells 0 and 1 twice (the trail
// allocation is at least 256 bytes) and returns 0, so the six tap-pair loads of
// an agent's three sensors are all in flight at once instead of one sample's
// loads waiting behind the previous sample's branch.
template <int S>
__device__ inline float sample(const void* t, uint32_t w, uint32_t h, float px, float py) {
  int x0 = static_cast<int>(__builtin_floorf(px));
  int y0 = static_cast<int>(__builtin_floorf(py));
  bool in = x0 >= 0 && static_cast<uint32_t>(x0) < w - 1u && y0 >= 0 &&
            static_cast<uint32_t>(y0) < h - 1u;
  float dx = px - static_cast<float>(x0);
  float dy = py - static_cast<float>(y0);
  uint32_t i00 = in ? static_cast<uint32_t>(y0) * w + static_cast<uint32_t>(x0) : 0u;
  uint32_t down = in ? w : 0u;
  float v00, v10, v01, v11;
  TrailRead<S>::load_pair(t, i00, v00, v10);
  TrailRead<S>::load_pair(t, i00 + down, v01, v11);
  float v0 = wgsl_mix(v00, v10, dx);
  float v1 = wgsl_mix(v01, v11, dx);
  float v = wgsl_mix(v0, v1, dy);
  return in ? v : 0.0f;
}

// Workgroups of one launch are dealt round-robin to the 8 XCDs; give XCD k a
// contiguous range of logical blocks so neighbouring agents share its L2.
__device__ inline uint32_t xcd_logical_block(uint32_t b, uint32_t nb) {
  uint32_t xcd = b & 7u, idx = b >> 3;
  uint32_t per = nb >> 3, rem = nb & 7u;
  return xcd < rem ? xcd * (per + 1u) + idx : rem * (per + 1u) + (xcd - rem) * per + idx;
}

struct AgentArgs {
  AgentSoA a;
  void* trail;
  const DevParams* p;
  uint32_t w, h, n;
  int xcd_remap;
};

template <int S>
__global__ __launch_bounds__(kAgentBlock) void agents_kernel(AgentArgs args) {
  uint32_t blk = args.xcd_remap ? xcd_logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
  uint32_t j = blk * kAgentBlock + threadIdx.x;
  if (j >= args.n) return;
  const DevParams p = *args.p;
  const uint32_t w = args.w, h = args.h;

  float x = args.a.x[j], y = args.a.y[j], angle = args.a.th[j], speed = args.a.v[j];

  float sp = wgsl_clamp(speed, p.speed_min, p.speed_max);  // :72
  if (sp != speed) args.a.v[j] = sp;
  speed = sp;

  float al = angle - p.sensor_angle, ar = angle + p.sensor_angle;  // :75-77
  float sl, cl, sr, cr, sc, cc;
  slm_sincosf(al, &sl, &cl);
  slm_sincosf(ar, &sr, &cr);
  slm_sincosf(angle, &sc, &cc);
  float sd = p.sensor_dist;
  float lx = x + sd * cl, ly = y + sd * sl;  // :79-90
  float rx = x + sd * cr, ry = y + sd * sr;
  float cx = x + sd * cc, cy = y + sd * sc;
  float lv = sample<S>(args.trail, w, h, lx, ly);  // :93-95
  float rv = sample<S>(args.trail, w, h, rx, ry);
  float cv = sample<S>(args.trail, w, h, cx, cy);

  if (cv > lv && cv > rv) {  // :98-112
  } else if (lv > rv) {
    float target = angle - kTau;
    float diff = target - angle;
    angle += wgsl_min(p.turn, __builtin_fabsf(diff)) * wgsl_sign(diff);
  } else if (rv > lv) {
    float target = angle + kTau;
    float diff = target - angle;
    angle += wgsl_min(p.turn, __builtin_fabsf(diff)) * wgsl_sign(diff);
  }

  // :115-118.  With zero jitter the term is (finite) * 0 = +-0, which leaves
  // the angle unchanged, so the hash is skipped.
  if (p.jitter != 0.0f) {
    uint32_t index = args.a.id ? args.a.id[j] : j;  // original index (S6)
    float hx = static_cast<float>(static_cast<int32_t>(index)) * 12.9898f;
    hx = hx + x * 78.233f;
    hx = hx + y * 37.719f;
    float v = slm_sinf(hx) * 43758.5453f;
    float rnd = v - __builtin_floorf(v);
    angle += (rnd * 2.0f - 1.0f) * p.jitter;
  }

  angle = wgsl_mod(angle, kTau);  // :121-122
  if (angle < 0.0f) angle = angle + kTau;

  float move = speed * kTimeStep;  // :125-127
  float ms, mc;
  slm_sincosf(angle, &ms, &mc);
  x = x + move * mc;
  y = y + move * ms;

  float fw = static_cast<float>(w), fh = static_cast<float>(h);  // :130-133
  x = wgsl_mod(x, fw);
  if (x < 0.0f) x = x + fw;
  y = wgsl_mod(y, fh);
  if (y < 0.0f) y = y + fh;

  int dx = static_cast<int>(x), dy = static_cast<int>(y);  // :136-141
  if (dx >= 0 && static_cast<uint32_t>(dx) < w && dy >= 0 && static_cast<uint32_t>(dy) < h)
    TrailRead<S>::mark(args.trail, static_cast<uint32_t>(dy) * w + static_cast<uint32_t>(dx));

  args.a.x[j] = x;  // :144
  args.a.y[j] = y;
  args.a.th[j] = angle;
}

// ---- initialisation ------------------------------------------------------
// SplitMix64 (Steele, Lea & Flood 2014): output n of the generator seeded
// with `seed`; uniforms are its top 24 bits times 2^-24.
__device__ inline uint64_t splitmix64_at(uint64_t seed, uint64_t n) {
  uint64_t z = seed + (n + 1u) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ inline float uniform24(uint64_t seed, uint64_t n) {
  return static_cast<float>(static_cast<uint32_t>(splitmix64_at(seed, n) >> 40)) *
         (1.0f / 16777216.0f);
}

// src/main.rs:273-280, agent i drawing generator outputs 4i..4i+3.
__global__ void init_agents_kernel(AgentSoA a, uint32_t n, uint64_t seed, float fw, float fh,
                                   float smin, float range) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float pi = 3.14159265358979323846f;  // std::f32::consts::PI
  uint64_t k = 4ull * i;
  a.x[i] = uniform24(seed, k) * fw;
  a.y[i] = uniform24(seed, k + 1) * fh;
  a.th[i] = (uniform24(seed, k + 2) * 2.0f) * pi;
  a.v[i] = smin + uniform24(seed, k + 3) * range;
  if (a.id) a.id[i] = i;
}

// src/main.rs:132-137; the draw is keyed by the original index.
__global__ void reassign_speeds_kernel(AgentSoA a, uint32_t n, uint64_t seed, float smin,
                                       float range) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t i = a.id ? a.id[j] : j;
  a.v[j] = smin + uniform24(seed, i) * range;
}

// src/main.rs:993-994
__global__ void scale_positions_kernel(AgentSoA a, uint32_t n, float sx, float sy) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  a.x[j] *= sx;
  a.y[j] *= sy;
}

__global__ void pack_agents_kernel(AgentSoA a, uint32_t n, float4* aos) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t i = a.id ? a.id[j] : j;
  aos[i] = make_float4(a.x[j], a.y[j], a.th[j], a.v[j]);
}

__global__ void unpack_agents_kernel(AgentSoA a, uint32_t n, const float4* aos) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 r = aos[i];
  a.x[i] = r.x;
  a.y[i] = r.y;
  a.th[i] = r.z;
  a.v[i] = r.w;
  if (a.id) a.id[i] = i;
}

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

}  // namespace

void launch_agents(const StepGeom& g, const AgentSoA& a, void* trail, const DevParams* p,
                   hipStream_t stream) {
  if (g.agents == 0) return;
  AgentArgs args{a, trail, p, g.width, g.height, g.agents, g.xcd_remap};
  uint32_t nb = blocks_for(g.agents, kAgentBlock);
  switch (g.storage) {
    case SLM_STORAGE_BF16:
      agents_kernel<SLM_STORAGE_BF16><<<nb, kAgentBlock, 0, stream>>>(args);
      break;
    case SLM_STORAGE_F16:
      agents_kernel<SLM_STORAGE_F16><<<nb, kAgentBlock, 0, stream>>>(args);
      break;
    default:
      agents_kernel<SLM_STORAGE_F32><<<nb, kAgentBlock, 0, stream>>>(args);
      break;
  }
}

void launch_init_agents(const AgentSoA& a, uint32_t n, uint64_t seed, uint32_t width,
                        uint32_t height, float speed_min, float speed_max, hipStream_t stream) {
  if (n == 0) return;
  float range = speed_max - speed_min;
  init_agents_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(
      a, n, seed, static_cast<float>(width), static_cast<float>(height), speed_min, range);
}

void launch_reassign_speeds(const AgentSoA& a, uint32_t n, uint64_t stream_seed,
                            float speed_min, float speed_max, hipStream_t stream) {
  if (n == 0) return;
  float range = speed_max - speed_min;
  reassign_speeds_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(a, n, stream_seed, speed_min,
                                                                   range);
}

void launch_scale_positions(const AgentSoA& a, uint32_t n, float sx, float sy,
                            hipStream_t stream) {
  if (n == 0) return;
  scale_positions_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(a, n, sx, sy);
}

void launch_pack_agents(const AgentSoA& a, uint32_t n, float* aos, hipStream_t stream) {
  if (n == 0) return;
  pack_agents_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(a, n, reinterpret_cast<float4*>(aos));
}

void launch_unpack_agents(const AgentSoA& a, uint32_t n, const float* aos, hipStream_t stream) {
  if (n == 0) return;
  unpack_agents_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(
      a, n, reinterpret_cast<const float4*>(aos));
}

}  // namespace slm
