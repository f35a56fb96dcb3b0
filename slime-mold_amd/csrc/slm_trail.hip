// slm_trail.hip - the fused trail pass: deposit marks + decay + 3x3 box diffuse
// (reference src/compute.wgsl:140, :147-161, :163-195), plus the small trail
// utilities behind the C ABI (parameter upload, f32 <-> storage conversion,
// f64 total mass).
//
// One read and one write of the trail per cell and step (Jacobi: trail_in ->
// trail_out).  For each cell the pass computes, in f32:
//   T1 = marked ? clamp(|T| + deposit, 0, 1) : |T|      (compute.wgsl:140)
//   T2 = max(T1 - decay_k, 0)                            (compute.wgsl:160)
//   out = mix(T2, (sum of the 3x3 toroidal T2 window)/9, rate)  (:176-194)
// with the window summed in the shader's order (dy outer, dx inner, one
// rounding per add) so the result is the reference's bit for bit.
//
// Strip kernel (row length a multiple of 16 bytes): each lane owns one 16-byte
// column strip (4 f32 or 8 16-bit cells) and walks a tile of rows keeping the
// rows above, at and below in registers; the left/right neighbour cells come
// from the adjacent lanes, and only lanes at a wave edge or at the wrap load a
// halo cell.  Rows of other lengths use a one-cell-per-lane kernel.
#include "slm_internal.h"

namespace slm {
namespace {

constexpr int kTrailBlock = 256;
constexpr int kWavesPerBlock = kTrailBlock / 64;

__device__ inline uint32_t xcd_logical_block(uint32_t b, uint32_t nb) {
  uint32_t xcd = b & 7u, idx = b >> 3;
  uint32_t per = nb >> 3, rem = nb & 7u;
  return xcd < rem ? xcd * (per + 1u) + idx : rem * (per + 1u) + (xcd - rem) * per + idx;
}

// ---- storage formats -------------------------------------------------------
// Elements are handled as raw bits; the sign bit is the deposit mark.
template <int S>
struct Codec;

template <>
struct Codec<SLM_STORAGE_F32> {
  using elem = uint32_t;
  static constexpr int kVec = 4;
  __device__ static float value(uint32_t b) { return __uint_as_float(b & 0x7fffffffu); }
  __device__ static bool marked(uint32_t b) { return (b >> 31) != 0u; }
  __device__ static uint32_t encode(float v) { return __float_as_uint(v); }
  __device__ static uint32_t lane(uint4 r, int k) {
    return k == 0 ? r.x : k == 1 ? r.y : k == 2 ? r.z : r.w;
  }
  __device__ static uint4 pack(const uint32_t (&e)[4]) { return make_uint4(e[0], e[1], e[2], e[3]); }
};

// Round-to-nearest-even f32 -> bfloat16 (finite inputs).
__device__ inline uint32_t bf16_rne(float v) {
  uint32_t b = __float_as_uint(v);
  return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}

template <>
struct Codec<SLM_STORAGE_BF16> {
  using elem = uint16_t;
  static constexpr int kVec = 8;
  __device__ static float value(uint32_t h) { return __uint_as_float((h & 0x7fffu) << 16); }
  __device__ static bool marked(uint32_t h) { return (h & 0x8000u) != 0u; }
  __device__ static uint32_t encode(float v) { return bf16_rne(v); }
  __device__ static uint32_t lane(uint4 r, int k) {
    uint32_t w = (k >> 1) == 0 ? r.x : (k >> 1) == 1 ? r.y : (k >> 1) == 2 ? r.z : r.w;
    return (k & 1) ? (w >> 16) : (w & 0xffffu);
  }
  __device__ static uint4 pack(const uint32_t (&e)[8]) {
    return make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16),
                      e[6] | (e[7] << 16));
  }
};

template <>
struct Codec<SLM_STORAGE_F16> {
  using elem = uint16_t;
  static constexpr int kVec = 8;
  __device__ static float value(uint32_t h) {
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h & 0x7fffu)));
  }
  __device__ static bool marked(uint32_t h) { return (h & 0x8000u) != 0u; }
  __device__ static uint32_t encode(float v) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
  }
  __device__ static uint32_t lane(uint4 r, int k) { return Codec<SLM_STORAGE_BF16>::lane(r, k); }
  __device__ static uint4 pack(const uint32_t (&e)[8]) { return Codec<SLM_STORAGE_BF16>::pack(e); }
};

// Deposit (S3) and decay of one stored element.
template <int S>
__device__ inline float deposit_decay(uint32_t raw, const DevParams& p) {
  float t = Codec<S>::value(raw);
  if (Codec<S>::marked(raw)) {
    float d = t + p.deposit;
    d = d > 0.0f ? d : 0.0f;
    t = d < 1.0f ? d : 1.0f;
  }
  float u = t - p.decay_k;
  return u > 0.0f ? u : 0.0f;
}

struct TrailArgs {
  const void* in;
  void* out;
  const DevParams* p;
  uint32_t w, h;
  uint32_t tile_rows;  // rows per wave tile
  uint32_t ntx;        // wave tiles across a row
  uint32_t ntiles;     // wave tiles in the grid
  int xcd_remap;
};

template <int S>
struct RawRow {
  uint4 v;
  uint32_t left, right;
};

template <int S>
__device__ inline void issue_row(RawRow<S>& r, const typename Codec<S>::elem* row, uint32_t c,
                                 bool valid, bool need_left, bool need_right, uint32_t lcol,
                                 uint32_t rcol) {
  r.v = valid ? *reinterpret_cast<const uint4*>(row + c) : make_uint4(0u, 0u, 0u, 0u);
  r.left = need_left ? static_cast<uint32_t>(row[lcol]) : 0u;
  r.right = need_right ? static_cast<uint32_t>(row[rcol]) : 0u;
}

// T2 values of one row strip with its left and right neighbours:
// t[0] = column c-1, t[1..V] = columns c..c+V-1, t[V+1] = column c+V.
template <int S>
__device__ inline void finish_row(float (&t)[Codec<S>::kVec + 2], const RawRow<S>& r,
                                  bool need_left, bool need_right, const DevParams& p) {
  constexpr int V = Codec<S>::kVec;
#pragma unroll
  for (int k = 0; k < V; k++) t[k + 1] = deposit_decay<S>(Codec<S>::lane(r.v, k), p);
  float from_left = __shfl_up(t[V], 1);
  float from_right = __shfl_down(t[1], 1);
  t[0] = need_left ? deposit_decay<S>(r.left, p) : from_left;
  t[V + 1] = need_right ? deposit_decay<S>(r.right, p) : from_right;
}

template <int S>
__global__ __launch_bounds__(kTrailBlock) void trail_strip_kernel(TrailArgs a) {
  using C = Codec<S>;
  using elem = typename C::elem;
  constexpr int V = C::kVec;
  const DevParams p = *a.p;
  uint32_t blk = a.xcd_remap ? xcd_logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
  uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint32_t tile = blk * kWavesPerBlock + wave;
  if (tile >= a.ntiles) return;  // uniform per wave
  uint32_t tx = tile % a.ntx, ty = tile / a.ntx;
  const uint32_t w = a.w, h = a.h;
  uint32_t c = (tx * 64u + lane) * V;
  bool valid = c < w;
  bool need_left = valid && lane == 0u;
  bool need_right = valid && (lane == 63u || c + V >= w);
  uint32_t lcol = (c == 0u ? w : c) - 1u;
  uint32_t rcol = c + V >= w ? c + V - w : c + V;
  uint32_t y0 = ty * a.tile_rows;
  uint32_t y1 = min(y0 + a.tile_rows, h);
  const elem* in = static_cast<const elem*>(a.in);
  elem* out = static_cast<elem*>(a.out);
  auto row_ptr = [&](uint32_t r) { return in + static_cast<size_t>(r) * w; };
  auto wrap_up = [&](uint32_t r) { return r == 0u ? h - 1u : r - 1u; };
  auto wrap_down = [&](uint32_t r) { return r + 1u >= h ? r + 1u - h : r + 1u; };

  float prev[V + 2], cur[V + 2], next[V + 2];
  RawRow<S> rp, rc, rn;
  issue_row<S>(rp, row_ptr(wrap_up(y0)), c, valid, need_left, need_right, lcol, rcol);
  issue_row<S>(rc, row_ptr(y0), c, valid, need_left, need_right, lcol, rcol);
  uint32_t ynext = wrap_down(y0);
  issue_row<S>(rn, row_ptr(ynext), c, valid, need_left, need_right, lcol, rcol);
  finish_row<S>(prev, rp, need_left, need_right, p);
  finish_row<S>(cur, rc, need_left, need_right, p);

  for (uint32_t y = y0; y < y1; y++) {
    // row y+1 is in flight; start row y+2 before consuming it
    RawRow<S> ra;
    bool more = y + 1u < y1;
    uint32_t yafter = wrap_down(ynext);
    if (more) issue_row<S>(ra, row_ptr(yafter), c, valid, need_left, need_right, lcol, rcol);
    finish_row<S>(next, rn, need_left, need_right, p);
    if (valid) {
      uint32_t e[V];
#pragma unroll
      for (int k = 0; k < V; k++) {
        float s = prev[k] + prev[k + 1];
        s = s + prev[k + 2];
        s = s + cur[k];
        s = s + cur[k + 1];
        s = s + cur[k + 2];
        s = s + next[k];
        s = s + next[k + 1];
        s = s + next[k + 2];
        float avg = s / 9.0f;
        float l = cur[k + 1] * p.one_minus_rate;
        float r = avg * p.rate;
        e[k] = C::encode(l + r);
      }
      *reinterpret_cast<uint4*>(out + static_cast<size_t>(y) * w + c) = C::pack(e);
    }
#pragma unroll
    for (int k = 0; k < V + 2; k++) {
      prev[k] = cur[k];
      cur[k] = next[k];
    }
    if (more) rn = ra;
    ynext = yafter;
  }
}

// One cell per lane, any row length.
template <int S>
__global__ __launch_bounds__(kTrailBlock) void trail_cell_kernel(TrailArgs a) {
  using C = Codec<S>;
  using elem = typename C::elem;
  const DevParams p = *a.p;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kTrailBlock + threadIdx.x;
  const uint32_t w = a.w, h = a.h;
  if (i >= static_cast<uint64_t>(w) * h) return;
  uint32_t x = static_cast<uint32_t>(i % w), y = static_cast<uint32_t>(i / w);
  const elem* in = static_cast<const elem*>(a.in);
  uint32_t xs[3] = {x == 0u ? w - 1u : x - 1u, x, x + 1u >= w ? x + 1u - w : x + 1u};
  uint32_t ys[3] = {y == 0u ? h - 1u : y - 1u, y, y + 1u >= h ? y + 1u - h : y + 1u};
  float s = 0.0f, centre = 0.0f;
#pragma unroll
  for (int dy = 0; dy < 3; dy++) {
#pragma unroll
    for (int dx = 0; dx < 3; dx++) {
      float t = deposit_decay<S>(in[static_cast<size_t>(ys[dy]) * w + xs[dx]], p);
      if (dy == 1 && dx == 1) centre = t;
      s = (dy == 0 && dx == 0) ? t : s + t;
    }
  }
  float avg = s / 9.0f;
  float l = centre * p.one_minus_rate;
  float r = avg * p.rate;
  static_cast<elem*>(a.out)[i] = static_cast<elem>(C::encode(l + r));
}

__global__ void set_params_kernel(DevParams* dst, DevParams src) { *dst = src; }

template <int S>
__global__ void decode_trail_kernel(const void* t, uint64_t cells, float* out) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  out[i] = Codec<S>::value(static_cast<const typename Codec<S>::elem*>(t)[i]);
}

template <int S>
__global__ void encode_trail_kernel(const float* in, uint64_t cells, void* t) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  static_cast<typename Codec<S>::elem*>(t)[i] =
      static_cast<typename Codec<S>::elem>(Codec<S>::encode(in[i]));
}

template <int S>
__global__ __launch_bounds__(256) void mass_partial_kernel(const void* t, uint64_t cells,
                                                           double* partials) {
  __shared__ double red[256];
  const typename Codec<S>::elem* e = static_cast<const typename Codec<S>::elem*>(t);
  double acc = 0.0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < cells;
       i += static_cast<uint64_t>(gridDim.x) * 256)
    acc += static_cast<double>(Codec<S>::value(e[i]));
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(s)) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void mass_final_kernel(const double* partials, int n,
                                                         double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(s)) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <int S>
void launch_trail_t(const StepGeom& g, const void* in, void* out, const DevParams* p,
                    hipStream_t stream) {
  constexpr int V = Codec<S>::kVec;
  TrailArgs a{in, out, p, g.width, g.height, 0, 0, 0, g.xcd_remap};
  if (g.width % V == 0) {
    a.ntx = (g.width + 64u * V - 1u) / (64u * V);
    // Rows per wave tile: the longest of 64..4 that still gives >= 8192 waves.
    uint32_t rows = 64;
    while (rows > 4 && static_cast<uint64_t>(a.ntx) * ((g.height + rows - 1) / rows) < 8192u)
      rows >>= 1;
    a.tile_rows = rows;
    a.ntiles = a.ntx * ((g.height + rows - 1u) / rows);
    uint32_t nb = (a.ntiles + kWavesPerBlock - 1u) / kWavesPerBlock;
    trail_strip_kernel<S><<<nb, kTrailBlock, 0, stream>>>(a);
  } else {
    uint64_t cells = static_cast<uint64_t>(g.width) * g.height;
    uint32_t nb = static_cast<uint32_t>((cells + kTrailBlock - 1) / kTrailBlock);
    trail_cell_kernel<S><<<nb, kTrailBlock, 0, stream>>>(a);
  }
}

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

}  // namespace

DevParams make_dev_params(const slm_params& s) {
  DevParams d;
  d.decay_k = s.decay_factor * 0.001f;
  d.jitter = s.agent_jitter;
  d.speed_min = s.agent_speed_min;
  d.speed_max = s.agent_speed_max;
  d.turn = s.agent_turn_speed;
  d.sensor_angle = s.agent_sensor_angle;
  d.sensor_dist = s.agent_sensor_distance;
  float r = s.diffusion_rate > 0.0f ? s.diffusion_rate : 0.0f;
  d.rate = r < 1.0f ? r : 1.0f;
  d.one_minus_rate = 1.0f - d.rate;
  d.deposit = s.pheromone_deposition_amount;
  d.blur_radius = s.blur_radius;
  d.blur_sigma = s.blur_sigma;
  return d;
}

void launch_trail(const StepGeom& g, const void* in, void* out, const DevParams* p,
                  hipStream_t stream) {
  switch (g.storage) {
    case SLM_STORAGE_BF16:
      launch_trail_t<SLM_STORAGE_BF16>(g, in, out, p, stream);
      break;
    case SLM_STORAGE_F16:
      launch_trail_t<SLM_STORAGE_F16>(g, in, out, p, stream);
      break;
    default:
      launch_trail_t<SLM_STORAGE_F32>(g, in, out, p, stream);
      break;
  }
}

void launch_set_params(DevParams* dst, const DevParams& src, hipStream_t stream) {
  set_params_kernel<<<1, 1, 0, stream>>>(dst, src);
}

void launch_decode_trail(const void* t, int storage, uint64_t cells, float* out,
                         hipStream_t stream) {
  if (cells == 0) return;
  uint32_t nb = blocks_for(cells, 256);
  if (storage == SLM_STORAGE_BF16)
    decode_trail_kernel<SLM_STORAGE_BF16><<<nb, 256, 0, stream>>>(t, cells, out);
  else if (storage == SLM_STORAGE_F16)
    decode_trail_kernel<SLM_STORAGE_F16><<<nb, 256, 0, stream>>>(t, cells, out);
  else
    decode_trail_kernel<SLM_STORAGE_F32><<<nb, 256, 0, stream>>>(t, cells, out);
}

void launch_encode_trail(const float* in, int storage, uint64_t cells, void* t,
                         hipStream_t stream) {
  if (cells == 0) return;
  uint32_t nb = blocks_for(cells, 256);
  if (storage == SLM_STORAGE_BF16)
    encode_trail_kernel<SLM_STORAGE_BF16><<<nb, 256, 0, stream>>>(in, cells, t);
  else if (storage == SLM_STORAGE_F16)
    encode_trail_kernel<SLM_STORAGE_F16><<<nb, 256, 0, stream>>>(in, cells, t);
  else
    encode_trail_kernel<SLM_STORAGE_F32><<<nb, 256, 0, stream>>>(in, cells, t);
}

void launch_total_mass(const void* t, int storage, uint64_t cells, double* partials, double* out,
                       hipStream_t stream) {
  if (storage == SLM_STORAGE_BF16)
    mass_partial_kernel<SLM_STORAGE_BF16><<<kMassBlocks, 256, 0, stream>>>(t, cells, partials);
  else if (storage == SLM_STORAGE_F16)
    mass_partial_kernel<SLM_STORAGE_F16><<<kMassBlocks, 256, 0, stream>>>(t, cells, partials);
  else
    mass_partial_kernel<SLM_STORAGE_F32><<<kMassBlocks, 256, 0, stream>>>(t, cells, partials);
  mass_final_kernel<<<1, 256, 0, stream>>>(partials, kMassBlocks, out);
}

}  // namespace slm
