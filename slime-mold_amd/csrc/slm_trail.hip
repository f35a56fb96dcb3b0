// slm_trail.hip - the fused trail pass: deposit marks + decay + 3x3 box diffuse
// (reference src/compute.wgsl:140, :147-161, :163-195), its opt-in variant
// with a separable Gaussian in place of the box (gauss_kernel, below), plus the
// small trail utilities behind the C ABI (parameter upload, f32 <-> storage
// conversion, f64 total mass).
//
// One read and one write of the trail per cell and step (Jacobi: trail_in ->
// trail_out).  For each cell the pass computes, in f32:
//   T1 = marked ? clamp(|T| + deposit, 0, 1) : |T|      (compute.wgsl:140)
//   T2 = max(T1 - decay_k, 0)                            (compute.wgsl:160)
//   out = mix(T2, (sum of the 3x3 toroidal T2 window)/9, rate)  (:176-194)
// with the window summed in the shader's order (dy outer, dx inner, one
// rounding per add) so the result is the reference's bit for bit.
//
// Work is cut into wave-sized units over a pool of instances.  Strip unit (row
// length a multiple of 16 bytes): each lane owns one 16-byte column strip (4
// f32 or 8 16-bit cells) and walks `tile_rows` rows keeping the rows above, at
// and below in registers; the left/right neighbour cells come from the
// adjacent lanes, and only lanes at a wave edge or at the wrap load a halo
// cell.  Rows of other lengths use 64-cell units, one cell per lane.
#include <cmath>
#include <type_traits>

#include "slm_internal.h"

namespace slm {
namespace {

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
  __device__ static uint4 pack_values(const float (&v)[4]) {
    return make_uint4(__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]),
                      __float_as_uint(v[3]));
  }
};

// Round-to-nearest-even f32 -> bfloat16: gfx950's v_cvt_pk_bf16_f32.
__device__ inline uint32_t bf16_rne(float v) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v));
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t bf16x2_rne(float a, float b) {
  f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
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
  __device__ static uint4 pack_values(const float (&v)[8]) {
    return make_uint4(bf16x2_rne(v[0], v[1]), bf16x2_rne(v[2], v[3]), bf16x2_rne(v[4], v[5]),
                      bf16x2_rne(v[6], v[7]));
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
  __device__ static uint4 pack_values(const float (&v)[8]) {
    uint32_t e[8];
#pragma unroll
    for (int k = 0; k < 8; k++) e[k] = encode(v[k]);
    return pack(e);
  }
};

// bfloat16 with deterministic stochastic rounding (the trail pass's stores
// only; reads and other conversions are those of bfloat16): the 16 discarded
// bits are added to 16 pseudo-random bits and the sum truncated, so a value
// rounds up with probability (discarded bits) / 2^16 - unbiased, where
// round-to-nearest-even drifts a long run's mass by several percent.  The
// random bits mix the cell index and the value, so every cell and step draws
// differently: x = (cell * 0x9e3779b1) ^ bits, then twice x ^= x >> 16 and a
// multiply of x's low 24 bits by an odd 24-bit constant (0xebca6b, 0x9e3779),
// mod 2^32; the draw is bits 16..31.  The xor-shifts keep neighbouring cells'
// draws independent (a single multiplicative hash of the linear cell sequence
// correlates them); the 24-bit products are full-rate v_mul_u32_u24, where a
// 32-bit multiply is quarter-rate.
constexpr int kStorageBf16Sr = 3;

// The first multiply of the draw, cell * kSrK1 (mod 2^32), is linear in the
// cell: callers that store consecutive cells, or rows a fixed stride apart,
// step it with one add per cell instead of a quarter-rate 32-bit multiply
// (the hashed cell `hc` below); the bits are the same.
constexpr uint32_t kSrK1 = 0x9e3779b1u;
constexpr uint32_t kSrK2 = 0xebca6bu, kSrK3 = 0x9e3779u;  // 24 bits each
__device__ inline uint32_t bf16_sr_hashed(float v, uint32_t hc) {
  uint32_t b = __float_as_uint(v);  // v >= 0 and finite
  uint32_t x = hc ^ b;
  x ^= x >> 16;
  x = __umul24(x, kSrK2);
  x ^= x >> 16;
  x = __umul24(x, kSrK3);
  return (b + (x >> 16)) >> 16;
}

template <>
struct Codec<kStorageBf16Sr> : Codec<SLM_STORAGE_BF16> {
  __device__ static uint32_t encode_hashed(float v, uint32_t hc) { return bf16_sr_hashed(v, hc); }
  __device__ static uint4 pack_values_hashed(const float (&v)[8], uint32_t hc) {
    uint32_t e[8];
#pragma unroll
    for (int k = 0; k < 8; k++) e[k] = bf16_sr_hashed(v[k], hc + static_cast<uint32_t>(k) * kSrK1);
    return make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16),
                      e[6] | (e[7] << 16));
  }
};

// Stores of the trail pass, given the hashed instance-local index of the
// first cell, hc = cell * kSrK1 (only stochastic rounding reads it).
template <int S>
__device__ inline uint4 pack_out(const float (&v)[Codec<S>::kVec], uint32_t hc) {
  if constexpr (S == kStorageBf16Sr)
    return Codec<S>::pack_values_hashed(v, hc);
  else
    return Codec<S>::pack_values(v);
}
template <int S>
__device__ inline uint32_t encode_out(float v, uint32_t hc) {
  if constexpr (S == kStorageBf16Sr)
    return Codec<S>::encode_hashed(v, hc);
  else
    return Codec<S>::encode(v);
}

// Deposit (S3) and decay of one stored element.  clamp(d, 0, 1) is one
// v_med3_f32: d = |T| + deposit is finite and never -0 (|T| >= +0), so the
// median of (d, 0, 1) is exactly the clamp.
template <int S>
__device__ inline float deposit_decay(uint32_t raw, const DevParams& p) {
  float t = Codec<S>::value(raw);
  if (Codec<S>::marked(raw)) t = __builtin_amdgcn_fmed3f(t + p.deposit, 0.0f, 1.0f);
  float u = t - p.decay_k;
  return u > 0.0f ? u : 0.0f;
}

// sum / 9.0f with IEEE round-to-nearest, as a reciprocal multiply and one
// fused correction step: equal to the division for every finite sum >= 0
// (checked exhaustively over all 2^31 non-negative floats), about a third of
// the division's instructions.
__device__ inline float div9(float sum) {
  const float inv9 = 0x1.c71c72p-4f;  // nearest float to 1/9
  float q = sum * inv9;
  float r = __builtin_fmaf(-q, 9.0f, sum);
  return __builtin_fmaf(r, inv9, q);
}

// How mix(centre, avg, rate) (compute.wgsl:193-194) reduces for this rate:
// rate 1 gives centre*0 + avg*1 = avg and rate 0 gives centre*1 + avg*0 =
// centre exactly (centre and avg are finite and >= 0), so those two skip the
// products (and rate 0 the window sum).
enum MixMode { kMixGeneral = 0, kMixAverage = 1, kMixCentre = 2 };

template <int M>
__device__ inline float diffuse_out(float centre, float sum, const DevParams& p) {
  if (M == kMixCentre) return centre;
  float avg = div9(sum);
  if (M == kMixAverage) return avg;
  float l = centre * p.one_minus_rate;
  float r = avg * p.rate;
  return l + r;
}

// A strip row's 16-byte store and load, optionally non-temporal (streamed
// past the caches: faster when the map does not fit the Infinity Cache).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kStreamStores = 1, kStreamLoads = 2;  // TrailArgs::streaming bits
__device__ inline void store_row(void* dst, uint4 v, bool streaming) {
  if (streaming) {
    u32x4 x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<u32x4*>(dst));
  } else {
    *reinterpret_cast<uint4*>(dst) = v;
  }
}

template <int S>
struct RawRow {
  uint4 v;
  uint32_t left, right;
};

template <int S>
__device__ inline void issue_row(RawRow<S>& r, const typename Codec<S>::elem* row, uint32_t c,
                                 bool valid, bool need_left, bool need_right, uint32_t lcol,
                                 uint32_t rcol, bool streaming) {
  if (!valid) {
    r.v = make_uint4(0u, 0u, 0u, 0u);
  } else if (streaming) {
    u32x4 x = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(row + c));
    r.v = make_uint4(x.x, x.y, x.z, x.w);
  } else {
    r.v = *reinterpret_cast<const uint4*>(row + c);
  }
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

// One strip unit: 64*V columns x tile_rows rows of instance `in`, streamed
// with kAhead rows in flight: the rows above and at the current output row
// are decoded in registers, and a ring of kAhead raw rows holds the loads of
// the rows below; each slot is refilled as soon as it is consumed.
constexpr int kAhead = 4;

template <int S, int M>
__device__ inline void strip_unit(const typename Codec<S>::elem* in, typename Codec<S>::elem* out,
                                  const Instance& inst, const DevParams& p, uint32_t tile,
                                  uint32_t lane, int streaming) {
  const bool nt_load = (streaming & kStreamLoads) != 0;
  constexpr int V = Codec<S>::kVec;
  const uint32_t w = inst.width, h = inst.height;
  uint32_t tx = tile % inst.ntx, ty = tile / inst.ntx;
  uint32_t c = (tx * 64u + lane) * V;
  bool valid = c < w;
  bool need_left = valid && lane == 0u;
  bool need_right = valid && (lane == 63u || c + V >= w);
  uint32_t lcol = (c == 0u ? w : c) - 1u;
  uint32_t rcol = c + V >= w ? c + V - w : c + V;
  const uint32_t y0 = ty * inst.tile_rows;
  const uint32_t y1 = min(y0 + inst.tile_rows, h);
  // row r of the unit's range y0 - 1 .. y1, wrapped onto the torus
  auto row_ptr = [&](uint32_t r) {
    uint32_t y = r == 0u ? h - 1u : (r - 1u >= h ? r - 1u - h : r - 1u);
    return in + static_cast<size_t>(y) * w;
  };  // r counts from y0 - 1 + 1: row_ptr(y + 1) is row y
  float prev[V + 2], cur[V + 2], next[V + 2];
  {
    RawRow<S> rp, rc;
    issue_row<S>(rp, row_ptr(y0), c, valid, need_left, need_right, lcol, rcol, nt_load);
    issue_row<S>(rc, row_ptr(y0 + 1u), c, valid, need_left, need_right, lcol, rcol, nt_load);
    finish_row<S>(prev, rp, need_left, need_right, p);
    finish_row<S>(cur, rc, need_left, need_right, p);
  }
  // slot j holds row y + 1 + j while the block of output rows starting at y runs
  RawRow<S> ring[kAhead];
#pragma unroll
  for (int j = 0; j < kAhead; j++)
    if (y0 + 1u + j <= y1)
      issue_row<S>(ring[j], row_ptr(y0 + 2u + j), c, valid, need_left, need_right, lcol, rcol,
                   nt_load);
  for (uint32_t y = y0; y < y1; y += kAhead) {
#pragma unroll
    for (int j = 0; j < kAhead; j++) {
      const uint32_t yy = y + j;
      if (yy < y1) {  // uniform
        finish_row<S>(next, ring[j], need_left, need_right, p);  // row yy + 1
        if (yy + 1u + kAhead <= y1)
          issue_row<S>(ring[j], row_ptr(yy + 2u + kAhead), c, valid, need_left, need_right, lcol,
                       rcol, nt_load);
        if (valid) {
          float o[V];
#pragma unroll
          for (int k = 0; k < V; k++) {
            float sum = 0.0f;
            if (M != kMixCentre) {
              sum = prev[k] + prev[k + 1];
              sum = sum + prev[k + 2];
              sum = sum + cur[k];
              sum = sum + cur[k + 1];
              sum = sum + cur[k + 2];
              sum = sum + next[k];
              sum = sum + next[k + 1];
              sum = sum + next[k + 2];
            }
            o[k] = diffuse_out<M>(cur[k + 1], sum, p);
          }
          store_row(out + static_cast<size_t>(yy) * w + c, pack_out<S>(o, (yy * w + c) * kSrK1),
                    (streaming & kStreamStores) != 0);
        }
#pragma unroll
        for (int k = 0; k < V + 2; k++) {
          prev[k] = cur[k];
          cur[k] = next[k];
        }
      }
    }
  }
}

// The same strip unit with its tile height R fixed at compile time and all
// R + 2 rows (tile plus the halo row above and below) loaded before any is
// used, so each wave keeps (R + 2) KiB-sized row loads in flight; taller tiles
// re-read fewer halo rows (R + 2 loads per R rows) without losing the memory
// parallelism that short tiles get from having more waves.
template <int S, int M, int R>
__device__ inline void strip_unit_preload(const typename Codec<S>::elem* in,
                                          typename Codec<S>::elem* out, const Instance& inst,
                                          const DevParams& p, uint32_t tile, uint32_t lane,
                                          int streaming) {
  const bool nt_load = (streaming & kStreamLoads) != 0;
  constexpr int V = Codec<S>::kVec;
  const uint32_t w = inst.width, h = inst.height;
  uint32_t tx = tile % inst.ntx, ty = tile / inst.ntx;
  uint32_t c = (tx * 64u + lane) * V;
  bool valid = c < w;
  bool need_left = valid && lane == 0u;
  bool need_right = valid && (lane == 63u || c + V >= w);
  uint32_t lcol = (c == 0u ? w : c) - 1u;
  uint32_t rcol = c + V >= w ? c + V - w : c + V;
  const uint32_t y0 = ty * R;
  const uint32_t rows = min(static_cast<uint32_t>(R), h - y0);  // uniform
  RawRow<S> raw[R + 2];
#pragma unroll
  for (int r = 0; r < R + 2; r++) {
    // raw[r] holds row y0 + r - 1, wrapped; rows past the tile's last + 1 are unused
    if (static_cast<uint32_t>(r) <= rows + 1u) {
      uint32_t y = y0 + static_cast<uint32_t>(r);
      y = y == 0u ? h - 1u : y - 1u;
      if (y >= h) y -= h;
      issue_row<S>(raw[r], in + static_cast<size_t>(y) * w, c, valid, need_left, need_right,
                   lcol, rcol, nt_load);
    }
  }
  float prev[V + 2], cur[V + 2], next[V + 2];
  finish_row<S>(prev, raw[0], need_left, need_right, p);
  finish_row<S>(cur, raw[1], need_left, need_right, p);
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (static_cast<uint32_t>(r) < rows) {
      finish_row<S>(next, raw[r + 2], need_left, need_right, p);
      if (valid) {
        float o[V];
#pragma unroll
        for (int k = 0; k < V; k++) {
          float sum = 0.0f;
          if (M != kMixCentre) {
            sum = prev[k] + prev[k + 1];
            sum = sum + prev[k + 2];
            sum = sum + cur[k];
            sum = sum + cur[k + 1];
            sum = sum + cur[k + 2];
            sum = sum + next[k];
            sum = sum + next[k + 1];
            sum = sum + next[k + 2];
          }
          o[k] = diffuse_out<M>(cur[k + 1], sum, p);
        }
        store_row(out + static_cast<size_t>(y0 + r) * w + c,
                  pack_out<S>(o, ((y0 + r) * w + c) * kSrK1), (streaming & kStreamStores) != 0);
      }
#pragma unroll
      for (int k = 0; k < V + 2; k++) {
        prev[k] = cur[k];
        cur[k] = next[k];
      }
    }
  }
}

// R = 0: streaming strip units of any height; R = 4, 8 or 16: every strip
// instance of the launch has R-row tiles, loaded whole.
template <int S, int M, int R>
__device__ inline void strip_dispatch(const typename Codec<S>::elem* in,
                                      typename Codec<S>::elem* out, const Instance& inst,
                                      const DevParams& p, uint32_t tile, uint32_t lane,
                                      int streaming) {
  if (R == 0)
    strip_unit<S, M>(in, out, inst, p, tile, lane, streaming);
  else
    strip_unit_preload<S, M, (R == 0 ? 8 : R)>(in, out, inst, p, tile, lane, streaming);
}

// One 64-cell unit, one cell per lane, any row length.
template <int S>
__device__ inline void cell_unit(const typename Codec<S>::elem* in, typename Codec<S>::elem* out,
                                 const Instance& inst, const DevParams& p, uint32_t unit,
                                 uint32_t lane) {
  const uint32_t w = inst.width, h = inst.height;
  uint64_t i = static_cast<uint64_t>(unit) * 64u + lane;
  if (i >= static_cast<uint64_t>(w) * h) return;
  uint32_t x = static_cast<uint32_t>(i % w), y = static_cast<uint32_t>(i / w);
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
  out[i] = static_cast<typename Codec<S>::elem>(
      encode_out<S>(diffuse_out<kMixGeneral>(centre, s, p), static_cast<uint32_t>(i) * kSrK1));
}

struct TrailArgs {
  const void* in;
  void* out;
  const Instance* inst;
  const DevParams* params;
  const uint32_t* block_inst;
  uint32_t count, total_units;
  int xcd_remap;
  int streaming;  // kStreamStores | kStreamLoads: non-temporal strip row stores / loads
  uint32_t* claim;  // persistent launch: claim counters (nullptr: one workgroup per block)
  uint32_t blocks;  // logical workgroups of the pass
};

// One logical workgroup `blk` (kTrailWavesPerBlock wave units) of the pass.
template <int S, int R>
__device__ inline void trail_block(const TrailArgs& a, const Instance* __restrict__ insts,
                                   const DevParams* __restrict__ params,
                                   const uint32_t* __restrict__ block_inst, uint32_t blk) {
  using elem = typename Codec<S>::elem;
  uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint32_t lane = threadIdx.x & 63u;
  uint32_t u = blk * kTrailWavesPerBlock + wave;
  if (u >= a.total_units) return;  // uniform per wave
  uint32_t ii = block_inst[blk];
  while (u >= insts[ii].unit_off + insts[ii].units) ii++;
  const Instance inst = insts[ii];
  const DevParams p = params[ii];
  const elem* in = static_cast<const elem*>(a.in) + inst.trail_off;
  elem* out = static_cast<elem*>(a.out) + inst.trail_off;
  uint32_t unit = u - inst.unit_off;
  if (inst.ntx == 0)
    cell_unit<S>(in, out, inst, p, unit, lane);
  else if (p.rate == 1.0f)
    strip_dispatch<S, kMixAverage, R>(in, out, inst, p, unit, lane, a.streaming);
  else if (p.rate == 0.0f)
    strip_dispatch<S, kMixCentre, R>(in, out, inst, p, unit, lane, a.streaming);
  else
    strip_dispatch<S, kMixGeneral, R>(in, out, inst, p, unit, lane, a.streaming);
}

// Tables as restrict arguments: their wave-uniform loads become scalar loads.
// With a.claim set the launch is persistent (persistent_blocks).
template <int S, int R>
__global__ __launch_bounds__(kTrailBlock) void trail_kernel(
    TrailArgs a, const Instance* __restrict__ insts, const DevParams* __restrict__ params,
    const uint32_t* __restrict__ block_inst) {
  __shared__ uint32_t s_claim;
  if (!a.claim) {
    uint32_t blk = a.xcd_remap ? xcd_logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
    trail_block<S, R>(a, insts, params, block_inst, blk);
    return;
  }
  persistent_blocks(a.claim, a.blocks, &s_claim,
                    [&](uint32_t blk) { trail_block<S, R>(a, insts, params, block_inst, blk); });
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

// ---- display colour map (reference src/display.wgsl:29-87) ---------------
// One thread per texture pixel.  The fit (scale, offsets) is uniform and
// computed on the host with the shader's f32 operations; each pixel then maps
// to a cell by f32 division and truncation, looks its clamped intensity up in
// the 768-byte R[256] G[256] B[256] table, and stores RGBA bytes (alpha 255).
// The shader stores f32(byte) / 255 to an rgba8unorm texture, which converts
// back to the same byte, so the bytes are the table's.  Pixels outside the
// fitted grid are black.  kDecode: the cell's value is deposit_decay of its
// raw element (the marked buffer between the agent and trail passes).
template <int S, bool kDecode>
__global__ __launch_bounds__(256) void render_kernel(RenderArgs a, LutBytes lut) {
  __shared__ uint32_t s_lut[192];
  if (threadIdx.x < 192) s_lut[threadIdx.x] = lut.word[threadIdx.x];
  __syncthreads();
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256u + threadIdx.x;
  if (i >= static_cast<uint64_t>(a.tex_w) * a.tex_h) return;
  uint32_t px = static_cast<uint32_t>(i % a.tex_w), py = static_cast<uint32_t>(i / a.tex_w);
  float fx = (static_cast<float>(px) - a.off_x) / a.scale;
  float fy = (static_cast<float>(py) - a.off_y) / a.scale;
  uint32_t rgba = 0xff000000u;
  if (fx >= 0.0f && fx < a.sim_w && fy >= 0.0f && fy < a.sim_h) {
    uint64_t idx = static_cast<uint64_t>(static_cast<int>(fy)) * a.width +
                   static_cast<uint64_t>(static_cast<int>(fx));
    uint32_t raw = static_cast<const typename Codec<S>::elem*>(a.trail)[idx];
    float t = kDecode ? deposit_decay<S>(raw, *a.decode) : Codec<S>::value(raw);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    uint32_t k = static_cast<uint32_t>(t * 255.0f);
    auto byte = [&](uint32_t b) { return (s_lut[b >> 2] >> ((b & 3u) * 8u)) & 0xffu; };
    rgba |= byte(k) | (byte(k + 256u) << 8) | (byte(k + 512u) << 16);
  }
  a.out[i] = rgba;
}

// partials[inst * kMassBlocks + b]: f64 sum of a fixed stride of inst's cells.
template <int S>
__global__ __launch_bounds__(256) void mass_partial_kernel(const void* t, const Instance* insts,
                                                           double* partials) {
  __shared__ double red[256];
  const Instance in = insts[blockIdx.y];
  const typename Codec<S>::elem* e =
      static_cast<const typename Codec<S>::elem*>(t) + in.trail_off;
  uint64_t cells = static_cast<uint64_t>(in.width) * in.height;
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
  if (threadIdx.x == 0) partials[static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void mass_final_kernel(const double* partials, int n,
                                                         double* out) {
  __shared__ double red[256];
  const double* p = partials + static_cast<size_t>(blockIdx.x) * n;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += p[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(s)) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <int S>
void launch_trail_t(const Pool& pool, const void* in, void* out, hipStream_t stream) {
  const uint32_t nb = blocks_for(pool.total_units, kTrailWavesPerBlock);
  TrailArgs a{in,         out,         pool.inst,      pool.params, pool.unit_block_inst,
              pool.count, pool.total_units, pool.xcd_remap, pool.streaming, pool.claim_trail, nb};
  auto go = [&](auto* kernel) {
    const uint32_t grid = a.claim ? persistent_grid(kernel, nb, kTrailBlock) : nb;
    kernel<<<grid, kTrailBlock, 0, stream>>>(a, a.inst, a.params, a.block_inst);
  };
  if (pool.preload_rows == 4)
    go(trail_kernel<S, 4>);
  else if (pool.preload_rows == 8)
    go(trail_kernel<S, 8>);
  else if (pool.preload_rows == 16)
    go(trail_kernel<S, 16>);
  else
    go(trail_kernel<S, 0>);
}


// ---- Gaussian diffuse (SLM_FLAG_GAUSSIAN_DIFFUSE) ------------------------------
// One 256-thread workgroup per tile of kGaussTW x gauss_th(storage) cells.  The tile's
// deposited and decayed cells plus an R-cell toroidal halo on every side are
// staged in LDS (as 16-byte chunks where the tile lies inside a row of whole
// chunks, else one cell per lane: see gauss_tile); thread c then runs the
// horizontal pass for column c over every row of the
// tile and its halo into registers, and the vertical pass from them.  Both
// passes start from the product with the first tap's weight and add the others
// in ascending tap order with one fused multiply-add each, so the result is
// the oracle's bit for bit.  LDS is sized for the radius class RM of the pool
// (gauss_class of its largest radius); the engine recaptures its graphs when
// slm_set_params moves a radius to another class.
struct GaussArgs {
  const void* in;
  void* out;
  uint32_t count;  // instances
};

// tile rows of storage S in radius class RM (Codec<S>: the stochastic-rounding
// format is 16-bit)
template <int S, int RM>
constexpr int gauss_rows_of() {
  return static_cast<int>(sizeof(typename Codec<S>::elem) == 4 ? gauss_th(SLM_STORAGE_F32, RM)
                                                                : gauss_th(SLM_STORAGE_BF16, RM));
}

// v mod n for |v| < 2^25 and 1 <= n <= 2^24 (grid sides are at most 2^24).
__device__ inline uint32_t wrap_index(int32_t v, uint32_t n) {
  int32_t m = v % static_cast<int32_t>(n);
  return static_cast<uint32_t>(m < 0 ? m + static_cast<int32_t>(n) : m);
}

// v mod n for -n <= v < 2n.
__device__ inline uint32_t wrap_near(int32_t v, uint32_t n) {
  const int32_t m = static_cast<int32_t>(n);
  return static_cast<uint32_t>(v < 0 ? v + m : (v >= m ? v - m : v));
}

// deposit_decay for the Gaussian pass's staging.  kSat: deposit >= 1 (every
// preset), so a marked cell's clamp(|T| + deposit, 0, 1) is exactly 1 and the
// cell becomes sat = max(1 - decay, 0), computed once with the same two
// operations; an unmarked cell is max(|T| - decay, 0) as always.  Same bits,
// one select instead of an add and a select per staged cell.
template <int S, bool kSat>
__device__ inline float gauss_stage(uint32_t raw, const DevParams& p, float sat) {
  if constexpr (kSat) {
    const float u = Codec<S>::value(raw) - p.decay_k;
    const float d = u > 0.0f ? u : 0.0f;
    return Codec<S>::marked(raw) ? sat : d;
  } else {
    return deposit_decay<S>(raw, p);
  }
}

// One tile at radius R (a compile-time constant, so every loop unrolls and the
// weights sit in registers).  A tile whose 256 columns lie inside the row, in
// a row of whole 16-byte chunks, moves 16 bytes per lane: the rows are loaded
// as chunks (plus single halo cells), and the results are staged in LDS and
// stored as chunks.  Other tiles load and store one cell per lane.  Both give
// the same bits.
template <int S, int R, bool kSat, int kCols, bool kPipe, int kGaussTH>
__device__ inline void gauss_tile(float (*s_t)[kCols],
                                  const typename Codec<S>::elem* src, typename Codec<S>::elem* dst,
                                  uint32_t W, uint32_t H, uint32_t x0, uint32_t y0,
                                  const DevParams& p, const float* wts) {
  using elem = typename Codec<S>::elem;
  constexpr int rows = kGaussTH + 2 * R;
  constexpr int cols = static_cast<int>(kGaussTW) + 2 * R;
  constexpr int V = static_cast<int>(16 / sizeof(elem));  // cells per 16-byte chunk
  constexpr int CPR = static_cast<int>(kGaussTW) / V;     // chunks per tile row
  float w[R + 1];
#pragma unroll
  for (int i = 0; i <= R; i++) w[i] = wts[i];
  const float sat_u = 1.0f - p.decay_k;
  const float sat = sat_u > 0.0f ? sat_u : 0.0f;
  const int c = static_cast<int>(threadIdx.x);
  const int32_t ix0 = static_cast<int32_t>(x0), iy0 = static_cast<int32_t>(y0);
  const bool vec = W % V == 0 && x0 + kGaussTW <= W;  // block-uniform
  if (vec) {
    const bool near = H >= kGaussTH + R;  // rows y0-R .. y0+TH+R-1 wrap at most once
    constexpr int kChunks = rows * CPR;
    constexpr int kPer = (kChunks + static_cast<int>(kGaussTW) - 1) / static_cast<int>(kGaussTW);
    constexpr int kHalo = rows * 2 * R;
    constexpr int kHaloPer = (kHalo + static_cast<int>(kGaussTW) - 1) / static_cast<int>(kGaussTW);
    uint4 raw[kPer];
    uint32_t hraw[kHaloPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int q = c + k * static_cast<int>(kGaussTW);
      raw[k] = make_uint4(0u, 0u, 0u, 0u);
      if (q < kChunks) {
        const int r = q / CPR, j = q % CPR;
        const uint32_t gy = near ? wrap_near(iy0 - R + r, H) : wrap_index(iy0 - R + r, H);
        raw[k] = *reinterpret_cast<const uint4*>(src + static_cast<uint64_t>(gy) * W + x0 + j * V);
      }
    }
#pragma unroll
    for (int k = 0; k < kHaloPer; k++) {
      const int q = c + k * static_cast<int>(kGaussTW);
      hraw[k] = 0u;
      if (q < kHalo) {
        const int r = q / (2 * R), jj = q % (2 * R);
        const int off = jj < R ? jj - R : static_cast<int>(kGaussTW) + jj - R;
        const uint32_t gy = near ? wrap_near(iy0 - R + r, H) : wrap_index(iy0 - R + r, H);
        hraw[k] = src[static_cast<uint64_t>(gy) * W + wrap_near(ix0 + off, W)];
      }
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int q = c + k * static_cast<int>(kGaussTW);
      if (q < kChunks) {
        const int r = q / CPR, j = q % CPR;
#pragma unroll
        for (int e = 0; e < V; e++)
          s_t[r][R + j * V + e] = gauss_stage<S, kSat>(Codec<S>::lane(raw[k], e), p, sat);
      }
    }
#pragma unroll
    for (int k = 0; k < kHaloPer; k++) {
      const int q = c + k * static_cast<int>(kGaussTW);
      if (q < kHalo) {
        const int r = q / (2 * R), jj = q % (2 * R);
        s_t[r][jj < R ? jj : static_cast<int>(kGaussTW) + jj] = gauss_stage<S, kSat>(hraw[k], p, sat);
      }
    }
  } else {
    const uint32_t gx0 = wrap_index(ix0 + c - R, W);
    const bool second = c + static_cast<int>(kGaussTW) < cols;
    const uint32_t gx1 = wrap_index(ix0 + c + static_cast<int32_t>(kGaussTW) - R, W);
    // every row's loads are issued before any is used
    uint32_t raw0[rows], raw1[rows];
    uint32_t gy = wrap_index(iy0 - R, H);
#pragma unroll
    for (int r = 0; r < rows; r++) {
      const elem* row = src + static_cast<uint64_t>(gy) * W;
      raw0[r] = row[gx0];
      raw1[r] = second ? static_cast<uint32_t>(row[gx1]) : 0u;
      gy = gy + 1u == H ? 0u : gy + 1u;
    }
#pragma unroll
    for (int r = 0; r < rows; r++) {
      s_t[r][c] = gauss_stage<S, kSat>(raw0[r], p, sat);
      if (second) s_t[r][c + kGaussTW] = gauss_stage<S, kSat>(raw1[r], p, sat);
    }
  }
  __syncthreads();
  // the horizontal pass of the thread's own column, every row of the tile
  // and its halo, stays in registers for the vertical pass
  float h[rows];
  if constexpr (kPipe) {
    // one ds_read_b32 per tap at an immediate offset from a single base,
    // one row ahead of the arithmetic: paired into ds_read2_b32 (8-bit
    // offsets, which cannot span a 1 KiB row) the reads cost two address
    // adds per row, 7% of this VALU-bound pass.  Relaxed atomic loads are
    // never paired; nothing else orders them.
    auto ld = [&](int r, int i) {
      return __uint_as_float(__atomic_load_n(reinterpret_cast<uint32_t*>(&s_t[r][c + i]),
                                             __ATOMIC_RELAXED));
    };
    float cur[2 * R + 1], nxt[2 * R + 1];
#pragma unroll
    for (int i = 0; i <= 2 * R; i++) cur[i] = ld(0, i);
#pragma unroll
    for (int r = 0; r < rows; r++) {
      if (r + 1 < rows) {
#pragma unroll
        for (int i = 0; i <= 2 * R; i++) nxt[i] = ld(r + 1, i);
      }
      float sum = w[R] * cur[0];
#pragma unroll
      for (int i = 1; i <= 2 * R; i++) sum = __builtin_fmaf(w[i < R ? R - i : i - R], cur[i], sum);
      h[r] = sum;
#pragma unroll
      for (int i = 0; i <= 2 * R; i++) cur[i] = nxt[i];
    }
  } else {
#pragma unroll
    for (int r = 0; r < rows; r++) {
      float sum = w[R] * s_t[r][c];
#pragma unroll
      for (int i = 1; i <= 2 * R; i++)
        sum = __builtin_fmaf(w[i < R ? R - i : i - R], s_t[r][c + i], sum);
      h[r] = sum;
    }
  }
  const uint32_t x = x0 + static_cast<uint32_t>(c);
  // hashed cell index of (x, y) for stochastic rounding, stepped a row at a time
  const uint32_t hstep = W * kSrK1;
  uint32_t hc = (y0 * W + x) * kSrK1;
  if (vec) {
    uint32_t o[kGaussTH];
#pragma unroll
    for (int r = 0; r < static_cast<int>(kGaussTH); r++) {
      float g = w[R] * h[r];
#pragma unroll
      for (int j = 1; j <= 2 * R; j++) g = __builtin_fmaf(w[j < R ? R - j : j - R], h[r + j], g);
      float l = s_t[r + R][c + R] * p.one_minus_rate;
      float rr = g * p.rate;
      o[r] = encode_out<S>(l + rr, hc);  // rows past the map are never stored
      hc += hstep;
    }
    __syncthreads();  // every read of s_t is done: it becomes the output stage
    elem* s_o = reinterpret_cast<elem*>(&s_t[0][0]);  // [kGaussTH][kGaussTW]
#pragma unroll
    for (int r = 0; r < static_cast<int>(kGaussTH); r++)
      s_o[r * static_cast<int>(kGaussTW) + c] = static_cast<elem>(o[r]);
    __syncthreads();
    constexpr int kOutPer = static_cast<int>(kGaussTH) * CPR / static_cast<int>(kGaussTW);
#pragma unroll
    for (int k = 0; k < kOutPer; k++) {
      const int q = c + k * static_cast<int>(kGaussTW);
      const int r = q / CPR, j = q % CPR;
      const uint32_t y = y0 + static_cast<uint32_t>(r);
      if (y < H)
        *reinterpret_cast<uint4*>(dst + static_cast<uint64_t>(y) * W + x0 + j * V) =
            *reinterpret_cast<const uint4*>(s_o + r * static_cast<int>(kGaussTW) + j * V);
    }
    return;
  }
  if (x >= W) return;
#pragma unroll
  for (int r = 0; r < static_cast<int>(kGaussTH); r++) {
    const uint32_t y = y0 + static_cast<uint32_t>(r);
    if (y >= H) break;
    float g = w[R] * h[r];
#pragma unroll
    for (int j = 1; j <= 2 * R; j++) g = __builtin_fmaf(w[j < R ? R - j : j - R], h[r + j], g);
    float centre = s_t[r + R][c + R];
    float l = centre * p.one_minus_rate;
    float rr = g * p.rate;
    dst[y * W + x] = static_cast<elem>(encode_out<S>(l + rr, hc));
    hc += hstep;
  }
}

template <int S, int RM>
__global__ __launch_bounds__(kGaussTW) void gauss_kernel(GaussArgs a,
                                                         const Instance* __restrict__ insts,
                                                         const DevParams* __restrict__ params) {
  constexpr int kCols = static_cast<int>(kGaussTW) + 2 * RM;
  constexpr int kTH = gauss_rows_of<S, RM>();
  __shared__ __align__(16) float s_t[kTH + 2 * RM][kCols];
  using elem = typename Codec<S>::elem;
  const uint32_t t = blockIdx.x;
  uint32_t lo = 0, hi = a.count - 1u;  // last instance whose tiles start at or before t
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1u) >> 1;
    if (insts[mid].gtile_off <= t)
      lo = mid;
    else
      hi = mid - 1u;
  }
  const Instance in = insts[lo];
  const DevParams p = params[lo];
  const uint32_t lt = t - in.gtile_off;
  const uint32_t x0 = (lt % in.gtx) * kGaussTW, y0 = (lt / in.gtx) * kTH;
  const elem* src = static_cast<const elem*>(a.in) + in.trail_off;
  elem* dst = static_cast<elem*>(a.out) + in.trail_off;
  const float* wts = params[lo].gauss_w;
  // 1..RM (the host sizes RM for the pool's largest radius); any other value
  // writes nothing
  auto run = [&](auto sat) {
    constexpr bool kSat = decltype(sat)::value;
    auto tile = [&](auto r) {
      constexpr int R = decltype(r)::value;
      // the pipelined horizontal reads in the 16-bit radius classes 2 and 4
      // (radius 4: 333 -> 324 us on 16384^2 bf16); with every radius in one
      // kernel they had raised its VGPRs to 214
      constexpr bool kPipe = RM <= 4 && sizeof(elem) == 2;
      if constexpr (R <= RM)
        gauss_tile<S, R, kSat, kCols, kPipe, kTH>(s_t, src, dst, in.width, in.height, x0, y0, p,
                                                  wts);
    };
    switch (p.gauss_r) {
      case 1: tile(std::integral_constant<int, 1>{}); break;
      case 2: tile(std::integral_constant<int, 2>{}); break;
      case 3: tile(std::integral_constant<int, 3>{}); break;
      case 4: tile(std::integral_constant<int, 4>{}); break;
      case 5: tile(std::integral_constant<int, 5>{}); break;
      case 6: tile(std::integral_constant<int, 6>{}); break;
      case 7: tile(std::integral_constant<int, 7>{}); break;
      case 8: tile(std::integral_constant<int, 8>{}); break;
      default: break;
    }
  };
  // deposit >= 1 (every preset) takes the one-select staging (same bits)
  if (p.deposit >= 1.0f)
    run(std::true_type{});
  else
    run(std::false_type{});
}

template <int S>
void launch_gauss_t(const Pool& pool, const void* in, void* out, hipStream_t stream) {
  GaussArgs a{in, out, pool.count};
  // f32 tiles keep the largest LDS tile: at radius 2 the smaller one (seven
  // workgroups per CU instead of four) measured 3-4% slower on the HBM-bound
  // f32 pass, while the VALU-bound 16-bit pass gains 11% from a fourth
  auto go = [&](auto* kernel) {
    kernel<<<pool.total_gtiles, kGaussTW, 0, stream>>>(a, pool.inst, pool.params);
  };
  if constexpr (sizeof(typename Codec<S>::elem) == 4) {
    go(gauss_kernel<S, kGaussMaxR>);
  } else {
    if (pool.gauss_rclass <= 2)
      go(gauss_kernel<S, 2>);
    else if (pool.gauss_rclass <= 4)
      go(gauss_kernel<S, 4>);
    else
      go(gauss_kernel<S, kGaussMaxR>);
  }
}

}  // namespace

void plan_trail(Instance& in, int storage, uint32_t target_waves, uint32_t fixed_rows) {
  uint32_t v = storage == SLM_STORAGE_F32 ? 4u : 8u;
  if (in.width % v == 0) {
    in.ntx = (in.width + 64u * v - 1u) / (64u * v);
    // rows per strip unit: fixed, or the longest of 64..4 that still gives
    // target_waves units
    uint32_t rows = fixed_rows ? fixed_rows : 64u;
    while (!fixed_rows && rows > 4 &&
           static_cast<uint64_t>(in.ntx) * ((in.height + rows - 1) / rows) < target_waves)
      rows >>= 1;
    in.tile_rows = rows;
    in.units = in.ntx * ((in.height + rows - 1u) / rows);
  } else {
    in.ntx = 0;
    in.tile_rows = 0;
    in.units = blocks_for(static_cast<uint64_t>(in.width) * in.height, 64u);
  }
}

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
  d.gauss_r = gauss_weights(s.blur_radius, s.blur_sigma, d.gauss_w);
  return d;
}

int gauss_weights(float blur_radius, float blur_sigma, float* w) {
  for (int i = 0; i <= kGaussMaxR; i++) w[i] = 0.0f;
  if (!(blur_radius >= 0.5f && blur_radius < kGaussMaxR + 0.5f) || !(blur_sigma > 0.0f) ||
      !std::isfinite(blur_sigma))
    return 0;
  const int r = static_cast<int>(std::floor(static_cast<double>(blur_radius) + 0.5));
  const double s2 = 2.0 * static_cast<double>(blur_sigma) * blur_sigma;
  double e[kGaussMaxR + 1];
  double sum = 0.0;
  for (int i = 0; i <= r; i++) {
    e[i] = std::exp(-static_cast<double>(i * i) / s2);
    sum += i == 0 ? e[i] : 2.0 * e[i];
  }
  for (int i = 0; i <= r; i++) w[i] = static_cast<float>(e[i] / sum);
  return r;
}

void launch_trail(const Pool& pool, const void* in, void* out, hipStream_t stream) {
  if (pool.gaussian) {
    if (pool.total_gtiles == 0) return;
    switch (pool.storage) {
      case SLM_STORAGE_BF16:
        if (pool.stochastic_rounding)
          launch_gauss_t<kStorageBf16Sr>(pool, in, out, stream);
        else
          launch_gauss_t<SLM_STORAGE_BF16>(pool, in, out, stream);
        break;
      case SLM_STORAGE_F16:
        launch_gauss_t<SLM_STORAGE_F16>(pool, in, out, stream);
        break;
      default:
        launch_gauss_t<SLM_STORAGE_F32>(pool, in, out, stream);
    }
    return;
  }
  if (pool.total_units == 0) return;
  switch (pool.storage) {
    case SLM_STORAGE_BF16:
      if (pool.stochastic_rounding)
        launch_trail_t<kStorageBf16Sr>(pool, in, out, stream);
      else
        launch_trail_t<SLM_STORAGE_BF16>(pool, in, out, stream);
      break;
    case SLM_STORAGE_F16:
      launch_trail_t<SLM_STORAGE_F16>(pool, in, out, stream);
      break;
    default:
      launch_trail_t<SLM_STORAGE_F32>(pool, in, out, stream);
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

void render_fit(uint32_t width, uint32_t height, uint32_t tex_w, uint32_t tex_h, RenderArgs& a) {
  // src/display.wgsl:48-72, in f32 as the shader evaluates it
  float sim_w = static_cast<float>(width), sim_h = static_cast<float>(height);
  float tw = static_cast<float>(tex_w), th = static_cast<float>(tex_h);
  float sim_aspect = sim_w / sim_h;
  float tex_aspect = tw / th;
  float scale, off_x = 0.0f, off_y = 0.0f;
  if (tex_aspect > sim_aspect) {
    scale = th / sim_h;
    off_x = (tw - sim_w * scale) * 0.5f;
  } else {
    scale = tw / sim_w;
    off_y = (th - sim_h * scale) * 0.5f;
  }
  a.width = width;
  a.sim_w = sim_w;
  a.sim_h = sim_h;
  a.tex_w = tex_w;
  a.tex_h = tex_h;
  a.scale = scale;
  a.off_x = off_x;
  a.off_y = off_y;
}

void launch_render(const RenderArgs& a, const LutBytes& lut, int storage, hipStream_t stream) {
  uint64_t px = static_cast<uint64_t>(a.tex_w) * a.tex_h;
  if (px == 0) return;
  uint32_t nb = blocks_for(px, 256);
  const bool dec = a.decode != nullptr;
  if (storage == SLM_STORAGE_BF16)
    dec ? render_kernel<SLM_STORAGE_BF16, true><<<nb, 256, 0, stream>>>(a, lut)
        : render_kernel<SLM_STORAGE_BF16, false><<<nb, 256, 0, stream>>>(a, lut);
  else if (storage == SLM_STORAGE_F16)
    dec ? render_kernel<SLM_STORAGE_F16, true><<<nb, 256, 0, stream>>>(a, lut)
        : render_kernel<SLM_STORAGE_F16, false><<<nb, 256, 0, stream>>>(a, lut);
  else
    dec ? render_kernel<SLM_STORAGE_F32, true><<<nb, 256, 0, stream>>>(a, lut)
        : render_kernel<SLM_STORAGE_F32, false><<<nb, 256, 0, stream>>>(a, lut);
}

void prime_persistent_trail() {
  auto each = [](auto s_tag) {
    constexpr int S = decltype(s_tag)::value;
    (void)persistent_grid(trail_kernel<S, 0>, 8, kTrailBlock);
    (void)persistent_grid(trail_kernel<S, 4>, 8, kTrailBlock);
    (void)persistent_grid(trail_kernel<S, 8>, 8, kTrailBlock);
    (void)persistent_grid(trail_kernel<S, 16>, 8, kTrailBlock);
  };
  each(std::integral_constant<int, SLM_STORAGE_F32>{});
  each(std::integral_constant<int, SLM_STORAGE_BF16>{});
  each(std::integral_constant<int, kStorageBf16Sr>{});
  each(std::integral_constant<int, SLM_STORAGE_F16>{});
}

void launch_total_mass(const Pool& pool, const void* t, double* partials, double* out,
                       hipStream_t stream) {
  if (pool.count == 0) return;
  dim3 grid(kMassBlocks, pool.count);
  if (pool.storage == SLM_STORAGE_BF16)
    mass_partial_kernel<SLM_STORAGE_BF16><<<grid, 256, 0, stream>>>(t, pool.inst, partials);
  else if (pool.storage == SLM_STORAGE_F16)
    mass_partial_kernel<SLM_STORAGE_F16><<<grid, 256, 0, stream>>>(t, pool.inst, partials);
  else
    mass_partial_kernel<SLM_STORAGE_F32><<<grid, 256, 0, stream>>>(t, pool.inst, partials);
  mass_final_kernel<<<pool.count, 256, 0, stream>>>(partials, kMassBlocks, out);
}

}  // namespace slm
