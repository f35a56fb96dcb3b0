// slm_sort.hip - periodic spatial re-sort of the agents (no counterpart in the
// reference, whose agents keep their creation order, src/main.rs:263-282).
//
// Agents are bucketed by the Morton (Z-order) code of the 2^k x 2^k cell bin
// they stand in, then scattered bin by bin into the other agent buffer, so a
// wave's lanes sense and deposit in nearby cells and share L2 lines.  The slot
// -> original index table id[] moves with them.  The order inside a bin comes
// from atomics and is not fixed, which is harmless: an agent's update depends
// only on its own state, its original index and the step-start trail, and the
// deposit marks commute, so any permutation gives the same result.
#include "slm_internal.h"

namespace slm {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanTile = kScanBlock * kScanPerThread;  // bins per scan block

__device__ inline uint32_t spread_bits(uint32_t v) {  // 16 bits -> even positions
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

struct BinGeom {
  uint32_t cell_log2;
  uint32_t nbx, nby;  // bins across / down
};

__device__ inline uint32_t bin_of(float x, float y, const BinGeom& b) {
  uint32_t bx = static_cast<uint32_t>(x) >> b.cell_log2;  // x in [0, width]
  uint32_t by = static_cast<uint32_t>(y) >> b.cell_log2;
  bx = bx < b.nbx ? bx : b.nbx - 1u;
  by = by < b.nby ? by : b.nby - 1u;
  return spread_bits(bx) | (spread_bits(by) << 1);
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Bins each agent and takes its rank inside the bin.  Lanes of a wave that share
// a bin are served by one atomic (agents re-sorted earlier arrive grouped by
// bin, so a wave usually holds one to four bins); ranks follow lane order.
__global__ void count_kernel(AgentSoA a, uint32_t n, BinGeom b, uint32_t* key, uint32_t* rank,
                             uint32_t* count) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  bool active = j < n;
  uint32_t k = active ? bin_of(a.x[j], a.y[j], b) : 0u;
  uint32_t lane = threadIdx.x & 63u;
  uint64_t below = (uint64_t(1) << lane) - 1u;
  uint64_t pending = __ballot(active);
  uint32_t r = 0;
  while (pending) {
    int leader = __builtin_ctzll(pending);
    uint32_t lk = __shfl(k, leader);
    uint64_t same = __ballot(active && k == lk);
    uint32_t base = 0;
    if (static_cast<int>(lane) == leader) base = atomicAdd(count + lk, __popcll(same));
    base = __shfl(base, leader);
    if (active && k == lk) r = base + __popcll(same & below);
    pending &= ~same;
  }
  if (active) {
    key[j] = k;
    rank[j] = r;
  }
}

// Exclusive scan of one tile of counts; writes the tile total to sums[tile].
__global__ __launch_bounds__(kScanBlock) void scan_tile_kernel(const uint32_t* in, uint32_t n,
                                                               uint32_t* out, uint32_t* sums) {
  __shared__ uint32_t sh[kScanBlock];
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
  uint32_t v[kScanPerThread];
  uint32_t local = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    v[k] = base + k < n ? in[base + k] : 0u;
    local += v[k];
  }
  sh[threadIdx.x] = local;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {  // inclusive Hillis-Steele
    uint32_t t = threadIdx.x >= static_cast<unsigned>(off) ? sh[threadIdx.x - off] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - local;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = sh[threadIdx.x];
}

// Adds the exclusive prefix of the tile totals (scanned by one block) to every
// tile; `ntiles` <= kScanTile.
__global__ __launch_bounds__(kScanBlock) void scan_add_kernel(uint32_t* out, uint32_t n,
                                                              const uint32_t* sums,
                                                              uint32_t ntiles) {
  __shared__ uint32_t prefix;
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (uint32_t t = 0; t < blockIdx.x && t < ntiles; t++) s += sums[t];
    prefix = s;
  }
  __syncthreads();
  uint32_t base = blockIdx.x * kScanTile;
  for (uint32_t k = threadIdx.x; k < kScanTile; k += kScanBlock)
    if (base + k < n) out[base + k] += prefix;
}

__global__ void scatter_kernel(AgentSoA src, AgentSoA dst, uint32_t n, const uint32_t* key,
                               const uint32_t* rank, const uint32_t* offset) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t d = offset[key[j]] + rank[j];
  dst.x[d] = src.x[j];
  dst.y[d] = src.y[j];
  dst.th[d] = src.th[j];
  dst.v[d] = src.v[j];
  dst.id[d] = src.id[j];
}

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

inline uint32_t next_pow2(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace

uint32_t sort_bins_for(uint32_t width, uint32_t height, uint32_t cell_log2) {
  uint32_t side = 1u << cell_log2;
  uint32_t nbx = (width + side - 1) / side, nby = (height + side - 1) / side;
  uint32_t p = next_pow2(nbx > nby ? nbx : nby);
  return p * p;
}

uint32_t sort_block_sums_for(uint32_t nbins) { return blocks_for(nbins, kScanTile); }

void launch_morton_sort(const StepGeom& g, const AgentSoA& src, const AgentSoA& dst,
                        const SortScratch& s, hipStream_t stream) {
  if (g.agents == 0) return;
  uint32_t side = 1u << s.cell_log2;
  BinGeom b{s.cell_log2, (g.width + side - 1) / side, (g.height + side - 1) / side};
  uint32_t ntiles = sort_block_sums_for(s.nbins);
  zero_u32_kernel<<<blocks_for(s.nbins, 256), 256, 0, stream>>>(s.count, s.nbins);
  count_kernel<<<blocks_for(g.agents, 256), 256, 0, stream>>>(src, g.agents, b, s.key, s.rank,
                                                             s.count);
  scan_tile_kernel<<<ntiles, kScanBlock, 0, stream>>>(s.count, s.nbins, s.offset, s.block_sums);
  scan_add_kernel<<<ntiles, kScanBlock, 0, stream>>>(s.offset, s.nbins, s.block_sums, ntiles);
  scatter_kernel<<<blocks_for(g.agents, 256), 256, 0, stream>>>(src, dst, g.agents, s.key, s.rank,
                                                               s.offset);
}

}  // namespace slm
