// slm_sort.hip - periodic spatial re-sort of the agents (no counterpart in the
// reference, whose agents keep their creation order, src/main.rs:263-282).
//
// Agents are bucketed by the Morton (Z-order) code of the 2^k x 2^k cell bin
// they stand in, then scattered bin by bin into the other agent buffer, so a
// wave's lanes sense and deposit in nearby cells and share L2 lines.  Each
// instance of a pool owns a contiguous range of bins, so agents stay inside
// their instance's slot range.  The slot -> original index table id[] moves
// with them.  The order inside a bin comes from atomics and is not fixed,
// which is harmless: an agent's update depends only on its own state, its
// original index and the step-start trail, and the deposit marks commute, so
// any permutation gives the same result.
#include "slm_internal.h"

namespace slm {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanPerThread = 4;
constexpr uint32_t kScanTile = kScanBlock * kScanPerThread;  // entries per scan block

__device__ inline uint32_t spread_bits(uint32_t v) {  // 16 bits -> even positions
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// float -> u32 of a value that may lie below 0 or beyond 2^32 (clamped).
__device__ inline uint32_t clamp_u32(float v) {
  v = __builtin_fmaxf(v, 0.0f);
  v = __builtin_fminf(v, 4294967040.0f);
  return static_cast<uint32_t>(v);
}

// The re-sort key of an agent at (x, y) with heading th: the Morton code of
// its 2^cell_log2 bin - with SLM_SORT_AT_SENSOR the bin holding the point its
// centre sensor reads, (x, y) + sensor_dist * (cos th, sin th) - followed by
// k = sort_key & SLM_SORT_HEADING_MASK bits of heading class.  The point only
// has to be near the true one, so it comes from the hardware sin/cos
// (v_sin/v_cos_f32, argument in revolutions).  The count and scatter kernels
// evaluate this function on the same loaded values, so they agree on every
// key.
__device__ inline uint32_t key_of(float x, float y, float th, const Instance& in,
                                  const DevParams& p, uint32_t cell_log2, uint32_t sort_key) {
  const uint32_t hl = sort_key & SLM_SORT_HEADING_MASK;
  const float rev = th * 0.159154943f;
  if (sort_key & SLM_SORT_AT_SENSOR) {
    x = __builtin_fmaf(p.sensor_dist, __builtin_amdgcn_cosf(rev), x);
    y = __builtin_fmaf(p.sensor_dist, __builtin_amdgcn_sinf(rev), y);
  }
  uint32_t bx = clamp_u32(x) >> cell_log2;  // x in [0, width] without the sensor offset
  uint32_t by = clamp_u32(y) >> cell_log2;
  bx = bx < in.nbx ? bx : in.nbx - 1u;
  by = by < in.nby ? by : in.nby - 1u;
  const uint32_t m = spread_bits(bx) | (spread_bits(by) << 1);
  const uint32_t h = clamp_u32(rev * static_cast<float>(1u << hl)) & ((1u << hl) - 1u);
  return in.bin_off + ((m << hl) | h);
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Bins each agent and takes its rank inside the bin, with as few global
// atomics as possible: agents re-sorted earlier arrive grouped by bin, so a
// block's agents usually span a handful of consecutive bins.  Each agent takes
// its rank in the block from an LDS atomic on its bin's counter, and one
// global atomic per distinct bin of the block reserves the block's range.
// (Dense instances put ~10^4 agents in one bin; a global atomic per wave and
// bin serialised ~10^2 times more.)  Blocks whose bins span a wider range fall
// back to one global atomic per wave and bin.  Ranks are not in a fixed order.
//
// A block bins kCountPer x 256 agents, each thread kCountPer of them (slots
// q*256 + threadIdx.x of the block's range), all loaded before any is used:
// the kernel is a chain of load -> barrier -> global atomic -> barrier per
// block, and with one agent per thread that latency, not bandwidth, set its
// time.
constexpr uint32_t kLocalBins = 2048;
constexpr int kCountPer = 4;
constexpr uint32_t kCountSlots = kCountPer * kAgentBlock;

__global__ __launch_bounds__(kAgentBlock) void count_kernel(
    AgentSoA a, const Instance* insts, const DevParams* params, const uint32_t* block_inst,
    uint32_t ninst, uint32_t total, int pos64, uint32_t cell_log2, uint32_t sort_key,
    int xcd_remap, uint32_t* rank, uint32_t* count) {
  __shared__ uint32_t s_min, s_max;
  __shared__ uint32_t s_cnt[kLocalBins];
  __shared__ uint32_t s_base[kLocalBins];
  // XCD-contiguous blocks, as in the scatter: neighbouring slots, their bins'
  // counters and their output ranges stay with one XCD's L2 (the re-sort of
  // the 256-instance mix measured 10% faster)
  const uint32_t base =
      (xcd_remap ? xcd_logical_block(blockIdx.x, gridDim.x) : blockIdx.x) * kCountSlots;
  uint32_t k[kCountPer];
  bool act[kCountPer];
  float px[kCountPer], py[kCountPer], pth[kCountPer];
  // the heading only enters refined keys: the plain key loads positions only
  const bool heading = sort_key != 0u;
#pragma unroll
  for (int q = 0; q < kCountPer; q++) {
    uint32_t j = base + q * kAgentBlock + threadIdx.x;
    // 64-bit test: near the 2^32 - 256 agent limit the last block's slots wrap
    act[q] = uint64_t(base) + q * kAgentBlock + threadIdx.x < total;
    k[q] = 0;
    px[q] = py[q] = pth[q] = 0.0f;
    if (act[q]) {
      px[q] = pos64 ? static_cast<float>(a.x64[j]) : a.x[j];
      py[q] = pos64 ? static_cast<float>(a.y64[j]) : a.y[j];
      if (heading) pth[q] = a.th[j];
    }
  }
#pragma unroll
  for (int q = 0; q < kCountPer; q++) {
    if (uint64_t(base) + q * kAgentBlock >= total) break;  // block-uniform; no block_inst entry
    uint32_t first = base + q * kAgentBlock;
    uint32_t last = min(first + kAgentBlock, total) - 1u;
    uint32_t j = first + threadIdx.x;
    for (uint32_t ii = block_inst[first / kAgentBlock];; ii++) {  // block-uniform walk
      const Instance in = insts[ii];
      uint32_t end = in.agent_off + in.agent_count;
      if (act[q] && j >= in.agent_off && j < end)
        k[q] = key_of(px[q], py[q], pth[q], in, params[ii], cell_log2, sort_key);
      if (last < end || ii + 1u >= ninst) break;
    }
  }
  if (threadIdx.x == 0) {
    s_min = 0xffffffffu;
    s_max = 0u;
  }
  __syncthreads();
  // wave min/max by butterfly, then one LDS atomic per wave (256 same-address
  // LDS atomics per block serialised and dominated the kernel)
  uint32_t wmin = 0xffffffffu, wmax = 0u;
#pragma unroll
  for (int q = 0; q < kCountPer; q++)
    if (act[q]) {
      wmin = min(wmin, k[q]);
      wmax = max(wmax, k[q]);
    }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    wmin = min(wmin, static_cast<uint32_t>(__shfl_xor(static_cast<int>(wmin), off)));
    wmax = max(wmax, static_cast<uint32_t>(__shfl_xor(static_cast<int>(wmax), off)));
  }
  if ((threadIdx.x & 63u) == 0u) {
    atomicMin(&s_min, wmin);
    atomicMax(&s_max, wmax);
  }
  __syncthreads();
  const uint32_t kmin = s_min, kmax = s_max;
  const bool local = kmax >= kmin && kmax - kmin < kLocalBins;  // block-uniform
  if (local)
    for (uint32_t i = threadIdx.x; i <= kmax - kmin; i += kAgentBlock) s_cnt[i] = 0u;
  __syncthreads();

  uint32_t r[kCountPer];
#pragma unroll
  for (int q = 0; q < kCountPer; q++) r[q] = 0;
  if (local) {
    // rank inside the block: one LDS atomic per agent (lanes of a bin take
    // distinct old values; the few bins of a block spread over the banks)
#pragma unroll
    for (int q = 0; q < kCountPer; q++)
      if (act[q]) r[q] = atomicAdd(&s_cnt[k[q] - kmin], 1u);
  } else {
    // wide key range: one global atomic per distinct bin of each wave
    uint32_t lane = threadIdx.x & 63u;
    uint64_t below = (uint64_t(1) << lane) - 1u;
#pragma unroll
    for (int q = 0; q < kCountPer; q++) {
      uint64_t pending = __ballot(act[q]);
      while (pending) {
        int leader = __builtin_ctzll(pending);
        uint32_t lk = __shfl(k[q], leader);
        uint64_t same = __ballot(act[q] && k[q] == lk);
        uint32_t b = 0;
        if (static_cast<int>(lane) == leader) b = atomicAdd(count + lk, __popcll(same));
        b = __shfl(b, leader);
        if (act[q] && k[q] == lk) r[q] = b + __popcll(same & below);
        pending &= ~same;
      }
    }
  }
  if (local) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i <= kmax - kmin; i += kAgentBlock) {
      uint32_t c = s_cnt[i];
      if (c) s_base[i] = atomicAdd(count + kmin + i, c);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kCountPer; q++)
      if (act[q]) r[q] += s_base[k[q] - kmin];
  }
#pragma unroll
  for (int q = 0; q < kCountPer; q++)
    if (act[q]) {
      uint32_t j = base + q * kAgentBlock + threadIdx.x;
      rank[j] = r[q];
    }
}

// Exclusive scan of one tile of `in` into `out` (in place allowed); writes
// the tile total to sums[tile].
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

// out[i] += prefix[tile of i] (the exclusively scanned tile totals).
__global__ __launch_bounds__(kScanBlock) void scan_add_kernel(uint32_t* out, uint32_t n,
                                                              const uint32_t* prefix) {
  uint32_t add = prefix[blockIdx.x];
  uint32_t base = blockIdx.x * kScanTile;
  for (uint32_t k = threadIdx.x; k < kScanTile; k += kScanBlock)
    if (base + k < n) out[base + k] += add;
}

// Moves each agent to offset[bin] + rank.  The bin is recomputed from the
// position the scatter reads anyway (the same float arithmetic as the count
// kernel's, so the same bin) instead of being stored by the count kernel and
// read back: 8 bytes per agent less traffic, ~15% of the re-sort.
__global__ __launch_bounds__(kAgentBlock) void scatter_kernel(
    AgentSoA src, AgentSoA dst, const Instance* insts, const DevParams* params,
    const uint32_t* block_inst, uint32_t ninst, uint32_t n, int pos64, uint32_t cell_log2,
    uint32_t sort_key, int xcd_remap, const uint32_t* rank, const uint32_t* offset) {
  const uint32_t blk = xcd_remap ? xcd_logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
  uint32_t first = blk * kAgentBlock;
  uint32_t last = min(first + kAgentBlock, n) - 1u;
  uint32_t j = first + threadIdx.x;
  bool act = j < n;
  double x64 = 0.0, y64 = 0.0;
  float x = 0.0f, y = 0.0f, th = 0.0f;
  if (act) {
    th = src.th[j];
    if (pos64) {
      x64 = src.x64[j];
      y64 = src.y64[j];
      x = static_cast<float>(x64);
      y = static_cast<float>(y64);
    } else {
      x = src.x[j];
      y = src.y[j];
    }
  }
  uint32_t k = 0;
  for (uint32_t ii = block_inst[blk];; ii++) {  // block-uniform walk
    const Instance in = insts[ii];
    uint32_t end = in.agent_off + in.agent_count;
    if (act && j >= in.agent_off && j < end) k = key_of(x, y, th, in, params[ii], cell_log2, sort_key);
    if (last < end || ii + 1u >= ninst) break;
  }
  if (!act) return;
  const uint32_t d = offset[k] + rank[j];
  if (d >= n) return;  // unreachable: count and scatter compute the same keys
  if (pos64) {
    dst.x64[d] = x64;
    dst.y64[d] = y64;
  } else {
    dst.x[d] = x;
    dst.y[d] = y;
  }
  dst.th[d] = th;
  dst.v[d] = src.v[j];
  dst.id[d] = src.id[j];
}

__global__ __launch_bounds__(kAgentBlock) void canonical_copy_kernel(
    AgentSoA src, AgentSoA dst, const Instance* insts, const uint32_t* block_inst, uint32_t total,
    int pos64) {
  uint32_t j = blockIdx.x * kAgentBlock + threadIdx.x;
  if (j >= total) return;
  uint32_t ii = block_inst[blockIdx.x];
  while (j >= insts[ii].agent_off + insts[ii].agent_count) ii++;  // ends: j < total
  uint32_t off = insts[ii].agent_off;
  uint32_t local = src.id ? src.id[j] : j - off;
  uint32_t d = off + local;
  if (pos64) {
    dst.x64[d] = src.x64[j];
    dst.y64[d] = src.y64[j];
  } else {
    dst.x[d] = src.x[j];
    dst.y[d] = src.y[j];
  }
  dst.th[d] = src.th[j];
  dst.v[d] = src.v[j];
  if (dst.id) dst.id[d] = local;
}

// Work items of each tile: ceil(agents / kItemAgents), 0 for an empty tile.
__global__ void tile_count_kernel(const Instance* insts, const uint32_t* tile_inst,
                                  uint32_t ntiles, uint32_t cell_log2, const uint32_t* offset,
                                  uint32_t* tile_items) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntiles) return;
  const Instance in = insts[tile_inst[g]];
  uint32_t t = g - in.tile_off;
  uint32_t per = 1u << (2u * (in.tile_log2 - cell_log2));  // bins per tile
  uint32_t b0 = in.bin_off + t * per;
  uint32_t n = offset[b0 + per] - offset[b0];
  tile_items[g] = (n + kItemAgents - 1u) / kItemAgents;
}

// Writes each tile's work items at its scanned offset; the last tile also
// writes the total.
__global__ void tile_items_kernel(const Instance* insts, const uint32_t* tile_inst,
                                  uint32_t ntiles, uint32_t cell_log2, const uint32_t* offset,
                                  const uint32_t* item_off, WorkItem* items, uint32_t* total) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntiles) return;
  uint32_t ii = tile_inst[g];
  const Instance in = insts[ii];
  uint32_t t = g - in.tile_off;
  uint32_t per = 1u << (2u * (in.tile_log2 - cell_log2));
  uint32_t b0 = in.bin_off + t * per;
  uint32_t begin = offset[b0], end = offset[b0 + per];
  uint32_t at = item_off[g];
  for (uint32_t s = begin; s < end; s += kItemAgents)
    items[at++] = WorkItem{ii, t, s, min(s + kItemAgents, end)};
  if (g == ntiles - 1u) *total = at;
}

inline uint32_t next_pow2(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Exclusive scan of in[0..n) into out, using the levels from `level` on.
void scan(const SortScratch& s, const uint32_t* in, uint32_t n, uint32_t* out, int level,
          hipStream_t stream) {
  uint32_t tiles = blocks_for(n, kScanTile);
  scan_tile_kernel<<<tiles, kScanBlock, 0, stream>>>(in, n, out, s.level[level]);
  if (tiles > 1) {
    scan(s, s.level[level], tiles, s.level[level], level + 1, stream);
    scan_add_kernel<<<tiles, kScanBlock, 0, stream>>>(out, n, s.level[level]);
  }
}

}  // namespace

uint64_t plan_bins(Instance& in, uint32_t cell_log2) {
  uint32_t side = 1u << cell_log2;
  in.nbx = (in.width + side - 1) / side;
  in.nby = (in.height + side - 1) / side;
  uint64_t p = next_pow2(in.nbx > in.nby ? in.nbx : in.nby);
  // a wide grid with small bins can need 2^32 or more; the caller rejects
  // counts beyond its limit (2^30), below which bin coordinates fit the
  // 16-bit Morton code
  in.nbins = static_cast<uint32_t>(std::min<uint64_t>(p * p, 0x80000000ull));
  return p * p;
}

uint32_t plan_scan(SortScratch& s, uint32_t nbins) {
  uint32_t n = nbins, total = 0;
  s.levels = 0;
  for (int k = 0; k < 3; k++) {
    n = blocks_for(n, kScanTile);
    s.level_len[k] = n;
    total += n;
    s.levels = k + 1;
    if (n <= 1) break;
  }
  return total;
}

void plan_tiles(Instance& in, uint32_t cell_log2, uint32_t tile_log2) {
  in.tile_log2 = tile_log2 > cell_log2 ? tile_log2 : cell_log2;
  in.tiles = in.nbins >> (2u * (in.tile_log2 - cell_log2));
  if (in.tiles == 0) {  // a whole grid smaller than one tile: one tile of all bins
    in.tile_log2 = cell_log2;
    while ((1u << (2u * (in.tile_log2 - cell_log2))) < in.nbins) in.tile_log2++;
    in.tiles = 1;
  }
}

void launch_morton_sort(const Pool& pool, const AgentSoA& src, const AgentSoA& dst,
                        const SortScratch& s, hipStream_t stream) {
  if (pool.total_agents == 0) return;
  zero_u32_kernel<<<blocks_for(pool.total_bins, 256), 256, 0, stream>>>(s.count, pool.total_bins);
  count_kernel<<<blocks_for(pool.total_agents, kCountSlots), kAgentBlock, 0, stream>>>(
      src, pool.inst, pool.params, pool.agent_block_inst, pool.count, pool.total_agents,
      pool.pos64, pool.sort_cell_log2, pool.sort_key, pool.xcd_remap, s.rank, s.count);
  scan(s, s.count, pool.total_bins, s.offset, 0, stream);
  scatter_kernel<<<blocks_for(pool.total_agents, kAgentBlock), kAgentBlock, 0, stream>>>(
      src, dst, pool.inst, pool.params, pool.agent_block_inst, pool.count, pool.total_agents,
      pool.pos64, pool.sort_cell_log2, pool.sort_key, pool.xcd_remap, s.rank, s.offset);
  if (pool.items) {  // work list of the tiled agent pass
    uint32_t nt = pool.total_tiles;
    tile_count_kernel<<<blocks_for(nt, 256), 256, 0, stream>>>(
        pool.inst, pool.tile_inst, nt, pool.sort_cell_log2, s.offset, s.tile_items);
    scan(s, s.tile_items, nt, s.tile_items, 0, stream);
    tile_items_kernel<<<blocks_for(nt, 256), 256, 0, stream>>>(
        pool.inst, pool.tile_inst, nt, pool.sort_cell_log2, s.offset, s.tile_items, pool.items,
        pool.item_total);
  }
}

void launch_canonical_copy(const Pool& pool, const AgentSoA& src, const AgentSoA& dst,
                           hipStream_t stream) {
  if (pool.total_agents == 0) return;
  canonical_copy_kernel<<<blocks_for(pool.total_agents, kAgentBlock), kAgentBlock, 0, stream>>>(
      src, dst, pool.inst, pool.agent_block_inst, pool.total_agents, pool.pos64);
}

}  // namespace slm
