// slm_api.hip - the C ABI declared in include/slimemold.h.
//
// Both handle types wrap one Engine: a pool of instances (a single simulation
// is a pool of one) laid out in shared device buffers -
//   trail[2]      every instance's trail, used alternately (the trail pass
//                 reads one and writes the other);
//   agents[2]     SoA agent arrays, instance by instance (the second set only
//                 when re-sorting, which scatters from one set to the other);
//   tables        the instance table, per-instance parameters and the
//                 block -> first-instance maps the kernels walk;
// plus re-sort scratch, f64 mass scratch and one non-blocking HIP stream.
// Steps are enqueued eagerly or as captured hipGraphs of `graph_chunk` steps;
// a graph is keyed by which trail and agent buffers are current when it
// starts, so at most four exist per handle.  Every entry point validates its
// arguments before any kernel is launched, so the kernels can rely on
// in-range shapes and finite, bounded positions.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "slm_internal.h"

using namespace slm;

namespace {

// A fixed per-thread buffer, so that recording an error never allocates (and
// so never throws, even when the error being recorded is a failed allocation).
thread_local char g_last_error[512] = "no error";

int fail(int code, const char* fmt, ...) {
  char buf[sizeof g_last_error];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  memcpy(g_last_error, buf, sizeof buf);
  return code;
}

// Converts the exception in flight into a status code.  Host-side containers
// (staging copies of a trail, graph tables, group tables) can throw
// std::bad_alloc; nothing may unwind through the C ABI, so every entry point
// below wraps its body in SLM_ENTRY_BEGIN / SLM_ENTRY_END.
int fail_exception() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return fail(SLM_ERR_OUT_OF_MEMORY, "host memory allocation failed");
  } catch (const std::exception& e) {
    return fail(SLM_ERR_HIP, "internal error: %s", e.what());
  } catch (...) {
    return fail(SLM_ERR_HIP, "internal error: unknown exception");
  }
}

#define SLM_ENTRY_BEGIN try {
#define SLM_ENTRY_END      \
  }                        \
  catch (...) {            \
    return fail_exception(); \
  }

#define SLM_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(SLM_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                \
  } while (0)

#define SLM_TRY(expr)         \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != SLM_OK) return rc_; \
  } while (0)

// Grid sides below 2^24: f32(width) is exact, and a row offset y * width with
// y < height is a 24-bit product (full-rate v_mul_u32_u24 in the agent pass).
constexpr uint32_t kMaxSide = (1u << 24) - 1u;
constexpr uint32_t kMaxTexSide = 1u << 24;
constexpr uint64_t kMaxCells = 1ull << 31;        // i32 cell index (compute.wgsl:21)
constexpr uint64_t kMaxPoolCells = 1ull << 40;
constexpr uint64_t kMaxAgents = 0xffffff00ull;    // u32 slot index, whole blocks
constexpr float kMaxSensorDistance = 4194304.0f;  // keeps sensor taps inside i32
constexpr uint64_t kMaxTotalBins = 1ull << 30;
constexpr float kMaxTrailValue = 0x1p100f;        // keeps every 3x3 window sum finite
constexpr uint32_t kTargetWaves = 8192;           // trail-pass units per pool
constexpr uint32_t kKnownFlags = SLM_FLAG_NO_XCD_REMAP | SLM_FLAG_F64_POSITIONS |
                                 SLM_FLAG_TILED_SENSING | SLM_FLAG_STOCHASTIC_ROUNDING |
                                 SLM_FLAG_GAUSSIAN_DIFFUSE | SLM_FLAG_WINDOW_SENSING |
                                 SLM_FLAG_PERSISTENT;

bool finite_params(const slm_params& p) {
  const float* f = &p.decay_factor;
  for (size_t i = 0; i < sizeof(slm_params) / sizeof(float); i++)
    if (!std::isfinite(f[i])) return false;
  return true;
}

// `flags`: the handle's SLM_FLAG_* (the Gaussian diffuse reads the blur fields).
int check_params(const slm_params* p, uint32_t flags) {
  if (!p) return fail(SLM_ERR_INVALID_ARGUMENT, "params is NULL");
  if (!finite_params(*p)) return fail(SLM_ERR_INVALID_ARGUMENT, "params must be finite");
  if (std::fabs(p->agent_sensor_distance) > kMaxSensorDistance)
    return fail(SLM_ERR_INVALID_ARGUMENT, "|agent_sensor_distance| must be <= 2^22");
  float w[kGaussMaxR + 1];
  if ((flags & SLM_FLAG_GAUSSIAN_DIFFUSE) && gauss_weights(p->blur_radius, p->blur_sigma, w) == 0)
    return fail(SLM_ERR_INVALID_ARGUMENT,
                "Gaussian diffuse needs round(blur_radius) in 1..%d and blur_sigma > 0", kGaussMaxR);
  return SLM_OK;
}

int check_grid(uint32_t w, uint32_t h) {
  if (w == 0 || h == 0) return fail(SLM_ERR_INVALID_ARGUMENT, "width and height must be >= 1");
  if (w > kMaxSide || h > kMaxSide)
    return fail(SLM_ERR_INVALID_ARGUMENT, "width and height must be < 2^24");
  if (static_cast<uint64_t>(w) * h > kMaxCells)
    return fail(SLM_ERR_INVALID_ARGUMENT, "width*height must be <= 2^31");
  return SLM_OK;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// Settings shared by every instance of an engine.
struct EngineConfig {
  int storage = SLM_STORAGE_F32;
  int device = 0;
  uint32_t resort_interval = 0;
  uint32_t graph_chunk = 0;
  uint32_t sort_cell_log2 = 4;
  uint32_t flags = 0;
  uint32_t agents_per_group = kAgentBlock;
  uint32_t trail_rows = 0;  // 0: chosen per instance by plan_trail
  uint32_t instance_groups = 1;
  uint32_t streaming = 0;
  uint32_t sort_key = 0;  // SLM_SORT_* refinement of the re-sort key
};

constexpr uint32_t kMaxInstanceGroups = 4096;
constexpr uint32_t kGroupStreams = 4;  // streams the groups of a batch are dealt to

int check_tuning(const slm_tuning& t) {
  if (t.sort_cell_log2 > 12) return fail(SLM_ERR_INVALID_ARGUMENT, "sort_cell_log2 must be <= 12");
  uint32_t g = t.agents_per_group;
  if (g < kAgentBlock || g > 16u * kAgentBlock || (g & (g - 1u)))
    return fail(SLM_ERR_INVALID_ARGUMENT, "agents_per_group must be 256, 512, 1024, 2048 or 4096");
  uint32_t r = t.trail_rows;
  if (r != 0 && (r < 4 || r > 64 || (r & (r - 1u))))
    return fail(SLM_ERR_INVALID_ARGUMENT, "trail_rows must be 0, 4, 8, 16, 32 or 64");
  if (t.tiled_sensing > 2) return fail(SLM_ERR_INVALID_ARGUMENT, "tiled_sensing must be 0, 1 or 2");
  if (t.instance_groups > kMaxInstanceGroups)
    return fail(SLM_ERR_INVALID_ARGUMENT, "instance_groups must be <= %u", kMaxInstanceGroups);
  if (t.streaming > 3) return fail(SLM_ERR_INVALID_ARGUMENT, "streaming must be 0..3");
  if (t.tiled_sensing == 1 && t.instance_groups > 1)
    return fail(SLM_ERR_UNSUPPORTED, "instance groups do not combine with tiled sensing");
  if ((t.sort_key & ~(SLM_SORT_HEADING_MASK | SLM_SORT_AT_SENSOR)) ||
      (t.sort_key & SLM_SORT_HEADING_MASK) > 5u)
    return fail(SLM_ERR_INVALID_ARGUMENT, "sort_key: heading bits 0..5, SLM_SORT_AT_SENSOR");
  if (t.tiled_sensing == 1 && t.sort_key)
    return fail(SLM_ERR_UNSUPPORTED, "tiled sensing needs the plain re-sort key");
  return SLM_OK;
}

int check_engine_config(const EngineConfig& c, const uint32_t* reserved, size_t nreserved) {
  if (c.storage < SLM_STORAGE_F32 || c.storage > SLM_STORAGE_F16)
    return fail(SLM_ERR_INVALID_ARGUMENT, "unknown storage %d", c.storage);
  if (c.sort_cell_log2 > 12) return fail(SLM_ERR_INVALID_ARGUMENT, "sort_cell_log2 must be <= 12");
  if (c.graph_chunk > 100000) return fail(SLM_ERR_INVALID_ARGUMENT, "graph_chunk must be <= 100000");
  if (c.flags & ~kKnownFlags) return fail(SLM_ERR_INVALID_ARGUMENT, "unknown flags");
  if ((c.flags & SLM_FLAG_STOCHASTIC_ROUNDING) && c.storage != SLM_STORAGE_BF16)
    return fail(SLM_ERR_UNSUPPORTED, "stochastic rounding needs bf16 storage");
  if ((c.flags & SLM_FLAG_TILED_SENSING) && (c.flags & SLM_FLAG_WINDOW_SENSING))
    return fail(SLM_ERR_INVALID_ARGUMENT, "tiled and windowed sensing are exclusive");
  for (size_t i = 0; i < nreserved; i++)
    if (reserved[i]) return fail(SLM_ERR_INVALID_ARGUMENT, "reserved fields must be zero");
  return SLM_OK;
}

// ---- snapshot format ----------------------------------------------------------
constexpr char kSnapMagic[8] = {'S', 'L', 'M', 'S', 'N', 'A', 'P', '1'};
constexpr uint32_t kSnapVersion = 1;

struct SnapHeader {  // 128 bytes, little-endian
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  uint32_t width, height;
  uint64_t agent_count;
  uint64_t seed;
  int32_t storage;
  uint32_t flags;  // SLM_FLAG_F64_POSITIONS
  uint64_t steps_done;
  uint64_t reassigns;
  slm_params params;
  uint32_t reserved;
  uint64_t agent_bytes;
  uint64_t trail_bytes;
};
static_assert(sizeof(SnapHeader) == 128, "snapshot header layout");

// FNV-1a over little-endian 64-bit words (the tail zero-padded).
uint64_t snapshot_checksum(const uint8_t* p, uint64_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, p + i, 8);
    h = (h ^ w) * 0x100000001b3ull;
  }
  if (i < n) {
    uint64_t w = 0;
    std::memcpy(&w, p + i, n - i);
    h = (h ^ w) * 0x100000001b3ull;
  }
  return h;
}

float decode_storage(uint32_t bits, int storage) {
  if (storage == SLM_STORAGE_F32) {
    float v;
    std::memcpy(&v, &bits, 4);
    return v;
  }
  if (storage == SLM_STORAGE_BF16) {
    uint32_t w = bits << 16;
    float v;
    std::memcpy(&v, &w, 4);
    return v;
  }
  _Float16 h;
  uint16_t b16 = static_cast<uint16_t>(bits);
  std::memcpy(&h, &b16, 2);
  return static_cast<float>(h);
}

// Structure and checksum of a snapshot buffer (not its shape).
int check_snapshot(const uint8_t* buf, uint64_t bytes, SnapHeader* out) {
  if (!buf) return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot buffer is NULL");
  if (bytes < 128 + 8) return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot too short");
  SnapHeader hd;
  std::memcpy(&hd, buf, sizeof hd);
  if (std::memcmp(hd.magic, kSnapMagic, 8) != 0)
    return fail(SLM_ERR_INVALID_ARGUMENT, "not a slime-mold snapshot");
  if (hd.version != kSnapVersion || hd.header_bytes != 128)
    return fail(SLM_ERR_UNSUPPORTED, "snapshot version %u not supported", hd.version);
  if (hd.storage < SLM_STORAGE_F32 || hd.storage > SLM_STORAGE_F16 ||
      (hd.flags & ~SLM_FLAG_F64_POSITIONS) || hd.reserved)
    return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot header is malformed");
  uint64_t n = hd.agent_count;
  uint64_t cells = static_cast<uint64_t>(hd.width) * hd.height;
  uint64_t want_agents = n * 16u + ((hd.flags & SLM_FLAG_F64_POSITIONS) ? n * 16u : 0u);
  uint64_t sb = hd.storage == SLM_STORAGE_F32 ? 4 : 2;
  if (n > kMaxAgents || hd.width == 0 || hd.height == 0 || hd.width > kMaxSide ||
      hd.height > kMaxSide || cells > kMaxCells || hd.agent_bytes != want_agents ||
      hd.trail_bytes != cells * sb || bytes != 128 + want_agents + cells * sb + 8)
    return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot sizes are inconsistent");
  uint64_t sum;
  std::memcpy(&sum, buf + bytes - 8, 8);
  if (sum != snapshot_checksum(buf, bytes - 8))
    return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot checksum mismatch");
  if (!finite_params(hd.params)) return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot params not finite");
  if (out) *out = hd;
  return SLM_OK;
}

struct InstanceSpec {
  uint32_t width, height;
  uint64_t agents;
  uint64_t seed;
  slm_params params;
};

class Engine {
 public:
  EngineConfig cfg;
  std::vector<InstanceSpec> spec;
  std::vector<Instance> inst;      // host copy of the device table
  std::vector<uint64_t> reassigns; // slm_reassign_speeds calls per instance
  hipStream_t stream = nullptr;
  uint64_t steps_done = 0;
  uint64_t device_bytes = 0;
  // A failure part-way through a re-layout (resize, agent count, tuning) can
  // leave buffers freed or tables out of step with them.  The status is kept
  // here and the handle then refuses every call but destroy, rather than
  // launch kernels on them.
  int broken = SLM_OK;

  ~Engine() { release(); }

  int create(const EngineConfig& c, std::vector<InstanceSpec> s) {
    cfg = c;
    spec = std::move(s);
    gauss_rclass_ = gauss_class_of_specs();  // the Gaussian tiles depend on it
    reassigns.assign(spec.size(), 0);
    SLM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    SLM_TRY(alloc_const());
    SLM_TRY(layout_agents());
    SLM_TRY(layout_trail());
    SLM_TRY(upload_params_all());
    for (size_t k = 0; k < spec.size(); k++) init_agents(k);
    return sync_checked("initialisation");
  }

  bool pos64() const { return (cfg.flags & SLM_FLAG_F64_POSITIONS) != 0; }
  bool sorting() const { return cfg.resort_interval > 0; }
  uint32_t count() const { return static_cast<uint32_t>(spec.size()); }
  uint64_t cells(size_t k) const { return static_cast<uint64_t>(spec[k].width) * spec[k].height; }

  Pool pool() const {
    Pool p;
    p.count = count();
    p.inst = inst_d_;
    p.params = params_d_;
    p.total_agents = total_agents_;
    p.total_units = total_units_;
    p.total_bins = total_bins_;
    p.agent_block_inst = agent_block_d_;
    p.unit_block_inst = unit_block_d_;
    p.storage = cfg.storage;
    p.pos64 = pos64() ? 1 : 0;
    p.xcd_remap = (cfg.flags & SLM_FLAG_NO_XCD_REMAP) ? 0 : 1;
    p.sort_cell_log2 = cfg.sort_cell_log2;
    p.sort_key = cfg.sort_key;
    p.total_tiles = total_tiles_;
    p.tile_inst = tile_inst_d_;
    // the re-sort builds the tiled pass's work list only when that pass is on
    // (turning it on re-lays the tables and waits for the next re-sort)
    p.items = (cfg.flags & SLM_FLAG_TILED_SENSING) ? items_d_ : nullptr;
    p.item_total = item_total_d_;
    p.max_items = max_items_;
    p.resort_interval = cfg.resort_interval;
    p.agents_per_group = cfg.agents_per_group;
    p.preload_rows = preload_rows_;
    p.stochastic_rounding = (cfg.flags & SLM_FLAG_STOCHASTIC_ROUNDING) ? 1 : 0;
    p.streaming = static_cast<int>(cfg.streaming);
    p.gaussian = (cfg.flags & SLM_FLAG_GAUSSIAN_DIFFUSE) ? 1 : 0;
    p.gauss_rclass = gauss_rclass_;
    p.window_sensing = (cfg.flags & SLM_FLAG_WINDOW_SENSING) ? 1 : 0;
    const bool persistent = (cfg.flags & SLM_FLAG_PERSISTENT) != 0;
    p.claim_agents = persistent ? claim_d_ : nullptr;
    p.claim_trail = persistent ? claim_d_ + kClaimWords : nullptr;
    p.total_gtiles = total_gtiles_;
    return p;
  }

  // ---- stepping -----------------------------------------------------------
  int step(uint32_t n) {
    uint32_t chunk = cfg.graph_chunk;
    while (n > 0) {
      if (chunk > 0 && n >= chunk) {
        SLM_TRY(run_chunk_graph());
        n -= chunk;
      } else {  // fewer than a chunk left, or no graphs: launch them directly
        const uint64_t s0 = steps_done;
        enqueue_steps(n, [&](uint32_t j) { return sort_due(s0 + j); });
        steps_done += n;
        n = 0;
      }
    }
    SLM_HIP(hipGetLastError());
    return SLM_OK;
  }

  // Steps with HIP events around every kernel.  Whole graph chunks replay
  // captured graphs that record the events between their kernels (the same
  // launch path as slm_step); other steps are launched eagerly, events for up
  // to kTimedBatch steps recorded before the host waits, so the stream never
  // drains between steps and host submission gaps stay out of kernel times.
  static constexpr int kEv = 5;  // sort begin/end, agents begin, agents end = trail begin, trail end
  int step_timed(uint32_t n, slm_timing* out) {
    std::memset(out, 0, sizeof *out);
    uint32_t chunk = cfg.graph_chunk;
    while (n > 0) {
      if (chunk > 0 && n >= chunk && !timed_graphs_off_) {
        int rc = run_timed_chunk(out);
        if (rc == kNoTimedGraph) {
          timed_graphs_off_ = true;  // this runtime cannot record events in graphs
          continue;
        }
        SLM_TRY(rc);
        n -= chunk;
      } else {
        uint32_t m = chunk > 0 && n >= chunk ? chunk : n;
        SLM_TRY(run_timed_eager(m, out));
        n -= m;
      }
    }
    SLM_HIP(hipGetLastError());
    return SLM_OK;
  }

  int sync() {
    SLM_HIP(hipStreamSynchronize(stream));
    return SLM_OK;
  }

  // ---- per-instance operations ---------------------------------------------
  int set_params(size_t k, const slm_params& p) {
    spec[k].params = p;
    launch_set_params(params_d_ + k, make_dev_params(p), stream);
    SLM_HIP(hipGetLastError());
    if (cfg.flags & SLM_FLAG_GAUSSIAN_DIFFUSE) {
      // a radius in another class needs the Gaussian pass with another LDS
      // tile (and, for 16-bit storage, another tile height, so another tile
      // plan), which captured chunks do not launch: wait for them, drop them
      const int c = gauss_class_of_specs();
      if (c != gauss_rclass_) {
        SLM_HIP(hipStreamSynchronize(stream));
        drop_graphs();
        const bool replan = gauss_th(cfg.storage, c) != gauss_th(cfg.storage, gauss_rclass_);
        gauss_rclass_ = c;
        if (replan) return relayout(layout_trail_tables());
      }
    }
    return SLM_OK;
  }

  int clear(size_t k) {
    SLM_HIP(hipMemsetAsync(trail_ptr(cur_, k), 0, cells(k) * storage_bytes(cfg.storage), stream));
    return SLM_OK;
  }

  int reassign_speeds(size_t k, float smin, float smax) {
    reassigns[k]++;
    uint64_t seed = spec[k].seed + reassigns[k] * 0xd1b54a32d192ed03ull;
    launch_reassign_speeds(inst[k], agents_[acur_], seed, smin, smax, stream);
    SLM_HIP(hipGetLastError());
    return SLM_OK;
  }

  int read_agents(size_t k, float* out) {
    uint64_t n = spec[k].agents;
    if (n == 0) return SLM_OK;
    void* staging = nullptr;
    SLM_TRY(scratch_alloc(&staging, n * 16));
    launch_pack_agents(inst[k], agents_[acur_], pos64() ? 1 : 0, static_cast<float*>(staging),
                       stream);
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipMemcpy(out, staging, n * 16, hipMemcpyDeviceToHost);
    (void)hipFree(staging);
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "read_agents failed: %s", hipGetErrorString(e));
    return SLM_OK;
  }

  int write_agents(size_t k, const float* in) {
    uint64_t n = spec[k].agents;
    float fw = static_cast<float>(spec[k].width), fh = static_cast<float>(spec[k].height);
    for (uint64_t i = 0; i < n; i++) {
      const float* r = in + 4 * i;
      if (!(r[0] >= 0.0f && r[0] <= fw && r[1] >= 0.0f && r[1] <= fh) || !std::isfinite(r[2]) ||
          !std::isfinite(r[3]))
        return fail(SLM_ERR_INVALID_ARGUMENT,
                    "agent %llu: position must lie in [0,width]x[0,height], angle and speed finite",
                    static_cast<unsigned long long>(i));
    }
    if (n == 0) return SLM_OK;
    SLM_TRY(sync());
    void* staging = nullptr;
    SLM_TRY(scratch_alloc(&staging, n * 16));
    hipError_t e = hipMemcpy(staging, in, n * 16, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_unpack_agents(inst[k], agents_[acur_], pos64() ? 1 : 0,
                           static_cast<const float*>(staging), stream);
      e = hipStreamSynchronize(stream);
    }
    (void)hipFree(staging);
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "write_agents failed: %s", hipGetErrorString(e));
    return SLM_OK;
  }

  int read_trail(size_t k, float* out) {
    SLM_TRY(sync());
    uint64_t c = cells(k);
    if (cfg.storage == SLM_STORAGE_F32) {
      SLM_HIP(hipMemcpy(out, trail_ptr(cur_, k), c * 4, hipMemcpyDeviceToHost));
      return SLM_OK;
    }
    void* staging = nullptr;
    SLM_TRY(scratch_alloc(&staging, c * 4));
    launch_decode_trail(trail_ptr(cur_, k), cfg.storage, c, static_cast<float*>(staging), stream);
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipMemcpy(out, staging, c * 4, hipMemcpyDeviceToHost);
    (void)hipFree(staging);
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "read_trail failed: %s", hipGetErrorString(e));
    return SLM_OK;
  }

  int write_trail(size_t k, const float* in) {
    uint64_t c = cells(k);
    std::vector<float> clean(in, in + c);
    for (uint64_t i = 0; i < c; i++) {
      if (!(clean[i] >= 0.0f) || !(clean[i] <= kMaxTrailValue))
        return fail(SLM_ERR_INVALID_ARGUMENT, "trail cell %llu must lie in [0, 2^100]",
                    static_cast<unsigned long long>(i));
      if (clean[i] == 0.0f) clean[i] = 0.0f;  // -0 -> +0: the sign bit is the deposit mark
    }
    SLM_TRY(sync());
    if (cfg.storage == SLM_STORAGE_F32) {
      SLM_HIP(hipMemcpy(trail_ptr(cur_, k), clean.data(), c * 4, hipMemcpyHostToDevice));
      return SLM_OK;
    }
    void* staging = nullptr;
    SLM_TRY(scratch_alloc(&staging, c * 4));
    hipError_t e = hipMemcpy(staging, clean.data(), c * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_encode_trail(static_cast<const float*>(staging), cfg.storage, c, trail_ptr(cur_, k),
                          stream);
      e = hipStreamSynchronize(stream);
    }
    (void)hipFree(staging);
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "write_trail failed: %s", hipGetErrorString(e));
    return SLM_OK;
  }

  int total_mass(double* out) {
    launch_total_mass(pool(), trail_[cur_], mass_partials_, mass_out_, stream);
    SLM_HIP(hipMemcpyAsync(out, mass_out_, count() * sizeof(double), hipMemcpyDeviceToHost,
                           stream));
    return sync();
  }

  // ---- single-simulation reshaping ----------------------------------------
  int resize(uint32_t w, uint32_t h) {  // pool of one
    SLM_TRY(sync());
    return relayout([&]() -> int {
      drop_graphs();
      float sx = static_cast<float>(w) / static_cast<float>(spec[0].width);
      float sy = static_cast<float>(h) / static_cast<float>(spec[0].height);
      launch_scale_positions(inst[0], agents_[acur_], pos64() ? 1 : 0, sx, sy, stream);
      SLM_TRY(sync());
      spec[0].width = w;
      spec[0].height = h;
      free_trail_layout();
      return layout_trail();
    }());
  }

  int set_agent_count(uint64_t n) {  // pool of one
    SLM_TRY(sync());
    return relayout([&]() -> int {
      drop_graphs();
      free_agent_layout();
      spec[0].agents = n;
      SLM_TRY(layout_agents());
      SLM_TRY(layout_trail_tables());
      init_agents(0);
      return sync_checked("agent re-initialisation");
    }());
  }

  // ---- display colour map -------------------------------------------------------
  // Renders instance k's current trail (after its last step's diffuse) into
  // `out`: device memory, or host memory through a short-lived staging buffer.
  int render(size_t k, const uint8_t* lut768, uint32_t tex_w, uint32_t tex_h, void* out,
             bool device_out) {
    return render_at(k, lut768, tex_w, tex_h, out, device_out, nullptr, 0);
  }

  // As render, but into host memory through `staging`, a device buffer of at
  // least tex_w * tex_h * 4 bytes the caller owns (nothing is allocated).
  int render_via(size_t k, const uint8_t* lut768, uint32_t tex_w, uint32_t tex_h, void* out,
                 void* staging) {
    RenderArgs a = render_args(k, tex_w, tex_h);
    a.trail = trail_ptr(cur_, k);
    a.out = static_cast<uint32_t*>(staging);
    launch_render(a, lut_of(lut768), cfg.storage, stream);
    uint64_t bytes = static_cast<uint64_t>(tex_w) * tex_h * 4u;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, staging, bytes, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "render failed: %s", hipGetErrorString(e));
    return SLM_OK;
  }

  // Enqueues n >= 1 steps (pool of one); the last runs eagerly with the
  // display between its agent pass and its trail pass, as the reference's
  // frame does (agents, decay, display, diffuse: src/main.rs:1163-1235): the
  // image shows the trail after the agents' deposits and the decay, before
  // the diffuse.  The render kernel reads the buffer the agent pass has just
  // marked and applies each sampled cell's deposit and decay itself.
  int step_render(uint32_t n, const uint8_t* lut768, uint32_t tex_w, uint32_t tex_h, void* out,
                  bool device_out) {
    SLM_TRY(step(n - 1));
    return render_at(0, lut768, tex_w, tex_h, out, device_out, params_d_, 1);
  }

  // ---- snapshots (single simulation) -------------------------------------------
  // Layout, little-endian: a 128-byte header (see slm_snapshot_size in
  // include/slimemold.h), the agents as n f32 records (x, y, angle, speed) in
  // original order, with f64 positions also n (x, y) f64 pairs, the trail as
  // width*height raw storage elements (row-major), and a 64-bit checksum.
  static constexpr uint64_t kSnapHeader = 128;
  uint64_t snap_agent_bytes() const {
    uint64_t n = spec[0].agents;
    return n * 16u + (pos64() ? n * 16u : 0u);
  }
  uint64_t snap_trail_bytes() const { return cells(0) * storage_bytes(cfg.storage); }
  uint64_t snapshot_bytes() const {
    return kSnapHeader + snap_agent_bytes() + snap_trail_bytes() + 8u;
  }

  int save_snapshot(uint8_t* buf) {
    const uint64_t n = spec[0].agents;
    SnapHeader hd;
    std::memset(&hd, 0, sizeof hd);
    std::memcpy(hd.magic, kSnapMagic, 8);
    hd.version = kSnapVersion;
    hd.header_bytes = kSnapHeader;
    hd.width = spec[0].width;
    hd.height = spec[0].height;
    hd.agent_count = n;
    hd.seed = spec[0].seed;
    hd.storage = cfg.storage;
    hd.flags = pos64() ? SLM_FLAG_F64_POSITIONS : 0u;
    hd.steps_done = steps_done;
    hd.reassigns = reassigns[0];
    hd.params = spec[0].params;
    hd.agent_bytes = snap_agent_bytes();
    hd.trail_bytes = snap_trail_bytes();
    std::memcpy(buf, &hd, sizeof hd);
    uint8_t* at = buf + kSnapHeader;
    std::vector<float> aos(n * 4);
    SLM_TRY(read_agents(0, aos.data()));
    std::memcpy(at, aos.data(), n * 16u);
    at += n * 16u;
    if (pos64() && n > 0) {
      void* staging = nullptr;
      SLM_TRY(scratch_alloc(&staging, n * 16u));
      launch_pack_positions64(inst[0], agents_[acur_], static_cast<double*>(staging), stream);
      hipError_t e = hipStreamSynchronize(stream);
      if (e == hipSuccess) e = hipMemcpy(at, staging, n * 16u, hipMemcpyDeviceToHost);
      (void)hipFree(staging);
      if (e != hipSuccess) return fail(SLM_ERR_HIP, "snapshot failed: %s", hipGetErrorString(e));
    }
    at += pos64() ? n * 16u : 0u;
    SLM_TRY(sync());
    SLM_HIP(hipMemcpy(at, trail_ptr(cur_, 0), snap_trail_bytes(), hipMemcpyDeviceToHost));
    at += snap_trail_bytes();
    uint64_t sum = snapshot_checksum(buf, snapshot_bytes() - 8u);
    std::memcpy(at, &sum, 8);
    return SLM_OK;
  }

  // `buf` has passed check_snapshot and matches this simulation's shape.
  int load_snapshot(const uint8_t* buf) {
    SnapHeader hd;
    std::memcpy(&hd, buf, sizeof hd);
    const uint64_t n = hd.agent_count;
    const uint8_t* at = buf + kSnapHeader;
    std::vector<float> aos(n * 4);
    std::memcpy(aos.data(), at, n * 16u);
    at += n * 16u;
    std::vector<double> xy;
    if (pos64()) {
      xy.resize(n * 2);
      std::memcpy(xy.data(), at, n * 16u);
      at += n * 16u;
      double fw = spec[0].width, fh = spec[0].height;
      for (uint64_t i = 0; i < n; i++)
        if (!(xy[2 * i] >= 0.0 && xy[2 * i] <= fw && xy[2 * i + 1] >= 0.0 && xy[2 * i + 1] <= fh))
          return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot agent %llu: f64 position out of range",
                      static_cast<unsigned long long>(i));
    }
    const uint64_t c = cells(0);
    const int sb = static_cast<int>(storage_bytes(cfg.storage));
    for (uint64_t i = 0; i < c; i++) {
      uint32_t bits = 0;
      std::memcpy(&bits, at + i * sb, sb);
      float v = decode_storage(bits, cfg.storage);
      if ((bits >> (8 * sb - 1)) != 0u || !(v >= 0.0f) || !(v <= kMaxTrailValue))
        return fail(SLM_ERR_INVALID_ARGUMENT, "snapshot trail cell %llu must lie in [0, 2^100]",
                    static_cast<unsigned long long>(i));
    }
    SLM_TRY(write_agents(0, aos.data()));  // validates positions, restores original order
    if (pos64() && n > 0) {
      void* staging = nullptr;
      SLM_TRY(scratch_alloc(&staging, n * 16u));
      hipError_t e = hipMemcpy(staging, xy.data(), n * 16u, hipMemcpyHostToDevice);
      if (e == hipSuccess) {
        launch_unpack_positions64(inst[0], agents_[acur_], static_cast<const double*>(staging),
                                  stream);
        e = hipStreamSynchronize(stream);
      }
      (void)hipFree(staging);
      if (e != hipSuccess) return fail(SLM_ERR_HIP, "snapshot load failed: %s", hipGetErrorString(e));
    }
    SLM_HIP(hipMemcpy(trail_ptr(cur_, 0), at, snap_trail_bytes(), hipMemcpyHostToDevice));
    spec[0].seed = hd.seed;
    reassigns[0] = hd.reassigns;
    steps_done = hd.steps_done;
    tiles_ready_ = false;
    SLM_TRY(set_params(0, hd.params));
    return sync_checked("snapshot load");
  }

  // ---- launch tuning ----------------------------------------------------------
  slm_tuning tuning() const {
    slm_tuning t;
    std::memset(&t, 0, sizeof t);
    t.resort_interval = cfg.resort_interval;
    t.sort_cell_log2 = cfg.sort_cell_log2;
    t.agents_per_group = cfg.agents_per_group;
    t.trail_rows = cfg.trail_rows;
    t.tiled_sensing = (cfg.flags & SLM_FLAG_TILED_SENSING)   ? 1u
                      : (cfg.flags & SLM_FLAG_WINDOW_SENSING) ? 2u
                                                              : 0u;
    t.instance_groups = cfg.instance_groups;
    t.streaming = cfg.streaming;
    t.sort_key = cfg.sort_key;
    return t;
  }

  // Keeps agents, trail, step count and parameters.
  int set_tuning(const slm_tuning& t) {
    if ((cfg.flags & SLM_FLAG_PERSISTENT) && t.instance_groups > 1)
      return fail(SLM_ERR_UNSUPPORTED, "persistent passes do not combine with instance groups");
    SLM_TRY(sync());
    if (t.resort_interval > 0 && bins_for(t.sort_cell_log2, t.sort_key) > kMaxTotalBins)
      return fail(SLM_ERR_INVALID_ARGUMENT, "too many re-sort bins: raise sort_cell_log2");
    if (!fits(t)) return fail(SLM_ERR_OUT_OF_MEMORY, "not enough device memory for this tuning");
    if (sorting() == (t.resort_interval > 0)) return apply_tuning(t);
    Snapshot snap;
    int rc = take_snapshot(snap, false);
    if (rc == SLM_OK) rc = apply_tuning(t);
    if (rc == SLM_OK) rc = relayout(restore_snapshot(snap));
    if (rc == SLM_OK) rc = relayout(sync_checked("re-tuning"));
    free_snapshot(snap);
    return rc;
  }

  // Coordinate descent over the tuning fields, each candidate timed on a copy
  // of the current state; the state is put back afterwards.
  int autotune(uint32_t trial_steps, slm_tuning* out) {
    SLM_TRY(sync());
    Snapshot snap;
    int rc = take_snapshot(snap, true);
    if (rc) {
      free_snapshot(snap);
      return rc;
    }
    const slm_tuning start = tuning();
    const uint64_t steps0 = steps_done;
    slm_tuning best = start;
    Trial bt{};
    trials_ = 0;
    rc = search(snap, trial_steps ? trial_steps : 1u, best, bt);
    int rc2 = put_tuning(rc == SLM_OK ? best : start, snap);
    // a tuning applied and the state restored in full: consistent again, even
    // if a trial failed part-way
    if (rc2 == SLM_OK) broken = SLM_OK;
    steps_done = steps0;
    free_snapshot(snap);
    if (rc) return rc;
    SLM_TRY(rc2);
    if (out) {
      *out = tuning();
      out->trials = trials_;
      out->step_ms = bt.step;
      out->agents_ms = bt.agents;
      out->trail_ms = bt.trail;
      out->sort_ms = bt.sort;
    }
    return SLM_OK;
  }

 private:
  static LutBytes lut_of(const uint8_t* lut768) {
    LutBytes lut;
    std::memcpy(lut.word, lut768, sizeof lut.word);
    return lut;
  }

  RenderArgs render_args(size_t k, uint32_t tex_w, uint32_t tex_h) const {
    RenderArgs a{};
    render_fit(spec[k].width, spec[k].height, tex_w, tex_h, a);
    return a;
  }

  // steps: 0 renders the current trail; 1 enqueues one step with the render
  // between its passes, decoding deposit and decay with `decode`.
  int render_at(size_t k, const uint8_t* lut768, uint32_t tex_w, uint32_t tex_h, void* out,
                bool device_out, const DevParams* decode, uint32_t steps) {
    RenderArgs a = render_args(k, tex_w, tex_h);
    const LutBytes lut = lut_of(lut768);
    uint64_t bytes = static_cast<uint64_t>(tex_w) * tex_h * 4u;
    void* staging = nullptr;
    if (!device_out) SLM_TRY(scratch_alloc(&staging, bytes));
    a.out = static_cast<uint32_t*>(device_out ? out : staging);
    a.decode = decode;
    auto draw = [&] {
      a.trail = trail_ptr(cur_, k);
      launch_render(a, lut, cfg.storage, stream);
    };
    if (steps) {
      enqueue_step(sort_due(steps_done), draw);
      steps_done++;
    } else {
      draw();
    }
    hipError_t e = hipGetLastError();
    if (!device_out) {
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e == hipSuccess) e = hipMemcpy(out, staging, bytes, hipMemcpyDeviceToHost);
      (void)hipFree(staging);
    }
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "render failed: %s", hipGetErrorString(e));
    return SLM_OK;
  }

  // Agents in original order (and optionally the current trail), in
  // short-lived buffers.
  struct Snapshot {
    AgentSoA a{};
    void* trail = nullptr;
  };
  struct Trial {
    double step, agents, trail, sort;  // ms per step, sorts amortised
    bool ok;
  };
  static constexpr uint32_t kTuneSpan = 40;  // warm-up and timing granule of a trial
  uint32_t trials_ = 0;

  int take_snapshot(Snapshot& s, bool with_trail) {
    uint64_t n = total_agents_;
    if (pos64()) {
      SLM_TRY(scratch_alloc(reinterpret_cast<void**>(&s.a.x64), n * 8));
      SLM_TRY(scratch_alloc(reinterpret_cast<void**>(&s.a.y64), n * 8));
    } else {
      SLM_TRY(scratch_alloc(reinterpret_cast<void**>(&s.a.x), n * 4));
      SLM_TRY(scratch_alloc(reinterpret_cast<void**>(&s.a.y), n * 4));
    }
    SLM_TRY(scratch_alloc(reinterpret_cast<void**>(&s.a.th), n * 4));
    SLM_TRY(scratch_alloc(reinterpret_cast<void**>(&s.a.v), n * 4));
    launch_canonical_copy(pool(), agents_[acur_], s.a, stream);
    if (with_trail) {
      size_t bytes = trail_elems_ * storage_bytes(cfg.storage);
      SLM_TRY(scratch_alloc(&s.trail, bytes));
      SLM_HIP(hipMemcpyAsync(s.trail, trail_[cur_], bytes, hipMemcpyDeviceToDevice, stream));
    }
    return sync_checked("snapshot");
  }

  int restore_snapshot(const Snapshot& s) {
    launch_canonical_copy(pool(), s.a, agents_[acur_], stream);
    if (s.trail)
      SLM_HIP(hipMemcpyAsync(trail_[cur_], s.trail, trail_elems_ * storage_bytes(cfg.storage),
                             hipMemcpyDeviceToDevice, stream));
    SLM_HIP(hipGetLastError());
    return SLM_OK;
  }

  void free_snapshot(Snapshot& s) {
    for (void* p : {static_cast<void*>(s.a.x), static_cast<void*>(s.a.y),
                    static_cast<void*>(s.a.x64), static_cast<void*>(s.a.y64),
                    static_cast<void*>(s.a.th), static_cast<void*>(s.a.v), s.trail})
      if (p) (void)hipFree(p);
    s = Snapshot{};
  }

  uint64_t bins_for(uint32_t cell_log2, uint32_t sort_key) const {
    uint64_t bins = 0;
    for (const Instance& i : inst) {
      Instance c = i;
      bins += plan_bins(c, cell_log2) << (sort_key & SLM_SORT_HEADING_MASK);
    }
    return bins;
  }

  // Whether switching to `t` fits the free device memory (counting the
  // snapshot an agent re-layout needs and not counting what it frees).
  bool fits(const slm_tuning& t) {
    bool sort = t.resort_interval > 0;
    uint64_t need = 64ull << 20;
    if (sort) {
      uint64_t bins = bins_for(t.sort_cell_log2, t.sort_key);
      if (bins > kMaxTotalBins) return false;
      need += bins * 12ull;
    }
    if (sort != sorting()) {
      uint64_t n = total_agents_;
      uint64_t pos = pos64() ? 16 : 8;
      need += n * ((pos + 8 + (sort ? 4 : 0)) * (sort ? 2 : 1) + (sort ? 4 : 0));  // new arrays
      need += n * (pos + 8);                                                      // snapshot
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return need <= free_b;
  }

  // Sets the tuning fields and rebuilds what depends on them.  Switching
  // re-sorting on or off reallocates the agent arrays, losing their contents:
  // the caller restores them.
  int apply_tuning(const slm_tuning& t) {
    return relayout([&]() -> int {
      bool new_agents = sorting() != (t.resort_interval > 0);
      drop_graphs();
      if (new_agents) free_agent_layout();
      cfg.resort_interval = t.resort_interval;
      cfg.sort_cell_log2 = t.sort_cell_log2;
      cfg.agents_per_group = t.agents_per_group;
      cfg.trail_rows = t.trail_rows;
      cfg.instance_groups = t.instance_groups ? t.instance_groups : 1u;
      cfg.streaming = t.streaming;
      cfg.sort_key = t.sort_key;
      cfg.flags &= ~(SLM_FLAG_TILED_SENSING | SLM_FLAG_WINDOW_SENSING);
      if (t.tiled_sensing == 1) cfg.flags |= SLM_FLAG_TILED_SENSING;
      if (t.tiled_sensing == 2) cfg.flags |= SLM_FLAG_WINDOW_SENSING;
      if (new_agents) SLM_TRY(layout_agents());
      return layout_trail_tables();
    }());
  }

  int relayout(int rc) {
    if (rc != SLM_OK && broken == SLM_OK) broken = rc;
    return rc;
  }

  int put_tuning(const slm_tuning& t, const Snapshot& snap) {
    SLM_TRY(apply_tuning(t));
    SLM_TRY(relayout(restore_snapshot(snap)));
    return relayout(sync_checked("autotune restore"));
  }

  // Times `t` over at least `steps` steps (whole re-sort intervals) from the
  // snapshot; r.ok is false when it does not fit in memory.
  int trial(const slm_tuning& t, const Snapshot& snap, uint32_t steps, Trial& r) {
    r = Trial{};
    if (!fits(t)) return SLM_OK;
    SLM_TRY(put_tuning(t, snap));
    // Every candidate warms up for the same kTuneSpan steps and is timed over
    // the same window (whole kTuneSpan multiples), so all see the same
    // evolving pattern (agent-pass time drifts by tens of percent over a few
    // dozen steps as trails form).  Candidate intervals divide kTuneSpan, so
    // the timed window starts with a re-sort of agents that drifted a whole
    // interval, as in a long run.
    steps_done = 0;
    enqueue_steps(kTuneSpan, [&](uint32_t j) { return sort_due(j); });
    steps_done = kTuneSpan;
    uint32_t n = (steps + kTuneSpan - 1) / kTuneSpan * kTuneSpan;
    // The step time is the device span of n steps captured as one graph, as
    // slm_step runs a chunk (instance groups concurrent); capture and
    // instantiation stay outside it.  The kernel times come from a further n
    // steps with events around each kernel (whole-pool launches in order).
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SLM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    enqueue_steps(n, [&](uint32_t j) { return sort_due(kTuneSpan + j); });
    hipError_t ce = hipStreamEndCapture(stream, &graph);
    if (ce == hipSuccess) ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (ce != hipSuccess) return fail(SLM_ERR_HIP, "trial capture failed: %s", hipGetErrorString(ce));
    hipEvent_t span[2] = {nullptr, nullptr};
    if (hipEventCreate(&span[0]) != hipSuccess || hipEventCreate(&span[1]) != hipSuccess) {
      if (span[0]) (void)hipEventDestroy(span[0]);
      (void)hipGraphExecDestroy(exec);
      return fail(SLM_ERR_HIP, "hipEventCreate failed");
    }
    slm_timing tm;
    (void)hipEventRecord(span[0], stream);
    hipError_t le = hipGraphLaunch(exec, stream);
    (void)hipEventRecord(span[1], stream);
    steps_done += n;
    int rc = SLM_OK;
    if (le != hipSuccess) rc = fail(SLM_ERR_HIP, "trial launch failed: %s", hipGetErrorString(le));
    if (rc == SLM_OK) rc = step_timed(n, &tm);
    float ms = 0.0f;
    if (rc == SLM_OK) {
      hipError_t e = hipEventSynchronize(span[1]);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, span[0], span[1]);
      if (e != hipSuccess) rc = fail(SLM_ERR_HIP, "trial timing failed: %s", hipGetErrorString(e));
    }
    (void)hipEventDestroy(span[0]);
    (void)hipEventDestroy(span[1]);
    (void)hipStreamSynchronize(stream);
    (void)hipGraphExecDestroy(exec);
    SLM_TRY(rc);
    r.agents = tm.agents_ms / n;
    r.trail = tm.trail_ms / n;
    r.sort = tm.sort_ms / n;
    r.step = static_cast<double>(ms) / n;
    r.ok = true;
    trials_++;
    return SLM_OK;
  }

  int search(const Snapshot& snap, uint32_t steps, slm_tuning& best, Trial& bt) {
    Trial r;
    SLM_TRY(trial(best, snap, steps, r));  // warm-up: clocks, code objects, caches
    SLM_TRY(trial(best, snap, steps, bt));
    if (!bt.ok) return fail(SLM_ERR_OUT_OF_MEMORY, "autotune: no memory for trials");
    const double kGain = 0.995;  // a change must save at least 0.5%
    auto take = [&](const slm_tuning& t) -> int {
      SLM_TRY(trial(t, snap, steps, r));
      if (r.ok && r.step < bt.step * kGain) {
        best = t;
        bt = r;
      }
      return SLM_OK;
    };
    // 1. trail tile rows, judged by the trail pass alone
    {
      const slm_tuning base = best;
      double cur = bt.trail;
      for (uint32_t rows : {4u, 8u, 16u, 32u, 64u}) {
        slm_tuning t = base;
        t.trail_rows = rows;
        SLM_TRY(trial(t, snap, steps, r));
        if (r.ok && r.trail < cur * kGain) {
          cur = r.trail;
          best = t;
          bt = r;
        }
      }
    }
    // 2. re-sort interval, 3. bin size, 4. tiled sensing, 5. agents per group
    for (uint32_t k : {0u, 5u, 10u, 20u, 40u}) {
      if (k == best.resort_interval) continue;
      slm_tuning t = best;
      t.resort_interval = k;
      SLM_TRY(take(t));
    }
    if (best.resort_interval) {
      const uint32_t cl = best.sort_cell_log2;
      for (uint32_t c : {cl - 1u, cl + 1u}) {
        if (cl == 0 || c < 2 || c > 8) continue;
        slm_tuning t = best;
        t.sort_cell_log2 = c;
        SLM_TRY(take(t));
      }
      // re-sort key: heading classes, then binning at the centre sensor
      for (uint32_t key : {3u, SLM_SORT_AT_SENSOR | 3u, SLM_SORT_AT_SENSOR | 5u}) {
        if (key == best.sort_key || best.tiled_sensing == 1) continue;
        slm_tuning t = best;
        t.sort_key = key;
        SLM_TRY(take(t));
      }
    }
    // sensing: global taps (0), tile work lists (1: needs re-sorting with the
    // plain key and one instance group), per-workgroup LDS windows (2)
    {
      const uint32_t cur = best.tiled_sensing;
      for (uint32_t m : {0u, 1u, 2u}) {
        if (m == cur || (m == 1u && (!best.resort_interval || best.instance_groups > 1 ||
                                     best.sort_key != 0u)))
          continue;
        slm_tuning t = best;
        t.tiled_sensing = m;
        SLM_TRY(take(t));
      }
    }
    for (uint32_t g : {256u, 512u, 1024u, 2048u, 4096u}) {
      if (g == best.agents_per_group) continue;
      slm_tuning t = best;
      t.agents_per_group = g;
      SLM_TRY(take(t));
    }
    // 6. instance groups (batches): powers of 4 up to the instance count, then
    // the neighbours (x2, /2) of the best
    std::vector<uint32_t> tried_groups;
    auto try_groups = [&](uint32_t g) -> int {
      const uint32_t cap = std::min<uint32_t>(count(), kMaxInstanceGroups);
      if (best.tiled_sensing == 1 || (cfg.flags & SLM_FLAG_PERSISTENT) || g < 1 || g > cap)
        return SLM_OK;
      for (uint32_t x : tried_groups)
        if (x == g) return SLM_OK;
      tried_groups.push_back(g);
      slm_tuning t = best;
      t.instance_groups = g;
      return take(t);
    };
    auto group_neighbours = [&]() -> int {
      const uint32_t b = best.instance_groups ? best.instance_groups : 1u;
      SLM_TRY(try_groups(b * 2));
      return try_groups(b / 2);
    };
    if (count() > 1) {
      tried_groups.push_back(best.instance_groups ? best.instance_groups : 1u);
      for (uint32_t g = 1; g <= std::min<uint32_t>(count(), kMaxInstanceGroups); g *= 4)
        SLM_TRY(try_groups(g));
      SLM_TRY(group_neighbours());
    }
    // 7. streaming (non-temporal) trail-pass loads and stores, judged by the
    // whole step: they also change what the next agent pass finds in the
    // caches; then the group count's neighbours again under the chosen mode
    {
      const slm_tuning base = best;
      for (uint32_t m = 0; m < 4; m++) {
        if (m == base.streaming) continue;
        slm_tuning t = base;
        t.streaming = m;
        SLM_TRY(take(t));
      }
      if (count() > 1 && best.streaming != base.streaming) {
        tried_groups.assign(1, best.instance_groups ? best.instance_groups : 1u);
        SLM_TRY(group_neighbours());
      }
    }
    return SLM_OK;
  }

  Instance* inst_d_ = nullptr;
  DevParams* params_d_ = nullptr;
  uint32_t* agent_block_d_ = nullptr;
  uint32_t* unit_block_d_ = nullptr;
  void* trail_[2] = {nullptr, nullptr};
  int cur_ = 0;
  AgentSoA agents_[2] = {};
  int acur_ = 0;
  SortScratch sort_{};
  uint32_t* scan_levels_ = nullptr;
  double* mass_partials_ = nullptr;
  double* mass_out_ = nullptr;
  uint32_t* claim_d_ = nullptr;  // persistent passes: agent counters, then trail counters
  uint32_t total_agents_ = 0, total_units_ = 0, total_bins_ = 0;
  uint32_t total_tiles_ = 0, max_items_ = 0;
  uint32_t total_gtiles_ = 0;  // Gaussian-pass tiles
  int gauss_rclass_ = kGaussMaxR;  // Gaussian-pass LDS radius class (gauss_class)
  uint32_t preload_rows_ = 0;
  uint32_t* tile_inst_d_ = nullptr;
  WorkItem* items_d_ = nullptr;
  uint32_t* item_total_d_ = nullptr;
  bool tiles_ready_ = false;  // the latest re-sort's work list matches the layout
  uint64_t trail_elems_ = 0;
  // Captured chunks.  A chunk's launches depend on the current trail and
  // agent buffers and on where the chunk starts in the re-sort schedule
  // (absolute steps 0, k, 2k, ..., as eager steps follow), so that is the key.
  struct ChunkKey {
    int cur, acur;
    uint32_t phase;  // steps_done % resort_interval at the chunk's start (0 without re-sorts)
    bool operator==(const ChunkKey& o) const {
      return cur == o.cur && acur == o.acur && phase == o.phase;
    }
  };
  static constexpr size_t kMaxCachedChunks = 16;  // per kind; the oldest goes first
  std::vector<std::pair<ChunkKey, hipGraphExec_t>> graphs_;
  std::vector<std::pair<void*, size_t>> allocs_;

  int dev_alloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      *p = nullptr;
      (void)hipGetLastError();
      return fail(e == hipErrorOutOfMemory ? SLM_ERR_OUT_OF_MEMORY : SLM_ERR_HIP,
                  "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    allocs_.push_back({*p, bytes});
    device_bytes += bytes;
    return SLM_OK;
  }

  template <class T>
  int dev_alloc(T** p, size_t bytes) {
    return dev_alloc(reinterpret_cast<void**>(p), bytes);
  }

  void dev_free(void* p) {
    if (!p) return;
    for (auto& a : allocs_)
      if (a.first == p) {
        device_bytes -= a.second;
        a.first = nullptr;
      }
    (void)hipFree(p);
  }

  // short-lived staging buffers (not counted in device_bytes)
  int scratch_alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 256);
    if (e != hipSuccess) {
      *p = nullptr;
      (void)hipGetLastError();
      return fail(e == hipErrorOutOfMemory ? SLM_ERR_OUT_OF_MEMORY : SLM_ERR_HIP,
                  "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return SLM_OK;
  }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipStream_t s : aux_) (void)hipStreamSynchronize(s);
    drop_graphs();
    for (auto& a : allocs_)
      if (a.first) (void)hipFree(a.first);
    allocs_.clear();
    groups_.clear();
    for (hipStream_t s : aux_) (void)hipStreamDestroy(s);
    aux_.clear();
    for (hipEvent_t e : join_ev_) (void)hipEventDestroy(e);
    join_ev_.clear();
    if (fork_ev_) (void)hipEventDestroy(fork_ev_);
    fork_ev_ = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  int sync_checked(const char* what) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    return SLM_OK;
  }

  void* trail_ptr(int buf, size_t k) const {
    return static_cast<char*>(trail_[buf]) + inst[k].trail_off * storage_bytes(cfg.storage);
  }

  int alloc_const() {
    inst.assign(spec.size(), Instance{});
    SLM_TRY(dev_alloc(&inst_d_, spec.size() * sizeof(Instance)));
    SLM_TRY(dev_alloc(&params_d_, spec.size() * sizeof(DevParams)));
    SLM_TRY(dev_alloc(&mass_partials_, spec.size() * kMassBlocks * sizeof(double)));
    SLM_TRY(dev_alloc(&mass_out_, spec.size() * sizeof(double)));
    // persistent passes' claim counters: zero here, reset by the kernels after
    SLM_TRY(dev_alloc(&claim_d_, 2 * kClaimWords * sizeof(uint32_t)));
    SLM_HIP(hipMemset(claim_d_, 0, 2 * kClaimWords * sizeof(uint32_t)));
    if (cfg.flags & SLM_FLAG_PERSISTENT) {
      prime_persistent_agents();
      prime_persistent_trail();
    }
    return SLM_OK;
  }

  int upload_params_all() {
    std::vector<DevParams> dp(spec.size());
    for (size_t k = 0; k < spec.size(); k++) dp[k] = make_dev_params(spec[k].params);
    SLM_HIP(hipMemcpy(params_d_, dp.data(), dp.size() * sizeof(DevParams), hipMemcpyHostToDevice));
    return SLM_OK;
  }

  // The Gaussian pass's radius class (gauss_class) for the largest radius of
  // the instances' current parameters.
  int gauss_class_of_specs() const {
    int rmax = 1;
    float w[kGaussMaxR + 1];
    for (const auto& sp : spec) {
      const int r = gauss_weights(sp.params.blur_radius, sp.params.blur_sigma, w);
      rmax = r > rmax ? r : rmax;
    }
    return gauss_class(rmax);
  }

  // Agent slots, agent arrays, re-sort per-slot scratch and the agent block map.
  int layout_agents() {
    uint64_t off = 0;
    for (size_t k = 0; k < spec.size(); k++) {
      inst[k].agent_off = static_cast<uint32_t>(off);
      inst[k].agent_count = static_cast<uint32_t>(spec[k].agents);
      off += spec[k].agents;
    }
    total_agents_ = static_cast<uint32_t>(off);
    uint64_t n = off;
    int sets = sorting() ? 2 : 1;
    for (int s = 0; s < sets; s++) {
      AgentSoA& a = agents_[s];
      if (pos64()) {
        SLM_TRY(dev_alloc(&a.x64, n * 8));
        SLM_TRY(dev_alloc(&a.y64, n * 8));
      } else {
        SLM_TRY(dev_alloc(&a.x, n * 4));
        SLM_TRY(dev_alloc(&a.y, n * 4));
      }
      SLM_TRY(dev_alloc(&a.th, n * 4));
      SLM_TRY(dev_alloc(&a.v, n * 4));
      if (sorting()) SLM_TRY(dev_alloc(&a.id, n * 4));
    }
    if (sorting()) SLM_TRY(dev_alloc(&sort_.rank, n * 4));
    acur_ = 0;
    // block -> first instance whose slots reach into the block
    uint32_t nb = blocks_for(n, kAgentBlock);
    std::vector<uint32_t> map(nb ? nb : 1, 0);
    size_t k = 0;
    for (uint32_t b = 0; b < nb; b++) {
      uint64_t first = static_cast<uint64_t>(b) * kAgentBlock;
      while (k + 1 < spec.size() &&
             static_cast<uint64_t>(inst[k].agent_off) + inst[k].agent_count <= first)
        k++;
      map[b] = static_cast<uint32_t>(k);
    }
    SLM_TRY(dev_alloc(&agent_block_d_, map.size() * 4));
    SLM_HIP(hipMemcpy(agent_block_d_, map.data(), map.size() * 4, hipMemcpyHostToDevice));
    return SLM_OK;
  }

  void free_agent_layout() {
    for (auto& a : agents_) {
      dev_free(a.x);
      dev_free(a.y);
      dev_free(a.x64);
      dev_free(a.y64);
      dev_free(a.th);
      dev_free(a.v);
      dev_free(a.id);
      a = AgentSoA{};
    }
    dev_free(sort_.rank);
    sort_.rank = nullptr;
    dev_free(agent_block_d_);
    agent_block_d_ = nullptr;
  }

  // Trail buffers, trail-pass units, re-sort bins and their tables.
  int layout_trail() {
    uint64_t off = 0;
    for (size_t k = 0; k < spec.size(); k++) {
      Instance& in = inst[k];
      in.width = spec[k].width;
      in.height = spec[k].height;
      in.trail_off = off;
      off += (cells(k) + 127u) & ~uint64_t(127u);  // 256-byte aligned starts
    }
    trail_elems_ = off + 128u;  // slack: tap pairs and 32-bit halo words stay inside
    for (int b = 0; b < 2; b++) {
      SLM_TRY(dev_alloc(&trail_[b], trail_elems_ * storage_bytes(cfg.storage)));
      SLM_HIP(hipMemsetAsync(trail_[b], 0, trail_elems_ * storage_bytes(cfg.storage), stream));
    }
    cur_ = 0;
    return layout_trail_tables();
  }

  // Units, bins and block maps (also redone when the agent layout changes).
  int layout_trail_tables() {
    uint64_t total_cells = 0;
    for (size_t k = 0; k < spec.size(); k++) total_cells += cells(k);
    uint64_t units = 0, bins = 0;
    for (size_t k = 0; k < spec.size(); k++) {
      Instance& in = inst[k];
      uint32_t target = static_cast<uint32_t>(
          std::max<uint64_t>(1, kTargetWaves * cells(k) / std::max<uint64_t>(1, total_cells)));
      plan_trail(in, cfg.storage, target, cfg.trail_rows);
      in.unit_off = static_cast<uint32_t>(units);
      units += in.units;
      uint32_t cl = cfg.sort_cell_log2;
      const uint64_t spatial = plan_bins(in, cl);
      plan_tiles(in, cl, kTileLog2);  // spatial bins: tiled sensing keeps the plain key
      const uint64_t keys = spatial << (cfg.sort_key & SLM_SORT_HEADING_MASK);
      in.nbins = static_cast<uint32_t>(std::min<uint64_t>(keys, kMaxTotalBins + 1ull));
      in.bin_off = static_cast<uint32_t>(std::min<uint64_t>(bins, kMaxTotalBins + 1ull));
      bins += keys;
    }
    uint64_t tiles = 0, gtiles = 0;
    for (auto& in : inst) {
      in.tile_off = static_cast<uint32_t>(tiles);
      tiles += in.tiles;
      in.gtx = (in.width + kGaussTW - 1u) / kGaussTW;
      in.gtile_off = static_cast<uint32_t>(gtiles);
      const uint32_t th = gauss_th(cfg.storage, gauss_rclass_);
      gtiles += static_cast<uint64_t>(in.gtx) * ((in.height + th - 1u) / th);
    }
    if (gtiles >= (1ull << 31)) return fail(SLM_ERR_INVALID_ARGUMENT, "too many trail tiles");
    total_gtiles_ = static_cast<uint32_t>(gtiles);
    tiles_ready_ = false;
    if (units >= (1ull << 32)) return fail(SLM_ERR_INVALID_ARGUMENT, "too many trail units");
    if (sorting() && bins > kMaxTotalBins)
      return fail(SLM_ERR_INVALID_ARGUMENT, "too many re-sort bins: raise sort_cell_log2");
    total_units_ = static_cast<uint32_t>(units);
    // whole-tile loads when every strip instance has the same 4-, 8- or 16-row tiles
    preload_rows_ = 0;
    bool first = true;
    for (const Instance& in : inst) {
      if (in.ntx == 0) continue;
      uint32_t r = in.tile_rows == 4u || in.tile_rows == 8u || in.tile_rows == 16u ? in.tile_rows : 0u;
      preload_rows_ = first || preload_rows_ == r ? r : 0u;
      if (!first && preload_rows_ == 0u) break;
      first = false;
    }
    total_bins_ = sorting() ? static_cast<uint32_t>(bins) : 0;
    SLM_HIP(hipMemcpy(inst_d_, inst.data(), inst.size() * sizeof(Instance), hipMemcpyHostToDevice));
    SLM_TRY(build_groups());

    dev_free(unit_block_d_);
    uint32_t nb = blocks_for(units, kTrailWavesPerBlock);
    std::vector<uint32_t> map(nb ? nb : 1, 0);
    size_t k = 0;
    for (uint32_t b = 0; b < nb; b++) {
      uint64_t first = static_cast<uint64_t>(b) * kTrailWavesPerBlock;
      while (k + 1 < spec.size() && static_cast<uint64_t>(inst[k].unit_off) + inst[k].units <= first)
        k++;
      map[b] = static_cast<uint32_t>(k);
    }
    SLM_TRY(dev_alloc(&unit_block_d_, map.size() * 4));
    SLM_HIP(hipMemcpy(unit_block_d_, map.data(), map.size() * 4, hipMemcpyHostToDevice));

    dev_free(sort_.count);
    dev_free(sort_.offset);
    dev_free(sort_.tile_items);
    dev_free(scan_levels_);
    dev_free(tile_inst_d_);
    dev_free(items_d_);
    dev_free(item_total_d_);
    sort_.count = sort_.offset = sort_.tile_items = scan_levels_ = nullptr;
    tile_inst_d_ = item_total_d_ = nullptr;
    items_d_ = nullptr;
    total_tiles_ = max_items_ = 0;
    if (sorting()) {
      SLM_TRY(dev_alloc(&sort_.count, total_bins_ * 4ull));
      SLM_TRY(dev_alloc(&sort_.offset, (total_bins_ + 1ull) * 4ull));
      // offset[total_bins] closes the last bin: the scan fills the rest
      SLM_HIP(hipMemcpy(sort_.offset + total_bins_, &total_agents_, 4, hipMemcpyHostToDevice));
      total_tiles_ = static_cast<uint32_t>(tiles);
      max_items_ = total_tiles_ + blocks_for(total_agents_, kItemAgents);
      std::vector<uint32_t> tmap(total_tiles_);
      for (size_t k = 0; k < spec.size(); k++)
        for (uint32_t t = 0; t < inst[k].tiles; t++) tmap[inst[k].tile_off + t] = static_cast<uint32_t>(k);
      SLM_TRY(dev_alloc(&tile_inst_d_, tmap.size() * 4ull));
      SLM_HIP(hipMemcpy(tile_inst_d_, tmap.data(), tmap.size() * 4ull, hipMemcpyHostToDevice));
      SLM_TRY(dev_alloc(&sort_.tile_items, total_tiles_ * 4ull));
      SLM_TRY(dev_alloc(&items_d_, max_items_ * sizeof(WorkItem)));
      SLM_TRY(dev_alloc(&item_total_d_, 4));
      SLM_HIP(hipMemset(item_total_d_, 0, 4));
      uint32_t lv = plan_scan(sort_, total_bins_);
      SLM_TRY(dev_alloc(&scan_levels_, lv * 4ull));
      uint32_t at = 0;
      for (int l = 0; l < 3; l++) {
        sort_.level[l] = l < sort_.levels ? scan_levels_ + at : nullptr;
        if (l < sort_.levels) at += sort_.level_len[l];
      }
    }
    return SLM_OK;
  }

  void free_trail_layout() {
    dev_free(trail_[0]);
    dev_free(trail_[1]);
    trail_[0] = trail_[1] = nullptr;
  }

  void init_agents(size_t k) {
    launch_init_agents(inst[k], agents_[acur_], pos64() ? 1 : 0, spec[k].seed,
                       spec[k].params.agent_speed_min, spec[k].params.agent_speed_max, stream);
  }

  bool sort_due(uint64_t step_index) const {
    return sorting() && step_index % cfg.resort_interval == 0;
  }

  // Tiled sensing (when asked for) once a re-sort has laid agents out by
  // tile; otherwise the global-memory pass.  Both give the same bits.
  void enqueue_agents(const Pool& p) {
    if (sorting() && tiles_ready_ && (cfg.flags & SLM_FLAG_TILED_SENSING))
      launch_agents_tiled(p, agents_[acur_], trail_[cur_], stream);
    else
      launch_agents(p, agents_[acur_], trail_[cur_], stream);
  }

  // One step; `between` runs after the agent pass is enqueued and before the
  // trail pass (the reference's display point).
  template <class Between>
  void enqueue_step(bool do_sort, Between&& between) {
    Pool p = pool();
    if (do_sort && sorting() && total_agents_ > 0) {
      launch_morton_sort(p, agents_[acur_], agents_[acur_ ^ 1], sort_, stream);
      acur_ ^= 1;
      tiles_ready_ = true;
    }
    enqueue_agents(p);
    between();
    launch_trail(p, trail_[cur_], trail_[cur_ ^ 1], stream);
    cur_ ^= 1;
  }
  void enqueue_step(bool do_sort) {
    enqueue_step(do_sort, [] {});
  }

  // ---- instance groups ---------------------------------------------------------
  // With instance_groups > 1 the instances are split into contiguous groups of
  // about equal cost (agents + cells / 8).  Between re-sorts the groups are
  // dealt round-robin to kGroupStreams streams, forked from and joined back
  // into the engine's stream (in a captured chunk: parallel branches of the
  // graph); each stream runs its groups' steps one group after another.  So
  // kernels of different groups run at once - one group's HBM-bound trail
  // pass beside another's gather-bound agent pass - and each kernel covers a
  // few instances, whose trail and agents stay in the caches better than a
  // whole-batch launch's.  On the 256-instance mix, 128 groups step 1.18x
  // faster than one.  A group has its own instance table rebased to its first
  // agent slot, trail-pass unit and trail element, and its own block maps;
  // its launches get agent and trail pointers offset to match, so the kernels
  // run on it unchanged.  The re-sort stays one whole-pool launch sequence on
  // the engine's stream.
  struct Group {
    uint32_t k0 = 0, k1 = 0;  // instances [k0, k1)
    uint32_t slot_base = 0, agents = 0, unit_base = 0, units = 0;
    uint64_t trail_base = 0;
    uint32_t gtile_base = 0, gtiles = 0;  // Gaussian-pass tiles
    uint32_t preload_rows = 0;
    const Instance* inst_d = nullptr;  // tables inside group_tables_
    const uint32_t* agent_block_d = nullptr;
    const uint32_t* unit_block_d = nullptr;
  };
  std::vector<Group> groups_;
  void* group_tables_ = nullptr;        // every group's tables, one allocation
  std::vector<hipStream_t> aux_;        // group streams 1.. (stream 0 is `stream`)
  hipEvent_t fork_ev_ = nullptr;
  std::vector<hipEvent_t> join_ev_;

  void free_groups() {
    dev_free(group_tables_);
    group_tables_ = nullptr;
    groups_.clear();
  }

  // First instance (group-local index) whose range [off, off + len) reaches
  // into each block of `per` items.
  static std::vector<uint32_t> block_map(const std::vector<Instance>& ins, uint64_t items,
                                         uint32_t per, bool units) {
    uint32_t nb = blocks_for(items, per);
    std::vector<uint32_t> map(nb ? nb : 1, 0);
    size_t k = 0;
    for (uint32_t b = 0; b < nb; b++) {
      uint64_t first = static_cast<uint64_t>(b) * per;
      while (k + 1 < ins.size()) {
        uint64_t end = units ? static_cast<uint64_t>(ins[k].unit_off) + ins[k].units
                             : static_cast<uint64_t>(ins[k].agent_off) + ins[k].agent_count;
        if (end > first) break;
        k++;
      }
      map[b] = static_cast<uint32_t>(k);
    }
    return map;
  }

  int build_groups() {
    free_groups();
    const uint32_t want = std::min<uint32_t>(cfg.instance_groups, count());
    if (want <= 1 || (cfg.flags & (SLM_FLAG_TILED_SENSING | SLM_FLAG_PERSISTENT))) return SLM_OK;
    std::vector<uint64_t> cost(count());
    uint64_t total = 0;
    for (size_t k = 0; k < count(); k++) {
      cost[k] = spec[k].agents + cells(k) / 8u + 1u;
      total += cost[k];
    }
    // cut after the instance where the running cost first reaches g/want of the total
    std::vector<uint32_t> start{0};
    uint64_t acc = 0;
    for (uint32_t k = 0; k + 1 < count() && start.size() < want; k++) {
      acc += cost[k];
      if (acc * want >= total * start.size()) start.push_back(k + 1);
    }
    start.push_back(count());
    // each group's rebased instance table and block maps, laid out in one host
    // blob (256-byte aligned pieces) and uploaded with one copy
    std::vector<uint8_t> blob;
    auto put = [&blob](const void* p, size_t bytes) {
      size_t at = (blob.size() + 255u) & ~size_t(255u);
      blob.resize(at + bytes);
      std::memcpy(blob.data() + at, p, bytes);
      return at;
    };
    std::vector<size_t> at_inst, at_agents, at_units;
    for (size_t g = 0; g + 1 < start.size(); g++) {
      Group G;
      G.k0 = start[g];
      G.k1 = start[g + 1];
      const Instance& a = inst[G.k0];
      const Instance& z = inst[G.k1 - 1];
      G.slot_base = a.agent_off;
      G.agents = z.agent_off + z.agent_count - a.agent_off;
      G.unit_base = a.unit_off;
      G.units = z.unit_off + z.units - a.unit_off;
      G.trail_base = a.trail_off;
      G.gtile_base = a.gtile_off;
      const uint32_t th = gauss_th(cfg.storage, gauss_rclass_);
      G.gtiles = z.gtile_off + z.gtx * ((z.height + th - 1u) / th) - a.gtile_off;
      std::vector<Instance> ins(inst.begin() + G.k0, inst.begin() + G.k1);
      bool first = true;
      for (Instance& in : ins) {
        in.agent_off -= G.slot_base;
        in.unit_off -= G.unit_base;
        in.trail_off -= G.trail_base;
        in.gtile_off -= G.gtile_base;
        if (in.ntx == 0) continue;
        uint32_t r = in.tile_rows == 4u || in.tile_rows == 8u || in.tile_rows == 16u ? in.tile_rows : 0u;
        G.preload_rows = first || G.preload_rows == r ? r : 0u;
        first = false;
      }
      at_inst.push_back(put(ins.data(), ins.size() * sizeof(Instance)));
      std::vector<uint32_t> am = block_map(ins, G.agents, kAgentBlock, false);
      at_agents.push_back(put(am.data(), am.size() * 4));
      std::vector<uint32_t> um = block_map(ins, G.units, kTrailWavesPerBlock, true);
      at_units.push_back(put(um.data(), um.size() * 4));
      groups_.push_back(G);
    }
    SLM_TRY(dev_alloc(&group_tables_, blob.size()));
    SLM_HIP(hipMemcpy(group_tables_, blob.data(), blob.size(), hipMemcpyHostToDevice));
    const uint8_t* base = static_cast<const uint8_t*>(group_tables_);
    for (size_t g = 0; g < groups_.size(); g++) {
      groups_[g].inst_d = reinterpret_cast<const Instance*>(base + at_inst[g]);
      groups_[g].agent_block_d = reinterpret_cast<const uint32_t*>(base + at_agents[g]);
      groups_[g].unit_block_d = reinterpret_cast<const uint32_t*>(base + at_units[g]);
    }
    // streams and fork/join events, kept until the engine is released
    const size_t lanes = std::min<size_t>(groups_.size(), kGroupStreams);
    while (aux_.size() + 1 < lanes) {
      hipStream_t s = nullptr;
      SLM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      aux_.push_back(s);
    }
    if (!fork_ev_) SLM_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    while (join_ev_.size() + 1 < lanes) {
      hipEvent_t e = nullptr;
      SLM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      join_ev_.push_back(e);
    }
    return SLM_OK;
  }

  Pool group_pool(const Group& g) const {
    Pool p = pool();
    p.count = g.k1 - g.k0;
    p.inst = g.inst_d;
    p.params = params_d_ + g.k0;
    p.total_agents = g.agents;
    p.total_units = g.units;
    p.agent_block_inst = g.agent_block_d;
    p.unit_block_inst = g.unit_block_d;
    p.preload_rows = g.preload_rows;
    p.total_gtiles = g.gtiles;
    p.total_bins = 0;
    p.total_tiles = 0;
    p.max_items = 0;
    return p;
  }

  static AgentSoA offset_agents(AgentSoA a, uint32_t base) {
    auto off = [&](auto*& p) {
      if (p) p += base;
    };
    off(a.x);
    off(a.y);
    off(a.x64);
    off(a.y64);
    off(a.th);
    off(a.v);
    off(a.id);
    return a;
  }

  void* trail_at(int buf, uint64_t elem) const {
    return static_cast<char*>(trail_[buf]) + elem * storage_bytes(cfg.storage);
  }

  // Enqueues n steps; sorts(j) says whether step j re-sorts first.  With
  // instance groups, each stretch between re-sorts forks onto the group
  // streams and joins back before the next re-sort (or the end).
  template <class SortAt>
  void enqueue_steps(uint32_t n, SortAt&& sorts) {
    if (groups_.size() <= 1) {
      for (uint32_t j = 0; j < n; j++) enqueue_step(sorts(j));
      return;
    }
    uint32_t j = 0;
    while (j < n) {
      if (sorts(j) && sorting() && total_agents_ > 0) {
        launch_morton_sort(pool(), agents_[acur_], agents_[acur_ ^ 1], sort_, stream);
        acur_ ^= 1;
        tiles_ready_ = true;
      }
      uint32_t m = 1;
      while (j + m < n && !sorts(j + m)) m++;
      // group g runs on stream g % L (0: the engine's stream)
      const size_t L = std::min<size_t>(groups_.size(), kGroupStreams);
      (void)hipEventRecord(fork_ev_, stream);
      for (size_t l = 1; l < L; l++) (void)hipStreamWaitEvent(aux_[l - 1], fork_ev_, 0);
      for (size_t g = 0; g < groups_.size(); g++) {
        const Group& G = groups_[g];
        const size_t l = g % L;
        hipStream_t s = l == 0 ? stream : aux_[l - 1];
        const Pool p = group_pool(G);
        const AgentSoA a = offset_agents(agents_[acur_], G.slot_base);
        for (uint32_t i = 0; i < m; i++) {
          int c = cur_ ^ static_cast<int>(i & 1u);
          launch_agents(p, a, trail_at(c, G.trail_base), s);
          launch_trail(p, trail_at(c, G.trail_base), trail_at(c ^ 1, G.trail_base), s);
        }
      }
      for (size_t l = 1; l < L; l++) {
        (void)hipEventRecord(join_ev_[l - 1], aux_[l - 1]);
        (void)hipStreamWaitEvent(stream, join_ev_[l - 1], 0);
      }
      cur_ ^= static_cast<int>(m & 1u);
      j += m;
    }
  }

  // Whether absolute step `step` starts with a re-sort.
  bool sorts_at(uint64_t step) const { return sort_due(step) && total_agents_ > 0; }

  ChunkKey chunk_key() const {
    return {cur_, acur_, sorting() ? static_cast<uint32_t>(steps_done % cfg.resort_interval) : 0u};
  }

  // Cached entry for `key` in `cache`, or nullptr.
  template <class T>
  static T* cached(std::vector<std::pair<ChunkKey, T>>& cache, const ChunkKey& key) {
    for (auto& e : cache)
      if (e.first == key) return &e.second;
    return nullptr;
  }

  int run_chunk_graph() {
    const uint32_t chunk = cfg.graph_chunk;
    const uint64_t s0 = steps_done;
    const ChunkKey key = chunk_key();
    hipGraphExec_t* found = cached(graphs_, key);
    hipGraphExec_t exec = found ? *found : nullptr;
    if (!exec) {
      if (graphs_.size() >= kMaxCachedChunks) {  // drop the oldest, once it has run
        SLM_HIP(hipStreamSynchronize(stream));
        (void)hipGraphExecDestroy(graphs_.front().second);
        graphs_.erase(graphs_.begin());
      }
      int cur0 = cur_, acur0 = acur_;
      hipGraph_t graph = nullptr;
      SLM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      enqueue_steps(chunk, [&](uint32_t j) { return sorts_at(s0 + j); });
      hipError_t e = hipStreamEndCapture(stream, &graph);
      cur_ = cur0;
      acur_ = acur0;
      if (e != hipSuccess) return fail(SLM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess)
        return fail(SLM_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
      graphs_.push_back({key, exec});
    }
    SLM_HIP(hipGraphLaunch(exec, stream));
    for (uint32_t j = 0; j < chunk; j++) {  // the buffer swaps the captured steps make
      if (sorts_at(s0 + j)) acur_ ^= 1;
      cur_ ^= 1;
    }
    steps_done += chunk;
    return SLM_OK;
  }

  // Enqueues one step with events e[0..kEv) around its kernels; returns
  // whether it re-sorted.
  bool enqueue_step_events(hipEvent_t* e, bool do_sort, const Pool& p) {
    bool sort = do_sort && sorting() && total_agents_ > 0;
    if (sort) {
      (void)hipEventRecord(e[0], stream);
      launch_morton_sort(p, agents_[acur_], agents_[acur_ ^ 1], sort_, stream);
      (void)hipEventRecord(e[1], stream);
      acur_ ^= 1;
      tiles_ready_ = true;
    }
    (void)hipEventRecord(e[2], stream);
    enqueue_agents(p);
    (void)hipEventRecord(e[3], stream);
    launch_trail(p, trail_[cur_], trail_[cur_ ^ 1], stream);
    (void)hipEventRecord(e[4], stream);
    cur_ ^= 1;
    return sort;
  }

  void add_times(const hipEvent_t* e, bool sorted, slm_timing* out) {
    float ms = 0.0f;
    if (sorted && hipEventElapsedTime(&ms, e[0], e[1]) == hipSuccess) {
      out->sort_ms += ms;
      out->sort_launches++;
    }
    if (total_agents_ > 0 && hipEventElapsedTime(&ms, e[2], e[3]) == hipSuccess) {
      out->agents_ms += ms;
      out->agents_launches++;
    }
    if (hipEventElapsedTime(&ms, e[3], e[4]) == hipSuccess) {
      out->trail_ms += ms;
      out->trail_launches++;
    }
  }

  static int create_events(std::vector<hipEvent_t>& ev, size_t n) {
    ev.clear();
    for (size_t i = 0; i < n; i++) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        destroy_events(ev);
        return fail(SLM_ERR_HIP, "hipEventCreate failed");
      }
      ev.push_back(e);
    }
    return SLM_OK;
  }

  static void destroy_events(std::vector<hipEvent_t>& ev) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    ev.clear();
  }

  int run_timed_eager(uint32_t n, slm_timing* out) {
    constexpr uint32_t kTimedBatch = 128;
    uint32_t batch = n < kTimedBatch ? n : kTimedBatch;
    std::vector<hipEvent_t> ev;
    SLM_TRY(create_events(ev, static_cast<size_t>(batch) * kEv));
    std::vector<uint8_t> sorted(batch);
    Pool p = pool();
    int rc = SLM_OK;
    for (uint32_t done = 0; done < n && rc == SLM_OK;) {
      uint32_t m = n - done < batch ? n - done : batch;
      for (uint32_t i = 0; i < m; i++) {
        sorted[i] = enqueue_step_events(&ev[static_cast<size_t>(i) * kEv], sort_due(steps_done), p);
        steps_done++;
      }
      hipError_t err = hipEventSynchronize(ev[static_cast<size_t>(m - 1) * kEv + 4]);
      if (err != hipSuccess) {
        rc = fail(SLM_ERR_HIP, "step failed: %s", hipGetErrorString(err));
        break;
      }
      for (uint32_t i = 0; i < m; i++) add_times(&ev[static_cast<size_t>(i) * kEv], sorted[i], out);
      done += m;
    }
    destroy_events(ev);
    return rc;
  }

  static constexpr int kNoTimedGraph = 1;  // capture with event records unsupported

  struct TimedChunk {
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;
    std::vector<uint8_t> sorted;
  };
  std::vector<std::pair<ChunkKey, TimedChunk>> timed_;
  bool timed_graphs_off_ = false;

  // A timed chunk is built node by node: HIP drops event records made during
  // stream capture, so each kernel group (re-sort, agents, trail) is captured
  // into a child graph and chained with explicit event-record nodes.
  int build_timed_chunk(TimedChunk& tc, uint32_t chunk) {
    SLM_TRY(create_events(tc.ev, static_cast<size_t>(chunk) * kEv));
    tc.sorted.assign(chunk, 0);
    const int cur0 = cur_, acur0 = acur_;
    const bool tiles0 = tiles_ready_;
    const Pool p = pool();
    hipGraph_t g = nullptr;
    hipError_t e = hipGraphCreate(&g, 0);
    hipGraphNode_t prev = nullptr;
    auto add_event = [&](hipEvent_t ev) {
      if (e != hipSuccess) return;
      hipGraphNode_t n = nullptr;
      e = hipGraphAddEventRecordNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, ev);
      if (e == hipSuccess) prev = n;
    };
    auto add_child = [&](auto&& enqueue) {
      if (e != hipSuccess) return;
      hipGraph_t sub = nullptr;
      e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return;
      enqueue();
      e = hipStreamEndCapture(stream, &sub);
      size_t nodes = 0;
      if (e == hipSuccess) e = hipGraphGetNodes(sub, nullptr, &nodes);
      if (e == hipSuccess && nodes > 0) {
        hipGraphNode_t n = nullptr;
        e = hipGraphAddChildGraphNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, sub);
        if (e == hipSuccess) prev = n;
      }
      if (sub) (void)hipGraphDestroy(sub);
    };
    for (uint32_t j = 0; j < chunk; j++) {
      hipEvent_t* ev = &tc.ev[static_cast<size_t>(j) * kEv];
      bool sort = sorts_at(steps_done + j);
      tc.sorted[j] = sort;
      if (sort) {
        add_event(ev[0]);
        add_child([&] { launch_morton_sort(p, agents_[acur_], agents_[acur_ ^ 1], sort_, stream); });
        add_event(ev[1]);
        acur_ ^= 1;
        tiles_ready_ = true;
      }
      add_event(ev[2]);
      add_child([&] { enqueue_agents(p); });
      add_event(ev[3]);
      add_child([&] { launch_trail(p, trail_[cur_], trail_[cur_ ^ 1], stream); });
      add_event(ev[4]);
      cur_ ^= 1;
    }
    cur_ = cur0;
    acur_ = acur0;
    tiles_ready_ = tiles0;
    if (e == hipSuccess) e = hipGraphInstantiate(&tc.exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      tc.exec = nullptr;
      destroy_events(tc.ev);
      return kNoTimedGraph;
    }
    return SLM_OK;
  }

  int run_timed_chunk(slm_timing* out) {
    uint32_t chunk = cfg.graph_chunk;
    const ChunkKey key = chunk_key();
    TimedChunk* found = cached(timed_, key);
    if (!found) {
      if (timed_.size() >= kMaxCachedChunks) {  // drop the oldest, once it has run
        SLM_HIP(hipStreamSynchronize(stream));
        free_timed(timed_.front().second);
        timed_.erase(timed_.begin());
      }
      TimedChunk fresh;
      SLM_TRY(build_timed_chunk(fresh, chunk));
      timed_.push_back({key, std::move(fresh)});
      found = &timed_.back().second;
    }
    TimedChunk& tc = *found;
    SLM_HIP(hipGraphLaunch(tc.exec, stream));
    for (uint32_t j = 0; j < chunk; j++) {  // the buffer swaps the captured steps make
      if (tc.sorted[j]) {
        acur_ ^= 1;
        tiles_ready_ = true;
      }
      cur_ ^= 1;
    }
    steps_done += chunk;
    hipError_t err = hipEventSynchronize(tc.ev[static_cast<size_t>(chunk - 1) * kEv + 4]);
    if (err != hipSuccess) return fail(SLM_ERR_HIP, "step failed: %s", hipGetErrorString(err));
    for (uint32_t j = 0; j < chunk; j++)
      add_times(&tc.ev[static_cast<size_t>(j) * kEv], tc.sorted[j] != 0, out);
    return SLM_OK;
  }

  static void free_timed(TimedChunk& tc) {
    if (tc.exec) (void)hipGraphExecDestroy(tc.exec);
    tc.exec = nullptr;
    destroy_events(tc.ev);
    tc.sorted.clear();
  }

  // Destroys every captured chunk.  Callers have waited for the engine's
  // stream, which every launch of them joins back into.
  void drop_graphs() {
    for (auto& e : graphs_) (void)hipGraphExecDestroy(e.second);
    graphs_.clear();
    for (auto& e : timed_) free_timed(e.second);
    timed_.clear();
  }
};

int make_engine(const EngineConfig& ec, std::vector<InstanceSpec> specs, Engine** out) {
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(SLM_ERR_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
  }
  if (ec.device < 0 || ec.device >= ndev)
    return fail(SLM_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", ec.device, ndev);
  DeviceGuard guard(ec.device);
  Engine* eng = new (std::nothrow) Engine();
  if (!eng) return fail(SLM_ERR_OUT_OF_MEMORY, "host allocation failed");
  int rc = eng->create(ec, std::move(specs));
  if (rc) {
    delete eng;
    return rc;
  }
  *out = eng;
  return SLM_OK;
}

}  // namespace

struct slm_sim {
  Engine* eng;
  slm_config cfg;
};

struct slm_batch {
  Engine* eng;
  slm_batch_config cfg;
};

extern "C" {

const char* slm_last_error(void) { return g_last_error; }

int slm_device_count(int* out_count) {
  SLM_ENTRY_BEGIN
  if (!out_count) return fail(SLM_ERR_INVALID_ARGUMENT, "out_count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out_count = 0;
    return fail(SLM_ERR_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
  }
  *out_count = n;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_config_default(slm_config* cfg) {
  SLM_ENTRY_BEGIN
  if (!cfg) return fail(SLM_ERR_INVALID_ARGUMENT, "cfg is NULL");
  std::memset(cfg, 0, sizeof *cfg);
  cfg->width = 1600;  // src/settings.rs:4-5
  cfg->height = 900;
  cfg->storage = SLM_STORAGE_F32;
  cfg->graph_chunk = 100;
  cfg->sort_cell_log2 = 4;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_params_default(slm_params* p) {
  SLM_ENTRY_BEGIN
  if (!p) return fail(SLM_ERR_INVALID_ARGUMENT, "params is NULL");
  p->decay_factor = 10.0f;  // src/settings.rs:9-27
  p->agent_jitter = 0.0f;
  p->agent_speed_min = 30.0f;
  p->agent_speed_max = 50.0f;
  p->agent_turn_speed = 0.43f;
  p->agent_sensor_angle = 0.3f;
  p->agent_sensor_distance = 20.0f;
  p->diffusion_rate = 1.0f;
  p->pheromone_deposition_amount = 1.0f;
  p->blur_radius = 2.0f;
  p->blur_sigma = 1.0f;
  return SLM_OK;
  SLM_ENTRY_END
}

// ---- single simulation ------------------------------------------------------

int slm_create(const slm_config* cfg, const slm_params* params, slm_sim** out) {
  SLM_ENTRY_BEGIN
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!cfg) return fail(SLM_ERR_INVALID_ARGUMENT, "cfg is NULL");
  SLM_TRY(check_params(params, cfg->flags));
  SLM_TRY(check_grid(cfg->width, cfg->height));
  if (cfg->agent_count > kMaxAgents)
    return fail(SLM_ERR_INVALID_ARGUMENT, "agent_count must be < 2^32 - 256");
  EngineConfig ec;
  ec.storage = cfg->storage;
  ec.device = cfg->device;
  ec.resort_interval = cfg->resort_interval;
  ec.graph_chunk = cfg->graph_chunk;
  ec.sort_cell_log2 = cfg->sort_cell_log2;
  ec.flags = cfg->flags;
  SLM_TRY(check_engine_config(ec, cfg->reserved, 6));
  std::vector<InstanceSpec> specs{{cfg->width, cfg->height, cfg->agent_count, cfg->seed, *params}};
  slm_sim* s = new (std::nothrow) slm_sim{nullptr, *cfg};
  if (!s) return fail(SLM_ERR_OUT_OF_MEMORY, "host allocation failed");
  int rc = make_engine(ec, std::move(specs), &s->eng);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return SLM_OK;
  SLM_ENTRY_END
}

#define SIM_HANDLE(s)                                                      \
  if (!(s) || !(s)->eng) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL"); \
  DeviceGuard guard_((s)->eng->cfg.device)

#define SIM_GUARD(s)                                                                  \
  SIM_HANDLE(s);                                                                      \
  if ((s)->eng->broken)                                                               \
  return fail((s)->eng->broken, "the simulation is unusable after a failed resize, agent " \
              "count or tuning change; destroy it")

int slm_destroy(slm_sim* s) {
  SLM_ENTRY_BEGIN
  SIM_HANDLE(s);
  delete s->eng;
  delete s;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_set_params(slm_sim* s, const slm_params* p) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SLM_TRY(check_params(p, s->eng->cfg.flags));
  return s->eng->set_params(0, *p);
  SLM_ENTRY_END
}

int slm_get_params(const slm_sim* s, slm_params* out) {
  SLM_ENTRY_BEGIN
  if (!s || !s->eng || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = s->eng->spec[0].params;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_step(slm_sim* s, uint32_t n) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  return s->eng->step(n);
  SLM_ENTRY_END
}

int slm_step_timed(slm_sim* s, uint32_t n, slm_timing* out) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  return s->eng->step_timed(n, out);
  SLM_ENTRY_END
}

int slm_synchronize(slm_sim* s) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  return s->eng->sync();
  SLM_ENTRY_END
}

int slm_clear(slm_sim* s) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  return s->eng->clear(0);
  SLM_ENTRY_END
}

int slm_resize(slm_sim* s, uint32_t w, uint32_t h) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SLM_TRY(check_grid(w, h));
  SLM_TRY(s->eng->resize(w, h));
  s->cfg.width = w;
  s->cfg.height = h;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_set_agent_count(slm_sim* s, uint64_t n) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (n > kMaxAgents) return fail(SLM_ERR_INVALID_ARGUMENT, "agent_count must be < 2^32 - 256");
  SLM_TRY(s->eng->set_agent_count(n));
  s->cfg.agent_count = n;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_reassign_speeds(slm_sim* s, float smin, float smax) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!std::isfinite(smin) || !std::isfinite(smax))
    return fail(SLM_ERR_INVALID_ARGUMENT, "speeds must be finite");
  return s->eng->reassign_speeds(0, smin, smax);
  SLM_ENTRY_END
}

static int check_n(uint64_t n, uint64_t want, const char* what) {
  if (n != want)
    return fail(SLM_ERR_INVALID_ARGUMENT, "%s: n (%llu) != expected (%llu)", what,
                static_cast<unsigned long long>(n), static_cast<unsigned long long>(want));
  return SLM_OK;
}

int slm_read_agents(slm_sim* s, float* out, uint64_t n) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(n, s->eng->spec[0].agents, "read_agents"));
  return s->eng->read_agents(0, out);
  SLM_ENTRY_END
}

int slm_write_agents(slm_sim* s, const float* in, uint64_t n) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!in) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(n, s->eng->spec[0].agents, "write_agents"));
  return s->eng->write_agents(0, in);
  SLM_ENTRY_END
}

int slm_read_trail(slm_sim* s, float* out, uint64_t cells) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(cells, s->eng->cells(0), "read_trail (cells != width*height)"));
  return s->eng->read_trail(0, out);
  SLM_ENTRY_END
}

int slm_write_trail(slm_sim* s, const float* in, uint64_t cells) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!in) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(cells, s->eng->cells(0), "write_trail (cells != width*height)"));
  return s->eng->write_trail(0, in);
  SLM_ENTRY_END
}

int slm_total_mass(slm_sim* s, double* out) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  return s->eng->total_mass(out);
  SLM_ENTRY_END
}

int slm_get_info(const slm_sim* s, slm_info* out) {
  SLM_ENTRY_BEGIN
  if (!s || !s->eng || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  std::memset(out, 0, sizeof *out);
  const Engine& e = *s->eng;
  out->width = e.spec[0].width;
  out->height = e.spec[0].height;
  out->agent_count = e.spec[0].agents;
  out->storage = e.cfg.storage;
  out->device = e.cfg.device;
  out->steps_done = e.steps_done;
  out->device_bytes = e.device_bytes;
  out->resort_interval = e.cfg.resort_interval;
  out->graph_chunk = e.cfg.graph_chunk;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_get_tuning(const slm_sim* s, slm_tuning* out) {
  SLM_ENTRY_BEGIN
  if (!s || !s->eng || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = s->eng->tuning();
  return SLM_OK;
  SLM_ENTRY_END
}

static void sync_config(slm_config& c, const Engine& e) {
  c.resort_interval = e.cfg.resort_interval;
  c.sort_cell_log2 = e.cfg.sort_cell_log2;
  c.flags = e.cfg.flags;
}

int slm_set_tuning(slm_sim* s, const slm_tuning* t) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!t) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_tuning(*t));
  int rc = s->eng->set_tuning(*t);
  sync_config(s->cfg, *s->eng);
  return rc;
  SLM_ENTRY_END
}

int slm_autotune(slm_sim* s, uint32_t trial_steps, slm_tuning* out) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  int rc = s->eng->autotune(trial_steps, out);
  sync_config(s->cfg, *s->eng);
  return rc;
  SLM_ENTRY_END
}

static int check_render(const uint8_t* lut, uint32_t tex_w, uint32_t tex_h, const void* out) {
  if (!lut || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  if (tex_w == 0 || tex_h == 0 || tex_w > kMaxTexSide || tex_h > kMaxTexSide ||
      static_cast<uint64_t>(tex_w) * tex_h > (1ull << 30))
    return fail(SLM_ERR_INVALID_ARGUMENT, "texture must be 1..2^24 on a side, <= 2^30 pixels");
  return SLM_OK;
}

int slm_render_rgba8(slm_sim* s, const uint8_t* lut, uint32_t tex_w, uint32_t tex_h,
                     uint8_t* rgba) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SLM_TRY(check_render(lut, tex_w, tex_h, rgba));
  return s->eng->render(0, lut, tex_w, tex_h, rgba, false);
  SLM_ENTRY_END
}

int slm_render_rgba8_device(slm_sim* s, const uint8_t* lut, uint32_t tex_w, uint32_t tex_h,
                            void* rgba_device) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SLM_TRY(check_render(lut, tex_w, tex_h, rgba_device));
  return s->eng->render(0, lut, tex_w, tex_h, rgba_device, true);
  SLM_ENTRY_END
}

// ---- rendering without allocation ----------------------------------------------
// slm_render_rgba8 stages its image in a device buffer it allocates and frees
// on every call.  slm_render_rgba8_noalloc stages it here instead: one fixed
// array per GPU, held until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory, so it is only ever used through the
// address hipGetSymbolAddress returns for the current device, looked up once
// per device (with that device current) and kept in g_scratch_ptr.
static __device__ __align__(256) uint32_t scratch[SLM_RENDER_NOALLOC_MAX_BYTES / 4];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// This device's copy of `scratch`; `dev` must be the current device.
int scratch_on(int dev, void** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(SLM_ERR_UNSUPPORTED, "device %d: no staging array beyond device %d", dev,
                kMaxScratchDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(SLM_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                  hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SLM_OK;
}
}  // namespace

int slm_render_rgba8_noalloc(slm_sim* s, const uint8_t* lut, uint32_t tex_w, uint32_t tex_h,
                             uint8_t* rgba) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SLM_TRY(check_render(lut, tex_w, tex_h, rgba));
  if (static_cast<uint64_t>(tex_w) * tex_h * 4u > SLM_RENDER_NOALLOC_MAX_BYTES)
    return fail(SLM_ERR_INVALID_ARGUMENT,
                "texture of %ux%u needs %llu bytes; slm_render_rgba8_noalloc stages at most "
                "%u (SLM_RENDER_NOALLOC_MAX_BYTES)",
                tex_w, tex_h, static_cast<unsigned long long>(tex_w) * tex_h * 4u,
                static_cast<unsigned>(SLM_RENDER_NOALLOC_MAX_BYTES));
  void* staging = nullptr;
  SLM_TRY(scratch_on(s->eng->cfg.device, &staging));
  return s->eng->render_via(0, lut, tex_w, tex_h, rgba, staging);
  SLM_ENTRY_END
}

int slm_step_render_rgba8(slm_sim* s, uint32_t n_steps, const uint8_t* lut, uint32_t tex_w,
                          uint32_t tex_h, uint8_t* rgba) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SLM_TRY(check_render(lut, tex_w, tex_h, rgba));
  if (n_steps == 0) return fail(SLM_ERR_INVALID_ARGUMENT, "n_steps must be >= 1");
  return s->eng->step_render(n_steps, lut, tex_w, tex_h, rgba, false);
  SLM_ENTRY_END
}

int slm_step_render_rgba8_device(slm_sim* s, uint32_t n_steps, const uint8_t* lut,
                                 uint32_t tex_w, uint32_t tex_h, void* rgba_device) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SLM_TRY(check_render(lut, tex_w, tex_h, rgba_device));
  if (n_steps == 0) return fail(SLM_ERR_INVALID_ARGUMENT, "n_steps must be >= 1");
  return s->eng->step_render(n_steps, lut, tex_w, tex_h, rgba_device, true);
  SLM_ENTRY_END
}

int slm_snapshot_size(const slm_sim* s, uint64_t* bytes) {
  SLM_ENTRY_BEGIN
  if (!s || !s->eng || !bytes) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  *bytes = s->eng->snapshot_bytes();
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_save_snapshot(slm_sim* s, void* buf, uint64_t bytes) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  if (!buf) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(bytes, s->eng->snapshot_bytes(), "save_snapshot (bytes != slm_snapshot_size)"));
  return s->eng->save_snapshot(static_cast<uint8_t*>(buf));
  SLM_ENTRY_END
}

int slm_load_snapshot(slm_sim* s, const void* buf, uint64_t bytes) {
  SLM_ENTRY_BEGIN
  SIM_GUARD(s);
  SnapHeader hd;
  SLM_TRY(check_snapshot(static_cast<const uint8_t*>(buf), bytes, &hd));
  const Engine& e = *s->eng;
  if (hd.width != e.spec[0].width || hd.height != e.spec[0].height ||
      hd.agent_count != e.spec[0].agents || hd.storage != e.cfg.storage ||
      ((hd.flags & SLM_FLAG_F64_POSITIONS) != 0) != e.pos64())
    return fail(SLM_ERR_INVALID_ARGUMENT,
                "snapshot shape (grid, agents, storage, positions) differs from the simulation's");
  SLM_TRY(check_params(&hd.params, e.cfg.flags));
  return s->eng->load_snapshot(static_cast<const uint8_t*>(buf));
  SLM_ENTRY_END
}

int slm_create_from_snapshot(const void* buf, uint64_t bytes, int32_t device, slm_sim** out) {
  SLM_ENTRY_BEGIN
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  SnapHeader hd;
  SLM_TRY(check_snapshot(static_cast<const uint8_t*>(buf), bytes, &hd));
  slm_config cfg;
  slm_config_default(&cfg);
  cfg.width = hd.width;
  cfg.height = hd.height;
  cfg.agent_count = hd.agent_count;
  cfg.seed = hd.seed;
  cfg.storage = hd.storage;
  cfg.device = device;
  cfg.flags = hd.flags;
  slm_sim* s = nullptr;
  SLM_TRY(slm_create(&cfg, &hd.params, &s));
  int rc = slm_load_snapshot(s, buf, bytes);
  if (rc) {
    char msg[sizeof g_last_error];
    memcpy(msg, g_last_error, sizeof msg);
    slm_destroy(s);
    return fail(rc, "%s", msg);
  }
  *out = s;
  return SLM_OK;
  SLM_ENTRY_END
}

// ---- batches ------------------------------------------------------------------

int slm_batch_config_default(slm_batch_config* cfg) {
  SLM_ENTRY_BEGIN
  if (!cfg) return fail(SLM_ERR_INVALID_ARGUMENT, "cfg is NULL");
  std::memset(cfg, 0, sizeof *cfg);
  cfg->storage = SLM_STORAGE_F32;
  cfg->graph_chunk = 100;
  cfg->sort_cell_log2 = 4;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_batch_create(const slm_batch_config* cfg, const slm_instance_desc* insts, uint32_t count,
                     slm_batch** out) {
  SLM_ENTRY_BEGIN
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!cfg || !insts) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count == 0) return fail(SLM_ERR_INVALID_ARGUMENT, "a batch needs at least one instance");
  if (count > 65535) return fail(SLM_ERR_INVALID_ARGUMENT, "a batch holds at most 65535 instances");
  EngineConfig ec;
  ec.storage = cfg->storage;
  ec.device = cfg->device;
  ec.resort_interval = cfg->resort_interval;
  ec.graph_chunk = cfg->graph_chunk;
  ec.sort_cell_log2 = cfg->sort_cell_log2;
  ec.flags = cfg->flags;
  SLM_TRY(check_engine_config(ec, cfg->reserved, 6));
  std::vector<InstanceSpec> specs;
  specs.reserve(count);
  uint64_t agents = 0, cells = 0;
  for (uint32_t k = 0; k < count; k++) {
    const slm_instance_desc& d = insts[k];
    int rc = check_grid(d.width, d.height);
    if (!rc) rc = check_params(&d.params, cfg->flags);
    if (rc) return fail(rc, "instance %u: %s", k, g_last_error);
    // bounded before the add, so the running total cannot wrap
    if (d.agent_count > kMaxAgents || agents + d.agent_count > kMaxAgents)
      return fail(SLM_ERR_INVALID_ARGUMENT, "instance %u: total agents must be < 2^32 - 256", k);
    agents += d.agent_count;
    cells += static_cast<uint64_t>(d.width) * d.height;
    specs.push_back({d.width, d.height, d.agent_count, d.seed, d.params});
  }
  if (agents > kMaxAgents) return fail(SLM_ERR_INVALID_ARGUMENT, "total agents must be < 2^32 - 256");
  if (cells > kMaxPoolCells) return fail(SLM_ERR_INVALID_ARGUMENT, "total cells must be <= 2^40");
  slm_batch* b = new (std::nothrow) slm_batch{nullptr, *cfg};
  if (!b) return fail(SLM_ERR_OUT_OF_MEMORY, "host allocation failed");
  int rc = make_engine(ec, std::move(specs), &b->eng);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return SLM_OK;
  SLM_ENTRY_END
}

#define BATCH_HANDLE(b)                                                        \
  if (!(b) || !(b)->eng) return fail(SLM_ERR_INVALID_ARGUMENT, "batch is NULL"); \
  DeviceGuard guard_((b)->eng->cfg.device)

#define BATCH_GUARD(b)                                                                     \
  BATCH_HANDLE(b);                                                                         \
  if ((b)->eng->broken)                                                                    \
  return fail((b)->eng->broken, "the batch is unusable after a failed tuning change; destroy it")

#define BATCH_INDEX(b, k)                                                                \
  if ((k) >= (b)->eng->count())                                                          \
  return fail(SLM_ERR_INVALID_ARGUMENT, "instance %u out of range (%u instances)", (k), \
              (b)->eng->count())

int slm_batch_destroy(slm_batch* b) {
  SLM_ENTRY_BEGIN
  BATCH_HANDLE(b);
  delete b->eng;
  delete b;
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_batch_step(slm_batch* b, uint32_t n) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  return b->eng->step(n);
  SLM_ENTRY_END
}

int slm_batch_step_timed(slm_batch* b, uint32_t n, slm_timing* out) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  return b->eng->step_timed(n, out);
  SLM_ENTRY_END
}

int slm_batch_synchronize(slm_batch* b) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  return b->eng->sync();
  SLM_ENTRY_END
}

int slm_batch_set_params(slm_batch* b, uint32_t k, const slm_params* p) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  BATCH_INDEX(b, k);
  SLM_TRY(check_params(p, b->eng->cfg.flags));
  return b->eng->set_params(k, *p);
  SLM_ENTRY_END
}

int slm_batch_clear(slm_batch* b, uint32_t k) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  BATCH_INDEX(b, k);
  return b->eng->clear(k);
  SLM_ENTRY_END
}

int slm_batch_read_agents(slm_batch* b, uint32_t k, float* out, uint64_t n) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  BATCH_INDEX(b, k);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(n, b->eng->spec[k].agents, "read_agents"));
  return b->eng->read_agents(k, out);
  SLM_ENTRY_END
}

int slm_batch_write_agents(slm_batch* b, uint32_t k, const float* in, uint64_t n) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  BATCH_INDEX(b, k);
  if (!in) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(n, b->eng->spec[k].agents, "write_agents"));
  return b->eng->write_agents(k, in);
  SLM_ENTRY_END
}

int slm_batch_read_trail(slm_batch* b, uint32_t k, float* out, uint64_t cells) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  BATCH_INDEX(b, k);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(cells, b->eng->cells(k), "read_trail (cells != width*height)"));
  return b->eng->read_trail(k, out);
  SLM_ENTRY_END
}

int slm_batch_write_trail(slm_batch* b, uint32_t k, const float* in, uint64_t cells) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  BATCH_INDEX(b, k);
  if (!in) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(cells, b->eng->cells(k), "write_trail (cells != width*height)"));
  return b->eng->write_trail(k, in);
  SLM_ENTRY_END
}

int slm_batch_total_mass(slm_batch* b, double* out, uint32_t count) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_n(count, b->eng->count(), "total_mass (count != instances)"));
  return b->eng->total_mass(out);
  SLM_ENTRY_END
}

static void sync_batch_config(slm_batch_config& c, const Engine& e) {
  c.resort_interval = e.cfg.resort_interval;
  c.sort_cell_log2 = e.cfg.sort_cell_log2;
  c.flags = e.cfg.flags;
}

int slm_batch_get_tuning(const slm_batch* b, slm_tuning* out) {
  SLM_ENTRY_BEGIN
  if (!b || !b->eng || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = b->eng->tuning();
  return SLM_OK;
  SLM_ENTRY_END
}

int slm_batch_set_tuning(slm_batch* b, const slm_tuning* t) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  if (!t) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  SLM_TRY(check_tuning(*t));
  int rc = b->eng->set_tuning(*t);
  sync_batch_config(b->cfg, *b->eng);
  return rc;
  SLM_ENTRY_END
}

int slm_batch_autotune(slm_batch* b, uint32_t trial_steps, slm_tuning* out) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  int rc = b->eng->autotune(trial_steps, out);
  sync_batch_config(b->cfg, *b->eng);
  return rc;
  SLM_ENTRY_END
}

int slm_batch_render_rgba8(slm_batch* b, uint32_t k, const uint8_t* lut, uint32_t tex_w,
                           uint32_t tex_h, uint8_t* rgba) {
  SLM_ENTRY_BEGIN
  BATCH_GUARD(b);
  BATCH_INDEX(b, k);
  SLM_TRY(check_render(lut, tex_w, tex_h, rgba));
  return b->eng->render(k, lut, tex_w, tex_h, rgba, false);
  SLM_ENTRY_END
}

int slm_batch_get_info(const slm_batch* b, slm_batch_info* out) {
  SLM_ENTRY_BEGIN
  if (!b || !b->eng || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  std::memset(out, 0, sizeof *out);
  const Engine& e = *b->eng;
  out->count = e.count();
  out->storage = e.cfg.storage;
  for (size_t k = 0; k < e.spec.size(); k++) {
    out->total_agents += e.spec[k].agents;
    out->total_cells += e.cells(k);
  }
  out->steps_done = e.steps_done;
  out->device_bytes = e.device_bytes;
  return SLM_OK;
  SLM_ENTRY_END
}

}  // extern "C"
