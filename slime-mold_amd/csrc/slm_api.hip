// slm_api.hip - the C ABI declared in include/slimemold.h.
//
// A handle owns: the parameter block (device), two trail buffers used
// alternately (the trail pass reads one and writes the other), the agents as
// SoA arrays (two sets when re-sorting, which scatters from one to the other),
// re-sort scratch, f64 mass scratch and one non-blocking HIP stream.  Steps are
// enqueued eagerly or as captured hipGraphs of `graph_chunk` steps; a graph is
// keyed by which trail and agent buffers are current when it starts, so at
// most four exist per handle.  Every entry point validates its arguments before
// any kernel is launched, so the kernels can rely on in-range shapes and
// finite, bounded positions.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "slm_internal.h"

using namespace slm;

namespace {

thread_local std::string g_last_error = "no error";

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define SLM_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(SLM_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                                  \
  } while (0)

constexpr uint32_t kMaxSide = 1u << 24;            // f32(width) exact
constexpr uint64_t kMaxCells = 1ull << 31;          // i32 cell index (compute.wgsl:21)
constexpr uint64_t kMaxAgents = 0xffffff00ull;      // u32 slot index, whole blocks
constexpr float kMaxSensorDistance = 4194304.0f;    // keeps sensor taps inside i32
constexpr uint32_t kMaxSortBins = 1u << 20;

bool finite_params(const slm_params& p) {
  const float* f = &p.decay_factor;
  for (size_t i = 0; i < sizeof(slm_params) / sizeof(float); i++)
    if (!std::isfinite(f[i])) return false;
  return true;
}

int check_params(const slm_params* p) {
  if (!p) return fail(SLM_ERR_INVALID_ARGUMENT, "params is NULL");
  if (!finite_params(*p)) return fail(SLM_ERR_INVALID_ARGUMENT, "params must be finite");
  if (std::fabs(p->agent_sensor_distance) > kMaxSensorDistance)
    return fail(SLM_ERR_INVALID_ARGUMENT, "|agent_sensor_distance| must be <= 2^22");
  return SLM_OK;
}

int check_grid(uint32_t w, uint32_t h) {
  if (w == 0 || h == 0) return fail(SLM_ERR_INVALID_ARGUMENT, "width and height must be >= 1");
  if (w > kMaxSide || h > kMaxSide)
    return fail(SLM_ERR_INVALID_ARGUMENT, "width and height must be <= 2^24");
  if (static_cast<uint64_t>(w) * h > kMaxCells)
    return fail(SLM_ERR_INVALID_ARGUMENT, "width*height must be <= 2^31");
  return SLM_OK;
}

struct Graph {
  hipGraphExec_t exec = nullptr;
};

}  // namespace

struct slm_sim {
  slm_config cfg{};
  slm_params params{};
  hipStream_t stream = nullptr;
  DevParams* d_params = nullptr;
  void* trail[2] = {nullptr, nullptr};
  int cur = 0;                 // trail[cur] holds the current trail
  AgentSoA agents[2] = {};     // agents[acur] is current; [1] only when re-sorting
  int acur = 0;
  SortScratch sort{};
  double* mass_partials = nullptr;
  double* mass_out = nullptr;
  uint64_t steps_done = 0;
  uint64_t reassign_calls = 0;
  uint64_t device_bytes = 0;
  Graph graphs[2][2];          // [trail parity][agent parity] at chunk start
  std::vector<void*> allocations;
};

namespace {

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

int dev_alloc(slm_sim* s, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? SLM_ERR_OUT_OF_MEMORY : SLM_ERR_HIP,
                "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  s->allocations.push_back(*p);
  s->device_bytes += bytes;
  return SLM_OK;
}

void dev_free(slm_sim* s, void* p, size_t bytes) {
  if (!p) return;
  for (auto& a : s->allocations)
    if (a == p) a = nullptr;
  (void)hipFree(p);
  s->device_bytes -= bytes ? bytes : 256;
}

uint64_t cells_of(const slm_sim* s) { return static_cast<uint64_t>(s->cfg.width) * s->cfg.height; }
// Rounded up to 256 bytes: the 16-bit deposit mark is a 32-bit atomic on the
// word holding the cell, which for the last cell of an odd-sized grid reaches
// two bytes past width*height*2.
uint64_t trail_bytes(const slm_sim* s) {
  return (cells_of(s) * storage_bytes(s->cfg.storage) + 255u) & ~uint64_t(255u);
}
bool sorting(const slm_sim* s) { return s->cfg.resort_interval > 0; }

StepGeom geom(const slm_sim* s) {
  StepGeom g;
  g.width = s->cfg.width;
  g.height = s->cfg.height;
  g.agents = static_cast<uint32_t>(s->cfg.agent_count);
  g.storage = s->cfg.storage;
  g.xcd_remap = (s->cfg.flags & SLM_FLAG_NO_XCD_REMAP) ? 0 : 1;
  return g;
}

void drop_graphs(slm_sim* s) {
  for (auto& row : s->graphs)
    for (auto& g : row)
      if (g.exec) {
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
      }
}

int alloc_trail(slm_sim* s) {
  for (int i = 0; i < 2; i++) {
    int rc = dev_alloc(s, &s->trail[i], trail_bytes(s));
    if (rc) return rc;
    SLM_HIP(hipMemsetAsync(s->trail[i], 0, trail_bytes(s), s->stream));
  }
  s->cur = 0;
  return SLM_OK;
}

void free_trail(slm_sim* s) {
  for (auto& t : s->trail) {
    dev_free(s, t, trail_bytes(s));
    t = nullptr;
  }
}

int alloc_agents(slm_sim* s) {
  uint64_t n = s->cfg.agent_count;
  int sets = sorting(s) ? 2 : 1;
  for (int k = 0; k < sets; k++) {
    AgentSoA& a = s->agents[k];
    int rc;
    if ((rc = dev_alloc(s, reinterpret_cast<void**>(&a.x), n * 4))) return rc;
    if ((rc = dev_alloc(s, reinterpret_cast<void**>(&a.y), n * 4))) return rc;
    if ((rc = dev_alloc(s, reinterpret_cast<void**>(&a.th), n * 4))) return rc;
    if ((rc = dev_alloc(s, reinterpret_cast<void**>(&a.v), n * 4))) return rc;
    if (sorting(s) && (rc = dev_alloc(s, reinterpret_cast<void**>(&a.id), n * 4))) return rc;
  }
  if (sorting(s)) {
    int rc;
    if ((rc = dev_alloc(s, reinterpret_cast<void**>(&s->sort.key), n * 4))) return rc;
    if ((rc = dev_alloc(s, reinterpret_cast<void**>(&s->sort.rank), n * 4))) return rc;
  }
  s->acur = 0;
  return SLM_OK;
}

void free_agents(slm_sim* s) {
  uint64_t n = s->cfg.agent_count;
  for (auto& a : s->agents) {
    dev_free(s, a.x, n * 4);
    dev_free(s, a.y, n * 4);
    dev_free(s, a.th, n * 4);
    dev_free(s, a.v, n * 4);
    dev_free(s, a.id, n * 4);
    a = AgentSoA{};
  }
  dev_free(s, s->sort.key, n * 4);
  dev_free(s, s->sort.rank, n * 4);
  s->sort.key = s->sort.rank = nullptr;
}

int alloc_bins(slm_sim* s) {
  if (!sorting(s)) return SLM_OK;
  uint32_t cl = s->cfg.sort_cell_log2;
  while (sort_bins_for(s->cfg.width, s->cfg.height, cl) > kMaxSortBins) cl++;
  s->sort.cell_log2 = cl;
  s->sort.nbins = sort_bins_for(s->cfg.width, s->cfg.height, cl);
  int rc;
  if ((rc = dev_alloc(s, reinterpret_cast<void**>(&s->sort.count), s->sort.nbins * 4ull))) return rc;
  if ((rc = dev_alloc(s, reinterpret_cast<void**>(&s->sort.offset), s->sort.nbins * 4ull))) return rc;
  if ((rc = dev_alloc(s, reinterpret_cast<void**>(&s->sort.block_sums),
                      sort_block_sums_for(s->sort.nbins) * 4ull)))
    return rc;
  return SLM_OK;
}

void free_bins(slm_sim* s) {
  dev_free(s, s->sort.count, s->sort.nbins * 4ull);
  dev_free(s, s->sort.offset, s->sort.nbins * 4ull);
  dev_free(s, s->sort.block_sums, sort_block_sums_for(s->sort.nbins) * 4ull);
  s->sort.count = s->sort.offset = s->sort.block_sums = nullptr;
}

void init_agents(slm_sim* s) {
  launch_init_agents(s->agents[s->acur], static_cast<uint32_t>(s->cfg.agent_count), s->cfg.seed,
                     s->cfg.width, s->cfg.height, s->params.agent_speed_min,
                     s->params.agent_speed_max, s->stream);
}

// Enqueue one step: optional re-sort, agents, trail pass.
void enqueue_step(slm_sim* s, bool do_sort) {
  StepGeom g = geom(s);
  if (do_sort && sorting(s) && g.agents > 0) {
    launch_morton_sort(g, s->agents[s->acur], s->agents[s->acur ^ 1], s->sort, s->stream);
    s->acur ^= 1;
  }
  launch_agents(g, s->agents[s->acur], s->trail[s->cur], s->d_params, s->stream);
  launch_trail(g, s->trail[s->cur], s->trail[s->cur ^ 1], s->d_params, s->stream);
  s->cur ^= 1;
}

bool sort_due(const slm_sim* s, uint64_t step_index) {
  return sorting(s) && step_index % s->cfg.resort_interval == 0;
}

int run_chunk_graph(slm_sim* s) {
  uint32_t chunk = s->cfg.graph_chunk;
  Graph& g = s->graphs[s->cur][s->acur];
  int cur0 = s->cur, acur0 = s->acur;
  if (!g.exec) {
    hipGraph_t graph = nullptr;
    SLM_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t j = 0; j < chunk; j++) enqueue_step(s, sorting(s) && j % s->cfg.resort_interval == 0);
    hipError_t e = hipStreamEndCapture(s->stream, &graph);
    int cur1 = s->cur, acur1 = s->acur;
    s->cur = cur0;
    s->acur = acur0;
    if (e != hipSuccess) return fail(SLM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g.exec = nullptr;
      return fail(SLM_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    (void)cur1;
    (void)acur1;
  }
  SLM_HIP(hipGraphLaunch(g.exec, s->stream));
  // replay the parity changes the captured steps make
  for (uint32_t j = 0; j < chunk; j++) {
    if (sorting(s) && j % s->cfg.resort_interval == 0 && s->cfg.agent_count > 0) s->acur ^= 1;
    s->cur ^= 1;
  }
  s->steps_done += chunk;
  return SLM_OK;
}

int sync_stream(slm_sim* s) {
  SLM_HIP(hipStreamSynchronize(s->stream));
  return SLM_OK;
}

void destroy_all(slm_sim* s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graphs(s);
  for (void* p : s->allocations)
    if (p) (void)hipFree(p);
  s->allocations.clear();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // namespace

extern "C" {

const char* slm_last_error(void) { return g_last_error.c_str(); }

int slm_device_count(int* out_count) {
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
}

int slm_config_default(slm_config* cfg) {
  if (!cfg) return fail(SLM_ERR_INVALID_ARGUMENT, "cfg is NULL");
  std::memset(cfg, 0, sizeof *cfg);
  cfg->width = 1600;  // src/settings.rs:4-5
  cfg->height = 900;
  cfg->storage = SLM_STORAGE_F32;
  cfg->graph_chunk = 100;
  cfg->sort_cell_log2 = 4;
  return SLM_OK;
}

int slm_params_default(slm_params* p) {
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
}

int slm_create(const slm_config* cfg, const slm_params* params, slm_sim** out) {
  if (!out) return fail(SLM_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!cfg) return fail(SLM_ERR_INVALID_ARGUMENT, "cfg is NULL");
  int rc = check_params(params);
  if (rc) return rc;
  if ((rc = check_grid(cfg->width, cfg->height))) return rc;
  if (cfg->agent_count > kMaxAgents)
    return fail(SLM_ERR_INVALID_ARGUMENT, "agent_count must be < 2^32 - 256");
  if (cfg->storage < SLM_STORAGE_F32 || cfg->storage > SLM_STORAGE_F16)
    return fail(SLM_ERR_INVALID_ARGUMENT, "unknown storage %d", cfg->storage);
  if (cfg->sort_cell_log2 > 12) return fail(SLM_ERR_INVALID_ARGUMENT, "sort_cell_log2 must be <= 12");
  if (cfg->graph_chunk > 100000) return fail(SLM_ERR_INVALID_ARGUMENT, "graph_chunk must be <= 100000");
  if (cfg->flags & ~SLM_FLAG_NO_XCD_REMAP) return fail(SLM_ERR_INVALID_ARGUMENT, "unknown flags");
  for (uint32_t r : cfg->reserved)
    if (r) return fail(SLM_ERR_INVALID_ARGUMENT, "reserved fields must be zero");
  int ndev = 0;
  if ((rc = slm_device_count(&ndev))) return rc;
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(SLM_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", cfg->device, ndev);

  DeviceGuard guard(cfg->device);
  slm_sim* s = new (std::nothrow) slm_sim();
  if (!s) return fail(SLM_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->cfg = *cfg;
  s->params = *params;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return fail(SLM_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
  }
  if ((rc = dev_alloc(s, reinterpret_cast<void**>(&s->d_params), sizeof(DevParams))) ||
      (rc = dev_alloc(s, reinterpret_cast<void**>(&s->mass_partials), kMassBlocks * sizeof(double))) ||
      (rc = dev_alloc(s, reinterpret_cast<void**>(&s->mass_out), sizeof(double))) ||
      (rc = alloc_trail(s)) || (rc = alloc_agents(s)) || (rc = alloc_bins(s))) {
    destroy_all(s);
    return rc;
  }
  launch_set_params(s->d_params, make_dev_params(s->params), s->stream);
  init_agents(s);
  e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    destroy_all(s);
    return fail(SLM_ERR_HIP, "initialisation failed: %s", hipGetErrorString(e));
  }
  *out = s;
  return SLM_OK;
}

int slm_destroy(slm_sim* s) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  DeviceGuard guard(s->cfg.device);
  destroy_all(s);
  return SLM_OK;
}

int slm_set_params(slm_sim* s, const slm_params* p) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  int rc = check_params(p);
  if (rc) return rc;
  DeviceGuard guard(s->cfg.device);
  s->params = *p;
  launch_set_params(s->d_params, make_dev_params(*p), s->stream);
  SLM_HIP(hipGetLastError());
  return SLM_OK;
}

int slm_get_params(const slm_sim* s, slm_params* out) {
  if (!s || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = s->params;
  return SLM_OK;
}

int slm_step(slm_sim* s, uint32_t n) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  DeviceGuard guard(s->cfg.device);
  uint32_t chunk = s->cfg.graph_chunk;
  while (n > 0) {
    if (chunk > 0 && n >= chunk) {
      int rc = run_chunk_graph(s);
      if (rc) return rc;
      n -= chunk;
    } else {
      enqueue_step(s, sort_due(s, s->steps_done));
      s->steps_done++;
      n--;
    }
  }
  SLM_HIP(hipGetLastError());
  return SLM_OK;
}

int slm_step_timed(slm_sim* s, uint32_t n, slm_timing* out) {
  if (!s || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  DeviceGuard guard(s->cfg.device);
  std::memset(out, 0, sizeof *out);
  hipEvent_t ev[6];
  for (auto& e : ev) SLM_HIP(hipEventCreate(&e));
  int rc = SLM_OK;
  StepGeom g = geom(s);
  for (uint32_t i = 0; i < n && rc == SLM_OK; i++) {
    bool sort = sort_due(s, s->steps_done) && g.agents > 0;
    if (sort) {
      (void)hipEventRecord(ev[0], s->stream);
      launch_morton_sort(g, s->agents[s->acur], s->agents[s->acur ^ 1], s->sort, s->stream);
      (void)hipEventRecord(ev[1], s->stream);
      s->acur ^= 1;
    }
    (void)hipEventRecord(ev[2], s->stream);
    launch_agents(g, s->agents[s->acur], s->trail[s->cur], s->d_params, s->stream);
    (void)hipEventRecord(ev[3], s->stream);
    launch_trail(g, s->trail[s->cur], s->trail[s->cur ^ 1], s->d_params, s->stream);
    (void)hipEventRecord(ev[4], s->stream);
    s->cur ^= 1;
    s->steps_done++;
    hipError_t e = hipEventSynchronize(ev[4]);
    if (e != hipSuccess) {
      rc = fail(SLM_ERR_HIP, "step failed: %s", hipGetErrorString(e));
      break;
    }
    float ms = 0.0f;
    if (sort && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) {
      out->sort_ms += ms;
      out->sort_launches++;
    }
    if (g.agents > 0 && hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess) {
      out->agents_ms += ms;
      out->agents_launches++;
    }
    if (hipEventElapsedTime(&ms, ev[3], ev[4]) == hipSuccess) {
      out->trail_ms += ms;
      out->trail_launches++;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  SLM_HIP(hipGetLastError());
  return SLM_OK;
}

int slm_synchronize(slm_sim* s) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  DeviceGuard guard(s->cfg.device);
  return sync_stream(s);
}

int slm_clear(slm_sim* s) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  DeviceGuard guard(s->cfg.device);
  SLM_HIP(hipMemsetAsync(s->trail[s->cur], 0, trail_bytes(s), s->stream));
  return SLM_OK;
}

int slm_resize(slm_sim* s, uint32_t w, uint32_t h) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  int rc = check_grid(w, h);
  if (rc) return rc;
  DeviceGuard guard(s->cfg.device);
  if ((rc = sync_stream(s))) return rc;
  drop_graphs(s);
  // src/main.rs:993-994: ratio computed in f32, then multiplied in
  float sx = static_cast<float>(w) / static_cast<float>(s->cfg.width);
  float sy = static_cast<float>(h) / static_cast<float>(s->cfg.height);
  launch_scale_positions(s->agents[s->acur], static_cast<uint32_t>(s->cfg.agent_count), sx, sy,
                         s->stream);
  if ((rc = sync_stream(s))) return rc;
  free_trail(s);
  free_bins(s);
  s->cfg.width = w;
  s->cfg.height = h;
  if ((rc = alloc_trail(s)) || (rc = alloc_bins(s))) return rc;
  return sync_stream(s);
}

int slm_set_agent_count(slm_sim* s, uint64_t n) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  if (n > kMaxAgents) return fail(SLM_ERR_INVALID_ARGUMENT, "agent_count must be < 2^32 - 256");
  DeviceGuard guard(s->cfg.device);
  int rc = sync_stream(s);
  if (rc) return rc;
  drop_graphs(s);
  free_agents(s);
  s->cfg.agent_count = n;
  if ((rc = alloc_agents(s))) return rc;
  init_agents(s);
  return sync_stream(s);
}

int slm_reassign_speeds(slm_sim* s, float smin, float smax) {
  if (!s) return fail(SLM_ERR_INVALID_ARGUMENT, "sim is NULL");
  if (!std::isfinite(smin) || !std::isfinite(smax))
    return fail(SLM_ERR_INVALID_ARGUMENT, "speeds must be finite");
  DeviceGuard guard(s->cfg.device);
  s->reassign_calls++;
  uint64_t seed = s->cfg.seed + s->reassign_calls * 0xd1b54a32d192ed03ull;
  launch_reassign_speeds(s->agents[s->acur], static_cast<uint32_t>(s->cfg.agent_count), seed, smin,
                         smax, s->stream);
  SLM_HIP(hipGetLastError());
  return SLM_OK;
}

int slm_read_agents(slm_sim* s, float* out, uint64_t n) {
  if (!s || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n != s->cfg.agent_count)
    return fail(SLM_ERR_INVALID_ARGUMENT, "n (%llu) != agent count (%llu)",
                static_cast<unsigned long long>(n),
                static_cast<unsigned long long>(s->cfg.agent_count));
  if (n == 0) return SLM_OK;
  DeviceGuard guard(s->cfg.device);
  void* staging = nullptr;
  int rc = dev_alloc(s, &staging, n * 16);
  if (rc) return rc;
  launch_pack_agents(s->agents[s->acur], static_cast<uint32_t>(n), static_cast<float*>(staging),
                     s->stream);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(out, staging, n * 16, hipMemcpyDeviceToHost);
  dev_free(s, staging, n * 16);
  if (e != hipSuccess) return fail(SLM_ERR_HIP, "read_agents failed: %s", hipGetErrorString(e));
  return SLM_OK;
}

int slm_write_agents(slm_sim* s, const float* in, uint64_t n) {
  if (!s || !in) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n != s->cfg.agent_count)
    return fail(SLM_ERR_INVALID_ARGUMENT, "n (%llu) != agent count (%llu)",
                static_cast<unsigned long long>(n),
                static_cast<unsigned long long>(s->cfg.agent_count));
  float fw = static_cast<float>(s->cfg.width), fh = static_cast<float>(s->cfg.height);
  for (uint64_t i = 0; i < n; i++) {
    const float* r = in + 4 * i;
    if (!(r[0] >= 0.0f && r[0] <= fw && r[1] >= 0.0f && r[1] <= fh) || !std::isfinite(r[2]) ||
        !std::isfinite(r[3]))
      return fail(SLM_ERR_INVALID_ARGUMENT,
                  "agent %llu: position must lie in [0,width]x[0,height], angle and speed finite",
                  static_cast<unsigned long long>(i));
  }
  if (n == 0) return SLM_OK;
  DeviceGuard guard(s->cfg.device);
  int rc = sync_stream(s);
  if (rc) return rc;
  void* staging = nullptr;
  if ((rc = dev_alloc(s, &staging, n * 16))) return rc;
  hipError_t e = hipMemcpy(staging, in, n * 16, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_unpack_agents(s->agents[s->acur], static_cast<uint32_t>(n),
                         static_cast<const float*>(staging), s->stream);
    e = hipStreamSynchronize(s->stream);
  }
  dev_free(s, staging, n * 16);
  if (e != hipSuccess) return fail(SLM_ERR_HIP, "write_agents failed: %s", hipGetErrorString(e));
  return SLM_OK;
}

int slm_read_trail(slm_sim* s, float* out, uint64_t cells) {
  if (!s || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  if (cells != cells_of(s)) return fail(SLM_ERR_INVALID_ARGUMENT, "cells != width*height");
  DeviceGuard guard(s->cfg.device);
  int rc = sync_stream(s);
  if (rc) return rc;
  if (s->cfg.storage == SLM_STORAGE_F32) {
    SLM_HIP(hipMemcpy(out, s->trail[s->cur], cells * 4, hipMemcpyDeviceToHost));
    return SLM_OK;
  }
  void* staging = nullptr;
  if ((rc = dev_alloc(s, &staging, cells * 4))) return rc;
  launch_decode_trail(s->trail[s->cur], s->cfg.storage, cells, static_cast<float*>(staging),
                      s->stream);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(out, staging, cells * 4, hipMemcpyDeviceToHost);
  dev_free(s, staging, cells * 4);
  if (e != hipSuccess) return fail(SLM_ERR_HIP, "read_trail failed: %s", hipGetErrorString(e));
  return SLM_OK;
}

int slm_write_trail(slm_sim* s, const float* in, uint64_t cells) {
  if (!s || !in) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  if (cells != cells_of(s)) return fail(SLM_ERR_INVALID_ARGUMENT, "cells != width*height");
  std::vector<float> clean(in, in + cells);
  for (uint64_t i = 0; i < cells; i++) {
    if (!(clean[i] >= 0.0f) || !std::isfinite(clean[i]))
      return fail(SLM_ERR_INVALID_ARGUMENT, "trail cell %llu must be finite and >= 0",
                  static_cast<unsigned long long>(i));
    if (clean[i] == 0.0f) clean[i] = 0.0f;  // -0 -> +0: the sign bit is the deposit mark
  }
  DeviceGuard guard(s->cfg.device);
  int rc = sync_stream(s);
  if (rc) return rc;
  if (s->cfg.storage == SLM_STORAGE_F32) {
    SLM_HIP(hipMemcpy(s->trail[s->cur], clean.data(), cells * 4, hipMemcpyHostToDevice));
    return SLM_OK;
  }
  void* staging = nullptr;
  if ((rc = dev_alloc(s, &staging, cells * 4))) return rc;
  hipError_t e = hipMemcpy(staging, clean.data(), cells * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_encode_trail(static_cast<const float*>(staging), s->cfg.storage, cells,
                        s->trail[s->cur], s->stream);
    e = hipStreamSynchronize(s->stream);
  }
  dev_free(s, staging, cells * 4);
  if (e != hipSuccess) return fail(SLM_ERR_HIP, "write_trail failed: %s", hipGetErrorString(e));
  return SLM_OK;
}

int slm_total_mass(slm_sim* s, double* out) {
  if (!s || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  DeviceGuard guard(s->cfg.device);
  launch_total_mass(s->trail[s->cur], s->cfg.storage, cells_of(s), s->mass_partials, s->mass_out,
                    s->stream);
  SLM_HIP(hipMemcpyAsync(out, s->mass_out, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  return sync_stream(s);
}

int slm_get_info(const slm_sim* s, slm_info* out) {
  if (!s || !out) return fail(SLM_ERR_INVALID_ARGUMENT, "NULL argument");
  std::memset(out, 0, sizeof *out);
  out->width = s->cfg.width;
  out->height = s->cfg.height;
  out->agent_count = s->cfg.agent_count;
  out->storage = s->cfg.storage;
  out->device = s->cfg.device;
  out->steps_done = s->steps_done;
  out->device_bytes = s->device_bytes;
  out->resort_interval = s->cfg.resort_interval;
  out->graph_chunk = s->cfg.graph_chunk;
  return SLM_OK;
}

}  // extern "C"
