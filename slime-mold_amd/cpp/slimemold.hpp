// slimemold.hpp - C++ host-side mirror of the reference's library interface,
// over the C ABI of libslimemold_hip.so (include/slimemold.h).
//
//   slime_mold::settings   constants and Settings     (src/settings.rs:3-70)
//   slime_mold::presets    Preset, PresetManager and
//                          init_preset_manager()      (src/presets.rs:6-155)
//   slime_mold::Simulation the compute dispatch of src/main.rs as one object:
//                          step (:1163-1235), clear (:908-913), resize
//                          (:921-1085), agent count (:682-791), speed
//                          reassignment (:101-145), parameter upload (:83-99)
//   slime_mold::Batch      many independent simulations advanced together
//
// Header-only; errors are thrown as slime_mold::Error (the C ABI itself never
// throws).  Field names and preset values are the reference's.
#pragma once

#include <algorithm>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/slimemold.h"

namespace slime_mold {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
  int code;
};

inline void check(int rc) {
  if (rc != SLM_OK) throw Error(rc, slm_last_error());
}

namespace settings {
// General settings (src/settings.rs:4-6)
constexpr unsigned DEFAULT_WIDTH = 1600;
constexpr unsigned DEFAULT_HEIGHT = 900;
constexpr bool DEFAULT_IS_FULLSCREEN = false;
// Agent settings (src/settings.rs:9-17)
constexpr size_t AGENT_COUNT = 10000000;
constexpr float AGENT_SPEED_MIN = 30.0f;
constexpr float AGENT_SPEED_MAX = 50.0f;
constexpr float AGENT_TURN_SPEED = 0.43f;
constexpr float AGENT_POSSIBLE_STARTING_HEADINGS_START = 0.0f;  // Range 0.0..360.0
constexpr float AGENT_POSSIBLE_STARTING_HEADINGS_END = 360.0f;
constexpr float DEPOSITION_AMOUNT = 1.0f;
constexpr float AGENT_JITTER = 0.0f;
constexpr float AGENT_SENSOR_ANGLE = 0.3f;
constexpr float AGENT_SENSOR_DISTANCE = 20.0f;
// Pheromone settings (src/settings.rs:19-27)
constexpr float DECAY_FACTOR = 10.0f;
constexpr float DIFFUSION_RATE = 1.0f;
constexpr float BLUR_RADIUS = 2.0f;
constexpr float BLUR_SIGMA = 1.0f;

// src/settings.rs:29-70
struct Settings {
  size_t agent_count = AGENT_COUNT;
  float agent_jitter = AGENT_JITTER;
  std::pair<float, float> agent_possible_starting_headings{AGENT_POSSIBLE_STARTING_HEADINGS_START,
                                                           AGENT_POSSIBLE_STARTING_HEADINGS_END};
  float agent_speed_max = AGENT_SPEED_MAX;
  float agent_speed_min = AGENT_SPEED_MIN;
  float agent_turn_speed = AGENT_TURN_SPEED;
  float pheromone_decay_factor = DECAY_FACTOR;
  float pheromone_diffusion_rate = DIFFUSION_RATE;
  float pheromone_deposition_amount = DEPOSITION_AMOUNT;
  bool window_fullscreen = DEFAULT_IS_FULLSCREEN;
  unsigned window_height = DEFAULT_HEIGHT;
  unsigned window_width = DEFAULT_WIDTH;
  float agent_sensor_angle = AGENT_SENSOR_ANGLE;
  float agent_sensor_distance = AGENT_SENSOR_DISTANCE;
  float blur_radius = BLUR_RADIUS;
  float blur_sigma = BLUR_SIGMA;

  // The step parameters as SimSizeUniform::new builds them (src/main.rs:48-67).
  slm_params to_params() const {
    slm_params p;
    p.decay_factor = pheromone_decay_factor;
    p.agent_jitter = agent_jitter;
    p.agent_speed_min = agent_speed_min;
    p.agent_speed_max = agent_speed_max;
    p.agent_turn_speed = agent_turn_speed;
    p.agent_sensor_angle = agent_sensor_angle;
    p.agent_sensor_distance = agent_sensor_distance;
    p.diffusion_rate = pheromone_diffusion_rate;
    p.pheromone_deposition_amount = pheromone_deposition_amount;
    p.blur_radius = blur_radius;
    p.blur_sigma = blur_sigma;
    return p;
  }
};
}  // namespace settings

namespace presets {
using settings::Settings;

struct Preset {
  std::string name;
  Settings settings;
};

class PresetManager {
 public:
  void add_preset(Preset p) { presets_.push_back(std::move(p)); }
  const Preset* get_preset(const std::string& name) const {
    for (const auto& p : presets_)
      if (p.name == name) return &p;
    return nullptr;
  }
  std::vector<std::string> get_preset_names() const {
    std::vector<std::string> out;
    for (const auto& p : presets_) out.push_back(p.name);
    return out;
  }

 private:
  std::vector<Preset> presets_;
};

// src/presets.rs:45-155: fields not listed keep Settings' defaults.
inline PresetManager init_preset_manager() {
  auto with = [](float jitter, float smin, float smax, float turn, float angle, float dist,
                 float deposit, float decay, float diffusion) {
    Settings s;
    s.agent_jitter = jitter;
    s.agent_speed_min = smin;
    s.agent_speed_max = smax;
    s.agent_turn_speed = turn;
    s.agent_sensor_angle = angle;
    s.agent_sensor_distance = dist;
    s.pheromone_deposition_amount = deposit;
    s.pheromone_decay_factor = decay;
    s.pheromone_diffusion_rate = diffusion;
    return s;
  };
  PresetManager m;
  m.add_preset({"Default", Settings{}});
  m.add_preset({"Sponge", with(0.0f, 20.0f, 30.0f, 0.43f, 0.3f, 20.0f, 1.0f, 1.0f, 1.0f)});
  m.add_preset(
      {"Firecracker Trees", with(0.1f, 60.0f, 60.0f, 1.47f, 0.3f, 20.0f, 1.0f, 10.0f, 1.0f)});
  m.add_preset({"Threads", with(0.0f, 70.0f, 80.0f, 0.02f, 0.3f, 20.0f, 1.0f, 10.0f, 0.1f)});
  Settings curls = with(5.0f, 70.0f, 80.0f, 0.05f, 0.3f, 20.0f, 1.0f, 75.0f, 0.1f);
  curls.agent_count = 3000000;
  m.add_preset({"Curls", curls});
  m.add_preset({"Waves", with(1.0f, 30.0f, 50.0f, 6.0f, 0.3f, 20.0f, 1.0f, 10.0f, 0.1f)});
  m.add_preset({"Snake", with(3.0f, 100.0f, 120.0f, 0.37f, 1.34f, 225.0f, 1.0f, 10.0f, 1.0f)});
  m.add_preset({"Mesh", with(3.0f, 100.0f, 120.0f, 6.0f, 1.57f, 225.0f, 1.0f, 10.0f, 1.0f)});
  return m;
}
}  // namespace presets

namespace lut {
// src/lut_manager.rs:4-18: 768 bytes, red[256] then green[256] then blue[256].
struct LutData {
  std::string name;
  std::vector<uint8_t> red, green, blue;

  static LutData from_bytes(const std::string& name, const std::vector<uint8_t>& data) {
    if (data.size() != 768) throw Error(SLM_ERR_INVALID_ARGUMENT, "Invalid LUT file size");
    return {name, std::vector<uint8_t>(data.begin(), data.begin() + 256),
            std::vector<uint8_t>(data.begin() + 256, data.begin() + 512),
            std::vector<uint8_t>(data.begin() + 512, data.end())};
  }
  void reverse() {
    std::reverse(red.begin(), red.end());
    std::reverse(green.begin(), green.end());
    std::reverse(blue.begin(), blue.end());
  }
  // The layout the renderer takes (src/main.rs:330-334).
  std::vector<uint8_t> bytes() const {
    std::vector<uint8_t> out(red);
    out.insert(out.end(), green.begin(), green.end());
    out.insert(out.end(), blue.begin(), blue.end());
    return out;
  }
};
}  // namespace lut

using lut::LutData;
using presets::init_preset_manager;
using presets::Preset;
using presets::PresetManager;
using settings::Settings;

// One simulation on one GPU.
class Simulation {
 public:
  Simulation(unsigned width, unsigned height, uint64_t agent_count, uint64_t seed,
             const Settings& s, slm_storage storage = SLM_STORAGE_F32,
             unsigned resort_interval = 0, unsigned graph_chunk = 100, int device = 0,
             unsigned flags = 0) {
    slm_config cfg;
    check(slm_config_default(&cfg));
    cfg.width = width;
    cfg.height = height;
    cfg.agent_count = agent_count;
    cfg.seed = seed;
    cfg.storage = storage;
    cfg.resort_interval = resort_interval;
    cfg.graph_chunk = graph_chunk;
    cfg.device = device;
    cfg.flags = flags;
    slm_params p = s.to_params();
    check(slm_create(&cfg, &p, &h_));
  }
  ~Simulation() {
    if (h_) slm_destroy(h_);
  }
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;
  Simulation(Simulation&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }

  slm_info info() const {
    slm_info i;
    check(slm_get_info(h_, &i));
    return i;
  }
  void set_settings(const Settings& s) {
    slm_params p = s.to_params();
    check(slm_set_params(h_, &p));
  }
  void step(uint32_t n = 1) { check(slm_step(h_, n)); }
  void synchronize() { check(slm_synchronize(h_)); }
  void clear() { check(slm_clear(h_)); }
  void resize(unsigned w, unsigned h) { check(slm_resize(h_, w, h)); }
  void set_agent_count(uint64_t n) { check(slm_set_agent_count(h_, n)); }
  void reassign_speeds(float smin, float smax) { check(slm_reassign_speeds(h_, smin, smax)); }
  std::vector<float> read_agents() {
    std::vector<float> a(info().agent_count * 4);
    check(slm_read_agents(h_, a.data(), a.size() / 4));
    return a;
  }
  void write_agents(const std::vector<float>& a) {
    check(slm_write_agents(h_, a.data(), a.size() / 4));
  }
  std::vector<float> read_trail() {
    slm_info i = info();
    std::vector<float> t(static_cast<size_t>(i.width) * i.height);
    check(slm_read_trail(h_, t.data(), t.size()));
    return t;
  }
  void write_trail(const std::vector<float>& t) { check(slm_write_trail(h_, t.data(), t.size())); }
  double total_mass() {
    double m = 0.0;
    check(slm_total_mass(h_, &m));
    return m;
  }
  slm_tuning tuning() const {
    slm_tuning t;
    check(slm_get_tuning(h_, &t));
    return t;
  }
  void set_tuning(const slm_tuning& t) { check(slm_set_tuning(h_, &t)); }
  slm_tuning autotune(uint32_t trial_steps = 20) {
    slm_tuning t;
    check(slm_autotune(h_, trial_steps, &t));
    return t;
  }
  // Display colour map: tex_h x tex_w RGBA8, row-major (src/display.wgsl:29-87).
  std::vector<uint8_t> render(const LutData& lut, uint32_t tex_w, uint32_t tex_h) {
    std::vector<uint8_t> table = lut.bytes(), out(static_cast<size_t>(tex_w) * tex_h * 4);
    check(slm_render_rgba8(h_, table.data(), tex_w, tex_h, out.data()));
    return out;
  }
  std::vector<uint8_t> snapshot() {
    uint64_t n = 0;
    check(slm_snapshot_size(h_, &n));
    std::vector<uint8_t> buf(n);
    check(slm_save_snapshot(h_, buf.data(), n));
    return buf;
  }
  void load_snapshot(const std::vector<uint8_t>& buf) {
    check(slm_load_snapshot(h_, buf.data(), buf.size()));
  }
  static Simulation from_snapshot(const std::vector<uint8_t>& buf, int device = 0) {
    slm_sim* h = nullptr;
    check(slm_create_from_snapshot(buf.data(), buf.size(), device, &h));
    return Simulation(h);
  }
  slm_sim* handle() { return h_; }

 private:
  explicit Simulation(slm_sim* h) : h_(h) {}
  slm_sim* h_ = nullptr;
};

// Independent simulations of uneven sizes advanced together.
class Batch {
 public:
  struct Instance {
    unsigned width, height;
    uint64_t agent_count, seed;
    Settings settings;
  };
  Batch(const std::vector<Instance>& instances, slm_storage storage = SLM_STORAGE_F32,
        unsigned resort_interval = 0, unsigned graph_chunk = 100, int device = 0,
        unsigned flags = 0) {
    slm_batch_config cfg;
    check(slm_batch_config_default(&cfg));
    cfg.storage = storage;
    cfg.resort_interval = resort_interval;
    cfg.graph_chunk = graph_chunk;
    cfg.device = device;
    cfg.flags = flags;
    std::vector<slm_instance_desc> d(instances.size());
    for (size_t k = 0; k < instances.size(); k++) {
      d[k].width = instances[k].width;
      d[k].height = instances[k].height;
      d[k].agent_count = instances[k].agent_count;
      d[k].seed = instances[k].seed;
      d[k].params = instances[k].settings.to_params();
    }
    check(slm_batch_create(&cfg, d.data(), static_cast<uint32_t>(d.size()), &h_));
    inst_ = instances;
  }
  ~Batch() {
    if (h_) slm_batch_destroy(h_);
  }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  void step(uint32_t n = 1) { check(slm_batch_step(h_, n)); }
  void synchronize() { check(slm_batch_synchronize(h_)); }
  void set_settings(uint32_t k, const Settings& s) {
    slm_params p = s.to_params();
    check(slm_batch_set_params(h_, k, &p));
  }
  std::vector<float> read_trail(uint32_t k) {
    std::vector<float> t(static_cast<size_t>(inst_.at(k).width) * inst_[k].height);
    check(slm_batch_read_trail(h_, k, t.data(), t.size()));
    return t;
  }
  std::vector<float> read_agents(uint32_t k) {
    std::vector<float> a(inst_.at(k).agent_count * 4);
    check(slm_batch_read_agents(h_, k, a.data(), a.size() / 4));
    return a;
  }
  slm_tuning tuning() const {
    slm_tuning t;
    check(slm_batch_get_tuning(h_, &t));
    return t;
  }
  void set_tuning(const slm_tuning& t) { check(slm_batch_set_tuning(h_, &t)); }
  slm_tuning autotune(uint32_t trial_steps = 20) {
    slm_tuning t;
    check(slm_batch_autotune(h_, trial_steps, &t));
    return t;
  }
  std::vector<double> total_mass() {
    std::vector<double> m(inst_.size());
    check(slm_batch_total_mass(h_, m.data(), static_cast<uint32_t>(m.size())));
    return m;
  }

 private:
  slm_batch* h_ = nullptr;
  std::vector<Instance> inst_;
};

}  // namespace slime_mold
