// tick_tuner.hpp — sizes the final renders' share of a served tick.
//
// A tick renders every admitted preview's whole pass plus slices (whole
// pixels) of the final renders' current passes, in one batched launch. With a
// fixed path budget a tick's duration follows the scenes' cost: ticks of
// cheap scenes are short (per-tick overhead, short late-bounce queues that do
// not fill the GPU), ticks of expensive ones long (previews wait behind
// them). The tuner instead aims each tick at a target duration: it keeps a
// running estimate of milliseconds per path for every scene, attributing each
// measured tick to its jobs in proportion to their ray counts, charges the
// previews' estimated cost first and gives the finals what is left — at
// least a quarter of the target, so a burst of previews cannot starve them.
#pragma once
#include <algorithm>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace rtm {

// Target duration of a tuned tick: a preview that arrives during a tick
// waits for it and then takes its own, ~2 x 7 ms, well inside a 50 ms
// preview latency goal.
constexpr double kTickTargetMs = 7.0;
// Target of a tick with no preview in it: a preview arriving meanwhile waits
// at most this long before its own tick, still inside 50 ms, and the finals
// lose less to tick boundaries (each one a host round trip with the GPU
// idle). Measured on the served workload, two runs each, worst load level's
// preview p99: 16 ms 43.0 / 42.0, 12 ms 40.5 / 39.0, 10 ms 44.8 / 41.5, at
// the same throughput.
constexpr double kFinalsTickTargetMs = 12.0;

class TickTuner {
 public:
  // fixed_paths > 0: a fixed budget of final paths per tick (no tuning);
  // 0: tuned to target_ms per tick, starting from fallback_paths per tick
  // for scenes not yet measured.
  TickTuner(uint32_t fixed_paths, double target_ms, uint32_t fallback_paths = 8u << 20)
      : fixed_(fixed_paths), target_ms_(target_ms), fallback_ms_per_path_(target_ms / fallback_paths) {}

  bool tuned() const { return fixed_ == 0; }

  // Start of a tick aimed at `target_ms` (0: the tuner's target).
  void begin(double target_ms = 0) {
    tick_target_ = target_ms > 0 ? target_ms : target_ms_;
    left_ = fixed_ ? (double)fixed_ : tick_target_;
    finals_ = false;
  }
  // A preview's whole pass (rendered whatever the budget).
  void charge_preview(const void* scene, uint64_t paths) {
    if (!fixed_) left_ -= cost(scene) * (double)paths;
  }
  // Paths a final of `scene` may take now (0: the tick is full).
  uint64_t final_paths(const void* scene) const {
    const double left = finals_left();
    if (left <= 0) return 0;
    return fixed_ ? (uint64_t)left : (uint64_t)(left / cost(scene));
  }
  void charge_final(const void* scene, uint64_t paths) {
    left_ = finals_left() - (fixed_ ? (double)paths : cost(scene) * (double)paths);
    finals_ = true;
  }

  // After the tick: every job's paths and rays, then the tick's wall time.
  void add_job(const void* scene, uint64_t paths, uint64_t rays) {
    obs_.push_back(Obs{scene, paths, rays});
    rays_ += rays;
  }
  void end(double tick_ms) {
    if (!fixed_ && rays_ > 0 && tick_ms > 0)
      for (const Obs& o : obs_) {
        if (!o.paths || !o.rays) continue;
        const double sample = tick_ms * ((double)o.rays / (double)rays_) / (double)o.paths;
        auto it = ms_per_path_.find(o.scene);
        if (it == ms_per_path_.end()) ms_per_path_.emplace(o.scene, sample);
        else it->second = 0.5 * it->second + 0.5 * sample;
      }
    obs_.clear();
    rays_ = 0;
  }
  // A scene left the server (its key may be reused).
  void forget(const void* scene) { ms_per_path_.erase(scene); }

 private:
  double finals_left() const { return fixed_ || finals_ ? left_ : std::max(left_, 0.25 * tick_target_); }
  double cost(const void* scene) const {
    auto it = ms_per_path_.find(scene);
    return it == ms_per_path_.end() ? fallback_ms_per_path_ : it->second;
  }
  struct Obs {
    const void* scene;
    uint64_t paths, rays;
  };
  uint32_t fixed_;
  double target_ms_, fallback_ms_per_path_;
  double tick_target_ = 0;
  double left_ = 0;
  bool finals_ = false;
  std::vector<Obs> obs_;
  uint64_t rays_ = 0;
  std::unordered_map<const void*, double> ms_per_path_;
};

}  // namespace rtm
