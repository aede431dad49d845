// renderer.hpp — GpuRenderer and RenderSession: the reference's Renderer /
// RenderSession surface (include/raytracer/rendering/renderer.hpp:25-103)
// with every pass rendered on the MI355X through the rtg C ABI.
//
//   * progressive passes of k = min(remaining, max(1, done), max_pass) samples
//     (renderer.cpp:66-69): 1, 1, 2, 4, 8, ... — same pass boundaries;
//   * on_pass runs on the calling thread between passes; without on_pass the
//     whole schedule is one rtg_render_passes call, whose passes overlap on
//     the device (bitwise the same framebuffer);
//   * cancel is polled between device chunks; a cancelled render reports
//     spp_done = the minimum per-pixel sample count;
//   * statistics use the reference's counting rules (rays = closest-hit
//     calls, shadow_rays = shadow rays cast).
#pragma once
#include <atomic>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "image_io.hpp"
#include "scene.hpp"

namespace rtm {

struct RenderProgress {
  int pass = 0;
  uint32_t spp_done = 0;
  int spp_target = 0;
  double elapsed_ms = 0;
  uint64_t rays = 0;  // closest-hit + shadow rays so far
  bool finished = false;
  bool cancelled = false;
};

struct RenderStatsSnapshot {
  uint64_t camera_samples = 0, rays = 0, shadow_rays = 0, rr_terminated = 0, invalid_samples = 0, passes = 0;
  uint64_t node_visits = 0, prim_tests = 0;
  double elapsed_ms = 0;  // wall time of render()
  double device_ms = 0;   // device time of the passes
  double mrays_per_sec() const { return elapsed_ms > 0 ? (double(rays + shadow_rays) / 1e3) / elapsed_ms : 0; }
  double grays_per_sec() const { return mrays_per_sec() / 1e3; }
};

struct RenderOptions {
  bool progressive = true;
  int max_pass_spp = 32;
  int32_t precision = RTG_PRECISION_F32;
  int device = 0;
};

class GpuRenderer;

// Per-pixel running sums and counts (rendering/framebuffer.hpp:16-50),
// fetched from the device on first access after each pass.
class Framebuffer {
 public:
  int width() const { return w_; }
  int height() const { return h_; }
  void mean(int x, int y, double out[3]) const;
  uint32_t samples(int x, int y) const;
  uint32_t min_samples() const;
  const std::vector<double>& sums() const;
  const std::vector<uint32_t>& counts() const;

 private:
  friend class GpuRenderer;
  void fetch() const;
  const GpuRenderer* owner_ = nullptr;
  int w_ = 0, h_ = 0;
  mutable bool valid_ = false;
  mutable std::vector<double> sum_;
  mutable std::vector<uint32_t> count_;
};

using ProgressFn = std::function<void(const Framebuffer&, const RenderProgress&)>;

class GpuRenderer {
 public:
  // Opens the device, uploads the scene and a zeroed accumulation buffer.
  explicit GpuRenderer(const CompiledScene& scene, RenderOptions opts = {});
  ~GpuRenderer();
  GpuRenderer(const GpuRenderer&) = delete;
  GpuRenderer& operator=(const GpuRenderer&) = delete;

  // `cancel` may be raised from any thread (renderer.hpp:56 of the
  // reference); it is polled between device chunks.
  RenderProgress render(const ProgressFn& on_pass = {}, const std::atomic<bool>* cancel = nullptr);

  const Framebuffer& framebuffer() const { return fb_; }
  Image8 image() const;  // device tone map + sRGB
  RenderStatsSnapshot stats() const { return stats_; }
  int thread_count() const { return 1; }  // one device stream
  std::string device_name() const;
  uint64_t device_bytes() const;

 private:
  friend class Framebuffer;
  void add_stats(const rtg_stats& st, uint32_t passes);
  const CompiledScene& scene_;
  RenderOptions opts_;
  rtg_context* ctx_ = nullptr;
  rtg_scene* dscene_ = nullptr;
  rtg_accum* acc_ = nullptr;
  Framebuffer fb_;
  RenderStatsSnapshot stats_;
};

// Asynchronous render job owning its scene (renderer.hpp:79-103): the
// destructor cancels and joins.
class RenderSession {
 public:
  explicit RenderSession(CompiledScene scene, RenderOptions opts = {});
  ~RenderSession();
  RenderSession(const RenderSession&) = delete;
  RenderSession& operator=(const RenderSession&) = delete;

  void start(ProgressFn on_pass = {});
  void cancel() { cancel_.store(true); }
  RenderProgress wait();
  void restart(CompiledScene scene, ProgressFn on_pass = {});
  bool running() const { return running_.load(); }
  const GpuRenderer& renderer() const { return *renderer_; }
  const std::string& error() const { return error_; }

 private:
  CompiledScene scene_;
  RenderOptions opts_;
  std::unique_ptr<GpuRenderer> renderer_;
  std::atomic<bool> cancel_{false};
  std::atomic<bool> running_{false};
  RenderProgress last_;
  std::string error_;
  std::thread thread_;
};

// Throws std::runtime_error carrying rtg_last_error() when st != RTG_OK.
void rtg_check(rtg_status st, const rtg_context* ctx, const char* what);

}  // namespace rtm
