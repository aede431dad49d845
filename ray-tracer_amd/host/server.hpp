// server.hpp — the served render loop: many concurrent render requests over
// scenes resident on one GPU, scheduled in ticks of batched passes.
//
// The reference renders one scene per process and request: the editor
// backend starts `rt - --stream` per job (web/backend/jobs.go:130-259) and
// each job renders alone (rendering/renderer.cpp:58-105). Here scenes stay
// resident and every tick renders one batched pass (rtg_render_batch) over
// all admitted requests:
//   * previews (small, few samples) render all their samples in the next
//     tick, so they never wait behind a large render;
//   * finals advance their progressive pass schedule (1, 1, 2, 4, ... capped
//     at max_pass, renderer.cpp:66-69) in whole-pixel ranges sized by the
//     tick tuner (tick_tuner.hpp: each tick aimed at kTickTargetMs from the
//     scenes' measured cost per path) or by a fixed `tick_paths` paths per
//     tick; a frame is resolved when a pass ends;
//   * requests are admitted first-come first-served while the HBM footprint
//     fits the memory budget: the resident scenes (BVH, primitives,
//     shading data), the context's path state and queues, and the
//     accumulation buffers of the admitted requests;
//   * a tick with previews is aimed at kTickTargetMs, a tick of finals only
//     at kFinalsTickTargetMs (tick_tuner.hpp); final images are read back
//     when done and hashed on a worker thread while the GPU goes on.
// A request's image is bitwise the image it gets when rendered alone: its
// samples are keyed on (seed, pixel, sample, bounce) and every pixel adds its
// samples in sample order, however passes are sliced and batched.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "rtg.h"
#include "scene.hpp"
#include "tick_tuner.hpp"

namespace rtm {

struct ServeRequest {
  int32_t scene = 0;  // resident scene index
  int32_t width = 128, height = 128, spp = 1, max_depth = 4;
  int32_t max_pass = 32;  // samples per progressive pass at most
  int32_t final_kind = 0; // 0 preview, 1 final
  uint64_t seed = 1;
  double arrival_ms = 0;  // from the start of the run
};

struct ServeResult {
  double arrival_ms = 0, start_ms = 0, first_frame_ms = 0, done_ms = 0;
  uint64_t rays = 0, shadow_rays = 0;
  uint32_t frames = 0, spp_done = 0;
  uint64_t image_hash = 0;  // FNV-1a over the final sums (64-bit words) then the counts
};

struct ServeStats {
  double wall_ms = 0;
  uint64_t rays = 0;  // closest-hit + shadow rays
  uint32_t ticks = 0, max_active = 0;
  uint64_t peak_accum_bytes = 0;
  double gpu_ms = 0;                  // wall time inside render calls
  uint64_t fixed_bytes = 0;           // resident scenes + the context's path state at the end
  uint64_t peak_footprint_bytes = 0;  // the largest footprint admission saw
};

class Server {
 public:
  // budget_bytes: HBM footprint (resident scenes, path state, accumulation
  // buffers of admitted requests; 0: 64 GiB);
  // tick_paths: paths of final renders per tick (0: tuned, each tick aimed
  // at kTickTargetMs).
  Server(int device, uint64_t budget_bytes, uint32_t tick_paths);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Uploads a compiled scene; returns its index.
  int add_scene(const CompiledScene& scene);
  // Replays requests (sorted by arrival) in real time: served = true for the
  // tick scheduler, false for the plain loop (one request at a time, each
  // through the single-render path). out[i] belongs to reqs[i].
  ServeStats run(const std::vector<ServeRequest>& reqs, bool served, std::vector<ServeResult>& out);
  static uint64_t accum_bytes(const ServeRequest& r) { return (uint64_t)r.width * r.height * 28; }
  void set_budget(uint64_t budget_bytes) { budget_ = budget_bytes ? budget_bytes : (64ull << 30); }
  // Resident scenes and the context's path state and queues now.
  uint64_t fixed_bytes() const;

 private:
  struct SceneSlot {
    rtg_scene* dev = nullptr;
    RenderSettings settings;
    CameraDesc camera;
    std::map<std::pair<int, int>, rtg_scene*> sized;  // the scene with its camera set up per image size
  };
  rtg_scene* scene_for(const ServeRequest& r);
  rtg_render_settings settings_of(const ServeRequest& r) const;
  // Accumulation buffers of small requests are kept for reuse (cleared on
  // the device) instead of being allocated per request.
  rtg_accum* acquire(const ServeRequest& r);
  void recycle(rtg_accum* acc, const ServeRequest& r);
  // Reads the final buffer back; its hash is computed on the worker thread
  // and written to *hash before run() returns.
  void finish(rtg_accum* acc, const ServeRequest& r, uint64_t* hash);
  void frame(rtg_accum* acc, const ServeRequest& r);       // tone map + readback of one frame

  rtg_context* ctx_ = nullptr;
  std::vector<SceneSlot> scenes_;
  uint64_t scene_bytes_ = 0;
  class Hasher;
  Hasher* hasher_ = nullptr;  // during run()
  uint64_t budget_;
  TickTuner tuner_;
  // Tuner key of a request: its scene and kind (previews and finals differ
  // in depth and size, so their costs per path are kept apart).
  static const void* tune_key(const ServeRequest& r) {
    return reinterpret_cast<const void*>(static_cast<uintptr_t>(r.scene) * 2 + (r.final_kind ? 1 : 0) + 1);
  }
  std::vector<uint8_t> rgb_;
  static constexpr uint64_t kPoolItemBytes = 256ull * 256 * 28, kPoolBytes = 1ull << 30;
  std::multimap<std::pair<int, int>, rtg_accum*> pool_;
  uint64_t pool_bytes_ = 0;
};

}  // namespace rtm
