// rt_host_api.cpp — the C ABI of include/rt_host.h.
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "renderer.hpp"
#include "rt_host.h"
#include "server.hpp"

using namespace rtm;

struct rth_scene {
  SceneDesc desc;
  CompiledScene compiled;
  rtg_scene_desc view{};
};

namespace {

thread_local std::string g_last_error;
std::string g_message;       // rt_message / rt_validate storage
std::mutex g_mu;             // guards the session below
std::unique_ptr<RenderSession> g_session;

void apply(SceneDesc& d, const rth_overrides* ov) {
  if (!ov) return;
  RenderSettings& R = d.render;
  if (ov->width > 0) R.width = ov->width;
  if (ov->height > 0) R.height = ov->height;
  if (ov->spp > 0) R.spp = ov->spp;
  if (ov->max_depth > 0) R.max_depth = ov->max_depth;
  if (ov->rr_min_depth >= 0) R.rr_min_depth = ov->rr_min_depth;
  if (ov->integrator >= 0) R.integrator = (Integrator)ov->integrator;
  if (ov->light_sampling >= 0) R.light_sampling = ov->light_sampling != 0;
  if (ov->use_bvh >= 0) R.use_bvh = ov->use_bvh != 0;
  if (ov->seed >= 0) R.seed = (uint64_t)ov->seed;
  if (ov->exposure > 0) R.exposure = ov->exposure;
}

SceneDesc load_checked(const char* json) {
  SceneDesc d = parse_scene(json ? json : "");
  if (const auto errs = validate_scene(d); !errs.empty()) throw SceneError(errs.front());
  return d;
}

}  // namespace

extern "C" {

const char* rt_validate(const char* json) {
  g_message.clear();
  try {
    load_checked(json);
  } catch (const std::exception& e) {
    g_message = e.what();
  }
  return g_message.c_str();
}

const char* rt_message(void) { return g_message.c_str(); }

int rt_threads(void) { return 1; }

int rt_running(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_session && g_session->running() ? 1 : 0;
}

void rt_cancel(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_session) g_session->cancel();
}

int rt_start(int32_t job, const char* json, int32_t max_pass, rt_frame_fn on_frame, rt_done_fn on_done, void* user) {
  g_message.clear();
  if (rt_running()) {
    g_message = "a render is still running";
    return 1;
  }
  try {
    const SceneDesc desc = load_checked(json);
    CompiledScene scene = compile_scene(desc);
    RenderOptions opts;
    if (max_pass > 0) opts.max_pass_spp = max_pass;
    std::lock_guard<std::mutex> lk(g_mu);
    g_session.reset();  // not running: joins immediately
    g_session = std::make_unique<RenderSession>(std::move(scene), opts);
    RenderSession* s = g_session.get();
    s->start([job, on_frame, on_done, user, s](const Framebuffer&, const RenderProgress& p) {
      const GpuRenderer& r = s->renderer();
      if (on_frame) {
        const Image8 img = r.image();
        std::vector<uint8_t> rgba((size_t)img.width * (size_t)img.height * 4);
        for (size_t i = 0; i < (size_t)img.width * (size_t)img.height; ++i) {
          rgba[4 * i] = img.rgb[3 * i];
          rgba[4 * i + 1] = img.rgb[3 * i + 1];
          rgba[4 * i + 2] = img.rgb[3 * i + 2];
          rgba[4 * i + 3] = 255;
        }
        rt_frame_info fi{p.pass, p.spp_done, p.spp_target, img.width, img.height, p.finished ? 1 : 0,
                         p.cancelled ? 1 : 0, 0, p.elapsed_ms, p.rays};
        on_frame(user, job, rgba.data(), &fi);
      }
      if (on_done && (p.finished || p.cancelled)) {
        const RenderStatsSnapshot st = r.stats();
        rt_done_info di{p.cancelled ? 1 : 0, p.spp_done, st.elapsed_ms, st.rays, st.shadow_rays,
                        st.invalid_samples, st.mrays_per_sec(), r.thread_count(), 0};
        on_done(user, job, &di);
      }
    });
  } catch (const std::exception& e) {
    g_message = e.what();
    return 1;
  }
  return 0;
}

void rt_wait(void) {
  RenderSession* s;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    s = g_session.get();
  }
  if (s) s->wait();
}

// ---------------------------------------------------------------------------
const char* rth_last_error(void) { return g_last_error.c_str(); }

int rth_scene_load(const char* json, const rth_overrides* ov, rth_scene** out) {
  try {
    auto s = std::make_unique<rth_scene>();
    s->desc = parse_scene(json ? json : "");
    apply(s->desc, ov);
    if (const auto errs = validate_scene(s->desc); !errs.empty()) throw SceneError(errs.front());
    s->compiled = compile_scene(s->desc);
    s->view = s->compiled.view();
    *out = s.release();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

void rth_scene_free(rth_scene* s) { delete s; }

const rtg_scene_desc* rth_scene_desc(const rth_scene* s) { return &s->view; }

void rth_scene_settings(const rth_scene* s, int32_t precision, rtg_render_settings* out) {
  *out = s->compiled.device_settings(precision);
}

void rth_scene_info(const rth_scene* s, uint64_t out[10]) {
  const SceneInfo& i = s->compiled.info;
  const uint64_t v[10] = {i.triangles, i.spheres,       i.planes,        i.degenerate_skipped,
                          i.lights,    i.bvh.nodes,     i.bvh.leaves,    i.bvh.max_depth,
                          (uint64_t)s->compiled.settings.spp, (uint64_t)s->compiled.settings.tone_map};
  std::memcpy(out, v, sizeof v);
}

double rth_scene_bvh_build_ms(const rth_scene* s) { return s->compiled.info.bvh.build_ms; }

int64_t rth_dump_scene(const char* json, char* out, uint64_t cap) {
  try {
    const std::string txt = scene_to_json(parse_scene(json ? json : "")).dump(-1);
    if (txt.size() + 1 > cap) return -(int64_t)(txt.size() + 1);
    std::memcpy(out, txt.c_str(), txt.size() + 1);
    return (int64_t)txt.size();
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 0;
  }
}

int rth_render(rth_scene* s, int32_t device, int32_t precision, int32_t progressive, int32_t max_pass, double* sum,
               uint32_t* count, uint8_t* rgb, rth_render_stats* stats) {
  try {
    RenderOptions opts;
    opts.device = device;
    opts.precision = precision;
    opts.progressive = progressive != 0;
    if (max_pass > 0) opts.max_pass_spp = max_pass;
    GpuRenderer r(s->compiled, opts);
    const RenderProgress p = r.render();
    const Framebuffer& fb = r.framebuffer();
    if (sum) std::memcpy(sum, fb.sums().data(), fb.sums().size() * sizeof(double));
    if (count) std::memcpy(count, fb.counts().data(), fb.counts().size() * sizeof(uint32_t));
    if (rgb) {
      const Image8 img = r.image();
      std::memcpy(rgb, img.rgb.data(), img.rgb.size());
    }
    if (stats) {
      const RenderStatsSnapshot st = r.stats();
      *stats = rth_render_stats{st.camera_samples, st.rays,      st.shadow_rays, st.rr_terminated,
                                st.invalid_samples, st.node_visits, st.prim_tests, st.passes,
                                st.elapsed_ms,      st.device_ms,   p.spp_done,    p.cancelled ? 1 : 0};
    }
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

// ---------------------------------------------------------------------------
struct rts_server {
  Server server;
  rts_server(int device, uint64_t budget, uint32_t tick) : server(device, budget, tick) {}
};

int rts_create(int32_t device, uint64_t budget_bytes, uint32_t tick_paths, rts_server** out) {
  if (!out) {
    g_last_error = "rts_create: out is NULL";
    return 1;
  }
  try {
    *out = new rts_server(device, budget_bytes, tick_paths);
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

int rts_add_scene(rts_server* s, const rth_scene* scene, int32_t* index) {
  if (!s || !scene) {
    g_last_error = "rts_add_scene: NULL argument";
    return 1;
  }
  try {
    const int i = s->server.add_scene(scene->compiled);
    if (index) *index = i;
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

int rts_run(rts_server* s, const rts_request* reqs, uint32_t n, int32_t served, rts_result* results,
            rts_stats* stats) {
  if (!s || (n && (!reqs || !results))) {
    g_last_error = "rts_run: NULL argument";
    return 1;
  }
  try {
    std::vector<ServeRequest> rq(n);
    for (uint32_t i = 0; i < n; ++i) {
      const rts_request& q = reqs[i];
      rq[i] = ServeRequest{q.scene, q.width, q.height, q.spp, q.max_depth, q.max_pass, q.final_kind, q.seed,
                           q.arrival_ms};
    }
    std::vector<ServeResult> res;
    const ServeStats st = s->server.run(rq, served != 0, res);
    for (uint32_t i = 0; i < n; ++i) {
      const ServeResult& r = res[i];
      results[i] = rts_result{r.arrival_ms, r.start_ms, r.first_frame_ms, r.done_ms, r.rays,
                              r.shadow_rays, r.frames,   r.spp_done,       r.image_hash};
    }
    if (stats)
      *stats = rts_stats{st.wall_ms,  st.rays,           st.ticks,      st.max_active, st.peak_accum_bytes,
                         st.gpu_ms,   st.fixed_bytes,    st.peak_footprint_bytes};
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

int rts_set_budget(rts_server* s, uint64_t budget_bytes) {
  if (!s) {
    g_last_error = "rts_set_budget: NULL server";
    return 1;
  }
  s->server.set_budget(budget_bytes);
  return 0;
}

void rts_free(rts_server* s) { delete s; }

}  // extern "C"
