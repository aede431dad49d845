// serve_stream.cpp — `rt --serve`: the served render loop behind the rt
// stream protocol.
//
// The editor backend starts one `rt - --stream` process per job and cancels
// the previous job when a new one starts (web/backend/jobs.go:130-259); each
// process parses, compiles and uploads its scene and renders alone. Here one
// long-running process takes jobs as JSON lines on stdin:
//   {"job": 7, "scene": {...}, "width": 240, "spp": 2, "max_pass": 1, ...}
//   {"cancel": 7}
// (overrides as the rt flags: width, height, spp, max_depth, seed, exposure,
// integrator, precision, max_pass, no_bvh, no_light_sampling; "preview": true
// renders each pass whole in the next tick; "supersede": true cancels every
// earlier job, the backend's one-job-at-a-time policy). Every job streams the
// rt CLI's lines (apps/rt_cli.cpp:162-174) with its id added:
//   FRAME {"job":7,"pass":...,"bytes":N}\n<N bytes of PNG>
//   DONE {"job":7,...}\n   ERROR {"job":7,"message":...}\n
// Jobs share ticks of batched passes (rtg_render_batch); a compiled scene
// stays resident, keyed by everything but its camera and render settings, so
// a job that only moves the camera or changes the sample count skips the
// parse-compile-BVH-upload path. A job's image is bitwise the one `rt`
// renders for the same scene and flags.
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <iostream>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "image_io.hpp"
#include "json.hpp"
#include "options.hpp"
#include "renderer.hpp"
#include "scene.hpp"
#include "tick_tuner.hpp"

namespace rtm {

namespace {

using Clock = std::chrono::steady_clock;
double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}
// ERROR {"job":id,"message":...}
std::string error_line(int64_t job, const std::string& message) {
  return LineFields().add("job", job).add("message", message).line("ERROR");
}

// Lines from stdin, read on their own thread; an empty optional marks EOF.
class LineQueue {
 public:
  void push(std::string s, bool eof) {
    std::lock_guard<std::mutex> lk(mu_);
    lines_.push_back(std::move(s));
    eof_ = eof_ || eof;
    cv_.notify_one();
  }
  // Takes every queued line; blocks while empty, unless `wait` is false.
  std::vector<std::string> take(bool wait, bool& eof) {
    std::unique_lock<std::mutex> lk(mu_);
    if (wait) cv_.wait(lk, [&] { return !lines_.empty() || eof_; });
    std::vector<std::string> out(lines_.begin(), lines_.end());
    lines_.clear();
    eof = eof_;
    return out;
  }

 private:
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::string> lines_;
  bool eof_ = false;
};

struct Resident {
  rtg_scene* dev = nullptr;
  uint64_t bytes = 0, last_use = 0;
  int users = 0;
};

struct Job {
  int64_t id = 0;
  bool preview = false, cancelled = false, scene_cached = false;
  std::string key;
  RenderSettings settings;
  int32_t precision = RTG_PRECISION_F32;
  uint32_t max_pass = 32, spp_done = 0, k = 0, cursor = 0, pass = 0;
  rtg_scene* view = nullptr;
  rtg_accum* acc = nullptr;
  Clock::time_point start;
  double prep_ms = 0;
  uint64_t camera_samples = 0, rays = 0, shadow_rays = 0, invalid = 0;
};

uint32_t pass_samples(uint32_t done, uint32_t spp, uint32_t max_pass) {
  return std::min({spp - done, std::max(1u, done), max_pass});  // renderer.cpp:66-69
}

class Service {
 public:
  Service(int device, bool supersede, uint32_t tick_paths, uint64_t cache_bytes, bool quiet)
      : supersede_(supersede), tuner_(tick_paths, kTickTargetMs), cache_bytes_(cache_bytes), quiet_(quiet) {
    rtg_check(rtg_open(device, &ctx_), nullptr, "rtg_open");
  }
  ~Service() {
    for (Job& j : jobs_) release(j);
    for (auto& kv : cache_) rtg_scene_release(kv.second.dev);
    rtg_close(ctx_);
  }

  void command(const std::string& line) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) return;
    int64_t id = -1;
    try {
      const Json req = Json::parse(line);
      if (!req.is_object()) throw std::runtime_error("a command is a JSON object");
      if (const Json* c = req.get("cancel")) {
        id = (int64_t)c->as_number();
        for (Job& j : jobs_)
          if (j.id == id) j.cancelled = true;
        return;
      }
      const Json* jid = req.get("job");
      if (!jid || !jid->is_number()) throw std::runtime_error("command without \"job\" or \"cancel\"");
      id = (int64_t)jid->as_number();
      start(id, req);
    } catch (const std::exception& e) {
      emit(error_line(id, e.what()));
    }
  }

  bool idle() const { return jobs_.empty(); }

  // One tick: cancelled jobs end; then one batched pass over the live jobs.
  void tick() {
    for (Job& j : jobs_)
      if (j.cancelled) finish(j, true);
    jobs_.remove_if([](const Job& j) { return j.acc == nullptr; });
    if (jobs_.empty()) return;
    std::vector<rtg_job> batch;
    std::vector<Job*> owner;
    std::vector<uint32_t> pixels;
    bool previews = false;
    for (const Job& j : jobs_) previews = previews || j.preview;
    tuner_.begin(previews ? kTickTargetMs : kFinalsTickTargetMs);
    for (int kind = 0; kind < 2; ++kind)  // previews first, whole
      for (Job& j : jobs_) {
        if (j.preview != (kind == 0)) continue;
        const uint32_t n_pix = (uint32_t)j.settings.width * (uint32_t)j.settings.height;
        uint32_t take = n_pix - j.cursor;
        if (!j.preview) {
          const uint64_t allow = tuner_.final_paths(tune_key(j));
          if (!allow) continue;
          take = (uint32_t)std::min<uint64_t>(take, std::max<uint64_t>(1, allow / j.k));
          tuner_.charge_final(tune_key(j), (uint64_t)take * j.k);
        } else {
          tuner_.charge_preview(tune_key(j), (uint64_t)take * j.k);
        }
        rtg_job b{};
        b.scene = j.view;
        b.acc = j.acc;
        b.settings = device_settings(j);
        b.s0 = j.spp_done;
        b.s1 = j.spp_done + j.k;
        b.p0 = j.cursor;
        b.p1 = j.cursor + take;
        batch.push_back(b);
        owner.push_back(&j);
        pixels.push_back(take);
      }
    std::vector<rtg_stats> st(batch.size());
    const auto tick0 = Clock::now();
    const rtg_status s = rtg_render_batch(ctx_, batch.data(), (uint32_t)batch.size(), st.data());
    if (s == RTG_OK) {
      for (size_t q = 0; q < batch.size(); ++q)
        tuner_.add_job(tune_key(*owner[q]), (uint64_t)pixels[q] * owner[q]->k, st[q].rays + st[q].shadow_rays);
      tuner_.end(ms_between(tick0, Clock::now()));
    }
    if (s != RTG_OK) {
      const std::string msg = rtg_last_error(ctx_);
      for (Job* j : owner) {
        emit(error_line(j->id, msg));
        release(*j);
      }
      jobs_.remove_if([](const Job& j) { return j.acc == nullptr; });
      return;
    }
    for (size_t q = 0; q < batch.size(); ++q) {
      Job& j = *owner[q];
      j.camera_samples += st[q].camera_samples;
      j.rays += st[q].rays;
      j.shadow_rays += st[q].shadow_rays;
      j.invalid += st[q].invalid_samples;
      j.cursor += pixels[q];
      if (j.cursor < (uint32_t)j.settings.width * (uint32_t)j.settings.height) continue;
      j.cursor = 0;
      j.spp_done += j.k;
      ++j.pass;
      const bool last = j.spp_done >= (uint32_t)j.settings.spp;
      frame(j, last, false);
      if (last) finish(j, false);
      else j.k = pass_samples(j.spp_done, (uint32_t)j.settings.spp, j.max_pass);
    }
    jobs_.remove_if([](const Job& j) { return j.acc == nullptr; });
  }

 private:
  static rtg_render_settings device_settings(const Job& j) {
    rtg_render_settings r{};
    r.width = j.settings.width;
    r.height = j.settings.height;
    r.max_depth = j.settings.max_depth;
    r.rr_min_depth = j.settings.rr_min_depth;
    r.integrator = (int32_t)j.settings.integrator;
    r.light_sampling = j.settings.light_sampling ? 1 : 0;
    r.use_bvh = j.settings.use_bvh ? 1 : 0;
    r.precision = j.precision;
    r.seed = j.settings.seed;
    return r;
  }

  void start(int64_t id, const Json& req) {
    const auto t0 = Clock::now();
    const Json* scene = req.get("scene");
    if (!scene || !scene->is_object()) throw std::runtime_error("job without a \"scene\" object");
    SceneDesc desc = parse_scene(scene->dump(-1));
    RenderSettings& R = desc.render;
    Options o;
    read_job_fields(req, o);  // the rt flags' overrides, as JSON keys
    o.apply(R);
    if (const auto errs = validate_scene(desc); !errs.empty()) throw SceneError(errs.front());
    Job j;
    j.id = id;
    j.settings = R;
    j.precision = o.precision;
    if (o.max_pass) j.max_pass = (uint32_t)std::max(1, *o.max_pass);
    j.preview = o.preview;
    if (o.supersede_job.value_or(supersede_))
      for (Job& other : jobs_) other.cancelled = true;
    // The resident scene: everything but the camera and the render settings.
    Json key = scene_to_json(desc);
    Json stripped = Json::object();
    for (const auto& m : key.members())
      if (m.first != "camera" && m.first != "render" && m.first != "name" && m.first != "description")
        stripped.set(m.first, m.second);
    j.key = stripped.dump(-1);
    auto it = cache_.find(j.key);
    j.scene_cached = it != cache_.end();
    if (!j.scene_cached) {
      const CompiledScene compiled = compile_scene(desc);
      const rtg_scene_desc view = compiled.view();
      Resident r;
      rtg_check(rtg_scene_create(ctx_, &view, &r.dev), ctx_, "rtg_scene_create");
      r.bytes = rtg_scene_bytes(r.dev);
      it = cache_.emplace(j.key, r).first;
      evict();
    }
    it->second.users++;
    it->second.last_use = ++clock_;
    const rtg_camera cam = setup_camera(desc.camera, R.width, R.height);
    rtg_check(rtg_scene_with_camera(ctx_, it->second.dev, &cam, &j.view), ctx_, "rtg_scene_with_camera");
    rtg_check(rtg_accum_create(ctx_, R.width, R.height, &j.acc), ctx_, "rtg_accum_create");
    j.k = pass_samples(0, (uint32_t)R.spp, j.max_pass);
    j.prep_ms = ms_between(t0, Clock::now());
    j.start = t0;
    jobs_.push_back(j);
  }

  // Drops least recently used scenes nobody renders while over the budget.
  void evict() {
    uint64_t total = 0;
    for (auto& kv : cache_) total += kv.second.bytes;
    while (total > cache_bytes_) {
      auto victim = cache_.end();
      for (auto it = cache_.begin(); it != cache_.end(); ++it)
        if (it->second.users == 0 && (victim == cache_.end() || it->second.last_use < victim->second.last_use))
          victim = it;
      if (victim == cache_.end()) return;
      total -= victim->second.bytes;
      rtg_scene_release(victim->second.dev);
      const char* key = reinterpret_cast<const char*>(&victim->second);
      tuner_.forget(key);
      tuner_.forget(key + 1);
      cache_.erase(victim);
    }
  }

  void frame(Job& j, bool finished, bool cancelled) {
    Image8 img;
    img.width = j.settings.width;
    img.height = j.settings.height;
    img.rgb.resize((size_t)img.width * img.height * 3);
    const bool normals = j.settings.integrator == Integrator::Normals;  // shown unmapped, as Renderer::image()
    rtg_check(rtg_resolve(ctx_, j.acc, normals ? 0 : (int32_t)j.settings.tone_map, normals ? 1.0 : j.settings.exposure,
                          img.rgb.data()),
              ctx_, "rtg_resolve");
    const std::vector<uint8_t> png = encode_png(img);
    const std::string hdr = LineFields()
                                .add("job", j.id)
                                .add("pass", j.pass)
                                .add("spp_done", j.spp_done)
                                .add("spp_target", j.settings.spp)
                                .add("elapsed_ms", ms_between(j.start, Clock::now()))
                                .add("rays", j.rays + j.shadow_rays)
                                .add("width", j.settings.width)
                                .add("height", j.settings.height)
                                .add("finished", finished)
                                .add("cancelled", cancelled)
                                .add("bytes", (uint64_t)png.size())
                                .line("FRAME");
    emit(hdr, &png);
  }

  // Ends a job: cancelled jobs send their last whole-pass frame first.
  void finish(Job& j, bool cancelled) {
    if (cancelled) frame(j, false, true);
    const double ms = ms_between(j.start, Clock::now());
    const double mrays = ms > 0 ? (double)(j.rays + j.shadow_rays) / 1e3 / ms : 0;
    emit(LineFields()
             .add("job", j.id)
             .add("cancelled", cancelled)
             .add("spp_done", j.spp_done)
             .add("elapsed_ms", ms)
             .add("rays", j.rays)
             .add("shadow_rays", j.shadow_rays)
             .add("mrays_per_sec", mrays)
             .add("grays_per_sec", mrays / 1e3)
             .add("invalid_samples", j.invalid)
             .add("threads", 1)
             .add("scene_cached", j.scene_cached)
             .add("scene_prep_ms", j.prep_ms)
             .line("DONE"));
    if (!quiet_)
      std::cerr << "job " << j.id << (cancelled ? " cancelled" : " done") << " after " << (long)ms << " ms\n";
    release(j);
  }

  void release(Job& j) {
    if (j.acc) rtg_accum_release(j.acc);
    if (j.view) rtg_scene_release(j.view);
    j.acc = nullptr;
    j.view = nullptr;
    auto it = cache_.find(j.key);
    if (it != cache_.end() && it->second.users > 0) it->second.users--;
  }

  void emit(const std::string& line, const std::vector<uint8_t>* payload = nullptr) {
    std::fwrite(line.data(), 1, line.size(), stdout);
    if (payload) std::fwrite(payload->data(), 1, payload->size(), stdout);
    std::fflush(stdout);
  }

  rtg_context* ctx_ = nullptr;
  bool supersede_;
  TickTuner tuner_;
  // Tuner key of a job: its resident scene and kind (previews and finals
  // differ in depth and size, so their costs per path are kept apart).
  const void* tune_key(const Job& j) const {
    return reinterpret_cast<const char*>(&cache_.at(j.key)) + (j.preview ? 0 : 1);
  }
  uint64_t cache_bytes_;
  bool quiet_;
  std::list<Job> jobs_;
  std::unordered_map<std::string, Resident> cache_;
  uint64_t clock_ = 0;
};

}  // namespace

int serve_main(int device, bool supersede, uint32_t tick_paths, uint64_t cache_bytes, bool quiet) {
  std::unique_ptr<Service> svc;
  try {
    svc = std::make_unique<Service>(device, supersede, tick_paths, cache_bytes ? cache_bytes : (8ull << 30), quiet);
  } catch (const std::exception& e) {
    std::cout << error_line(-1, e.what()) << std::flush;
    return 2;
  }
  LineQueue queue;
  std::thread reader([&]() {
    std::string line;
    while (std::getline(std::cin, line)) queue.push(line, false);
    queue.push("", true);
  });
  bool eof = false;
  while (true) {
    bool end = false;
    for (const std::string& l : queue.take(svc->idle() && !eof, end)) svc->command(l);
    eof = eof || end;
    if (svc->idle()) {
      if (eof) break;
      continue;
    }
    try {
      svc->tick();
    } catch (const std::exception& e) {
      std::cout << error_line(-1, e.what()) << std::flush;
      reader.detach();
      return 2;
    }
  }
  reader.join();
  return 0;
}

}  // namespace rtm
