// server.cpp — the served render loop (see server.hpp).
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

#include "renderer.hpp"

namespace rtm {

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

void wait_until(Clock::time_point t0, double ms) {
  std::this_thread::sleep_until(t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double, std::milli>(ms)));
}

// FNV-1a over 64-bit words (each array element widened to 64 bits).
template <class T>
uint64_t fnv1a_words(uint64_t h, const std::vector<T>& v) {
  for (const T& x : v) {
    uint64_t w = 0;
    std::memcpy(&w, &x, sizeof(T));
    h = (h ^ w) * 1099511628211ull;
  }
  return h;
}

// Samples of the next progressive pass (rendering/renderer.cpp:66-69).
uint32_t pass_samples(uint32_t done, uint32_t spp, uint32_t max_pass) {
  return std::min({spp - done, std::max(1u, done), max_pass});
}

}  // namespace

// Hashes final images on a worker thread, so the GPU goes on with the next
// tick (or request) while the host reduces a read-back buffer.
class Server::Hasher {
 public:
  Hasher() : worker_([this] { loop(); }) {}
  ~Hasher() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    worker_.join();
  }
  void submit(uint64_t* dst, std::vector<double> sum, std::vector<uint32_t> count) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      work_.push_back(Item{dst, std::move(sum), std::move(count)});
    }
    cv_.notify_all();
  }
  // Returns once every submitted image is hashed.
  void drain() {
    std::unique_lock<std::mutex> lk(mu_);
    idle_.wait(lk, [&] { return work_.empty() && !busy_; });
  }

 private:
  struct Item {
    uint64_t* dst;
    std::vector<double> sum;
    std::vector<uint32_t> count;
  };
  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
      cv_.wait(lk, [&] { return stop_ || !work_.empty(); });
      if (work_.empty()) return;
      Item it = std::move(work_.front());
      work_.pop_front();
      busy_ = true;
      lk.unlock();
      *it.dst = fnv1a_words(fnv1a_words(1469598103934665603ull, it.sum), it.count);
      lk.lock();
      busy_ = false;
      idle_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, idle_;
  std::deque<Item> work_;
  bool stop_ = false, busy_ = false;
  std::thread worker_;
};

Server::Server(int device, uint64_t budget_bytes, uint32_t tick_paths)
    : budget_(budget_bytes ? budget_bytes : (64ull << 30)), tuner_(tick_paths, kTickTargetMs) {
  rtg_check(rtg_open(device, &ctx_), nullptr, "rtg_open");
}

Server::~Server() {
  for (auto& kv : pool_) rtg_accum_release(kv.second);
  for (SceneSlot& s : scenes_) {
    for (auto& kv : s.sized) rtg_scene_release(kv.second);
    rtg_scene_release(s.dev);
  }
  rtg_close(ctx_);
}

int Server::add_scene(const CompiledScene& c) {
  SceneSlot slot;
  slot.settings = c.settings;
  slot.camera = c.camera_desc;
  const rtg_scene_desc v = c.view();
  rtg_check(rtg_scene_create(ctx_, &v, &slot.dev), ctx_, "rtg_scene_create");
  scene_bytes_ += rtg_scene_bytes(slot.dev);
  scenes_.push_back(std::move(slot));
  return (int)scenes_.size() - 1;
}

rtg_scene* Server::scene_for(const ServeRequest& r) {
  SceneSlot& s = scenes_.at((size_t)r.scene);
  if (r.width == s.settings.width && r.height == s.settings.height) return s.dev;
  const auto key = std::make_pair((int)r.width, (int)r.height);
  auto it = s.sized.find(key);
  if (it != s.sized.end()) return it->second;
  const rtg_camera cam = setup_camera(s.camera, r.width, r.height);
  rtg_scene* v = nullptr;
  rtg_check(rtg_scene_with_camera(ctx_, s.dev, &cam, &v), ctx_, "rtg_scene_with_camera");
  s.sized[key] = v;
  return v;
}

rtg_render_settings Server::settings_of(const ServeRequest& r) const {
  const RenderSettings& R = scenes_.at((size_t)r.scene).settings;
  rtg_render_settings o{};
  o.width = r.width;
  o.height = r.height;
  o.max_depth = r.max_depth;
  o.rr_min_depth = R.rr_min_depth;
  o.integrator = RTG_INTEGRATOR_PATH;
  o.light_sampling = R.light_sampling ? 1 : 0;
  o.use_bvh = R.use_bvh ? 1 : 0;
  o.precision = RTG_PRECISION_F32;
  o.seed = r.seed;
  return o;
}

rtg_accum* Server::acquire(const ServeRequest& r) {
  auto it = pool_.find(std::make_pair((int)r.width, (int)r.height));
  if (it != pool_.end()) {
    rtg_accum* a = it->second;
    pool_.erase(it);
    pool_bytes_ -= accum_bytes(r);
    rtg_check(rtg_accum_clear(ctx_, a), ctx_, "rtg_accum_clear");
    return a;
  }
  rtg_accum* a = nullptr;
  rtg_check(rtg_accum_create(ctx_, r.width, r.height, &a), ctx_, "rtg_accum_create");
  return a;
}

void Server::recycle(rtg_accum* a, const ServeRequest& r) {
  if (accum_bytes(r) <= kPoolItemBytes && pool_bytes_ + accum_bytes(r) <= kPoolBytes) {
    pool_.emplace(std::make_pair((int)r.width, (int)r.height), a);
    pool_bytes_ += accum_bytes(r);
  } else {
    rtg_accum_release(a);
  }
}

void Server::frame(rtg_accum* acc, const ServeRequest& r) {
  const RenderSettings& R = scenes_.at((size_t)r.scene).settings;
  rgb_.resize((size_t)r.width * r.height * 3);
  rtg_check(rtg_resolve(ctx_, acc, (int32_t)R.tone_map, R.exposure, rgb_.data()), ctx_, "rtg_resolve");
}

uint64_t Server::fixed_bytes() const { return scene_bytes_ + rtg_context_bytes(ctx_); }

void Server::finish(rtg_accum* acc, const ServeRequest& r, uint64_t* hash) {
  const size_t n = (size_t)r.width * r.height;
  std::vector<double> sum(n * 3);
  std::vector<uint32_t> count(n);
  rtg_check(rtg_accum_read(ctx_, acc, sum.data(), count.data()), ctx_, "rtg_accum_read");
  hasher_->submit(hash, std::move(sum), std::move(count));
}

ServeStats Server::run(const std::vector<ServeRequest>& reqs, bool served, std::vector<ServeResult>& out) {
  const size_t n = reqs.size();
  for (size_t i = 0; i < n; ++i) {
    const ServeRequest& r = reqs[i];
    if (r.scene < 0 || (size_t)r.scene >= scenes_.size()) throw std::invalid_argument("request: no such scene");
    if (r.width < 1 || r.height < 1 || r.width > 16384 || r.height > 16384 || r.spp < 1 || r.max_pass < 1)
      throw std::invalid_argument("request: bad size, spp or max_pass");
    if (r.max_depth < 1 || r.max_depth > 64) throw std::invalid_argument("request: max_depth must be in [1, 64]");
    if (i && r.arrival_ms < reqs[i - 1].arrival_ms) throw std::invalid_argument("requests must be sorted by arrival");
  }
  out.assign(n, ServeResult{});
  for (const ServeRequest& r : reqs) scene_for(r);  // camera set-ups before the clock starts
  ServeStats st;
  uint64_t used = 0;  // accumulation buffers of admitted requests
  Hasher hasher;
  hasher_ = &hasher;
  struct Detach {
    Server* s;
    ~Detach() { s->hasher_ = nullptr; }
  } detach{this};
  auto gpu_call = [&](auto&& f) {
    const auto c0 = Clock::now();
    f();
    st.gpu_ms += ms_since(c0);
  };
  const auto t0 = Clock::now();

  auto complete_pass = [&](size_t i, rtg_accum* acc, uint32_t k) {
    ServeResult& o = out[i];
    frame(acc, reqs[i]);
    o.spp_done += k;
    if (++o.frames == 1) o.first_frame_ms = ms_since(t0);
  };

  if (!served) {
    // The plain loop: first come first served, each request alone.
    for (size_t i = 0; i < n; ++i) {
      const ServeRequest& r = reqs[i];
      ServeResult& o = out[i];
      o.arrival_ms = r.arrival_ms;
      wait_until(t0, r.arrival_ms);
      o.start_ms = ms_since(t0);
      rtg_accum* acc = acquire(r);
      struct Back {
        Server* s;
        rtg_accum* a;
        const ServeRequest& r;
        ~Back() { s->recycle(a, r); }
      } back{this, acc, r};
      st.peak_accum_bytes = std::max(st.peak_accum_bytes, accum_bytes(r));
      st.peak_footprint_bytes = std::max(st.peak_footprint_bytes, fixed_bytes() + accum_bytes(r));
      st.max_active = 1;
      rtg_scene* sc = scene_for(r);
      const rtg_render_settings cfg = settings_of(r);
      for (uint32_t done = 0; done < (uint32_t)r.spp;) {
        const uint32_t k = pass_samples(done, (uint32_t)r.spp, (uint32_t)r.max_pass);
        rtg_stats s{};
        gpu_call([&] {
          rtg_check(rtg_render_pass(ctx_, sc, &cfg, done, done + k, acc, nullptr, &s), ctx_, "rtg_render_pass");
        });
        o.rays += s.rays;
        o.shadow_rays += s.shadow_rays;
        ++st.ticks;
        done += k;
        complete_pass(i, acc, k);
      }
      o.done_ms = ms_since(t0);
      finish(acc, r, &o.image_hash);
      st.rays += o.rays + o.shadow_rays;
    }
    st.wall_ms = ms_since(t0);
    hasher.drain();
    st.fixed_bytes = fixed_bytes();
    return st;
  }

  // The served loop.
  struct Live {
    size_t i;
    rtg_accum* acc;
    uint32_t k, cursor;  // current pass: samples, pixels done
  };
  std::vector<Live> live;
  std::vector<rtg_job> jobs;
  std::vector<uint32_t> job_pixels;
  std::vector<size_t> job_live;
  std::vector<rtg_stats> js;
  size_t next = 0;
  try {
  while (next < n || !live.empty()) {
    const double now = ms_since(t0);
    while (next < n && reqs[next].arrival_ms <= now) {
      const ServeRequest& r = reqs[next];
      // Admission by HBM footprint: resident scenes, path state and queues,
      // the admitted requests' accumulation buffers.
      const uint64_t footprint = fixed_bytes() + used + accum_bytes(r);
      if (!live.empty() && footprint > budget_) break;
      rtg_accum* acc = acquire(r);
      used += accum_bytes(r);
      st.peak_accum_bytes = std::max(st.peak_accum_bytes, used);
      st.peak_footprint_bytes = std::max(st.peak_footprint_bytes, footprint);
      out[next].arrival_ms = r.arrival_ms;
      out[next].start_ms = now;
      live.push_back(Live{next, acc, pass_samples(0, (uint32_t)r.spp, (uint32_t)r.max_pass), 0});
      ++next;
    }
    st.max_active = std::max(st.max_active, (uint32_t)live.size());
    if (live.empty()) {
      wait_until(t0, reqs[next].arrival_ms);
      continue;
    }
    // One tick: every preview's whole pass, then finals in admission order
    // within the tick's budget (whole pixels, at least one).
    jobs.clear();
    job_pixels.clear();
    job_live.clear();
    bool previews = false;
    for (const Live& L : live) previews = previews || reqs[L.i].final_kind == 0;
    tuner_.begin(previews ? kTickTargetMs : kFinalsTickTargetMs);
    for (int pass_kind = 0; pass_kind < 2; ++pass_kind)
      for (size_t li = 0; li < live.size(); ++li) {
        const Live& L = live[li];
        const ServeRequest& r = reqs[L.i];
        if (r.final_kind != pass_kind) continue;
        const uint32_t n_pix = (uint32_t)r.width * (uint32_t)r.height;
        uint32_t take = n_pix - L.cursor;
        if (r.final_kind) {
          const uint64_t allow = tuner_.final_paths(tune_key(r));
          if (!allow) continue;
          take = (uint32_t)std::min<uint64_t>(take, std::max<uint64_t>(1, allow / L.k));
          tuner_.charge_final(tune_key(r), (uint64_t)take * L.k);
        } else {
          tuner_.charge_preview(tune_key(r), (uint64_t)take * L.k);
        }
        rtg_job j{};
        j.scene = scene_for(r);
        j.acc = L.acc;
        j.settings = settings_of(r);
        j.s0 = out[L.i].spp_done;
        j.s1 = j.s0 + L.k;
        j.p0 = L.cursor;
        j.p1 = L.cursor + take;
        jobs.push_back(j);
        job_pixels.push_back(take);
        job_live.push_back(li);
      }
    js.assign(jobs.size(), rtg_stats{});
    const auto tick0 = std::chrono::steady_clock::now();
    rtg_check(rtg_render_batch(ctx_, jobs.data(), (uint32_t)jobs.size(), js.data()), ctx_, "rtg_render_batch");
    const double tick_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tick0).count();
    st.gpu_ms += tick_ms;
    ++st.ticks;
    for (size_t q = 0; q < jobs.size(); ++q) {
      const ServeRequest& rq = reqs[live[job_live[q]].i];
      tuner_.add_job(tune_key(rq), (uint64_t)job_pixels[q] * (jobs[q].s1 - jobs[q].s0),
                     js[q].rays + js[q].shadow_rays);
    }
    tuner_.end(tick_ms);
    for (size_t q = 0; q < jobs.size(); ++q) {
      Live& L = live[job_live[q]];
      const ServeRequest& r = reqs[L.i];
      ServeResult& o = out[L.i];
      o.rays += js[q].rays;
      o.shadow_rays += js[q].shadow_rays;
      L.cursor += job_pixels[q];
      if (L.cursor < (uint32_t)r.width * (uint32_t)r.height) continue;
      L.cursor = 0;
      complete_pass(L.i, L.acc, L.k);
      if (o.spp_done < (uint32_t)r.spp) {
        L.k = pass_samples(o.spp_done, (uint32_t)r.spp, (uint32_t)r.max_pass);
        continue;
      }
      o.done_ms = ms_since(t0);
      finish(L.acc, r, &o.image_hash);
      recycle(L.acc, r);
      L.acc = nullptr;
      used -= accum_bytes(r);
      st.rays += o.rays + o.shadow_rays;
    }
    live.erase(std::remove_if(live.begin(), live.end(), [](const Live& L) { return L.acc == nullptr; }), live.end());
  }
  } catch (...) {
    for (Live& L : live)
      if (L.acc) rtg_accum_release(L.acc);
    throw;
  }
  st.wall_ms = ms_since(t0);
  hasher.drain();
  st.fixed_bytes = fixed_bytes();
  return st;
}

}  // namespace rtm
