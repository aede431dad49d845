// renderer.cpp — GpuRenderer / RenderSession over the rtg C ABI.
#include "renderer.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <thread>

namespace rtm {

void rtg_check(rtg_status st, const rtg_context* ctx, const char* what) {
  if (st != RTG_OK) throw std::runtime_error(std::string(what) + ": " + rtg_last_error(ctx));
}

// ---------------------------------------------------------------- Framebuffer
void Framebuffer::fetch() const {
  if (valid_) return;
  const size_t n = (size_t)w_ * (size_t)h_;
  sum_.resize(n * 3);
  count_.resize(n);
  rtg_check(rtg_accum_read(owner_->ctx_, owner_->acc_, sum_.data(), count_.data()), owner_->ctx_, "rtg_accum_read");
  valid_ = true;
}
void Framebuffer::mean(int x, int y, double out[3]) const {
  fetch();
  const size_t i = (size_t)y * (size_t)w_ + (size_t)x;
  const uint32_t c = count_[i];
  for (int k = 0; k < 3; ++k) out[k] = c ? sum_[3 * i + (size_t)k] * (1.0 / (double)c) : 0.0;
}
uint32_t Framebuffer::samples(int x, int y) const {
  fetch();
  return count_[(size_t)y * (size_t)w_ + (size_t)x];
}
uint32_t Framebuffer::min_samples() const {
  fetch();
  return count_.empty() ? 0 : *std::min_element(count_.begin(), count_.end());
}
const std::vector<double>& Framebuffer::sums() const {
  fetch();
  return sum_;
}
const std::vector<uint32_t>& Framebuffer::counts() const {
  fetch();
  return count_;
}

// ---------------------------------------------------------------- GpuRenderer
GpuRenderer::GpuRenderer(const CompiledScene& scene, RenderOptions opts) : scene_(scene), opts_(opts) {
  rtg_check(rtg_open(opts_.device, &ctx_), nullptr, "rtg_open");
  const rtg_scene_desc d = scene_.view();
  rtg_status st = rtg_scene_create(ctx_, &d, &dscene_);
  if (st == RTG_OK) st = rtg_accum_create(ctx_, scene_.settings.width, scene_.settings.height, &acc_);
  if (st != RTG_OK) {
    const std::string msg = rtg_last_error(ctx_);
    rtg_scene_release(dscene_);
    rtg_close(ctx_);
    throw std::runtime_error("GPU setup: " + msg);
  }
  fb_.owner_ = this;
  fb_.w_ = scene_.settings.width;
  fb_.h_ = scene_.settings.height;
}

GpuRenderer::~GpuRenderer() {
  rtg_accum_release(acc_);
  rtg_scene_release(dscene_);
  rtg_close(ctx_);
}

std::string GpuRenderer::device_name() const {
  char name[128] = {0};
  int cus = 0;
  uint64_t mem = 0;
  rtg_device_info(ctx_, name, sizeof name, &cus, &mem);
  return std::string(name) + " (" + std::to_string(cus) + " CUs)";
}

uint64_t GpuRenderer::device_bytes() const { return rtg_scene_bytes(dscene_); }

void GpuRenderer::add_stats(const rtg_stats& st, uint32_t passes) {
  stats_.camera_samples += st.camera_samples;
  stats_.rays += st.rays;
  stats_.shadow_rays += st.shadow_rays;
  stats_.rr_terminated += st.rr_terminated;
  stats_.invalid_samples += st.invalid_samples;
  stats_.node_visits += st.node_visits;
  stats_.prim_tests += st.prim_tests;
  stats_.device_ms += st.elapsed_ms;
  stats_.passes += passes;
  fb_.valid_ = false;
}

namespace {
// The device driver polls a plain int flag between chunks (rtg.h); a
// caller's std::atomic<bool> request is copied into one by a watcher thread
// for the duration of a render.
class CancelMirror {
 public:
  explicit CancelMirror(const std::atomic<bool>* request) : request_(request) {
    if (!request_) return;
    watcher_ = std::thread([this] {
      while (!done_.load()) {
        if (request_->load()) {
          raised_.store(1);
          return;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
      }
    });
  }
  ~CancelMirror() {
    done_.store(true);
    if (watcher_.joinable()) watcher_.join();
  }
  const volatile int32_t* flag() const {
    return request_ ? reinterpret_cast<const volatile int32_t*>(&raised_) : nullptr;
  }

 private:
  const std::atomic<bool>* request_;
  std::atomic<bool> done_{false};
  std::atomic<int32_t> raised_{0};
  std::thread watcher_;
};
}  // namespace

RenderProgress GpuRenderer::render(const ProgressFn& on_pass, const std::atomic<bool>* cancel) {
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t target = (uint32_t)scene_.settings.spp;
  const rtg_render_settings cfg = scene_.device_settings(opts_.precision);
  const CancelMirror mirror(cancel);
  const volatile int32_t* flag = mirror.flag();
  RenderProgress prog;
  prog.spp_target = scene_.settings.spp;
  uint32_t done = 0;
  if (!on_pass) {
    // Nothing reads the framebuffer between passes: the whole schedule in
    // one call, whose passes run side by side on the device (bitwise the
    // same result as the loop below).
    std::vector<uint32_t> bounds{0};
    while (bounds.back() < target) {
      uint32_t k = target - bounds.back();
      if (opts_.progressive)
        k = std::min({k, std::max<uint32_t>(1, bounds.back()), (uint32_t)std::max(1, opts_.max_pass_spp)});
      bounds.push_back(bounds.back() + k);
    }
    rtg_stats st{};
    uint32_t passes = 0;
    const rtg_status rc = rtg_render_passes(ctx_, dscene_, &cfg, bounds.data(), (uint32_t)bounds.size() - 1, acc_,
                                            flag, &passes, &st);
    if (rc != RTG_OK && rc != RTG_ERR_CANCELLED) rtg_check(rc, ctx_, "rtg_render_passes");
    add_stats(st, passes);
    prog.pass = (int)passes;
    prog.elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats_.elapsed_ms = prog.elapsed_ms;
    prog.rays = stats_.rays + stats_.shadow_rays;
    if (rc == RTG_ERR_CANCELLED || (cancel && cancel->load())) {
      prog.cancelled = true;
      prog.spp_done = rtg_accum_min_samples(ctx_, acc_);
      return prog;
    }
    prog.spp_done = target;
    prog.finished = true;
    return prog;
  }
  while (done < target) {
    uint32_t k = target - done;
    if (opts_.progressive)
      k = std::min({k, std::max<uint32_t>(1, done), (uint32_t)std::max(1, opts_.max_pass_spp)});
    rtg_stats st{};
    const rtg_status rc = rtg_render_pass(ctx_, dscene_, &cfg, done, done + k, acc_, flag, &st);
    if (rc != RTG_OK && rc != RTG_ERR_CANCELLED) rtg_check(rc, ctx_, "rtg_render_pass");
    add_stats(st, 1);
    prog.pass++;
    prog.elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats_.elapsed_ms = prog.elapsed_ms;
    prog.rays = stats_.rays + stats_.shadow_rays;
    if (rc == RTG_ERR_CANCELLED || (cancel && cancel->load())) {
      prog.cancelled = true;
      prog.spp_done = rtg_accum_min_samples(ctx_, acc_);
      if (on_pass) on_pass(fb_, prog);
      return prog;
    }
    done += k;
    prog.spp_done = done;
    prog.finished = done >= target;
    if (on_pass) on_pass(fb_, prog);
  }
  return prog;
}

Image8 GpuRenderer::image() const {
  Image8 img;
  img.width = scene_.settings.width;
  img.height = scene_.settings.height;
  img.rgb.resize((size_t)img.width * (size_t)img.height * 3);
  // Renderer::image(): the normals view is shown unmapped.
  const bool normals = scene_.settings.integrator == Integrator::Normals;
  rtg_check(rtg_resolve(ctx_, acc_, normals ? 0 : (int32_t)scene_.settings.tone_map,
                        normals ? 1.0 : scene_.settings.exposure, img.rgb.data()),
            ctx_, "rtg_resolve");
  return img;
}

// ---------------------------------------------------------------- RenderSession
RenderSession::RenderSession(CompiledScene scene, RenderOptions opts) : scene_(std::move(scene)), opts_(opts) {
  renderer_ = std::make_unique<GpuRenderer>(scene_, opts_);
}

RenderSession::~RenderSession() {
  cancel();
  wait();
}

void RenderSession::start(ProgressFn on_pass) {
  wait();  // one job at a time
  cancel_.store(false);
  error_.clear();
  running_.store(true);
  renderer_.reset();
  renderer_ = std::make_unique<GpuRenderer>(scene_, opts_);
  thread_ = std::thread([this, cb = std::move(on_pass)] {
    try {
      last_ = renderer_->render(cb, &cancel_);
    } catch (const std::exception& e) {
      error_ = e.what();
      last_.cancelled = true;
    }
    running_.store(false);
  });
}

RenderProgress RenderSession::wait() {
  if (thread_.joinable()) thread_.join();
  return last_;
}

void RenderSession::restart(CompiledScene scene, ProgressFn on_pass) {
  cancel();
  wait();
  scene_ = std::move(scene);
  start(std::move(on_pass));
}

}  // namespace rtm
