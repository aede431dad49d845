// session_check — the reference's renderer-level tests, run against
// GpuRenderer / RenderSession on HIP device 0 (tests/test_render.cpp of the
// reference: progressive sample counts :211-221, cancellation and restart
// :223-255, RAII join :257-266), plus a cancel raised from another thread
// while the whole pass schedule runs in one device call.
//
//   session_check <assets dir>     exit 0 when every check holds
//
// Each check prints one line "ok <name> ..." or "FAIL <name>: ...".
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "renderer.hpp"

using namespace rtm;
using Clock = std::chrono::steady_clock;

namespace {

int g_failures = 0;
std::string g_assets;

void check(bool ok, const char* name, const std::string& detail) {
  std::printf("%s %s%s%s\n", ok ? "ok" : "FAIL", name, ok ? " " : ": ", detail.c_str());
  std::fflush(stdout);
  if (!ok) ++g_failures;
}

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

SceneDesc asset(const char* name, int w, int h, int spp) {
  SceneDesc d = load_scene(g_assets + "/" + name);
  d.render.width = w;
  d.render.height = h;
  d.render.spp = spp;
  return d;
}

bool all_means_finite(const Framebuffer& fb) {
  for (int y = 0; y < fb.height(); ++y)
    for (int x = 0; x < fb.width(); ++x) {
      double m[3];
      fb.mean(x, y, m);
      if (!std::isfinite(m[0]) || !std::isfinite(m[1]) || !std::isfinite(m[2])) return false;
    }
  return true;
}

// test_render.cpp:211-221: passes of 1, 1, 2, 4, 8, 4 samples at a cap of 8.
void progressive_counts() {
  const CompiledScene s = compile_scene(asset("first_light.json", 32, 32, 20));
  RenderOptions o;
  o.max_pass_spp = 8;
  GpuRenderer r(s, o);
  std::vector<uint32_t> spp;
  bool consistent = true;
  r.render([&](const Framebuffer& fb, const RenderProgress& p) {
    spp.push_back(p.spp_done);
    consistent = consistent && fb.min_samples() == p.spp_done;
  });
  std::string got;
  for (uint32_t v : spp) got += std::to_string(v) + " ";
  check(spp == std::vector<uint32_t>{1, 2, 4, 8, 16, 20} && consistent, "progressive_counts", got);
}

// test_render.cpp:223-255
void cancel_and_restart() {
  SceneDesc d = asset("cornell_box.json", 128, 128, 100000);
  RenderSession session(compile_scene(d));
  std::atomic<int> callbacks{0};
  session.start([&](const Framebuffer&, const RenderProgress&) { callbacks++; });
  std::this_thread::sleep_for(std::chrono::milliseconds(150));
  const bool was_running = session.running();
  const auto t0 = Clock::now();
  session.cancel();
  const RenderProgress p = session.wait();
  const double ms = ms_since(t0);
  check(was_running && p.cancelled && !p.finished && !session.running() && ms < 1000 && callbacks > 0 &&
            session.error().empty(),
        "cancel_within_1s",
        "stopped in " + std::to_string(ms) + " ms after " + std::to_string(callbacks.load()) + " passes, spp_done " +
            std::to_string(p.spp_done));
  check(all_means_finite(session.renderer().framebuffer()) &&
            session.renderer().framebuffer().min_samples() == p.spp_done,
        "cancelled_buffer_consistent", "min samples " + std::to_string(p.spp_done));
  d.render.spp = 4;
  d.render.width = 16;
  d.render.height = 16;
  session.restart(compile_scene(d));
  const RenderProgress p2 = session.wait();
  check(p2.finished && !p2.cancelled && p2.spp_done == 4, "restart_completes", "spp_done " + std::to_string(p2.spp_done));
}

// test_render.cpp:257-266
void destructor_joins() {
  const SceneDesc d = asset("cornell_box.json", 64, 64, 100000);
  const CompiledScene s = compile_scene(d);
  const auto t0 = Clock::now();
  {
    RenderSession session(s);
    session.start();
    std::this_thread::sleep_for(std::chrono::milliseconds(50));
  }
  const double ms = ms_since(t0);
  check(ms < 2000, "destructor_cancels_and_joins", std::to_string(ms) + " ms");
}

// A render with no per-pass observer runs its schedule in one device call;
// a std::atomic<bool> raised from another thread still stops it.
void cancel_one_call() {
  const CompiledScene s = compile_scene(asset("cornell_box.json", 256, 256, 100000));
  GpuRenderer r(s);
  std::atomic<bool> cancel{false};
  Clock::time_point raised;
  std::thread t([&] {
    std::this_thread::sleep_for(std::chrono::milliseconds(200));
    raised = Clock::now();
    cancel.store(true);
  });
  const RenderProgress p = r.render({}, &cancel);
  const double ms = ms_since(raised);
  t.join();
  check(p.cancelled && !p.finished && ms < 1000 && p.spp_done == r.framebuffer().min_samples() &&
            all_means_finite(r.framebuffer()),
        "cancel_one_call", "stopped " + std::to_string(ms) + " ms after the flag, spp_done " +
                               std::to_string(p.spp_done));
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: session_check <assets dir>\n");
    return 2;
  }
  g_assets = argv[1];
  try {
    progressive_counts();
    cancel_and_restart();
    destructor_joins();
    cancel_one_call();
  } catch (const std::exception& e) {
    check(false, "exception", e.what());
  }
  return g_failures ? 1 : 0;
}
