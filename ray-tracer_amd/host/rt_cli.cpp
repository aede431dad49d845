// rt — command-line front end, rendering on the MI355X.
//
//   rt scene.json -o out.png [overrides...]
//   rt - --stream < scene.json
//   rt --serve                 (jobs as JSON lines on stdin; serve_stream.cpp)
//
// Interface kept from the reference CLI (apps/rt_cli.cpp): the flags and
// override rules (options.cpp), the --stream protocol on stdout —
//   FRAME <json header>\n<png bytes>   after every progressive pass
//   DONE <json stats>\n                finished or cancelled
//   ERROR {"message":...}\n            the scene could not be loaded
// — diagnostics on stderr, and the exit codes: 1 usage, 2 scene or render
// error, 3 image write error, 130 cancelled (SIGINT / SIGTERM). DONE also
// carries grays_per_sec; --threads and --tile are accepted and ignored.
#include <atomic>
#include <csignal>
#include <cstdio>
#include <iostream>
#include <iterator>
#include <memory>
#include <optional>
#include <string>

#include "options.hpp"
#include "renderer.hpp"

namespace rtm {
int serve_main(int device, bool supersede, uint32_t tick_paths, uint64_t cache_bytes, bool quiet);
}

using namespace rtm;

namespace {

std::atomic<bool> g_cancel{false};
extern "C" void on_signal(int) { g_cancel.store(true); }

enum Exit { kOk = 0, kUsage = 1, kSceneError = 2, kWriteError = 3, kCancelled = 130 };

// One render of the command line: load, render, write, report.
class CliRender {
 public:
  explicit CliRender(const Options& o) : o_(o) {}

  int run() {
    if (const std::optional<int> early = load()) return *early;
    if (!o_.quiet) describe_scene();
    if (const std::optional<int> early = open_device()) return *early;
    RenderProgress final;
    try {
      final = render();
    } catch (const std::exception& e) {
      return fail(e.what(), "\nerror: ");
    }
    if (!o_.quiet) std::cerr << "\n";
    try {
      if (writes_output()) write_image(o_.output, renderer_->image());
    } catch (const std::exception& e) {
      std::cerr << "error: " << e.what() << "\n";
      return kWriteError;
    }
    report(final);
    return final.cancelled ? kCancelled : kOk;
  }

 private:
  // An ERROR line for the stream reader, the message for a person.
  int fail(const std::string& what, const char* prefix = "error: ") {
    if (o_.stream) std::cout << LineFields().add("message", what).line("ERROR") << std::flush;
    std::cerr << prefix << what << "\n";
    return kSceneError;
  }

  // Scene text → SceneDesc with overrides → checks → compiled scene. Returns
  // an exit code when the program ends here.
  std::optional<int> load() {
    try {
      if (o_.scene_path == "-") {
        const std::string text{std::istreambuf_iterator<char>(std::cin), std::istreambuf_iterator<char>()};
        desc_ = parse_scene(text);
      } else {
        desc_ = load_scene(o_.scene_path);
      }
      o_.apply(desc_.render);
      const std::vector<std::string> problems = validate_scene(desc_);
      if (!problems.empty()) throw SceneError(problems.front());
      if (o_.dump) {
        std::cout << scene_to_json(desc_).dump(2) << "\n";
        return kOk;
      }
      if (o_.validate_only) {
        std::cerr << "scene OK: " << desc_.objects.size() << " objects\n";
        return kOk;
      }
      scene_ = compile_scene(desc_);
    } catch (const std::exception& e) {
      return fail(e.what());
    }
    return std::nullopt;
  }

  void describe_scene() const {
    const SceneInfo& in = scene_.info;
    std::cerr << "scene '" << desc_.name << "': " << in.spheres << " spheres, " << in.triangles << " triangles, "
              << in.planes << " planes, " << in.lights << " lights";
    if (in.degenerate_skipped) std::cerr << ", " << in.degenerate_skipped << " degenerate skipped";
    std::cerr << "\nBVH: " << in.bvh.nodes << " nodes, depth " << in.bvh.max_depth << ", built in "
              << in.bvh.build_ms << " ms\n";
  }

  std::optional<int> open_device() {
    RenderOptions ro;
    ro.progressive = !o_.single_pass;
    if (o_.max_pass) ro.max_pass_spp = *o_.max_pass;
    ro.precision = o_.precision;
    ro.device = o_.device;
    try {
      renderer_ = std::make_unique<GpuRenderer>(scene_, ro);
    } catch (const std::exception& e) {
      return fail(e.what());
    }
    if (!o_.quiet) {
      const RenderSettings& s = scene_.settings;
      std::cerr << "rendering " << s.width << "x" << s.height << " @ " << s.spp << " spp on "
                << renderer_->device_name() << ", integrator=" << integrator_name(s.integrator)
                << ", precision=" << (o_.precision == RTG_PRECISION_F64 ? "f64" : "f32") << "\n";
    }
    return std::nullopt;
  }

  // A render nobody watches pass by pass (no stream, snapshots or progress
  // lines) runs its whole schedule in one device call.
  bool observed() const { return o_.stream || o_.save_every || !o_.snapshot_dir.empty() || !o_.quiet; }
  // With --stream the image goes to stdout unless -o names a file.
  bool writes_output() const { return !o_.stream || o_.output != "render.png"; }

  RenderProgress render() {
    if (!observed()) return renderer_->render({}, &g_cancel);
    return renderer_->render([this](const Framebuffer&, const RenderProgress& p) { after_pass(p); }, &g_cancel);
  }

  void after_pass(const RenderProgress& p) {
    if (o_.stream) {
      const std::vector<uint8_t> png = encode_png(renderer_->image());
      const std::string header = LineFields()
                                     .add("pass", p.pass)
                                     .add("spp_done", p.spp_done)
                                     .add("spp_target", p.spp_target)
                                     .add("elapsed_ms", p.elapsed_ms)
                                     .add("rays", p.rays)
                                     .add("width", scene_.settings.width)
                                     .add("height", scene_.settings.height)
                                     .add("finished", p.finished)
                                     .add("cancelled", p.cancelled)
                                     .add("bytes", (uint64_t)png.size())
                                     .line("FRAME");
      std::fwrite(header.data(), 1, header.size(), stdout);
      std::fwrite(png.data(), 1, png.size(), stdout);
      std::fflush(stdout);
    }
    if (o_.save_every) write_image(o_.output, renderer_->image());
    if (!o_.snapshot_dir.empty() && !p.cancelled) {
      char file[32];
      std::snprintf(file, sizeof file, "/spp_%05u.png", p.spp_done);
      write_image(o_.snapshot_dir + file, renderer_->image());
    }
    if (!o_.quiet)
      std::cerr << "\r  pass " << p.pass << ": " << p.spp_done << "/" << p.spp_target << " spp, " << (long)p.elapsed_ms
                << " ms" << std::flush;
  }

  void report(const RenderProgress& final) const {
    const RenderStatsSnapshot st = renderer_->stats();
    if (o_.stream)
      std::cout << LineFields()
                       .add("cancelled", final.cancelled)
                       .add("spp_done", final.spp_done)
                       .add("elapsed_ms", st.elapsed_ms)
                       .add("rays", st.rays)
                       .add("shadow_rays", st.shadow_rays)
                       .add("mrays_per_sec", st.mrays_per_sec())
                       .add("grays_per_sec", st.grays_per_sec())
                       .add("invalid_samples", st.invalid_samples)
                       .add("threads", renderer_->thread_count())
                       .line("DONE")
                << std::flush;
    if (o_.quiet) return;
    std::cerr << (final.cancelled ? "cancelled" : "done") << " in " << st.elapsed_ms << " ms, "
              << (st.rays + st.shadow_rays) << " rays (" << st.mrays_per_sec() << " Mrays/s)";
    if (st.invalid_samples) std::cerr << ", " << st.invalid_samples << " invalid samples";
    std::cerr << " -> " << (writes_output() ? o_.output : std::string("stdout")) << "\n";
  }

  Options o_;
  SceneDesc desc_;
  CompiledScene scene_;
  std::unique_ptr<GpuRenderer> renderer_;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << usage_text();
    return kUsage;
  }
  Options o;
  try {
    o = parse_command_line(argc, argv);
  } catch (const UsageError& e) {
    std::cerr << e.what() << "\n";
    if (e.show_usage) std::cerr << usage_text();
    return kUsage;
  }
  if (o.help) {
    std::cerr << usage_text();
    return kOk;
  }
  if (o.serve)
    return serve_main(o.device, o.supersede, (uint32_t)std::max(0, o.tick_paths), (uint64_t)std::max(0, o.cache_mb) << 20,
                      o.quiet);
  if (o.scene_path.empty()) {
    std::cerr << usage_text();
    return kUsage;
  }
  std::signal(SIGINT, on_signal);
  std::signal(SIGTERM, on_signal);
  return CliRender(o).run();
}
