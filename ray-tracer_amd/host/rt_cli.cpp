// rt — command-line front end, rendering on the MI355X.
//
//   rt scene.json -o out.png [overrides...]
//   rt - --stream < scene.json
//
// Flags, override semantics, the --stream protocol (FRAME <json>\n<png>,
// DONE <json>\n, ERROR <json>\n on stdout) and the exit codes (1 usage,
// 2 scene error, 3 write error, 130 cancelled) are those of the reference
// CLI (apps/rt_cli.cpp:6-10, 32-209). Added: --precision f32|f64 and
// --device N; DONE also reports grays_per_sec. --threads and --tile are
// accepted for compatibility (the GPU schedules its own work). --serve runs
// many jobs in one process (serve_stream.cpp).
#include <atomic>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <iterator>
#include <optional>
#include <string>

#include "renderer.hpp"

namespace rtm {
int serve_main(int device, bool supersede, uint32_t tick_paths, uint64_t cache_bytes, bool quiet);
}

using namespace rtm;

namespace {

std::atomic<int32_t> g_cancel{0};
extern "C" void on_signal(int) { g_cancel.store(1); }

void usage() {
  std::cerr << "usage: rt <scene.json | -> [options]\n"
               "  -o, --output FILE      output image (.png or .ppm), default render.png\n"
               "  --width N --height N   override resolution\n"
               "  --spp N                samples per pixel\n"
               "  --max-depth N          maximum bounces\n"
               "  --integrator NAME      path | direct | normals\n"
               "  --threads N            accepted for compatibility (ignored)\n"
               "  --tile N               accepted for compatibility (ignored)\n"
               "  --seed N               RNG seed\n"
               "  --exposure X           exposure multiplier\n"
               "  --no-bvh               brute-force intersection (reference)\n"
               "  --no-light-sampling    disable next-event estimation\n"
               "  --single-pass          render all samples in one pass\n"
               "  --max-pass N           cap samples per progressive pass (default 32)\n"
               "  --precision P          f32 (default) | f64 (reference-exact arithmetic and RNG)\n"
               "  --device N             HIP device (default 0)\n"
               "  --stream               write progressive PNG frames to stdout\n"
               "  --save-every-pass      rewrite the output image after every pass\n"
               "  --snapshot-dir DIR     also save each pass as DIR/spp_NNNNN.png\n"
               "  --validate             only parse + validate the scene\n"
               "  --dump-scene           print the normalised scene JSON and exit\n"
               "  -q, --quiet            no progress output\n"
               "  --serve                serve jobs given as JSON lines on stdin (no scene argument):\n"
               "                         {\"job\":N,\"scene\":{...},overrides...} or {\"cancel\":N}; streams\n"
               "                         FRAME/DONE/ERROR lines tagged with the job id\n"
               "  --supersede            (serve) a new job cancels the running ones\n"
               "  --tick-paths N         (serve) paths of large jobs per batched pass (default: tuned per tick)\n"
               "  --cache-mb N           (serve) resident compiled scenes, default 8192\n";
}

std::string json_str(const std::string& s) {
  Json j = Json::string(s);
  return j.dump(-1);
}

std::string fmt(double v) { return std::to_string(v); }

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    usage();
    return 1;
  }
  std::string scene_path, output = "render.png", snapshot_dir;
  std::optional<int> width, height, spp, max_depth, threads, tile, max_pass;
  std::optional<uint64_t> seed;
  std::optional<double> exposure;
  std::optional<std::string> integrator;
  bool no_bvh = false, no_nee = false, single_pass = false, stream = false, quiet = false;
  bool validate_only = false, dump = false, save_every = false, serve = false, supersede = false;
  int32_t precision = RTG_PRECISION_F32;
  int device = 0, tick_paths = 0, cache_mb = 0;

  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) {
        std::cerr << "missing value for " << a << "\n";
        std::exit(1);
      }
      return argv[++i];
    };
    auto next_int = [&]() {
      const std::string v = next();
      try {
        return std::stoi(v);
      } catch (...) {
        std::cerr << "invalid integer for " << a << ": " << v << "\n";
        std::exit(1);
      }
    };
    if (a == "-o" || a == "--output") output = next();
    else if (a == "--width") width = next_int();
    else if (a == "--height") height = next_int();
    else if (a == "--spp") spp = next_int();
    else if (a == "--max-depth") max_depth = next_int();
    else if (a == "--threads") threads = next_int();
    else if (a == "--tile") tile = next_int();
    else if (a == "--max-pass") max_pass = next_int();
    else if (a == "--seed") seed = (uint64_t)next_int();
    else if (a == "--exposure") exposure = std::stod(next());
    else if (a == "--integrator") integrator = next();
    else if (a == "--no-bvh") no_bvh = true;
    else if (a == "--no-light-sampling") no_nee = true;
    else if (a == "--single-pass") single_pass = true;
    else if (a == "--stream") stream = true;
    else if (a == "--save-every-pass") save_every = true;
    else if (a == "--snapshot-dir") snapshot_dir = next();
    else if (a == "--validate") validate_only = true;
    else if (a == "--dump-scene") dump = true;
    else if (a == "-q" || a == "--quiet") quiet = true;
    else if (a == "--device") device = next_int();
    else if (a == "--serve") serve = true;
    else if (a == "--supersede") supersede = true;
    else if (a == "--tick-paths") tick_paths = next_int();
    else if (a == "--cache-mb") cache_mb = next_int();
    else if (a == "--precision") {
      const std::string p = next();
      if (p == "f32") precision = RTG_PRECISION_F32;
      else if (p == "f64") precision = RTG_PRECISION_F64;
      else {
        std::cerr << "--precision: expected f32 or f64\n";
        return 1;
      }
    } else if (a == "-h" || a == "--help") {
      usage();
      return 0;
    } else if (!a.empty() && a[0] == '-' && a != "-") {
      std::cerr << "unknown option " << a << "\n";
      usage();
      return 1;
    } else {
      scene_path = a;
    }
  }
  if (serve) return rtm::serve_main(device, supersede, (uint32_t)std::max(0, tick_paths),
                                    (uint64_t)std::max(0, cache_mb) << 20, quiet);
  if (scene_path.empty()) {
    usage();
    return 1;
  }
  std::signal(SIGINT, on_signal);
  std::signal(SIGTERM, on_signal);

  SceneDesc desc;
  CompiledScene scene;
  try {
    if (scene_path == "-") {
      const std::string text{std::istreambuf_iterator<char>(std::cin), std::istreambuf_iterator<char>()};
      desc = parse_scene(text);
    } else {
      desc = load_scene(scene_path);
    }
    RenderSettings& R = desc.render;
    if (width) R.width = *width;
    if (height) R.height = *height;
    if (spp) R.spp = *spp;
    if (max_depth) R.max_depth = *max_depth;
    if (threads) R.threads = *threads;
    if (tile) R.tile_size = *tile;
    if (seed) R.seed = *seed;
    if (exposure) R.exposure = *exposure;
    if (no_bvh) R.use_bvh = false;
    if (no_nee) R.light_sampling = false;
    if (integrator) {
      if (*integrator == "path") R.integrator = Integrator::Path;
      else if (*integrator == "direct") R.integrator = Integrator::Direct;
      else if (*integrator == "normals") R.integrator = Integrator::Normals;
      else throw SceneError("--integrator: expected path, direct or normals");
    }
    if (const auto errs = validate_scene(desc); !errs.empty()) throw SceneError(errs.front());
    if (dump) {
      std::cout << scene_to_json(desc).dump(2) << "\n";
      return 0;
    }
    if (validate_only) {
      std::cerr << "scene OK: " << desc.objects.size() << " objects\n";
      return 0;
    }
    scene = compile_scene(desc);
  } catch (const std::exception& e) {
    if (stream) std::cout << "ERROR {\"message\":" << json_str(e.what()) << "}\n" << std::flush;
    std::cerr << "error: " << e.what() << "\n";
    return 2;
  }

  const SceneInfo& info = scene.info;
  if (!quiet)
    std::cerr << "scene '" << desc.name << "': " << info.spheres << " spheres, " << info.triangles << " triangles, "
              << info.planes << " planes, " << info.lights << " lights"
              << (info.degenerate_skipped ? ", " + std::to_string(info.degenerate_skipped) + " degenerate skipped" : "")
              << "\nBVH: " << info.bvh.nodes << " nodes, depth " << info.bvh.max_depth << ", built in "
              << info.bvh.build_ms << " ms\n";

  RenderOptions opts;
  opts.progressive = !single_pass;
  if (max_pass) opts.max_pass_spp = *max_pass;
  opts.precision = precision;
  opts.device = device;
  std::unique_ptr<GpuRenderer> renderer;
  try {
    renderer = std::make_unique<GpuRenderer>(scene, opts);
  } catch (const std::exception& e) {
    if (stream) std::cout << "ERROR {\"message\":" << json_str(e.what()) << "}\n" << std::flush;
    std::cerr << "error: " << e.what() << "\n";
    return 2;
  }
  if (!quiet)
    std::cerr << "rendering " << scene.settings.width << "x" << scene.settings.height << " @ " << scene.settings.spp
              << " spp on " << renderer->device_name() << ", integrator=" << integrator_name(scene.settings.integrator)
              << ", precision=" << (precision == RTG_PRECISION_F64 ? "f64" : "f32") << "\n";

  auto on_pass = [&](const Framebuffer&, const RenderProgress& p) {
    if (stream) {
      const auto png = encode_png(renderer->image());
      const std::string hdr = "FRAME {\"pass\":" + std::to_string(p.pass) + ",\"spp_done\":" + std::to_string(p.spp_done) +
                              ",\"spp_target\":" + std::to_string(p.spp_target) + ",\"elapsed_ms\":" + fmt(p.elapsed_ms) +
                              ",\"rays\":" + std::to_string(p.rays) + ",\"width\":" + std::to_string(scene.settings.width) +
                              ",\"height\":" + std::to_string(scene.settings.height) +
                              ",\"finished\":" + (p.finished ? "true" : "false") +
                              ",\"cancelled\":" + (p.cancelled ? "true" : "false") +
                              ",\"bytes\":" + std::to_string(png.size()) + "}\n";
      std::fwrite(hdr.data(), 1, hdr.size(), stdout);
      std::fwrite(png.data(), 1, png.size(), stdout);
      std::fflush(stdout);
    }
    if (save_every) write_image(output, renderer->image());
    if (!snapshot_dir.empty() && !p.cancelled) {
      char name[32];
      std::snprintf(name, sizeof name, "/spp_%05u.png", p.spp_done);
      write_image(snapshot_dir + name, renderer->image());
    }
    if (!quiet)
      std::cerr << "\r  pass " << p.pass << ": " << p.spp_done << "/" << p.spp_target << " spp, " << (long)p.elapsed_ms
                << " ms" << std::flush;
  };

  // Without an observer of the passes (no stream, snapshots or progress
  // lines) the renderer runs its schedule in one call, passes side by side.
  const bool observed = stream || save_every || !snapshot_dir.empty() || !quiet;
  RenderProgress final;
  try {
    final = observed ? renderer->render(on_pass, &g_cancel) : renderer->render({}, &g_cancel);
  } catch (const std::exception& e) {
    if (stream) std::cout << "ERROR {\"message\":" << json_str(e.what()) << "}\n" << std::flush;
    std::cerr << "\nerror: " << e.what() << "\n";
    return 2;
  }
  if (!quiet) std::cerr << "\n";
  try {
    if (!stream || output != "render.png") write_image(output, renderer->image());
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 3;
  }
  const RenderStatsSnapshot st = renderer->stats();
  const std::string stats_json =
      "{\"cancelled\":" + std::string(final.cancelled ? "true" : "false") + ",\"spp_done\":" + std::to_string(final.spp_done) +
      ",\"elapsed_ms\":" + fmt(st.elapsed_ms) + ",\"rays\":" + std::to_string(st.rays) +
      ",\"shadow_rays\":" + std::to_string(st.shadow_rays) + ",\"mrays_per_sec\":" + fmt(st.mrays_per_sec()) +
      ",\"grays_per_sec\":" + fmt(st.grays_per_sec()) + ",\"invalid_samples\":" + std::to_string(st.invalid_samples) +
      ",\"threads\":" + std::to_string(renderer->thread_count()) + "}";
  if (stream) std::cout << "DONE " << stats_json << "\n" << std::flush;
  if (!quiet)
    std::cerr << (final.cancelled ? "cancelled" : "done") << " in " << st.elapsed_ms << " ms, "
              << (st.rays + st.shadow_rays) << " rays (" << st.mrays_per_sec() << " Mrays/s)"
              << (st.invalid_samples ? ", " + std::to_string(st.invalid_samples) + " invalid samples" : "") << " -> "
              << (!stream || output != "render.png" ? output : "stdout") << "\n";
  return final.cancelled ? 130 : 0;
}
