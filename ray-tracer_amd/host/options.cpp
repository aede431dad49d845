// options.cpp — the option table behind `rt`'s command line and the job
// lines of `rt --serve`.
#include "options.hpp"

#include <cstring>

namespace rtm {

namespace {

enum class Arg { Flag, Int, Real, Text };

// A value as the command line or a job line gave it; `name` is the flag or
// key as written (error messages quote it).
struct Value {
  std::string name;
  bool on = true;
  double num = 0;
  std::string text;
};

struct Option {
  const char* flag;   // command-line name (nullptr: job lines only)
  const char* alias;  // short command-line name
  const char* key;    // job-line key (nullptr: command line only)
  Arg arg;
  const char* usage;  // left column of the usage text (nullptr: not listed)
  const char* help;   // right column; '\n' continues it on the next line
  void (*set)(Options&, const Value&);
};

int as_int(const Value& v) { return (int)v.num; }

// One row per option, in usage order. The command line parses integers with
// std::stoi (so "--seed -1" wraps to 2^64 - 1, as the reference's cast does);
// a job line takes them from JSON numbers.
const Option kOptions[] = {
    {"--output", "-o", nullptr, Arg::Text, "-o, --output FILE", "output image (.png or .ppm), default render.png",
     [](Options& o, const Value& v) { o.output = v.text; }},
    {"--width", nullptr, "width", Arg::Int, "--width N --height N", "override resolution",
     [](Options& o, const Value& v) { o.width = as_int(v); }},
    {"--height", nullptr, "height", Arg::Int, nullptr, nullptr,
     [](Options& o, const Value& v) { o.height = as_int(v); }},
    {"--spp", nullptr, "spp", Arg::Int, "--spp N", "samples per pixel",
     [](Options& o, const Value& v) { o.spp = as_int(v); }},
    {"--max-depth", nullptr, "max_depth", Arg::Int, "--max-depth N", "maximum bounces",
     [](Options& o, const Value& v) { o.max_depth = as_int(v); }},
    {"--integrator", nullptr, "integrator", Arg::Text, "--integrator NAME", "path | direct | normals",
     [](Options& o, const Value& v) { o.integrator = v.text; }},
    {"--threads", nullptr, nullptr, Arg::Int, "--threads N", "accepted for compatibility (ignored)",
     [](Options& o, const Value& v) { o.threads = as_int(v); }},
    {"--tile", nullptr, nullptr, Arg::Int, "--tile N", "accepted for compatibility (ignored)",
     [](Options& o, const Value& v) { o.tile = as_int(v); }},
    {"--seed", nullptr, "seed", Arg::Int, "--seed N", "RNG seed",
     [](Options& o, const Value& v) { o.seed = v.num < 0 ? (uint64_t)(int64_t)v.num : (uint64_t)v.num; }},
    {"--exposure", nullptr, "exposure", Arg::Real, "--exposure X", "exposure multiplier",
     [](Options& o, const Value& v) { o.exposure = v.num; }},
    {"--no-bvh", nullptr, "no_bvh", Arg::Flag, "--no-bvh", "brute-force intersection (reference)",
     [](Options& o, const Value& v) { o.no_bvh = o.no_bvh || v.on; }},
    {"--no-light-sampling", nullptr, "no_light_sampling", Arg::Flag, "--no-light-sampling",
     "disable next-event estimation",
     [](Options& o, const Value& v) { o.no_light_sampling = o.no_light_sampling || v.on; }},
    {"--single-pass", nullptr, nullptr, Arg::Flag, "--single-pass", "render all samples in one pass",
     [](Options& o, const Value&) { o.single_pass = true; }},
    {"--max-pass", nullptr, "max_pass", Arg::Int, "--max-pass N", "cap samples per progressive pass (default 32)",
     [](Options& o, const Value& v) { o.max_pass = as_int(v); }},
    {"--precision", nullptr, "precision", Arg::Text, "--precision P",
     "f32 (default) | f64 (reference-exact arithmetic and RNG)",
     [](Options& o, const Value& v) {
       if (v.text == "f32") o.precision = RTG_PRECISION_F32;
       else if (v.text == "f64") o.precision = RTG_PRECISION_F64;
       else throw std::runtime_error(v.name + ": expected f32 or f64");
     }},
    {"--device", nullptr, nullptr, Arg::Int, "--device N", "HIP device (default 0)",
     [](Options& o, const Value& v) { o.device = as_int(v); }},
    {"--stream", nullptr, nullptr, Arg::Flag, "--stream", "write progressive PNG frames to stdout",
     [](Options& o, const Value&) { o.stream = true; }},
    {"--save-every-pass", nullptr, nullptr, Arg::Flag, "--save-every-pass", "rewrite the output image after every pass",
     [](Options& o, const Value&) { o.save_every = true; }},
    {"--snapshot-dir", nullptr, nullptr, Arg::Text, "--snapshot-dir DIR", "also save each pass as DIR/spp_NNNNN.png",
     [](Options& o, const Value& v) { o.snapshot_dir = v.text; }},
    {"--validate", nullptr, nullptr, Arg::Flag, "--validate", "only parse + validate the scene",
     [](Options& o, const Value&) { o.validate_only = true; }},
    {"--dump-scene", nullptr, nullptr, Arg::Flag, "--dump-scene", "print the normalised scene JSON and exit",
     [](Options& o, const Value&) { o.dump = true; }},
    {"--quiet", "-q", nullptr, Arg::Flag, "-q, --quiet", "no progress output",
     [](Options& o, const Value&) { o.quiet = true; }},
    {"--serve", nullptr, nullptr, Arg::Flag, "--serve",
     "serve jobs given as JSON lines on stdin (no scene argument):\n"
     "{\"job\":N,\"scene\":{...},overrides...} or {\"cancel\":N}; streams\n"
     "FRAME/DONE/ERROR lines tagged with the job id",
     [](Options& o, const Value&) { o.serve = true; }},
    {"--supersede", nullptr, "supersede", Arg::Flag, "--supersede", "(serve) a new job cancels the running ones",
     [](Options& o, const Value& v) {
       o.supersede = v.on;
       o.supersede_job = v.on;
     }},
    {"--tick-paths", nullptr, nullptr, Arg::Int, "--tick-paths N",
     "(serve) paths of large jobs per batched pass (default: tuned per tick)",
     [](Options& o, const Value& v) { o.tick_paths = as_int(v); }},
    {"--cache-mb", nullptr, nullptr, Arg::Int, "--cache-mb N", "(serve) resident compiled scenes, default 8192",
     [](Options& o, const Value& v) { o.cache_mb = as_int(v); }},
    {nullptr, nullptr, "preview", Arg::Flag, nullptr, nullptr, [](Options& o, const Value& v) { o.preview = v.on; }},
};

const Option* find_flag(const char* a) {
  for (const Option& opt : kOptions)
    if ((opt.flag && std::strcmp(opt.flag, a) == 0) || (opt.alias && std::strcmp(opt.alias, a) == 0)) return &opt;
  return nullptr;
}

}  // namespace

void Options::apply(RenderSettings& R) const {
  // Every override replaces the scene file's value; the compatibility-only
  // ones (threads, tile) are recorded as the reference records them.
  const std::pair<const std::optional<int>*, int*> ints[] = {
      {&width, &R.width}, {&height, &R.height}, {&spp, &R.spp}, {&max_depth, &R.max_depth},
      {&threads, &R.threads}, {&tile, &R.tile_size}};
  for (const auto& [from, to] : ints)
    if (*from) *to = **from;
  if (seed) R.seed = *seed;
  if (exposure) R.exposure = *exposure;
  if (no_bvh) R.use_bvh = false;
  if (no_light_sampling) R.light_sampling = false;
  if (!integrator) return;
  static const std::pair<const char*, Integrator> kNames[] = {
      {"path", Integrator::Path}, {"direct", Integrator::Direct}, {"normals", Integrator::Normals}};
  for (const auto& [name, kind] : kNames)
    if (*integrator == name) {
      R.integrator = kind;
      return;
    }
  throw SceneError("--integrator: expected path, direct or normals");
}

Options parse_command_line(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    if (std::strcmp(a, "-h") == 0 || std::strcmp(a, "--help") == 0) {
      o.help = true;
      return o;
    }
    const Option* opt = find_flag(a);
    if (!opt) {
      if (a[0] == '-' && a[1] != '\0') throw UsageError(std::string("unknown option ") + a, true);
      o.scene_path = a;  // a path, or "-" for stdin
      continue;
    }
    Value v;
    v.name = a;
    if (opt->arg != Arg::Flag) {
      if (i + 1 >= argc) throw UsageError("missing value for " + v.name, false);
      v.text = argv[++i];
      try {
        if (opt->arg == Arg::Int) v.num = std::stoi(v.text);
        else if (opt->arg == Arg::Real) v.num = std::stod(v.text);
      } catch (const std::exception&) {
        throw UsageError(std::string(opt->arg == Arg::Int ? "invalid integer for " : "invalid number for ") + v.name +
                             ": " + v.text,
                         false);
      }
    }
    try {
      opt->set(o, v);
    } catch (const std::runtime_error& e) {
      throw UsageError(e.what(), false);
    }
  }
  return o;
}

void read_job_fields(const Json& job, Options& o) {
  for (const Option& opt : kOptions) {
    if (!opt.key) continue;
    const Json* j = job.get(opt.key);
    if (!j) continue;
    Value v;
    v.name = opt.key;
    switch (opt.arg) {
      case Arg::Flag:
        if (!j->is_bool()) continue;  // a non-boolean switch is ignored
        v.on = j->as_bool();
        break;
      case Arg::Int:
      case Arg::Real:
        if (!j->is_number()) throw SceneError(v.name + ": expected a number");
        v.num = j->as_number();
        break;
      case Arg::Text: v.text = j->as_string(); break;
    }
    opt.set(o, v);
  }
}

std::string usage_text() {
  constexpr size_t kColumn = 23;
  std::string s = "usage: rt <scene.json | -> [options]\n";
  for (const Option& opt : kOptions) {
    if (!opt.usage) continue;
    std::string left = opt.usage;
    left.resize(std::max(kColumn, left.size() + 1), ' ');
    s += "  " + left;
    for (const char* h = opt.help; *h; ++h) {
      s += *h;
      if (*h == '\n') s.append(kColumn + 2, ' ');
    }
    s += '\n';
  }
  return s;
}

}  // namespace rtm
