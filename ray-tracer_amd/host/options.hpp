// options.hpp — what a render request can ask for, read from the rt command
// line or from an `rt --serve` job line through one option table.
//
// The flag names, their meaning and the override rules are the reference
// CLI's interface (apps/rt_cli.cpp:32-54 flags, :71-104 parsing, :119-135
// overrides); --precision, --device and the --serve options are additions.
// A job line of `rt --serve` carries the same overrides as JSON keys
// (width, spp, max_depth, no_bvh, ...).
#pragma once
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>

#include "json.hpp"
#include "scene.hpp"

namespace rtm {

struct Options {
  // Render overrides, applied over the scene file's render block.
  std::optional<int> width, height, spp, max_depth, threads, tile, max_pass;
  std::optional<uint64_t> seed;
  std::optional<double> exposure;
  std::optional<std::string> integrator;
  bool no_bvh = false, no_light_sampling = false;
  int32_t precision = RTG_PRECISION_F32;
  // Command line only.
  std::string scene_path, output = "render.png", snapshot_dir;
  bool single_pass = false, stream = false, quiet = false, validate_only = false, dump = false,
       save_every = false, serve = false, supersede = false, help = false;
  int device = 0, tick_paths = 0, cache_mb = 0;
  // Job lines only: preview jobs render each pass whole in the next tick;
  // supersede cancels every earlier job.
  bool preview = false;
  std::optional<bool> supersede_job;

  // Writes the overrides into `R` (rt_cli.cpp:119-135). An unknown
  // integrator name is a scene error.
  void apply(RenderSettings& R) const;
};

// A malformed command line: the message, and whether the usage text follows.
class UsageError : public std::runtime_error {
 public:
  UsageError(const std::string& m, bool show_usage) : std::runtime_error(m), show_usage(show_usage) {}
  bool show_usage;
};

// Parses argv[1..argc). Throws UsageError; stops at -h / --help.
Options parse_command_line(int argc, char** argv);
// Reads the override keys of a `rt --serve` job line into `o`.
void read_job_fields(const Json& job, Options& o);
// The usage text of `rt`.
std::string usage_text();

// Builds the JSON object of a stream protocol line (FRAME / DONE / ERROR)
// field by field, in order: integers as integers, reals with six decimals,
// flags as true/false, text as JSON strings.
class LineFields {
 public:
  LineFields& add(const char* key, uint64_t v) { return raw(key, std::to_string(v)); }
  LineFields& add(const char* key, int64_t v) { return raw(key, std::to_string(v)); }
  LineFields& add(const char* key, int v) { return raw(key, std::to_string(v)); }
  LineFields& add(const char* key, uint32_t v) { return raw(key, std::to_string(v)); }
  LineFields& add(const char* key, double v) { return raw(key, std::to_string(v)); }
  LineFields& add(const char* key, bool v) { return raw(key, v ? "true" : "false"); }
  LineFields& add(const char* key, const std::string& v) { return raw(key, Json::string(v).dump(-1)); }
  // "<tag> {...}\n"
  std::string line(const char* tag) const { return std::string(tag) + " {" + body_ + "}\n"; }

 private:
  LineFields& raw(const char* key, const std::string& value) {
    if (!body_.empty()) body_ += ',';
    body_ += '"';
    body_ += key;
    body_ += "\":";
    body_ += value;
    return *this;
  }
  std::string body_;
};

}  // namespace rtm
