// scene.hpp — host-side scene model: the editable description read from
// JSON (SceneDesc, the format of src/io/scene_io.cpp) and the compiled,
// render-ready form handed to the device back end (Scene::compile,
// src/core/scene.cpp:31-137).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.hpp"
#include "rtg.h"

namespace rtm {

struct Vec3d {
  double x = 0, y = 0, z = 0;
  bool operator==(const Vec3d&) const = default;
};

enum class Integrator : int32_t { Path = RTG_INTEGRATOR_PATH, Direct = RTG_INTEGRATOR_DIRECT, Normals = RTG_INTEGRATOR_NORMALS };
enum class ToneMap : int32_t { Clamp = 0, Reinhard = 1, Aces = 2 };

// core/scene_desc.hpp:20-35 (defaults included).
struct RenderSettings {
  int width = 640, height = 360, spp = 64, max_depth = 8, rr_min_depth = 3;
  Integrator integrator = Integrator::Path;
  bool light_sampling = true;
  uint64_t seed = 1;
  ToneMap tone_map = ToneMap::Aces;
  double exposure = 1.0;
  int threads = 0;
  int tile_size = 16;
  bool use_bvh = true;
  bool operator==(const RenderSettings&) const = default;
};

struct CameraDesc {  // core/camera.hpp:9-17
  Vec3d look_from{0, 0, 0}, look_at{0, 0, -1}, vup{0, 1, 0};
  double vfov_deg = 40, aperture = 0, focus_dist = 0;
  bool operator==(const CameraDesc&) const = default;
};

struct BackgroundDesc {  // core/scene_desc.hpp:35-42
  bool solid = false;
  Vec3d top{0.5, 0.7, 1.0}, bottom{1, 1, 1};
  double intensity = 1;
  bool operator==(const BackgroundDesc&) const = default;
};

struct MaterialDesc {  // materials/material.hpp:14-56, flattened
  std::string name;
  uint32_t kind = RTG_MAT_LAMBERTIAN;
  Vec3d c0{0.5, 0.5, 0.5}, c1{0.5, 0.5, 0.5};
  double p0 = 0;
  bool two_sided = false;
  bool operator==(const MaterialDesc&) const = default;
};

enum class Shape { Sphere, Plane, Triangle, Quad, Box, Mesh };
enum class MeshSource { Icosphere, Torus, Obj };

struct ObjectDesc {  // core/scene_desc.hpp:44-75
  std::string name, material;
  Shape shape = Shape::Sphere;
  // sphere: a = centre, radius; plane: a = point, b = normal;
  // triangle: a b c; quad: a = corner, b = u, c = v; box: a = min, b = max.
  Vec3d a, b, c;
  double radius = 1, rotate_y = 0;
  // mesh
  MeshSource source = MeshSource::Icosphere;
  int subdivisions = 3, segments = 48, segments_v = 0;  // segments_v 0 = segments/2
  double major_radius = 1, minor_radius = 0.35;
  std::string path;
  bool smooth = true;
  Vec3d scale{1, 1, 1}, translate{0, 0, 0};
  bool operator==(const ObjectDesc&) const = default;
};

struct PointLightDesc {
  std::string name;
  Vec3d position, intensity{10, 10, 10};
  bool operator==(const PointLightDesc&) const = default;
};

struct SceneDesc {
  std::string name = "untitled";
  CameraDesc camera;
  RenderSettings render;
  BackgroundDesc background;
  std::vector<MaterialDesc> materials;
  std::vector<ObjectDesc> objects;
  std::vector<PointLightDesc> lights;
  std::string base_dir;  // not serialised
  bool same_as(const SceneDesc& o) const {
    return name == o.name && camera == o.camera && render == o.render && background == o.background &&
           materials == o.materials && objects == o.objects && lights == o.lights;
  }
};

// Raised for any problem in a scene description; the message names the
// offending field, e.g. "objects[2].radius: must be > 0".
class SceneError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

SceneDesc scene_from_json(const Json& j);
SceneDesc parse_scene(const std::string& text);
SceneDesc load_scene(const std::string& path);
Json scene_to_json(const SceneDesc& s);
std::vector<std::string> validate_scene(const SceneDesc& s);
const char* integrator_name(Integrator i);
const char* tone_map_name(ToneMap t);

// ---------------------------------------------------------------------------
struct BvhStats {  // geometry/bvh.hpp:28-32
  uint64_t nodes = 0, leaves = 0, max_depth = 0, max_leaf_size = 0;
  double build_ms = 0, sah_cost = 0;
};

struct SceneInfo {  // core/scene.hpp:25-28
  uint64_t triangles = 0, spheres = 0, planes = 0, degenerate_skipped = 0, lights = 0;
  BvhStats bvh;
};

// Render-ready scene: exactly the arrays rtg_scene_create() consumes.
struct CompiledScene {
  RenderSettings settings;
  CameraDesc camera_desc;  // to set the camera up for another image size
  rtg_camera camera{};
  rtg_background background{};
  std::vector<rtg_material> materials;
  std::vector<rtg_primitive> prims;
  std::vector<rtg_plane> planes;
  std::vector<rtg_point_light> point_lights;
  std::vector<rtg_light_ref> lights;
  std::vector<uint8_t> is_sampled_light;
  std::vector<rtg_bvh_node> nodes;
  std::vector<uint32_t> prim_indices;
  SceneInfo info;

  rtg_scene_desc view() const;
  rtg_render_settings device_settings(int32_t precision) const;
};

CompiledScene compile_scene(const SceneDesc& d);

// Binned-SAH BVH over per-primitive boxes (lo xyz, hi xyz). Leaves with
// more than 0xFFFF primitives throw.
void build_bvh(const std::vector<double>& boxes, int max_leaf, std::vector<rtg_bvh_node>& nodes,
               std::vector<uint32_t>& prim_indices, BvhStats& stats);

rtg_camera setup_camera(const CameraDesc& d, int width, int height);

}  // namespace rtm
