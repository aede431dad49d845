// scene_io.cpp — JSON <-> SceneDesc with field-path errors and semantic
// validation. The accepted keys, defaults and error texts are the scene
// format's contract (src/io/scene_io.cpp:101-478), so they are kept as the
// reference states them; a torus may additionally carry "segments_v".
#include <cmath>
#include <filesystem>
#include <fstream>
#include <sstream>

#include "scene.hpp"

namespace rtm {

const char* integrator_name(Integrator i) {
  switch (i) {
    case Integrator::Direct: return "direct";
    case Integrator::Normals: return "normals";
    default: return "path";
  }
}
const char* tone_map_name(ToneMap t) {
  switch (t) {
    case ToneMap::Clamp: return "clamp";
    case ToneMap::Reinhard: return "reinhard";
    default: return "aces";
  }
}

namespace {

// A JSON object being read, and its path in the document ("objects[2]",
// "render", "" for the root) for error messages.
struct Scope {
  Scope(const Json& j, std::string p) : node(j), path(std::move(p)) {
    if (!node.is_object()) throw SceneError((path.empty() ? std::string("scene") : path) + ": expected an object");
  }
  std::string where(const std::string& key) const { return path.empty() ? key : path + "." + key; }
  const Json* find(const std::string& key) const { return node.get(key); }
  const Json& node;
  std::string path;
};

// Conversion of a present value to T; `at` names it in the error.
template <class T>
struct As;
template <>
struct As<double> {
  static double from(const Json& v, const std::string& at) {
    if (!v.is_number()) throw SceneError(at + ": expected a number");
    return v.as_number();
  }
};
template <>
struct As<int> {  // a whole number of magnitude at most 2e9
  static int from(const Json& v, const std::string& at) {
    if (!v.is_number() || v.as_number() != std::floor(v.as_number()) || std::fabs(v.as_number()) > 2e9)
      throw SceneError(at + ": expected an integer");
    return (int)v.as_number();
  }
};
template <>
struct As<bool> {
  static bool from(const Json& v, const std::string& at) {
    if (!v.is_bool()) throw SceneError(at + ": expected true/false");
    return v.as_bool();
  }
};
template <>
struct As<std::string> {
  static std::string from(const Json& v, const std::string& at) {
    if (!v.is_string()) throw SceneError(at + ": expected a string");
    return v.as_string();
  }
};
template <>
struct As<Vec3d> {  // [x, y, z], or one number for all three
  static Vec3d from(const Json& v, const std::string& at) {
    if (v.is_number()) return {v.as_number(), v.as_number(), v.as_number()};
    if (!v.is_array() || v.items().size() != 3) throw SceneError(at + ": expected [x, y, z]");
    double c[3];
    for (size_t i = 0; i < 3; ++i) {
      if (!v.items()[i].is_number()) throw SceneError(at + ": expected [x, y, z] of numbers");
      c[i] = v.items()[i].as_number();
    }
    return {c[0], c[1], c[2]};
  }
};

// An optional field (the fallback when absent) and a required one.
template <class T>
T opt(const Scope& s, const std::string& key, const T& fallback) {
  const Json* v = s.find(key);
  return v ? As<T>::from(*v, s.where(key)) : fallback;
}
template <class T>
T req(const Scope& s, const std::string& key) {
  const Json* v = s.find(key);
  if (!v) throw SceneError(s.where(key) + ": required");
  return As<T>::from(*v, s.where(key));
}

MaterialDesc read_material(const std::string& name, const Scope& f) {
  MaterialDesc m;
  m.name = name;
  const std::string type = req<std::string>(f, "type");
  if (type == "lambertian" || type == "diffuse") {
    m.kind = RTG_MAT_LAMBERTIAN;
    m.c0 = opt<Vec3d>(f, "albedo", {0.5, 0.5, 0.5});
    m.c1 = opt<Vec3d>(f, "albedo2", m.c0);
    m.p0 = opt<double>(f, "checker_scale", 0);
  } else if (type == "metal") {
    m.kind = RTG_MAT_METAL;
    m.c0 = opt<Vec3d>(f, "albedo", {0.9, 0.9, 0.9});
    m.c1 = {};
    m.p0 = opt<double>(f, "roughness", 0);
  } else if (type == "dielectric" || type == "glass") {
    m.kind = RTG_MAT_DIELECTRIC;
    m.p0 = opt<double>(f, "ior", 1.5);
    m.c0 = opt<Vec3d>(f, "tint", {1, 1, 1});
    m.c1 = {};
  } else if (type == "emissive" || type == "light") {
    m.kind = RTG_MAT_EMISSIVE;
    m.c0 = opt<Vec3d>(f, "color", {1, 1, 1});
    m.c1 = {};
    m.p0 = opt<double>(f, "intensity", 1);
    m.two_sided = opt<bool>(f, "two_sided", false);
  } else {
    throw SceneError(f.where("type") + ": unknown material type '" + type + "'");
  }
  return m;
}

ObjectDesc read_object(const Scope& f) {
  ObjectDesc o;
  const std::string type = req<std::string>(f, "type");
  o.name = opt<std::string>(f, "name", type);
  o.material = req<std::string>(f, "material");
  if (type == "sphere") {
    o.shape = Shape::Sphere;
    o.a = req<Vec3d>(f, "center");
    o.radius = opt<double>(f, "radius", 1);
  } else if (type == "plane") {
    o.shape = Shape::Plane;
    o.a = opt<Vec3d>(f, "point", {});
    o.b = opt<Vec3d>(f, "normal", {0, 1, 0});
  } else if (type == "triangle") {
    o.shape = Shape::Triangle;
    o.a = req<Vec3d>(f, "v0");
    o.b = req<Vec3d>(f, "v1");
    o.c = req<Vec3d>(f, "v2");
  } else if (type == "quad") {
    o.shape = Shape::Quad;
    o.a = req<Vec3d>(f, "corner");
    o.b = req<Vec3d>(f, "u");
    o.c = req<Vec3d>(f, "v");
  } else if (type == "box") {
    o.shape = Shape::Box;
    o.a = req<Vec3d>(f, "min");
    o.b = req<Vec3d>(f, "max");
    o.rotate_y = opt<double>(f, "rotate_y", 0);
  } else if (type == "mesh") {
    o.shape = Shape::Mesh;
    const std::string src = opt<std::string>(f, "source", "icosphere");
    if (src == "icosphere") o.source = MeshSource::Icosphere;
    else if (src == "torus") o.source = MeshSource::Torus;
    else if (src == "obj") o.source = MeshSource::Obj;
    else throw SceneError(f.where("source") + ": expected icosphere, torus or obj");
    o.subdivisions = opt<int>(f, "subdivisions", o.subdivisions);
    o.major_radius = opt<double>(f, "major_radius", o.major_radius);
    o.minor_radius = opt<double>(f, "minor_radius", o.minor_radius);
    o.segments = opt<int>(f, "segments", o.segments);
    o.segments_v = opt<int>(f, "segments_v", 0);
    o.path = opt<std::string>(f, "path", "");
    o.smooth = opt<bool>(f, "smooth", true);
    o.scale = opt<Vec3d>(f, "scale", {1, 1, 1});
    o.rotate_y = opt<double>(f, "rotate_y", 0);
    o.translate = opt<Vec3d>(f, "translate", {0, 0, 0});
  } else {
    throw SceneError(f.where("type") + ": unknown object type '" + type + "'");
  }
  return o;
}

Json vj(const Vec3d& v) {
  Json a = Json::array();
  a.push(Json::number(v.x));
  a.push(Json::number(v.y));
  a.push(Json::number(v.z));
  return a;
}

}  // namespace

SceneDesc scene_from_json(const Json& root) {
  SceneDesc s;
  const Scope r(root, "");
  static const char* known[] = {"name", "camera", "render", "background", "materials", "objects", "lights", "$schema",
                                "description"};
  for (const auto& kv : root.members()) {
    bool ok = false;
    for (const char* k : known) ok = ok || kv.first == k;
    if (!ok) throw SceneError("scene: unknown top-level key '" + kv.first + "'");
  }
  s.name = opt<std::string>(r, "name", s.name);
  if (const Json* cj = r.find("camera")) {
    const Scope c(*cj, "camera");
    s.camera.look_from = opt<Vec3d>(c, "look_from", s.camera.look_from);
    s.camera.look_at = opt<Vec3d>(c, "look_at", s.camera.look_at);
    s.camera.vup = opt<Vec3d>(c, "vup", s.camera.vup);
    s.camera.vfov_deg = opt<double>(c, "vfov", s.camera.vfov_deg);
    s.camera.aperture = opt<double>(c, "aperture", 0);
    s.camera.focus_dist = opt<double>(c, "focus_dist", 0);
  }
  if (const Json* rj = r.find("render")) {
    const Scope c(*rj, "render");
    RenderSettings& R = s.render;
    R.width = opt<int>(c, "width", R.width);
    R.height = opt<int>(c, "height", R.height);
    R.spp = opt<int>(c, "spp", R.spp);
    R.max_depth = opt<int>(c, "max_depth", R.max_depth);
    R.rr_min_depth = opt<int>(c, "rr_min_depth", R.rr_min_depth);
    R.light_sampling = opt<bool>(c, "light_sampling", R.light_sampling);
    R.seed = (uint64_t)opt<int>(c, "seed", (int)R.seed);
    R.exposure = opt<double>(c, "exposure", R.exposure);
    R.threads = opt<int>(c, "threads", R.threads);
    R.tile_size = opt<int>(c, "tile_size", R.tile_size);
    R.use_bvh = opt<bool>(c, "use_bvh", R.use_bvh);
    const std::string integ = opt<std::string>(c, "integrator", integrator_name(R.integrator));
    if (integ == "path") R.integrator = Integrator::Path;
    else if (integ == "direct") R.integrator = Integrator::Direct;
    else if (integ == "normals") R.integrator = Integrator::Normals;
    else throw SceneError("render.integrator: expected path, direct or normals");
    const std::string tm = opt<std::string>(c, "tone_map", tone_map_name(R.tone_map));
    if (tm == "clamp") R.tone_map = ToneMap::Clamp;
    else if (tm == "reinhard") R.tone_map = ToneMap::Reinhard;
    else if (tm == "aces") R.tone_map = ToneMap::Aces;
    else throw SceneError("render.tone_map: expected clamp, reinhard or aces");
  }
  if (const Json* bj = r.find("background")) {
    const Scope b(*bj, "background");
    const std::string kind = opt<std::string>(b, "type", "gradient");
    if (kind == "gradient") s.background.solid = false;
    else if (kind == "solid") s.background.solid = true;
    else throw SceneError("background.type: expected gradient or solid");
    s.background.top = opt<Vec3d>(b, kind == "solid" ? "color" : "top", s.background.top);
    s.background.bottom = opt<Vec3d>(b, "bottom", s.background.bottom);
    s.background.intensity = opt<double>(b, "intensity", 1);
  }
  if (const Json* mj = r.find("materials")) {
    if (!mj->is_object()) throw SceneError("materials: expected an object of name -> material");
    for (const auto& kv : mj->members()) s.materials.push_back(read_material(kv.first, Scope(kv.second, "materials." + kv.first)));
  }
  if (const Json* oj = r.find("objects")) {
    if (!oj->is_array()) throw SceneError("objects: expected an array");
    const auto& a = oj->items();
    for (size_t i = 0; i < a.size(); ++i) s.objects.push_back(read_object(Scope(a[i], "objects[" + std::to_string(i) + "]")));
  }
  if (const Json* lj = r.find("lights")) {
    if (!lj->is_array()) throw SceneError("lights: expected an array");
    const auto& a = lj->items();
    for (size_t i = 0; i < a.size(); ++i) {
      const Scope l(a[i], "lights[" + std::to_string(i) + "]");
      const std::string type = opt<std::string>(l, "type", "point");
      if (type != "point")
        throw SceneError(l.where("type") + ": only 'point' lights are supported (use emissive materials for area lights)");
      PointLightDesc p;
      p.name = opt<std::string>(l, "name", "light" + std::to_string(i));
      p.position = req<Vec3d>(l, "position");
      p.intensity = opt<Vec3d>(l, "intensity", {10, 10, 10});
      s.lights.push_back(p);
    }
  }
  const auto errors = validate_scene(s);
  if (!errors.empty()) {
    std::string msg = errors.front();
    if (errors.size() > 1) msg += " (and " + std::to_string(errors.size() - 1) + " more problem(s))";
    throw SceneError(msg);
  }
  return s;
}

std::vector<std::string> validate_scene(const SceneDesc& s) {
  std::vector<std::string> e;
  auto finite3 = [](const Vec3d& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); };
  auto len = [](double x, double y, double z) { return std::sqrt(x * x + y * y + z * z); };
  auto finite = [&](const Vec3d& v, const std::string& at) {
    if (!finite3(v)) e.push_back(at + ": must be finite");
  };
  auto colour = [&](const Vec3d& v, const std::string& at) {
    finite(v, at);
    if (v.x < 0 || v.y < 0 || v.z < 0) e.push_back(at + ": colour components must be >= 0");
  };
  const RenderSettings& R = s.render;
  if (R.width < 1 || R.width > 8192) e.push_back("render.width: must be in [1, 8192]");
  if (R.height < 1 || R.height > 8192) e.push_back("render.height: must be in [1, 8192]");
  if (R.spp < 1 || R.spp > 1000000) e.push_back("render.spp: must be in [1, 1000000]");
  if (R.max_depth < 1 || R.max_depth > 1000) e.push_back("render.max_depth: must be in [1, 1000]");
  if (R.rr_min_depth < 0) e.push_back("render.rr_min_depth: must be >= 0");
  if (R.threads < 0 || R.threads > 1024) e.push_back("render.threads: must be in [0, 1024]");
  if (R.tile_size < 1 || R.tile_size > 1024) e.push_back("render.tile_size: must be in [1, 1024]");
  if (!(R.exposure > 0) || !std::isfinite(R.exposure)) e.push_back("render.exposure: must be > 0");

  const CameraDesc& c = s.camera;
  finite(c.look_from, "camera.look_from");
  finite(c.look_at, "camera.look_at");
  if (len(c.look_at.x - c.look_from.x, c.look_at.y - c.look_from.y, c.look_at.z - c.look_from.z) < 1e-9)
    e.push_back("camera: look_from and look_at must differ");
  if (!(c.vfov_deg > 0 && c.vfov_deg < 180)) e.push_back("camera.vfov: must be in (0, 180)");
  if (c.aperture < 0) e.push_back("camera.aperture: must be >= 0");
  if (c.focus_dist < 0) e.push_back("camera.focus_dist: must be >= 0");
  colour(s.background.top, "background.top");
  colour(s.background.bottom, "background.bottom");

  for (size_t i = 0; i < s.materials.size(); ++i) {
    const MaterialDesc& m = s.materials[i];
    const std::string at = "materials." + m.name;
    if (m.name.empty()) e.push_back("materials: material names must be non-empty");
    for (size_t k = 0; k < i; ++k)
      if (s.materials[k].name == m.name) e.push_back(at + ": duplicate material name");
    switch (m.kind) {
      case RTG_MAT_LAMBERTIAN:
        colour(m.c0, at + ".albedo");
        colour(m.c1, at + ".albedo2");
        if (m.p0 < 0) e.push_back(at + ".checker_scale: must be >= 0");
        break;
      case RTG_MAT_METAL:
        colour(m.c0, at + ".albedo");
        if (m.p0 < 0 || m.p0 > 1) e.push_back(at + ".roughness: must be in [0, 1]");
        break;
      case RTG_MAT_DIELECTRIC:
        if (!(m.p0 >= 1.0 && m.p0 <= 4.0)) e.push_back(at + ".ior: must be in [1, 4]");
        colour(m.c0, at + ".tint");
        break;
      default:
        colour(m.c0, at + ".color");
        if (!(m.p0 >= 0) || !std::isfinite(m.p0)) e.push_back(at + ".intensity: must be >= 0");
    }
  }
  auto find_material = [&](const std::string& n) {
    for (const auto& m : s.materials)
      if (m.name == n) return true;
    return false;
  };
  for (size_t i = 0; i < s.objects.size(); ++i) {
    const ObjectDesc& o = s.objects[i];
    const std::string at = "objects[" + std::to_string(i) + "]";
    if (!find_material(o.material)) e.push_back(at + ".material: unknown material '" + o.material + "'");
    switch (o.shape) {
      case Shape::Sphere:
        finite(o.a, at + ".center");
        if (!(o.radius > 0) || !std::isfinite(o.radius)) e.push_back(at + ".radius: must be > 0");
        break;
      case Shape::Plane:
        finite(o.a, at + ".point");
        finite(o.b, at + ".normal");
        if (len(o.b.x, o.b.y, o.b.z) < 1e-9) e.push_back(at + ".normal: must be non-zero");
        break;
      case Shape::Triangle: {
        finite(o.a, at + ".v0");
        finite(o.b, at + ".v1");
        finite(o.c, at + ".v2");
        const double ux = o.b.x - o.a.x, uy = o.b.y - o.a.y, uz = o.b.z - o.a.z;
        const double vx = o.c.x - o.a.x, vy = o.c.y - o.a.y, vz = o.c.z - o.a.z;
        if (len(uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx) < 1e-12)
          e.push_back(at + ": triangle is degenerate");
        break;
      }
      case Shape::Quad: {
        finite(o.a, at + ".corner");
        finite(o.b, at + ".u");
        finite(o.c, at + ".v");
        const Vec3d& u = o.b;
        const Vec3d& v = o.c;
        if (len(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x) < 1e-12)
          e.push_back(at + ": quad is degenerate (u parallel to v)");
        break;
      }
      case Shape::Box:
        finite(o.a, at + ".min");
        finite(o.b, at + ".max");
        if (!(o.b.x > o.a.x && o.b.y > o.a.y && o.b.z > o.a.z))
          e.push_back(at + ": box max must exceed min on every axis");
        break;
      case Shape::Mesh:
        if (o.subdivisions < 0 || o.subdivisions > 7) e.push_back(at + ".subdivisions: must be in [0, 7]");
        if (o.segments < 3 || o.segments > 2048) e.push_back(at + ".segments: must be in [3, 2048]");
        if (o.segments_v != 0 && (o.segments_v < 3 || o.segments_v > 2048))
          e.push_back(at + ".segments_v: must be in [3, 2048]");
        if (!(o.major_radius > 0) || !(o.minor_radius > 0)) e.push_back(at + ": torus radii must be > 0");
        if (o.source == MeshSource::Obj && o.path.empty()) e.push_back(at + ".path: required for obj meshes");
        finite(o.scale, at + ".scale");
        finite(o.translate, at + ".translate");
        if (std::fabs(o.scale.x * o.scale.y * o.scale.z) < 1e-12) e.push_back(at + ".scale: components must be non-zero");
        break;
    }
  }
  for (size_t i = 0; i < s.lights.size(); ++i) {
    finite(s.lights[i].position, "lights[" + std::to_string(i) + "].position");
    colour(s.lights[i].intensity, "lights[" + std::to_string(i) + "].intensity");
  }
  return e;
}

Json scene_to_json(const SceneDesc& s) {
  Json j = Json::object();
  j.set("name", Json::string(s.name));
  Json c = Json::object();
  c.set("look_from", vj(s.camera.look_from));
  c.set("look_at", vj(s.camera.look_at));
  c.set("vup", vj(s.camera.vup));
  c.set("vfov", Json::number(s.camera.vfov_deg));
  c.set("aperture", Json::number(s.camera.aperture));
  c.set("focus_dist", Json::number(s.camera.focus_dist));
  j.set("camera", c);
  const RenderSettings& R = s.render;
  Json r = Json::object();
  r.set("width", Json::number(R.width));
  r.set("height", Json::number(R.height));
  r.set("spp", Json::number(R.spp));
  r.set("max_depth", Json::number(R.max_depth));
  r.set("rr_min_depth", Json::number(R.rr_min_depth));
  r.set("integrator", Json::string(integrator_name(R.integrator)));
  r.set("light_sampling", Json::boolean(R.light_sampling));
  r.set("seed", Json::number((double)R.seed));
  r.set("tone_map", Json::string(tone_map_name(R.tone_map)));
  r.set("exposure", Json::number(R.exposure));
  r.set("threads", Json::number(R.threads));
  r.set("tile_size", Json::number(R.tile_size));
  r.set("use_bvh", Json::boolean(R.use_bvh));
  j.set("render", r);
  Json b = Json::object();
  if (s.background.solid) {
    b.set("type", Json::string("solid"));
    b.set("color", vj(s.background.top));
  } else {
    b.set("type", Json::string("gradient"));
    b.set("top", vj(s.background.top));
    b.set("bottom", vj(s.background.bottom));
  }
  b.set("intensity", Json::number(s.background.intensity));
  j.set("background", b);
  Json mats = Json::object();
  for (const MaterialDesc& m : s.materials) {
    Json mj = Json::object();
    switch (m.kind) {
      case RTG_MAT_LAMBERTIAN:
        mj.set("type", Json::string("lambertian"));
        mj.set("albedo", vj(m.c0));
        if (m.p0 > 0) {
          mj.set("albedo2", vj(m.c1));
          mj.set("checker_scale", Json::number(m.p0));
        }
        break;
      case RTG_MAT_METAL:
        mj.set("type", Json::string("metal"));
        mj.set("albedo", vj(m.c0));
        mj.set("roughness", Json::number(m.p0));
        break;
      case RTG_MAT_DIELECTRIC:
        mj.set("type", Json::string("dielectric"));
        mj.set("ior", Json::number(m.p0));
        mj.set("tint", vj(m.c0));
        break;
      default:
        mj.set("type", Json::string("emissive"));
        mj.set("color", vj(m.c0));
        mj.set("intensity", Json::number(m.p0));
        if (m.two_sided) mj.set("two_sided", Json::boolean(true));
    }
    mats.set(m.name, mj);
  }
  j.set("materials", mats);
  Json objs = Json::array();
  for (const ObjectDesc& o : s.objects) {
    Json oj = Json::object();
    switch (o.shape) {
      case Shape::Sphere:
        oj.set("type", Json::string("sphere"));
        oj.set("name", Json::string(o.name));
        oj.set("center", vj(o.a));
        oj.set("radius", Json::number(o.radius));
        break;
      case Shape::Plane:
        oj.set("type", Json::string("plane"));
        oj.set("name", Json::string(o.name));
        oj.set("point", vj(o.a));
        oj.set("normal", vj(o.b));
        break;
      case Shape::Triangle:
        oj.set("type", Json::string("triangle"));
        oj.set("name", Json::string(o.name));
        oj.set("v0", vj(o.a));
        oj.set("v1", vj(o.b));
        oj.set("v2", vj(o.c));
        break;
      case Shape::Quad:
        oj.set("type", Json::string("quad"));
        oj.set("name", Json::string(o.name));
        oj.set("corner", vj(o.a));
        oj.set("u", vj(o.b));
        oj.set("v", vj(o.c));
        break;
      case Shape::Box:
        oj.set("type", Json::string("box"));
        oj.set("name", Json::string(o.name));
        oj.set("min", vj(o.a));
        oj.set("max", vj(o.b));
        oj.set("rotate_y", Json::number(o.rotate_y));
        break;
      case Shape::Mesh:
        oj.set("type", Json::string("mesh"));
        oj.set("name", Json::string(o.name));
        if (o.source == MeshSource::Icosphere) {
          oj.set("source", Json::string("icosphere"));
          oj.set("subdivisions", Json::number(o.subdivisions));
        } else if (o.source == MeshSource::Torus) {
          oj.set("source", Json::string("torus"));
          oj.set("major_radius", Json::number(o.major_radius));
          oj.set("minor_radius", Json::number(o.minor_radius));
          oj.set("segments", Json::number(o.segments));
          if (o.segments_v) oj.set("segments_v", Json::number(o.segments_v));
        } else {
          oj.set("source", Json::string("obj"));
          oj.set("path", Json::string(o.path));
        }
        oj.set("smooth", Json::boolean(o.smooth));
        oj.set("scale", vj(o.scale));
        oj.set("rotate_y", Json::number(o.rotate_y));
        oj.set("translate", vj(o.translate));
        break;
    }
    oj.set("material", Json::string(o.material));
    objs.push(oj);
  }
  j.set("objects", objs);
  Json lights = Json::array();
  for (const PointLightDesc& l : s.lights) {
    Json lj = Json::object();
    lj.set("type", Json::string("point"));
    lj.set("name", Json::string(l.name));
    lj.set("position", vj(l.position));
    lj.set("intensity", vj(l.intensity));
    lights.push(lj);
  }
  j.set("lights", lights);
  return j;
}

SceneDesc parse_scene(const std::string& text) {
  try {
    return scene_from_json(Json::parse(text));
  } catch (const JsonError& e) {
    throw SceneError(e.what());
  }
}

SceneDesc load_scene(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw SceneError("cannot open scene file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  SceneDesc s = parse_scene(ss.str());
  s.base_dir = std::filesystem::path(path).parent_path().string();
  return s;
}

}  // namespace rtm
