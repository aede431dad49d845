// compile.cpp — SceneDesc -> CompiledScene (src/core/scene.cpp:31-137):
// materials resolved by name; bounded primitives emitted in object order
// (quad -> 2 triangles, box -> 12, mesh -> one per face) with degenerate
// triangles (area < 1e-14) dropped; planes kept apart; lights = point lights
// then emissive primitives; BVH with leaves of 4. Procedural meshes follow
// src/geometry/mesh.cpp (torus with segments_v when given, else segments/2).
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <unordered_map>

#include "scene.hpp"
#include "vec.hpp"

namespace rtm {

namespace {

D3 d3(const Vec3d& v) { return {v.x, v.y, v.z}; }

struct Mesh {
  std::vector<D3> pos, nrm;
  std::vector<std::array<uint32_t, 3>> faces;
};

// mesh.cpp:67-88
Mesh make_torus(double major, double minor, int segments_u, int segments_v) {
  Mesh m;
  const int su = std::max(3, segments_u), sv = std::max(3, segments_v);
  m.pos.reserve((size_t)su * sv);
  m.nrm.reserve((size_t)su * sv);
  for (int i = 0; i < su; ++i) {
    const double u = 2 * kPi * i / su;
    const D3 dir{std::cos(u), 0, std::sin(u)};
    for (int j = 0; j < sv; ++j) {
      const double v = 2 * kPi * j / sv;
      const D3 n = std::cos(v) * dir + D3{0, std::sin(v), 0};
      m.pos.push_back(major * dir + minor * n);
      m.nrm.push_back(n);
    }
  }
  auto at = [&](int i, int j) { return (uint32_t)(((i % su) * sv) + (j % sv)); };
  m.faces.reserve((size_t)su * sv * 2);
  for (int i = 0; i < su; ++i)
    for (int j = 0; j < sv; ++j) {
      m.faces.push_back({at(i, j), at(i, j + 1), at(i + 1, j)});
      m.faces.push_back({at(i + 1, j), at(i, j + 1), at(i + 1, j + 1)});
    }
  return m;
}

// mesh.cpp:45-65: 4-way subdivision of the icosahedron, edge midpoints
// projected to the unit sphere; normals are the positions.
Mesh make_icosphere(int subdivisions) {
  Mesh m;
  const double t = (1 + std::sqrt(5.0)) / 2;
  const D3 base[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (const D3& p : base) m.pos.push_back(normalize(p));
  m.faces = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
             {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
             {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  for (int s = 0; s < subdivisions; ++s) {
    std::unordered_map<uint64_t, uint32_t> mid_of;
    mid_of.reserve(m.faces.size() * 2);
    auto mid = [&](uint32_t a, uint32_t b) {
      const uint64_t key = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
      auto it = mid_of.find(key);
      if (it != mid_of.end()) return it->second;
      m.pos.push_back(normalize(m.pos[a] + m.pos[b]));
      const uint32_t id = (uint32_t)(m.pos.size() - 1);
      mid_of.emplace(key, id);
      return id;
    };
    std::vector<std::array<uint32_t, 3>> next;
    next.reserve(m.faces.size() * 4);
    for (const auto& f : m.faces) {
      const uint32_t a = mid(f[0], f[1]);
      const uint32_t b = mid(f[1], f[2]);
      const uint32_t c = mid(f[2], f[0]);
      next.push_back({f[0], a, c});
      next.push_back({f[1], b, a});
      next.push_back({f[2], c, b});
      next.push_back({a, b, c});
    }
    m.faces.swap(next);
  }
  m.nrm = m.pos;
  return m;
}

// mesh.cpp:20-31 (rotation about +Y after scaling, then translation;
// normals by the inverse transpose).
struct Transform {
  D3 scale{1, 1, 1};
  double rot_deg = 0;
  D3 translate{};
  D3 point(D3 p) const {
    const D3 s = p * scale;
    const double a = rot_deg * kPi / 180, c = std::cos(a), sn = std::sin(a);
    return D3{c * s.x + sn * s.z, s.y, -sn * s.x + c * s.z} + translate;
  }
  D3 normal(D3 n) const {
    const D3 s{n.x / scale.x, n.y / scale.y, n.z / scale.z};
    const double a = rot_deg * kPi / 180, c = std::cos(a), sn = std::sin(a);
    return normalize_or(D3{c * s.x + sn * s.z, s.y, -sn * s.x + c * s.z}, n);
  }
};

}  // namespace

// core/camera.hpp:29-51
rtg_camera setup_camera(const CameraDesc& d, int width, int height) {
  const double aspect = double(width) / double(height);
  const D3 origin = d3(d.look_from);
  const D3 view = d3(d.look_at) - origin;
  const D3 w = normalize_or(-view, D3{0, 0, 1});
  D3 up = normalize_or(d3(d.vup), D3{0, 1, 0});
  if (std::fabs(dot(up, w)) > 0.999999) up = std::fabs(w.y) < 0.9 ? D3{0, 1, 0} : D3{0, 0, 1};
  const D3 u = normalize(cross(up, w));
  const D3 v = cross(w, u);
  const double focus = d.focus_dist > 0 ? d.focus_dist : std::fmax(length(view), 1e-3);
  const double theta = std::clamp(d.vfov_deg, 0.1, 179.0) * kPi / 180;
  const double vp_h = 2 * std::tan(theta / 2) * focus;
  const double vp_w = vp_h * aspect;
  const D3 vp_u = vp_w * u;
  const D3 vp_v = -vp_h * v;
  rtg_camera c{};
  c.origin = to_rtg(origin);
  c.du = to_rtg(vp_u / width);
  c.dv = to_rtg(vp_v / height);
  c.corner = to_rtg(origin - focus * w - vp_u / 2 - vp_v / 2);
  c.u = to_rtg(u);
  c.v = to_rtg(v);
  c.w = to_rtg(w);
  c.lens_radius = d.aperture / 2;
  return c;
}

CompiledScene compile_scene(const SceneDesc& desc) {
  CompiledScene s;
  s.settings = desc.render;
  s.camera_desc = desc.camera;
  s.camera = setup_camera(desc.camera, desc.render.width, desc.render.height);
  s.background.kind = desc.background.solid ? RTG_BG_SOLID : RTG_BG_GRADIENT;
  s.background.top = to_rtg(d3(desc.background.top));
  s.background.bottom = to_rtg(d3(desc.background.bottom));
  s.background.intensity = desc.background.intensity;
  for (const MaterialDesc& m : desc.materials) {
    rtg_material r{};
    r.kind = m.kind;
    r.two_sided = m.two_sided ? 1u : 0u;
    r.c0 = to_rtg(d3(m.c0));
    if (m.kind == RTG_MAT_LAMBERTIAN) r.c1 = to_rtg(d3(m.c1));
    r.p0 = m.p0;
    s.materials.push_back(r);
  }
  auto material_id = [&](const std::string& name) -> uint32_t {
    for (size_t i = 0; i < desc.materials.size(); ++i)
      if (desc.materials[i].name == name) return (uint32_t)i;
    throw std::runtime_error("unknown material '" + name + "'");
  };
  auto add_triangle = [&](D3 a, D3 b, D3 c, const D3* n, uint32_t mat) {
    if (0.5 * length(cross(b - a, c - a)) < 1e-14) {  // shapes.hpp:160-161
      s.info.degenerate_skipped++;
      return;
    }
    rtg_primitive p{};
    p.kind = RTG_PRIM_TRIANGLE;
    p.material = mat;
    p.v0 = to_rtg(a);
    p.v1 = to_rtg(b);
    p.v2 = to_rtg(c);
    if (n) {
      p.has_vertex_normals = 1;
      p.n0 = to_rtg(n[0]);
      p.n1 = to_rtg(n[1]);
      p.n2 = to_rtg(n[2]);
    }
    s.prims.push_back(p);
    s.info.triangles++;
  };

  for (const ObjectDesc& o : desc.objects) {
    const uint32_t mat = material_id(o.material);
    switch (o.shape) {
      case Shape::Sphere: {
        rtg_primitive p{};
        p.kind = RTG_PRIM_SPHERE;
        p.material = mat;
        p.v0 = to_rtg(d3(o.a));
        p.v1 = {o.radius, 0, 0};
        s.prims.push_back(p);
        s.info.spheres++;
        break;
      }
      case Shape::Plane:
        s.planes.push_back({to_rtg(d3(o.a)), to_rtg(normalize_or(d3(o.b), D3{0, 1, 0})), mat, 0});
        s.info.planes++;
        break;
      case Shape::Triangle: add_triangle(d3(o.a), d3(o.b), d3(o.c), nullptr, mat); break;
      case Shape::Quad: {
        const D3 c = d3(o.a), u = d3(o.b), v = d3(o.c);
        add_triangle(c, c + u, c + u + v, nullptr, mat);
        add_triangle(c, c + u + v, c + v, nullptr, mat);
        break;
      }
      case Shape::Box: {
        const D3 lo = d3(o.a), hi = d3(o.b);
        const Transform xf{{1, 1, 1}, o.rotate_y, 0.5 * (lo + hi)};
        const D3 h = 0.5 * (hi - lo);
        auto corner = [&](double sx, double sy, double sz) { return xf.point(D3{sx * h.x, sy * h.y, sz * h.z}); };
        const D3 v[8] = {corner(-1, -1, -1), corner(1, -1, -1), corner(1, 1, -1), corner(-1, 1, -1),
                         corner(-1, -1, 1),  corner(1, -1, 1),  corner(1, 1, 1),  corner(-1, 1, 1)};
        // Faces as corner quads ordered so u x v points outward.
        static const int faces[6][4] = {{4, 5, 6, 7}, {1, 0, 3, 2}, {5, 1, 2, 6},
                                        {0, 4, 7, 3}, {7, 6, 2, 3}, {0, 1, 5, 4}};
        for (const auto& f : faces) {
          add_triangle(v[f[0]], v[f[1]], v[f[2]], nullptr, mat);
          add_triangle(v[f[0]], v[f[2]], v[f[3]], nullptr, mat);
        }
        break;
      }
      case Shape::Mesh: {
        Mesh m;
        if (o.source == MeshSource::Torus) {
          const int sv = o.segments_v > 0 ? o.segments_v : std::max(3, o.segments / 2);
          m = make_torus(o.major_radius, o.minor_radius, o.segments, sv);
        } else if (o.source == MeshSource::Icosphere) {
          m = make_icosphere(o.subdivisions);
        } else {
          throw std::runtime_error("obj meshes are not supported by this build (" + o.path + ")");
        }
        const Transform xf{d3(o.scale), o.rotate_y, d3(o.translate)};
        const bool smooth = o.smooth && !m.nrm.empty();
        s.prims.reserve(s.prims.size() + m.faces.size());
        for (const auto& f : m.faces) {
          const D3 a = xf.point(m.pos[f[0]]), b = xf.point(m.pos[f[1]]), c = xf.point(m.pos[f[2]]);
          if (smooth) {
            const D3 n[3] = {xf.normal(m.nrm[f[0]]), xf.normal(m.nrm[f[1]]), xf.normal(m.nrm[f[2]])};
            add_triangle(a, b, c, n, mat);
          } else {
            add_triangle(a, b, c, nullptr, mat);
          }
        }
        break;
      }
    }
  }
  for (size_t i = 0; i < desc.lights.size(); ++i) {
    s.lights.push_back({RTG_LIGHT_POINT, (uint32_t)s.point_lights.size()});
    s.point_lights.push_back({to_rtg(d3(desc.lights[i].position)), to_rtg(d3(desc.lights[i].intensity))});
  }
  s.is_sampled_light.assign(s.prims.size(), 0);
  for (size_t i = 0; i < s.prims.size(); ++i) {
    const rtg_material& m = s.materials[s.prims[i].material];
    const D3 em = from_rtg(m.c0) * m.p0;
    if (m.kind == RTG_MAT_EMISSIVE && !(em.x == 0 && em.y == 0 && em.z == 0)) {
      s.lights.push_back({RTG_LIGHT_EMITTER, (uint32_t)i});
      s.is_sampled_light[i] = 1;
    }
  }
  s.info.lights = s.lights.size();
  // Primitive bounds (shapes.hpp:68-71, :152-159).
  std::vector<double> boxes(6 * s.prims.size());
  for (size_t i = 0; i < s.prims.size(); ++i) {
    const rtg_primitive& p = s.prims[i];
    D3 lo, hi;
    if (p.kind == RTG_PRIM_SPHERE) {
      const double r = std::fabs(p.v1.x);
      lo = from_rtg(p.v0) - D3{r, r, r};
      hi = from_rtg(p.v0) + D3{r, r, r};
    } else {
      lo = vmin(vmin(D3{INFINITY, INFINITY, INFINITY}, from_rtg(p.v0)), vmin(from_rtg(p.v1), from_rtg(p.v2)));
      hi = vmax(vmax(D3{-INFINITY, -INFINITY, -INFINITY}, from_rtg(p.v0)), vmax(from_rtg(p.v1), from_rtg(p.v2)));
      lo = lo - D3{1e-9, 1e-9, 1e-9};
      hi = hi + D3{1e-9, 1e-9, 1e-9};
    }
    std::memcpy(&boxes[6 * i], &lo, sizeof lo);
    std::memcpy(&boxes[6 * i + 3], &hi, sizeof hi);
  }
  build_bvh(boxes, 4, s.nodes, s.prim_indices, s.info.bvh);
  return s;
}

rtg_scene_desc CompiledScene::view() const {
  rtg_scene_desc d{};
  d.prims = prims.data();
  d.n_prims = (uint32_t)prims.size();
  d.is_sampled_light = is_sampled_light.data();
  d.planes = planes.data();
  d.n_planes = (uint32_t)planes.size();
  d.materials = materials.data();
  d.n_materials = (uint32_t)materials.size();
  d.point_lights = point_lights.data();
  d.n_point_lights = (uint32_t)point_lights.size();
  d.lights = lights.data();
  d.n_lights = (uint32_t)lights.size();
  d.nodes = nodes.data();
  d.n_nodes = (uint32_t)nodes.size();
  d.prim_indices = prim_indices.data();
  d.n_prim_indices = (uint32_t)prim_indices.size();
  d.camera = camera;
  d.background = background;
  return d;
}

rtg_render_settings CompiledScene::device_settings(int32_t precision) const {
  rtg_render_settings r{};
  r.width = settings.width;
  r.height = settings.height;
  r.max_depth = settings.max_depth;
  r.rr_min_depth = settings.rr_min_depth;
  r.integrator = (int32_t)settings.integrator;
  r.light_sampling = settings.light_sampling ? 1 : 0;
  r.use_bvh = settings.use_bvh ? 1 : 0;
  r.precision = precision;
  r.seed = settings.seed;
  return r;
}

}  // namespace rtm
