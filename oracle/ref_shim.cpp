// ref_shim.cpp — a C ABI over the REFERENCE renderer, compiled together with
// the reference's own sources (left where they lie, never copied) into
// oracle/_ref/librtref.so by oracle/Makefile. Test infrastructure only: it
// lets the tests and the CPU baseline call the reference itself.
//
// The one place this file does more than forward calls is the compile path
// for torus meshes whose JSON carries the "segments_v" extension key (the
// reference's Scene::compile always uses segments/2, scene.cpp:87). There it
// assembles the Scene from the reference's own make_torus, MeshTransform,
// Triangle, Bvh::build; the tests check that path against Scene::compile on
// scenes where segments_v == segments/2.
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "../include/rtg.h"
#include "raytracer/core/scene.hpp"
#include "raytracer/geometry/mesh.hpp"
#include "raytracer/io/json.hpp"
#include "raytracer/io/scene_io.hpp"
#include "raytracer/rendering/framebuffer.hpp"
#include "raytracer/rendering/integrator.hpp"
#include "raytracer/rendering/renderer.hpp"

using namespace rt;

namespace {

struct RefScene {
  SceneDesc desc;
  Scene scene;
  std::vector<rtg_primitive> prims;
  std::vector<rtg_plane> planes;
  std::vector<rtg_material> materials;
  std::vector<rtg_point_light> point_lights;
  std::vector<rtg_light_ref> lights;
  std::vector<rtg_bvh_node> nodes;
  std::vector<uint32_t> prim_indices;
  std::vector<uint8_t> sampled;
  rtg_scene_desc out{};
};

rtg_vec3 V3(const Vec3& v) { return {v.x, v.y, v.z}; }

void export_scene(RefScene& r) {
  const Scene& s = r.scene;
  for (const auto& p : s.primitives) {
    rtg_primitive q{};
    if (const auto* sp = std::get_if<Sphere>(&p)) {
      q.kind = RTG_PRIM_SPHERE;
      q.material = sp->material;
      q.v0 = V3(sp->center);
      q.v1 = {sp->radius, 0, 0};
    } else {
      const auto& t = std::get<Triangle>(p);
      q.kind = RTG_PRIM_TRIANGLE;
      q.material = t.material;
      q.has_vertex_normals = t.has_vertex_normals ? 1u : 0u;
      q.v0 = V3(t.v0); q.v1 = V3(t.v1); q.v2 = V3(t.v2);
      if (t.has_vertex_normals) { q.n0 = V3(t.n0); q.n1 = V3(t.n1); q.n2 = V3(t.n2); }
    }
    r.prims.push_back(q);
  }
  for (const auto& pl : s.planes) r.planes.push_back({V3(pl.point), V3(pl.normal), pl.material, 0});
  for (const auto& m : s.materials) {
    rtg_material q{};
    if (const auto* l = std::get_if<Lambertian>(&m)) {
      q.kind = RTG_MAT_LAMBERTIAN; q.c0 = V3(l->albedo.a); q.c1 = V3(l->albedo.b); q.p0 = l->albedo.checker_scale;
    } else if (const auto* me = std::get_if<Metal>(&m)) {
      q.kind = RTG_MAT_METAL; q.c0 = V3(me->albedo); q.p0 = me->roughness;
    } else if (const auto* d = std::get_if<Dielectric>(&m)) {
      q.kind = RTG_MAT_DIELECTRIC; q.c0 = V3(d->tint); q.p0 = d->ior;
    } else {
      const auto& e = std::get<Emissive>(m);
      q.kind = RTG_MAT_EMISSIVE; q.c0 = V3(e.color); q.p0 = e.intensity; q.two_sided = e.two_sided;
    }
    r.materials.push_back(q);
  }
  for (const auto& pl : s.point_lights) r.point_lights.push_back({V3(pl.position), V3(pl.intensity)});
  for (const auto& l : s.lights)
    r.lights.push_back({l.kind == LightRef::Kind::Point ? (uint32_t)RTG_LIGHT_POINT : (uint32_t)RTG_LIGHT_EMITTER, l.index});
  for (const auto& n : s.bvh.nodes()) {
    rtg_bvh_node q{};
    q.lo = V3(n.bounds.min); q.hi = V3(n.bounds.max);
    q.offset = n.offset; q.count = n.count; q.axis = n.axis;
    r.nodes.push_back(q);
  }
  r.prim_indices = s.bvh.prim_indices();
  r.sampled = s.is_sampled_light;
  rtg_scene_desc& o = r.out;
  o.prims = r.prims.data(); o.n_prims = (uint32_t)r.prims.size();
  o.is_sampled_light = r.sampled.data();
  o.planes = r.planes.data(); o.n_planes = (uint32_t)r.planes.size();
  o.materials = r.materials.data(); o.n_materials = (uint32_t)r.materials.size();
  o.point_lights = r.point_lights.data(); o.n_point_lights = (uint32_t)r.point_lights.size();
  o.lights = r.lights.data(); o.n_lights = (uint32_t)r.lights.size();
  o.nodes = r.nodes.data(); o.n_nodes = (uint32_t)r.nodes.size();
  o.prim_indices = r.prim_indices.data(); o.n_prim_indices = (uint32_t)r.prim_indices.size();
  // Camera corner/du/dv are private in the reference: the tests rebuild the
  // camera from ref_scene_camera_desc() and pin it against ref_camera_rays().
  const Camera& c = s.camera;
  o.camera.origin = V3(c.origin());
  o.camera.u = V3(c.u()); o.camera.v = V3(c.v()); o.camera.w = V3(c.w());
  o.background.kind = s.background.kind == Background::Kind::Solid ? RTG_BG_SOLID : RTG_BG_GRADIENT;
  o.background.top = V3(s.background.top); o.background.bottom = V3(s.background.bottom);
  o.background.intensity = s.background.intensity;
}

// Scene::compile with "segments_v" honoured on torus meshes (see header).
Scene compile_with_segments_v(const SceneDesc& desc, const std::vector<int>& seg_v) {
  SceneDesc stripped = desc;
  stripped.objects.clear();
  Scene s = Scene::compile(stripped);  // settings, camera, background, materials
  s.primitives.clear();
  s.planes.clear();
  s.lights.clear();
  s.point_lights.clear();
  auto mat_id = [&](const std::string& n) -> MaterialId {
    for (std::size_t i = 0; i < desc.materials.size(); ++i)
      if (desc.materials[i].name == n) return (MaterialId)i;
    throw std::runtime_error("unknown material '" + n + "'");
  };
  auto add = [&](Triangle t) {
    if (t.degenerate()) { s.info.degenerate_skipped++; return; }
    s.primitives.emplace_back(t);
  };
  for (std::size_t oi = 0; oi < desc.objects.size(); ++oi) {
    const auto& obj = desc.objects[oi];
    const MaterialId mat = mat_id(obj.material);
    if (const auto* m = std::get_if<MeshDesc>(&obj.shape); m && m->source == MeshDesc::Source::Torus && seg_v[oi] > 0) {
      TriangleMesh mesh = make_torus(m->major_radius, m->minor_radius, m->segments, seg_v[oi]);
      const MeshTransform xf{m->scale, m->rotate_y, m->translate};
      for (const auto& f : mesh.faces) {
        Triangle t;
        t.v0 = xf.apply_point(mesh.positions[f[0]]);
        t.v1 = xf.apply_point(mesh.positions[f[1]]);
        t.v2 = xf.apply_point(mesh.positions[f[2]]);
        if (m->smooth) {
          t.n0 = xf.apply_normal(mesh.normals[f[0]]);
          t.n1 = xf.apply_normal(mesh.normals[f[1]]);
          t.n2 = xf.apply_normal(mesh.normals[f[2]]);
          t.has_vertex_normals = true;
        }
        t.material = mat;
        add(t);
      }
      continue;
    }
    // Any other object: compile it alone and splice its output in.
    SceneDesc one = stripped;
    one.objects = {obj};
    Scene part = Scene::compile(one);
    for (auto& p : part.primitives) s.primitives.push_back(p);
    for (auto& pl : part.planes) s.planes.push_back(pl);
    s.info.degenerate_skipped += part.info.degenerate_skipped;
  }
  for (const auto& l : desc.lights) {
    s.lights.push_back({LightRef::Kind::Point, (uint32_t)s.point_lights.size()});
    s.point_lights.push_back({l.position, l.intensity});
  }
  s.is_sampled_light.assign(s.primitives.size(), 0);
  for (std::size_t i = 0; i < s.primitives.size(); ++i) {
    const MaterialId m = std::visit([](const auto& p) { return p.material; }, s.primitives[i]);
    if (!emission(s.materials[m]).is_zero()) {
      s.lights.push_back({LightRef::Kind::Emitter, (uint32_t)i});
      s.is_sampled_light[i] = 1;
    }
  }
  std::vector<AABB> bounds;
  for (const auto& p : s.primitives) bounds.push_back(std::visit([](const auto& q) { return q.bounds(); }, p));
  s.bvh.build(bounds, 4);
  s.info.bvh = s.bvh.stats();
  return s;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

typedef struct ref_overrides {
  int32_t width, height, spp, max_depth, rr_min_depth, integrator, light_sampling, use_bvh;
  int64_t seed;
  int32_t force_segments_v_path;  // 1: use the segments_v compile path even without the key
  int32_t pad;
} ref_overrides;

const char* ref_last_error() { return g_err.c_str(); }

void* ref_scene_from_json(const char* text, const ref_overrides* ov) {
  try {
    auto r = std::make_unique<RefScene>();
    r->desc = parse_scene(text);
    auto& R = r->desc.render;
    if (ov) {
      if (ov->width > 0) R.width = ov->width;
      if (ov->height > 0) R.height = ov->height;
      if (ov->spp > 0) R.spp = ov->spp;
      if (ov->max_depth > 0) R.max_depth = ov->max_depth;
      if (ov->rr_min_depth >= 0) R.rr_min_depth = ov->rr_min_depth;
      if (ov->integrator >= 0) R.integrator = (Integrator)ov->integrator;
      if (ov->light_sampling >= 0) R.light_sampling = ov->light_sampling != 0;
      if (ov->use_bvh >= 0) R.use_bvh = ov->use_bvh != 0;
      if (ov->seed >= 0) R.seed = (std::uint64_t)ov->seed;
    }
    // segments_v extension: read straight from the JSON document.
    const Json j = Json::parse(text);
    std::vector<int> seg_v(r->desc.objects.size(), 0);
    bool any = false;
    if (const Json* objs = j.find("objects"))
      for (std::size_t i = 0; i < objs->as_array().size() && i < seg_v.size(); ++i)
        if (const Json* sv = objs->as_array()[i].find("segments_v")) { seg_v[i] = (int)sv->as_number(); any = true; }
    if (ov && ov->force_segments_v_path) {
      for (std::size_t i = 0; i < seg_v.size(); ++i)
        if (const auto* m = std::get_if<MeshDesc>(&r->desc.objects[i].shape); m && m->source == MeshDesc::Source::Torus && seg_v[i] == 0)
          seg_v[i] = std::max(3, m->segments / 2);
      any = true;
    }
    r->scene = any ? compile_with_segments_v(r->desc, seg_v) : Scene::compile(r->desc);
    export_scene(*r);
    return r.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void ref_scene_free(void* h) { delete static_cast<RefScene*>(h); }

const rtg_scene_desc* ref_scene_desc(void* h) { return &static_cast<RefScene*>(h)->out; }

void ref_scene_info(void* h, uint64_t* out /* 8 */) {
  const auto& i = static_cast<RefScene*>(h)->scene.info;
  out[0] = i.triangles; out[1] = i.spheres; out[2] = i.planes; out[3] = i.degenerate_skipped; out[4] = i.lights;
  out[5] = i.bvh.nodes; out[6] = i.bvh.leaves; out[7] = i.bvh.max_depth;
}

// CameraDesc + resolution: {look_from, look_at, vup, vfov, aperture, focus, w, h}.
void ref_scene_camera_desc(void* h, double* out /* 14 */) {
  const auto& d = static_cast<RefScene*>(h)->desc;
  const CameraDesc& c = d.camera;
  const double v[14] = {c.look_from.x, c.look_from.y, c.look_from.z, c.look_at.x, c.look_at.y, c.look_at.z,
                        c.vup.x, c.vup.y, c.vup.z, c.vfov_deg, c.aperture, c.focus_dist,
                        (double)d.render.width, (double)d.render.height};
  std::memcpy(out, v, sizeof v);
}

// Reference camera rays: generate(i, j, sx, sy) for a list of pixels (lens
// draws taken from for_sample(seed, 0, k)).
void ref_camera_rays(void* h, const double* px /* n*4: i j sx sy */, uint64_t n, uint64_t seed, double* out /* n*6 */) {
  const Camera& c = static_cast<RefScene*>(h)->scene.camera;
  for (uint64_t k = 0; k < n; ++k) {
    Rng rng = Rng::for_sample(seed, 0, k);
    const Ray r = c.generate((int)px[4 * k], (int)px[4 * k + 1], px[4 * k + 2], px[4 * k + 3], rng);
    out[6 * k + 0] = r.origin.x; out[6 * k + 1] = r.origin.y; out[6 * k + 2] = r.origin.z;
    out[6 * k + 3] = r.direction.x; out[6 * k + 4] = r.direction.y; out[6 * k + 5] = r.direction.z;
  }
}

void ref_trace_batch(void* h, int mode, int use_bvh, const double* rays, uint64_t n, rtg_hit* hits, uint64_t* nv,
                     uint64_t* pt) {
  Scene& s = static_cast<RefScene*>(h)->scene;
  const bool saved = s.settings.use_bvh;
  s.settings.use_bvh = use_bvh != 0;
  TraversalCounters ctr;
  for (uint64_t i = 0; i < n; ++i) {
    const double* q = rays + 8 * i;
    const Ray r{{q[0], q[1], q[2]}, {q[3], q[4], q[5]}};
    rtg_hit out{kInfinity, 0xFFFFFFFFu, 0};
    if (mode == 0) {
      HitRecord rec;
      if (s.intersect(r, Interval(q[6], q[7]), rec, &ctr)) {
        out.t = rec.t;
        out.prim = rec.primitive;
        out.flags = 1u | (rec.is_plane ? 2u : 0u) | (rec.front_face ? 4u : 0u);
      }
    } else {
      out.flags = s.occluded(r, q[7], &ctr) ? 1u : 0u;
    }
    hits[i] = out;
  }
  s.settings.use_bvh = saved;
  if (nv) *nv += ctr.node_visits;
  if (pt) *pt += ctr.prim_tests;
}

// render_tile's per-sample loop (renderer.cpp:31-51) over rows [y0, y1),
// samples [s0, s1), through the reference's public radiance().
void ref_render_rows(void* h, uint32_t s0, uint32_t s1, int y0, int y1, double* sum, uint32_t* count, uint64_t* stats /* 5 */) {
  const Scene& s = static_cast<RefScene*>(h)->scene;
  const auto& cfg = s.settings;
  PathStats st;
  uint64_t samples = 0;
  for (int y = y0; y < y1; ++y)
    for (int x = 0; x < cfg.width; ++x) {
      const auto pixel = static_cast<std::uint64_t>(y) * static_cast<std::uint64_t>(cfg.width) + static_cast<std::uint64_t>(x);
      for (std::uint32_t smp = s0; smp < s1; ++smp) {
        Rng rng = Rng::for_sample(cfg.seed, pixel, smp);
        const Real jx = rng.uniform();
        const Real jy = rng.uniform();
        const Ray r = s.camera.generate(x, y, jx, jy, rng);
        Color L = radiance(s, r, rng, st);
        if (!L.is_finite() || L.x < 0 || L.y < 0 || L.z < 0) { st.invalid_samples++; L = Color{}; }
        const std::size_t i = (std::size_t)pixel;
        sum[3 * i] += L.x; sum[3 * i + 1] += L.y; sum[3 * i + 2] += L.z;
        count[i]++;
        ++samples;
      }
    }
  if (stats) {
    stats[0] += samples; stats[1] += st.rays; stats[2] += st.shadow_rays; stats[3] += st.rr_terminated;
    stats[4] += st.invalid_samples;
  }
}

// Same loop as ref_render_rows, also summing each sample's squared radiance
// per channel (sumsq), so the fixtures can state each pixel's sample
// variance; the reference Framebuffer keeps only sum and count
// (framebuffer.hpp:48-49).
void ref_render_moments(void* h, uint32_t s0, uint32_t s1, int y0, int y1, double* sum, double* sumsq, uint32_t* count,
                        uint64_t* stats /* 5 */) {
  const Scene& s = static_cast<RefScene*>(h)->scene;
  const auto& cfg = s.settings;
  PathStats st;
  uint64_t samples = 0;
  for (int y = y0; y < y1; ++y)
    for (int x = 0; x < cfg.width; ++x) {
      const auto pixel = static_cast<std::uint64_t>(y) * static_cast<std::uint64_t>(cfg.width) + static_cast<std::uint64_t>(x);
      const std::size_t i = (std::size_t)pixel;
      for (std::uint32_t smp = s0; smp < s1; ++smp) {
        Rng rng = Rng::for_sample(cfg.seed, pixel, smp);
        const Real jx = rng.uniform();
        const Real jy = rng.uniform();
        const Ray r = s.camera.generate(x, y, jx, jy, rng);
        Color L = radiance(s, r, rng, st);
        if (!L.is_finite() || L.x < 0 || L.y < 0 || L.z < 0) { st.invalid_samples++; L = Color{}; }
        sum[3 * i] += L.x; sum[3 * i + 1] += L.y; sum[3 * i + 2] += L.z;
        sumsq[3 * i] += L.x * L.x; sumsq[3 * i + 1] += L.y * L.y; sumsq[3 * i + 2] += L.z * L.z;
        count[i]++;
        ++samples;
      }
    }
  if (stats) {
    stats[0] += samples; stats[1] += st.rays; stats[2] += st.shadow_rays; stats[3] += st.rr_terminated;
    stats[4] += st.invalid_samples;
  }
}

// The reference Renderer end to end (thread pool, tiles, progressive
// passes): the CPU baseline. Writes the per-pixel mean if `mean` is set.
int ref_render_full(void* h, int threads, int progressive, int max_pass, double* mean, uint64_t* stats /* 5 */,
                    double* elapsed_ms) {
  try {
    Scene& s = static_cast<RefScene*>(h)->scene;
    s.settings.threads = threads;
    Renderer rr(s, RenderOptions{progressive != 0, max_pass > 0 ? max_pass : 32});
    rr.render();
    const auto st = rr.stats();
    if (stats) {
      stats[0] = st.camera_samples; stats[1] = st.rays; stats[2] = st.shadow_rays; stats[3] = st.rr_terminated;
      stats[4] = st.invalid_samples;
    }
    if (elapsed_ms) *elapsed_ms = st.elapsed_ms;
    if (mean) {
      const Framebuffer& fb = rr.framebuffer();
      for (int y = 0; y < fb.height(); ++y)
        for (int x = 0; x < fb.width(); ++x) {
          const Color c = fb.mean(x, y);
          const std::size_t i = (std::size_t)y * fb.width() + x;
          mean[3 * i] = c.x; mean[3 * i + 1] = c.y; mean[3 * i + 2] = c.z;
        }
    }
    return rr.thread_count();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void ref_rng_stream(uint64_t seed, uint64_t pixel, uint64_t sample, uint32_t* out, int n) {
  Rng r = Rng::for_sample(seed, pixel, sample);
  for (int i = 0; i < n; ++i) out[i] = r.next_u32();
}

int ref_bvh_build(const double* bounds, uint32_t n, int max_leaf, rtg_bvh_node* nodes, uint32_t* n_nodes,
                  uint32_t* prim_out, uint32_t* n_prim) {
  try {
    std::vector<AABB> b(n);
    for (uint32_t i = 0; i < n; ++i)
      b[i] = AABB{{bounds[6 * i], bounds[6 * i + 1], bounds[6 * i + 2]}, {bounds[6 * i + 3], bounds[6 * i + 4], bounds[6 * i + 5]}};
    Bvh bvh;
    bvh.build(b, max_leaf);
    for (std::size_t i = 0; i < bvh.nodes().size(); ++i) {
      const auto& q = bvh.nodes()[i];
      nodes[i] = rtg_bvh_node{V3(q.bounds.min), V3(q.bounds.max), q.offset, q.count, q.axis, 0};
    }
    *n_nodes = (uint32_t)bvh.nodes().size();
    for (std::size_t i = 0; i < bvh.prim_indices().size(); ++i) prim_out[i] = bvh.prim_indices()[i];
    *n_prim = (uint32_t)bvh.prim_indices().size();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

uint32_t ref_make_torus(double major, double minor, int su, int sv, double* pos, double* nrm, uint32_t* faces) {
  const TriangleMesh m = make_torus(major, minor, su, sv);
  for (std::size_t i = 0; i < m.positions.size(); ++i) {
    pos[3 * i] = m.positions[i].x; pos[3 * i + 1] = m.positions[i].y; pos[3 * i + 2] = m.positions[i].z;
    nrm[3 * i] = m.normals[i].x; nrm[3 * i + 1] = m.normals[i].y; nrm[3 * i + 2] = m.normals[i].z;
  }
  for (std::size_t i = 0; i < m.faces.size(); ++i)
    for (int k = 0; k < 3; ++k) faces[3 * i + k] = m.faces[i][k];
  return (uint32_t)m.faces.size();
}

uint8_t ref_linear_to_srgb8(double v) { return linear_to_srgb8(v); }
void ref_tone_map(const double* in, int op, double exposure, double* out) {
  const Color c = tone_map(Color{in[0], in[1], in[2]}, (ToneMap)op, exposure);
  out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

int ref_sample_direct_light(void* h, const double* o, const double* d, uint64_t seed, double* out) {
  const Scene& s = static_cast<RefScene*>(h)->scene;
  HitRecord rec;
  if (!s.intersect(Ray{{o[0], o[1], o[2]}, {d[0], d[1], d[2]}}, Interval(1e-6, kInfinity), rec)) return 0;
  Rng rng(seed);
  PathStats st;
  const Color L = sample_direct_light(s, rec, s.material(rec.material), rng, st);
  out[0] = L.x; out[1] = L.y; out[2] = L.z;
  return 1;
}

// Reference JSON round trip (scene_to_json) for boundary checks.
int ref_dump_scene(const char* text, char* out, uint64_t cap) {
  try {
    const std::string s = scene_to_json(parse_scene(text)).dump(-1);
    if (s.size() + 1 > cap) return -(int)(s.size() + 1);
    std::memcpy(out, s.c_str(), s.size() + 1);
    return (int)s.size();
  } catch (const std::exception& e) {
    g_err = e.what();
    return 0;
  }
}

// parse + validate only: "" or the reference's error message.
int ref_validate(const char* text, char* out, uint64_t cap) {
  std::string msg;
  try {
    const SceneDesc d = parse_scene(text);
    if (const auto errs = validate_scene(d); !errs.empty()) msg = errs.front();
  } catch (const std::exception& e) {
    msg = e.what();
  }
  std::strncpy(out, msg.c_str(), cap - 1);
  out[cap - 1] = 0;
  return msg.empty() ? 0 : 1;
}

}  // extern "C"
