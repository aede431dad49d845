// rtg_shade.hpp — hit finalisation, materials, light sampling, camera and
// background on the device, for both precision modes.
//
// Reference: materials src/materials/material.cpp:14-109, light sampling
// src/rendering/integrator.cpp:14-91, finalize geometry/shapes.hpp:49-58 and
// :129-142, camera core/camera.hpp:54-62, background src/core/scene.cpp:10-14.
// Random draws are taken in the reference's order; the FP64 mode feeds them
// from the reference's PCG32 stream, the FP32 mode from the counter hash.
#pragma once
#include "rtg_rng.hpp"
#include "rtg_trace.hpp"

namespace rtg {

// ---- random draw sources ----
struct PcgDraws {  // FP64: the reference stream
  Pcg32 pcg;
  __device__ __forceinline__ double next() { return pcg.uniform(); }
  __device__ __forceinline__ uint32_t below(uint32_t n) { return (uint32_t)(pcg.uniform() * n); }
};

struct HashDraws {  // FP32: pure function of (key, bounce, index)
  uint2 key;
  uint32_t bounce;
  uint32_t idx;
  uint4 cur;
  __device__ __forceinline__ HashDraws(uint2 k, uint32_t b) : key(k), bounce(b), idx(0), cur(make_uint4(0, 0, 0, 0)) {}
  __device__ __forceinline__ float next() {
    const uint32_t j = idx & 3u;
    if (j == 0) cur = draw4(key, bounce, idx >> 2);
    ++idx;
    const uint32_t u = j == 0 ? cur.x : (j == 1 ? cur.y : (j == 2 ? cur.z : cur.w));
    return u01(u);
  }
  __device__ __forceinline__ uint32_t below(uint32_t n) {
    const uint32_t k = (uint32_t)(next() * (float)n);
    return k < n ? k : n - 1;
  }
};

// ---- sampling (math/sampling.hpp) ----
template <class T, class R>
__device__ __forceinline__ V3<T> random_unit_vector(R& rng) {  // :30-35
  const T z = T(1) - T(2) * rng.next();
  const T r = r_sqrt(r_fmax(T(0), T(1) - z * z));
  const T phi = T(2) * Consts<T>::pi * rng.next();
  T s, c;
  r_sincos(phi, &s, &c);
  return V3<T>{r * c, r * s, z};
}
template <class T, class R>
__device__ __forceinline__ void random_in_unit_disk(R& rng, T& x, T& y) {  // :38-42
  const T r = r_sqrt(rng.next());
  const T phi = T(2) * Consts<T>::pi * rng.next();
  T s, c;
  r_sincos(phi, &s, &c);
  x = r * c;
  y = r * s;
}
template <class T, class R>
__device__ __forceinline__ V3<T> cosine_hemisphere(R& rng) {  // :53-57
  T x, y;
  random_in_unit_disk<T>(rng, x, y);
  const T z = r_sqrt(r_fmax(T(0), T(1) - x * x - y * y));
  return V3<T>{x, y, z};
}
template <class T, class R>
__device__ __forceinline__ V3<T> uniform_cone(R& rng, T cos_max) {  // :62-67
  const T cos_t = T(1) - rng.next() * (T(1) - cos_max);
  const T sin_t = r_sqrt(r_fmax(T(0), T(1) - cos_t * cos_t));
  const T phi = T(2) * Consts<T>::pi * rng.next();
  T s, c;
  r_sincos(phi, &s, &c);
  return V3<T>{sin_t * c, sin_t * s, cos_t};
}

// ---- hit record (core/hit.hpp) ----
template <class T>
struct HitRec {
  V3<T> p, normal, gn;
  bool front_face;
  uint32_t material;
  uint32_t prim;  // primitive or plane index
  bool is_plane;
};

template <class T>
__device__ __forceinline__ void set_face_normal(HitRec<T>& rec, V3<T> d, V3<T> outward) {
  rec.front_face = dot(d, outward) < 0;
  rec.gn = rec.front_face ? outward : -outward;
  rec.normal = rec.gn;
}

// FP64 finalize from the reference primitive array.
__device__ inline void finalize64(const DevScene& S, V3<double> o, V3<double> d, double t, double b1, double b2,
                                  uint32_t prim, int plane, HitRec<double>& rec) {
  rec.p = o + d * t;
  if (prim == 0xFFFFFFFFu) {
    const rtg_plane& pl = S.planes[plane];
    set_face_normal(rec, d, d3(pl.normal));
    rec.material = pl.material;
    rec.prim = (uint32_t)plane;
    rec.is_plane = true;
    return;
  }
  const rtg_primitive& p = S.prims[prim];
  rec.prim = prim;
  rec.is_plane = false;
  rec.material = p.material;
  if (p.kind == RTG_PRIM_SPHERE) {
    set_face_normal(rec, d, (rec.p - d3(p.v0)) * (1.0 / p.v1.x));
    return;
  }
  const V3<double> v0 = d3(p.v0);
  const V3<double> ng = normalize(cross(d3(p.v1) - v0, d3(p.v2) - v0));
  set_face_normal(rec, d, ng);
  if (p.has_vertex_normals) {
    const double w0 = 1 - b1 - b2;
    const V3<double> ns = normalize_or(d3(p.n0) * w0 + d3(p.n1) * b1 + d3(p.n2) * b2, ng);
    rec.normal = rec.front_face ? ns : -ns;
    if (dot(rec.normal, rec.gn) <= 0) rec.normal = rec.gn;
  }
}

// FP32 finalize from shade32.
__device__ inline void finalize32(const DevScene& S, V3<float> o, V3<float> d, float t, float b1, float b2,
                                  uint32_t prim, int plane, HitRec<float>& rec) {
  rec.p = o + d * t;
  if (prim == 0xFFFFFFFFu) {
    const float4 n = S.plane32[2 * plane + 1];
    set_face_normal(rec, d, V3<float>{n.x, n.y, n.z});
    rec.material = __float_as_uint(S.plane32[2 * plane].w);
    rec.prim = (uint32_t)plane;
    rec.is_plane = true;
    return;
  }
  const float4* sh = S.shade32 + (size_t)prim * kShade32Stride;
  const float4 s0 = sh[0];
  const uint32_t w = __float_as_uint(s0.w);
  rec.prim = prim;
  rec.is_plane = false;
  rec.material = w & 0x3FFFFFFFu;
  if (w & kShadeSphere) {
    const float4 s1 = sh[1];
    set_face_normal(rec, d, (rec.p - V3<float>{s0.x, s0.y, s0.z}) * s1.y);
    return;
  }
  const V3<float> ng{s0.x, s0.y, s0.z};
  set_face_normal(rec, d, ng);
  if (w & kShadeSmooth) {
    const float4 s1 = sh[1], s2 = sh[2];
    const float s3x = reinterpret_cast<const float*>(sh + 3)[0];  // the record's last used word
    const float w0 = 1.0f - b1 - b2;
    const V3<float> n0{s1.x, s1.y, s1.z}, n1{s1.w, s2.x, s2.y}, n2{s2.z, s2.w, s3x};
    const V3<float> ns = normalize_or(n0 * w0 + n1 * b1 + n2 * b2, ng);
    rec.normal = rec.front_face ? ns : -ns;
    if (dot(rec.normal, rec.gn) <= 0.0f) rec.normal = rec.gn;
  }
}

// ---- materials ----
template <class T>
struct Mat {
  uint32_t kind;
  bool two_sided;
  V3<T> c0, c1;
  T p0;
};
__device__ __forceinline__ Mat<double> load_mat64(const DevScene& S, uint32_t id) {
  const rtg_material& m = S.materials[id];
  return Mat<double>{m.kind, m.two_sided != 0, d3(m.c0), d3(m.c1), m.p0};
}
// The second float4 (c1, p0) is read only where it is used: every kind but
// an untextured Lambertian (kind word bit 9 marks a checker texture), whose
// albedo is c0 alone (material.cpp:14-24 with p0 <= 0).
__device__ __forceinline__ Mat<float> load_mat32(const DevScene& S, uint32_t id) {
  const float4 a = S.mat32[2 * id];
  const uint32_t k = __float_as_uint(a.w);
  float4 b{0.0f, 0.0f, 0.0f, 0.0f};
  if ((k & 0xFFu) != RTG_MAT_LAMBERTIAN || (k & 512u)) b = S.mat32[2 * id + 1];
  return Mat<float>{k & 0xFFu, ((k >> 8) & 1u) != 0, V3<float>{a.x, a.y, a.z}, V3<float>{b.x, b.y, b.z}, b.w};
}
template <class T>
__device__ __forceinline__ Mat<T> load_mat(const DevScene& S, uint32_t id);
template <>
__device__ __forceinline__ Mat<double> load_mat<double>(const DevScene& S, uint32_t id) {
  return load_mat64(S, id);
}
template <>
__device__ __forceinline__ Mat<float> load_mat<float>(const DevScene& S, uint32_t id) {
  return load_mat32(S, id);
}

template <class T>
__device__ __forceinline__ V3<T> texture_sample(const Mat<T>& m, V3<T> p) {  // material.cpp:14-24
  if (m.p0 <= 0) return m.c0;
  const T k = T(1e-4);
  const long long parity = (long long)r_floor(p.x * m.p0 + k) + (long long)r_floor(p.y * m.p0 + k) +
                           (long long)r_floor(p.z * m.p0 + k);
  return (parity & 1) ? m.c1 : m.c0;
}
template <class T>
__device__ __forceinline__ T schlick(T cos_theta, T eta) {  // material.cpp:26-31
  T r0 = (T(1) - eta) / (T(1) + eta);
  r0 *= r0;
  const T m = T(1) - cos_theta;
  return r0 + (T(1) - r0) * m * m * m * m * m;
}
template <class T>
__device__ __forceinline__ V3<T> emitted(const Mat<T>& m, bool front) {  // material.cpp:86-90
  if (m.kind == RTG_MAT_EMISSIVE && (front || m.two_sided)) return m.c0 * m.p0;
  return V3<T>{0, 0, 0};
}
template <class T>
__device__ __forceinline__ V3<T> eval_brdf_cos(const Mat<T>& m, const HitRec<T>& rec, V3<T> wi) {  // :94-100
  if (m.kind != RTG_MAT_LAMBERTIAN) return V3<T>{0, 0, 0};
  const T cos_t = dot(rec.normal, wi);
  if (cos_t <= 0 || dot(rec.gn, wi) <= 0) return V3<T>{0, 0, 0};
  return texture_sample(m, rec.p) * (cos_t * Consts<T>::inv_pi);
}

template <class T>
struct Scatter {
  V3<T> dir, weight;
  bool specular;
};

template <class T, class R>
__device__ __forceinline__ bool scatter(const Mat<T>& m, V3<T> in_dir, const HitRec<T>& rec, R& rng,
                                        Scatter<T>& out) {  // material.cpp:33-74
  if (m.kind == RTG_MAT_LAMBERTIAN) {
    const Onb<T> b = make_onb(rec.normal);
    out.dir = to_world(b, cosine_hemisphere<T>(rng));
    if (dot(out.dir, rec.gn) <= 0) return false;
    out.weight = texture_sample(m, rec.p);
    out.specular = false;
    return true;
  }
  if (m.kind == RTG_MAT_METAL) {
    const V3<T> unit_in = normalize(in_dir);
    V3<T> d = reflect(unit_in, rec.normal);
    if (m.p0 > 0) d = normalize(normalize(d) + random_unit_vector<T>(rng) * m.p0);
    if (dot(d, rec.gn) <= 0) return false;
    out.dir = d;
    out.weight = m.c0;
    out.specular = true;
    return true;
  }
  if (m.kind == RTG_MAT_DIELECTRIC) {
    const T eta = rec.front_face ? (T(1) / m.p0) : m.p0;
    const V3<T> unit_in = normalize(in_dir);
    const T cos_i = r_fmin(dot(-unit_in, rec.normal), T(1));
    V3<T> refracted;
    const bool can = refract(unit_in, rec.normal, eta, &refracted);
    if (!can || rng.next() < schlick(cos_i, eta)) out.dir = reflect(unit_in, rec.normal);
    else out.dir = normalize(refracted);
    out.weight = m.c0;
    out.specular = true;
    return true;
  }
  return false;
}

template <class T>
__device__ __forceinline__ V3<T> background(const DevScene& S, V3<T> dir);
template <>
__device__ __forceinline__ V3<double> background<double>(const DevScene& S, V3<double> dir) {  // scene.cpp:10-14
  const rtg_background& b = S.bg;
  if (b.kind == RTG_BG_SOLID) return d3(b.top) * b.intensity;
  const double t = 0.5 * (normalize(dir).y + 1.0);
  return (d3(b.bottom) * (1 - t) + d3(b.top) * t) * b.intensity;
}
template <>
__device__ __forceinline__ V3<float> background<float>(const DevScene& S, V3<float> dir) {
  const V3<float> top{S.bg32[0], S.bg32[1], S.bg32[2]};
  if (S.bg_solid) return top * S.bg32[6];
  const V3<float> bot{S.bg32[3], S.bg32[4], S.bg32[5]};
  const float t = 0.5f * (normalize(dir).y + 1.0f);
  return (bot * (1.0f - t) + top * t) * S.bg32[6];
}

// ---- light sampling (integrator.cpp:14-91) up to the shadow ray ----
template <class T>
struct ShadowReq {
  V3<T> o, d;
  T tmax;
  V3<T> f_rop;  // f * radiance_over_pdf
};

template <class T>
__device__ __forceinline__ bool finish_light(const Mat<T>& mat, const HitRec<T>& rec, V3<T> wi, T dist, V3<T> rop,
                                             ShadowReq<T>& req) {
  const V3<T> f = eval_brdf_cos(mat, rec, wi);
  if (is_zero(f) || is_zero(rop)) return false;
  const V3<T> origin = offset_ray_origin(rec.p, rec.gn, wi);
  const V3<T> q = rec.p + wi * dist;
  T stmax = r_sqrt(len2(q - origin)) - T(4) * spawn_offset(q);
  if (stmax < 0) stmax = 0;
  req.o = origin;
  req.d = wi;
  req.tmax = stmax;
  req.f_rop = f * rop;
  return true;
}

template <class R>
__device__ inline bool sample_light64(const DevScene& S, uint32_t li, const HitRec<double>& rec,
                                      const Mat<double>& mat, R& rng, ShadowReq<double>& req) {
  const rtg_light_ref L = S.lights[li];
  const V3<double> p = rec.p;
  V3<double> wi, rop;
  double dist = 0;
  if (L.kind == RTG_LIGHT_POINT) {
    const rtg_point_light& pl = S.point_lights[L.index];
    const V3<double> d = d3(pl.position) - p;
    const double d2 = len2(d);
    if (d2 < 1e-16) return false;
    dist = r_sqrt(d2);
    wi = d * (1.0 / dist);
    rop = d3(pl.intensity) * (1.0 / d2);
  } else {
    const rtg_primitive& prim = S.prims[L.index];
    const Mat<double> lm = load_mat64(S, prim.material);
    const V3<double> em = lm.c0 * lm.p0;
    if (prim.kind == RTG_PRIM_SPHERE) {
      const V3<double> c = d3(prim.v0);
      const double radius = prim.v1.x;
      const V3<double> oc = c - p;
      const double d2 = len2(oc);
      const double r2 = radius * radius;
      if (d2 <= r2 * (1 + 1e-9)) {
        if (!lm.two_sided) return false;
        const V3<double> n_l = random_unit_vector<double>(rng);
        const V3<double> d = (c + n_l * radius) - p;
        const double dd2 = len2(d);
        if (dd2 < 1e-16) return false;
        dist = r_sqrt(dd2);
        wi = d * (1.0 / dist);
        const double cos_l = r_fabs(dot(n_l, wi));
        if (cos_l < 1e-9) return false;
        const double area = 4 * Consts<double>::pi * radius * radius;
        rop = em * (cos_l * area / dd2);
      } else {
        const double x = r2 / d2;
        const double omc = x / (1 + r_sqrt(1 - x));
        const double cos_max = 1 - omc;
        const Onb<double> b = make_onb(oc * (1.0 / r_sqrt(d2)));
        wi = to_world(b, uniform_cone<double>(rng, cos_max));
        double th;
        if (!sphere_hit64(c, radius, p, wi, 0.0, __builtin_inf(), th)) return false;
        dist = th;
        const double pdf = 1 / (2 * Consts<double>::pi * omc);
        rop = em * (1.0 / pdf);
      }
    } else {
      const double su = r_sqrt(rng.next());
      const double b1 = 1 - su, b2 = rng.next() * su;
      const V3<double> v0 = d3(prim.v0), v1 = d3(prim.v1), v2 = d3(prim.v2);
      const V3<double> q = v0 * (1 - b1 - b2) + v1 * b1 + v2 * b2;
      const V3<double> d = q - p;
      const double d2 = len2(d);
      if (d2 < 1e-16) return false;
      dist = r_sqrt(d2);
      wi = d * (1.0 / dist);
      const V3<double> cr = cross(v1 - v0, v2 - v0);
      const V3<double> nl = normalize(cr);
      const double cos_signed = -dot(nl, wi);
      if (cos_signed <= 0 && !lm.two_sided) return false;
      const double cos_l = r_fabs(cos_signed);
      if (cos_l < 1e-9) return false;
      const double area = 0.5 * r_sqrt(len2(cr));
      const double pdf = d2 / (cos_l * area);
      rop = em * (1.0 / pdf);
    }
  }
  return finish_light(mat, rec, wi, dist, rop, req);
}

// light32 record (kLight32Stride float4):
//   point:    {pos, kind} {intensity, 0}
//   sphere:   {centre, kind} {r, r^2, area, two_sided} {emission, 0}
//   triangle: {v0, kind} {e1, area} {e2, two_sided} {emission, 0} {nl, 0}
template <class R>
__device__ inline bool sample_light32(const DevScene& S, uint32_t li, const HitRec<float>& rec,
                                      const Mat<float>& mat, R& rng, ShadowReq<float>& req) {
  const float4* L = S.light32 + (size_t)li * kLight32Stride;
  const float4 l0 = L[0], l1 = L[1];
  const uint32_t kind = __float_as_uint(l0.w);
  const V3<float> p = rec.p;
  V3<float> wi, rop;
  float dist;
  if (kind == kL32Point) {
    const V3<float> d = V3<float>{l0.x, l0.y, l0.z} - p;
    const float d2 = len2(d);
    if (d2 < 1e-16f) return false;
    const float inv = __builtin_amdgcn_rsqf(d2);
    dist = d2 * inv;
    wi = d * inv;
    rop = V3<float>{l1.x, l1.y, l1.z} * r_rcp(d2);
  } else if (kind == kL32Sphere) {
    const float4 l2 = L[2];
    const V3<float> c{l0.x, l0.y, l0.z};
    const V3<float> em{l2.x, l2.y, l2.z};
    const V3<float> oc = c - p;
    const float d2 = len2(oc);
    const float r2 = l1.y;
    if (d2 <= r2 * (1.0f + 1e-9f)) {
      if (l1.w == 0.0f) return false;
      const V3<float> n_l = random_unit_vector<float>(rng);
      const V3<float> d = (c + n_l * l1.x) - p;
      const float dd2 = len2(d);
      if (dd2 < 1e-16f) return false;
      const float inv = __builtin_amdgcn_rsqf(dd2);
      dist = dd2 * inv;
      wi = d * inv;
      const float cos_l = r_fabs(dot(n_l, wi));
      if (cos_l < 1e-9f) return false;
      rop = em * (cos_l * l1.z / dd2);
    } else {
      const float x = r2 / d2;
      const float omc = x / (1.0f + r_sqrt(1.0f - x));
      const float cos_max = 1.0f - omc;
      const Onb<float> b = make_onb(oc * __builtin_amdgcn_rsqf(d2));
      wi = to_world(b, uniform_cone<float>(rng, cos_max));
      // Sphere::intersect from p along wi over (0, inf).
      const V3<float> ocw = c - p;
      const float h = dot(wi, ocw);
      const V3<float> v = ocw - wi * h;
      const float disc = r2 - len2(v);
      if (disc < 0.0f) return false;
      const float q = h + r_copysign(r_sqrt(disc), h);
      float t0 = (len2(ocw) - r2) / q, t1 = q;
      if (t0 > t1) {
        const float t = t0;
        t0 = t1;
        t1 = t;
      }
      float t = t0;
      if (!(t > 0.0f)) {
        t = t1;
        if (!(t > 0.0f)) return false;
      }
      dist = t;
      rop = em * (2.0f * Consts<float>::pi * omc);
    }
  } else {
    const float4 l2 = L[2], l3 = L[3], l4 = L[4];
    const float su = r_sqrt(rng.next());
    const float b1 = 1.0f - su, b2 = rng.next() * su;
    const V3<float> q = V3<float>{l0.x, l0.y, l0.z} + V3<float>{l1.x, l1.y, l1.z} * b1 + V3<float>{l2.x, l2.y, l2.z} * b2;
    const V3<float> d = q - p;
    const float d2 = len2(d);
    if (d2 < 1e-16f) return false;
    const float inv = __builtin_amdgcn_rsqf(d2);
    dist = d2 * inv;
    wi = d * inv;
    const float cos_signed = -dot(V3<float>{l4.x, l4.y, l4.z}, wi);
    if (cos_signed <= 0.0f && l2.w == 0.0f) return false;
    const float cos_l = r_fabs(cos_signed);
    if (cos_l < 1e-9f) return false;
    rop = V3<float>{l3.x, l3.y, l3.z} * (cos_l * l1.w / d2);
  }
  return finish_light(mat, rec, wi, dist, rop, req);
}

// ---- camera (core/camera.hpp:54-62) ----
template <class R>
__device__ __forceinline__ void camera64(const DevScene& S, int i, int j, double sx, double sy, R& rng,
                                         V3<double>& o, V3<double>& d) {
  const rtg_camera& c = S.cam;
  const V3<double> target = d3(c.corner) + d3(c.du) * (i + sx) + d3(c.dv) * (j + sy);
  o = d3(c.origin);
  if (c.lens_radius > 0) {
    double dx, dy;
    random_in_unit_disk<double>(rng, dx, dy);
    o = d3(c.origin) + (d3(c.u) * dx + d3(c.v) * dy) * c.lens_radius;
  }
  d = normalize(target - o);
}
template <class R>
__device__ __forceinline__ void camera32(const DevScene& S, int i, int j, float sx, float sy, R& rng, V3<float>& o,
                                         V3<float>& d) {
  const float* c = S.cam32;
  const V3<float> org{c[0], c[1], c[2]};
  V3<float> rel = V3<float>{c[3], c[4], c[5]} + V3<float>{c[6], c[7], c[8]} * ((float)i + sx) +
                  V3<float>{c[9], c[10], c[11]} * ((float)j + sy);
  o = org;
  if (c[18] > 0.0f) {
    float dx, dy;
    random_in_unit_disk<float>(rng, dx, dy);
    const V3<float> off = (V3<float>{c[12], c[13], c[14]} * dx + V3<float>{c[15], c[16], c[17]} * dy) * c[18];
    o = org + off;
    rel = rel - off;
  }
  d = normalize(rel);
}

}  // namespace rtg
