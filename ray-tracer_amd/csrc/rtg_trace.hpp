// rtg_trace.hpp — ray/primitive tests and BVH traversal on the device.
//
// Semantics follow Scene::intersect / occluded / intersect_brute
// (src/core/scene.cpp:139-234) and Bvh::closest / any
// (include/raytracer/geometry/bvh.hpp:75-132):
//   * planes are tested first, outside the BVH, and bound t_max;
//   * closest-hit keeps the first strictly closer hit; any-hit returns on
//     the first hit;
//   * node_visits counts node boxes tested, prim_tests primitive tests.
// FP64 functions are the reference algorithm on the reference layout (the
// near child first by the sign of the direction on the node's split axis)
// and reproduce its hits bit for bit. FP32 functions run on the two-child
// layout: both child boxes (half floats in the scene's node frame) are
// tested from one 32-byte record against the current closest t, the child
// with the smaller entry distance is entered first (any-hit: the larger),
// and only the other child's reference is pushed; a popped child is entered
// without a second box test (its box was already tested, and any primitive
// in it still has to beat the closest t). The FP32 walk
// parks leaves (see Walk), so a ray may test a leaf after nodes that follow
// it in depth-first order; every such choice depends on the ray alone.
#pragma once
#include <stddef.h>

#include "rtg_math.hpp"
#include "rtg_scene.hpp"

namespace rtg {

struct Counters {
  uint32_t visits, tests;
};

// LDS stack: entry k of this lane lives at ref[k * kStackStride] (one
// column per thread of a traversal workgroup, which has kStackStride threads).
// The FP32 stack holds child refs only (4 bytes per entry): half the LDS of
// a {ref, entry distance} stack doubles the resident waves, which is worth
// more than culling far entries on pop (measured on the torus workload:
// closest-hit 40.4 -> 29.0 ms per render for 5 % more node visits). A far
// child popped after a closer hit is visited again; its boxes lie inside the
// pushed box, so it cannot change the result.
constexpr uint32_t kStackStride = 128;
struct LdsStack {
  uint32_t* ref;
};

// =========================== FP64 (exact) ===============================

__device__ __forceinline__ V3<double> d3(rtg_vec3 v) { return V3<double>{v.x, v.y, v.z}; }

// aabb.hpp:57-66
__device__ __forceinline__ bool slab64(double lo, double hi, double o, double inv, double& tmin, double& tmax) {
  double t0 = (lo - o) * inv;
  double t1 = (hi - o) * inv;
  if (t0 != t0 || t1 != t1) return true;
  if (t0 > t1) {
    const double t = t0;
    t0 = t1;
    t1 = t;
  }
  tmin = t0 > tmin ? t0 : tmin;
  tmax = t1 < tmax ? t1 : tmax;
  return tmin <= tmax;
}
__device__ __forceinline__ bool box64(const rtg_bvh_node& n, V3<double> o, V3<double> inv, double tmin, double tmax) {
  return slab64(n.lo.x, n.hi.x, o.x, inv.x, tmin, tmax) && slab64(n.lo.y, n.hi.y, o.y, inv.y, tmin, tmax) &&
         slab64(n.lo.z, n.hi.z, o.z, inv.z, tmin, tmax);
}

// shapes.hpp:25-46
__device__ __forceinline__ bool sphere_hit64(V3<double> c, double radius, V3<double> o, V3<double> d, double tmin,
                                             double tmax, double& t_out) {
  const V3<double> oc = c - o;
  const double a = len2(d);
  if (a == 0) return false;
  const double h = dot(d, oc);
  const double cc = len2(oc) - radius * radius;
  const V3<double> v = oc - d * (h / a);
  const double disc = a * (radius * radius - len2(v));
  if (disc < 0) return false;
  const double sq = r_sqrt(disc);
  const double q = h + r_copysign(sq, h);
  double t0 = cc / q, t1 = q / a;
  if (q == 0) { t0 = t1 = h / a; }
  if (t0 > t1) {
    const double t = t0;
    t0 = t1;
    t1 = t;
  }
  double t = t0;
  if (!(tmin < t && t < tmax)) {
    t = t1;
    if (!(tmin < t && t < tmax)) return false;
  }
  t_out = t;
  return true;
}

// shapes.hpp:110-127
__device__ __forceinline__ bool tri_hit64(V3<double> v0, V3<double> v1, V3<double> v2, V3<double> o, V3<double> d,
                                          double tmin, double tmax, double& t_out, double& b1o, double& b2o) {
  const V3<double> e1 = v1 - v0, e2 = v2 - v0;
  const V3<double> pvec = cross(d, e2);
  const double det = dot(e1, pvec);
  if (r_fabs(det) < 1e-14) return false;
  const double inv_det = 1.0 / det;
  const V3<double> tvec = o - v0;
  const double b1 = dot(tvec, pvec) * inv_det;
  if (b1 < 0 || b1 > 1) return false;
  const V3<double> qvec = cross(tvec, e1);
  const double b2 = dot(d, qvec) * inv_det;
  if (b2 < 0 || b1 + b2 > 1) return false;
  const double t = dot(e2, qvec) * inv_det;
  if (!(tmin < t && t < tmax)) return false;
  t_out = t;
  b1o = b1;
  b2o = b2;
  return true;
}

// A primitive's kind and vertices (the first 88 bytes of rtg_primitive),
// loaded together. Read field by field, the vertex loads would be issued
// after the kind test, a second wait for memory per primitive. The array's
// elements are 160 bytes apart from a 256-byte aligned base, so the 16-byte
// loads are aligned; the empty asm statement keeps the last load ahead of
// the branch.
struct Prim64 {
  uint32_t kind;
  V3<double> v0, v1, v2;
};
static_assert(sizeof(rtg_primitive) == 160 && offsetof(rtg_primitive, v0) == 16 && offsetof(rtg_primitive, v1) == 40 &&
                  offsetof(rtg_primitive, v2) == 64,
              "load_prim64 reads the documented rtg_primitive layout");
typedef double NativeD2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ Prim64 load_prim64(const rtg_primitive* p) {
  const auto* q = (const __attribute__((address_space(1))) NativeD2*)p;
  const uint32_t kind = *(const __attribute__((address_space(1))) uint32_t*)p;
  const NativeD2 a = q[1], b = q[2], c = q[3], e = q[4];
  const double f = ((const __attribute__((address_space(1))) double*)p)[10];
  asm volatile("" ::"v"(f));
  return Prim64{kind, {a.x, a.y, b.x}, {b.y, c.x, c.y}, {e.x, e.y, f}};
}

__device__ __forceinline__ bool prim_hit64(const Prim64& p, V3<double> o, V3<double> d, double tmin, double tmax,
                                           double& t, double& b1, double& b2) {
  if (p.kind == RTG_PRIM_SPHERE) return sphere_hit64(p.v0, p.v1.x, o, d, tmin, tmax, t);
  return tri_hit64(p.v0, p.v1, p.v2, o, d, tmin, tmax, t, b1, b2);
}

// shapes.hpp:82-86 (t only); returns the plane index hit last (planes loop
// keeps the last closer one) or -1.
__device__ __forceinline__ int planes64(const DevScene& S, V3<double> o, V3<double> d, double tmin, double& tmax) {
  int hit = -1;
  for (uint32_t i = 0; i < S.n_planes; ++i) {
    const rtg_plane& pl = S.planes[i];
    const V3<double> n = d3(pl.normal);
    const double denom = dot(n, d);
    if (r_fabs(denom) < 1e-12) continue;
    const double t = dot(d3(pl.point) - o, n) / denom;
    if (!(tmin < t && t < tmax)) continue;
    tmax = t;
    hit = (int)i;
  }
  return hit;
}

struct Hit64 {
  double t, b1, b2;
  uint32_t prim;  // 0xFFFFFFFF none
  int plane;      // -1 none
};

// Scene::intersect (scene.cpp:176-198) without finalize.
__device__ inline Hit64 closest64(const DevScene& S, bool use_bvh, V3<double> o, V3<double> d, double tmin,
                                  double tmax, LdsStack stk, Counters& c) {
  Hit64 h;
  h.plane = planes64(S, o, d, tmin, tmax);
  h.t = tmax;
  h.prim = 0xFFFFFFFFu;
  h.b1 = h.b2 = 0;
  if (!use_bvh) {  // scene.cpp:200-215
    for (uint32_t i = 0; i < S.n_prims; ++i) {
      ++c.tests;
      double th, b1, b2;
      if (prim_hit64(load_prim64(S.prims + i), o, d, tmin, h.t, th, b1, b2)) {
        h.t = th;
        h.b1 = b1;
        h.b2 = b2;
        h.prim = i;
      }
    }
    if (h.prim == 0xFFFFFFFFu && h.plane >= 0) h.t = tmax;
    return h;
  }
  if (S.n_nodes == 0) return h;
  const V3<double> inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
  const bool neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
  uint32_t sp = 0;  // the stack holds one pushed sibling per ancestor: at most the BVH depth
  uint32_t idx = 0;
  double tm = tmax;
  for (;;) {
    const rtg_bvh_node n = S.nodes[idx];
    ++c.visits;
    if (box64(n, o, inv, tmin, tm)) {
      if (n.count > 0) {
        // The next primitive's index is loaded before the current one is
        // tested (the last reloads its own).
        uint32_t pi = S.prim_indices[n.offset];
        for (uint32_t i = 0; i < n.count; ++i) {
          const uint32_t pn = S.prim_indices[n.offset + (i + 1 < n.count ? i + 1 : i)];
          ++c.tests;
          double th, b1, b2;
          if (prim_hit64(load_prim64(S.prims + pi), o, d, tmin, tm, th, b1, b2)) {
            h.t = th;
            h.b1 = b1;
            h.b2 = b2;
            h.prim = pi;
            tm = th;
          }
          pi = pn;
        }
      } else {
        if (neg[n.axis]) {
          stk.ref[sp++ * kStackStride] = idx + 1;
          idx = n.offset;
        } else {
          stk.ref[sp++ * kStackStride] = n.offset;
          idx = idx + 1;
        }
        continue;
      }
    }
    if (sp == 0) break;
    idx = stk.ref[--sp * kStackStride];
  }
  return h;
}

// Scene::occluded (scene.cpp:217-234); t_min is the caller's (1e-6 there).
__device__ inline bool any64(const DevScene& S, bool use_bvh, V3<double> o, V3<double> d, double tmin, double tmax,
                             LdsStack stk, Counters& c) {
  for (uint32_t i = 0; i < S.n_planes; ++i) {
    const rtg_plane& pl = S.planes[i];
    const V3<double> n = d3(pl.normal);
    const double denom = dot(n, d);
    if (r_fabs(denom) < 1e-12) continue;
    const double t = dot(d3(pl.point) - o, n) / denom;
    if (tmin < t && t < tmax) return true;
  }
  double th, b1, b2;
  if (!use_bvh) {
    for (uint32_t i = 0; i < S.n_prims; ++i) {
      ++c.tests;
      if (prim_hit64(load_prim64(S.prims + i), o, d, tmin, tmax, th, b1, b2)) return true;
    }
    return false;
  }
  if (S.n_nodes == 0) return false;
  const V3<double> inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
  uint32_t sp = 0;
  uint32_t idx = 0;
  for (;;) {
    const rtg_bvh_node n = S.nodes[idx];
    ++c.visits;
    if (box64(n, o, inv, tmin, tmax)) {
      if (n.count > 0) {
        uint32_t pi = S.prim_indices[n.offset];
        for (uint32_t i = 0; i < n.count; ++i) {
          const uint32_t pn = S.prim_indices[n.offset + (i + 1 < n.count ? i + 1 : i)];
          ++c.tests;
          if (prim_hit64(load_prim64(S.prims + pi), o, d, tmin, tmax, th, b1, b2)) return true;
          pi = pn;
        }
      } else {
        stk.ref[sp++ * kStackStride] = n.offset;
        idx = idx + 1;
        continue;
      }
    }
    if (sp == 0) break;
    idx = stk.ref[--sp * kStackStride];
  }
  return false;
}

// =========================== FP32 (fast) ================================

// Slab distances are one fma each: the distance to the plane x = b is
// fma(b, 1/d, -q) with q = o * (1/d). That rounds with a relative error
// below 1.5 * 2^-23 of the distance (1/d and the fma) plus an absolute error
// below |q| * 2^-24 (the rounding of q), which does not shrink with the
// distance: at the Cornell box's 555-unit scale it is ~1e-4 for a box a few
// units away. So a slab's lower plane is measured from q moved by a margin
// m = |q| * 2^-21 one way and its upper plane from q moved the other way (lq
// for a box's lo planes, hq for its hi planes; the sign of 1/d decides which
// plane is entered first): every entry moves down and every exit up by at
// least 4x the absolute error, at no cost per box. The relative part is
// covered by widening the exit distance (box32).
// A ray for the FP32 walks: origin and direction (primitive tests), and
// its slab terms in the scene's node frame (node boxes are half floats of
// (b - frame_o) / frame_s): a box plane h lies at t = fma(h, i, -q) with
// i = frame_s / d and q = (o - frame_o) / d, q moved by a margin per side.
struct Ray32 {
  float ox, oy, oz, dx, dy, dz;
  float ix, iy, iz;     // frame_s / d
  float lqx, lqy, lqz;  // offsets of the lo planes
  float hqx, hqy, hqz;  // offsets of the hi planes
};

__device__ __forceinline__ float safe_inv(float d) {
  // |d| < 1e-30 is replaced by +-1e-30 so slab distances stay finite (no
  // 0 * inf); boxes still bound the ray conservatively.
  const float a = r_fabs(d) < 1e-30f ? r_copysign(1e-30f, d) : d;
  return 1.0f / a;
}

// The margin covers the rounding of q (the origin's offset from the frame
// and the product, each within half an ulp of q) with room to spare (8 ulp).
__device__ __forceinline__ Ray32 make_ray32(const DevScene& S, float ox, float oy, float oz, float dx, float dy,
                                            float dz) {
  Ray32 r;
  r.ox = ox; r.oy = oy; r.oz = oz;
  r.dx = dx; r.dy = dy; r.dz = dz;
  const float ix = safe_inv(dx), iy = safe_inv(dy), iz = safe_inv(dz);
  r.ix = ix * S.frame_s; r.iy = iy * S.frame_s; r.iz = iz * S.frame_s;
  const float qx = (ox - S.frame_o[0]) * ix, qy = (oy - S.frame_o[1]) * iy, qz = (oz - S.frame_o[2]) * iz;
  constexpr float k = 0x1p-21f;
  const float mx = r_copysign(r_fabs(qx) * k, ix), my = r_copysign(r_fabs(qy) * k, iy),
              mz = r_copysign(r_fabs(qz) * k, iz);
  r.lqx = qx + mx; r.lqy = qy + my; r.lqz = qz + mz;
  r.hqx = qx - mx; r.hqy = qy - my; r.hqz = qz - mz;
  return r;
}

// Slab test of one box in the node frame; returns the entry distance in tn.
// Plane distances are fma(bound, frame_s/d, -offset) with the ray's margins
// folded into the offsets (Ray32); the far distance is also widened by 5e-7
// relative, so rounding never culls a box the ray touches.
__device__ __forceinline__ bool box32(const Ray32& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                      float tmin, float tmax, float& tn) {
  const float t0x = __builtin_fmaf(lx, r.ix, -r.lqx), t1x = __builtin_fmaf(hx, r.ix, -r.hqx);
  const float t0y = __builtin_fmaf(ly, r.iy, -r.lqy), t1y = __builtin_fmaf(hy, r.iy, -r.hqy);
  const float t0z = __builtin_fmaf(lz, r.iz, -r.lqz), t1z = __builtin_fmaf(hz, r.iz, -r.hqz);
  const float nx = r_fmin(t0x, t1x), fx = r_fmax(t0x, t1x);
  const float ny = r_fmin(t0y, t1y), fy = r_fmax(t0y, t1y);
  const float nz = r_fmin(t0z, t1z), fz = r_fmax(t0z, t1z);
  tn = r_fmax(r_fmax(nx, ny), r_fmax(nz, tmin));
  const float tf = r_fmin(r_fmin(fx, fy), fz) * 1.0000005f;
  return tn <= r_fmin(tf, tmax);
}

// Loads from scene arrays through global-address-space pointers. A batch
// reads its jobs' scene pointers from a table in memory, so the compiler
// cannot tell they are global and would emit flat loads, which count on the
// LDS wait counter too: every traversal-stack access would then wait for the
// node loads in flight (measured: batched closest-hit 38 vs 26 ms per torus
// render). Loading a native vector through an address_space(1) pointer makes
// them global loads in every kernel.
typedef float NativeF4 __attribute__((ext_vector_type(4)));
typedef uint32_t NativeU4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_global(const float4* p) {
  const NativeF4 v = *(const __attribute__((address_space(1))) NativeF4*)p;
  return float4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ uint4 ld_global(const uint4* p) {
  const NativeU4 v = *(const __attribute__((address_space(1))) NativeU4*)p;
  return uint4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ uint32_t ld_global(const uint32_t* p) {
  return *(const __attribute__((address_space(1))) uint32_t*)p;
}

// One primitive slot's nine floats (36 bytes: three dword-aligned vec3
// loads), loaded together so that a triangle test waits for memory once.
// Spheres do not read the third vec3 (their slot pads it), and the compiler
// would sink its load into the triangle branch, after the wait for the first
// two: a second round trip per triangle. The empty asm statement takes it as
// an operand before any branch (measured: closest-hit 3 % faster on the
// torus workload).
typedef float NativeF3 __attribute__((ext_vector_type(3)));
typedef NativeF3 NativeF3u __attribute__((aligned(4)));
struct Slot32 {
  NativeF3 a, b, c;
};
__device__ __forceinline__ Slot32 load_slot32(const float* __restrict__ slot32, uint32_t k) {
  // A 32-bit byte offset (k < 2^25 slots of 36 bytes) from the array base:
  // one address add per load group.
  // (a vec3 type is 16 bytes in memory: the vectors sit 12 bytes apart)
  typedef const __attribute__((address_space(1))) NativeF3u* G3;
  const char* p = reinterpret_cast<const char*>(slot32) + k * (uint32_t)(kSlot32Floats * sizeof(float));
  const Slot32 s{*(G3)p, *(G3)(p + 12), *(G3)(p + 24)};
  asm volatile("" ::"v"(s.c.x));
  return s;
}

// Slot test (triangle: Möller–Trumbore on precomputed edges; sphere: the
// stable quadratic). Writes t, barycentrics when closer than tmax.
__device__ __forceinline__ bool slot_test32(const Slot32& s, const Ray32& r, float tmin, float tmax, float& t_out,
                                            float& b1o, float& b2o) {
  const NativeF3 a = s.a, b = s.b, c = s.c;
  if (__float_as_uint(b.x) != kSlotSphere) {
    const V3<float> d{r.dx, r.dy, r.dz};
    const V3<float> e1{b.x, b.y, b.z}, e2{c.x, c.y, c.z};
    const V3<float> pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    const float inv_det = r_rcp(det);
    const V3<float> tvec{r.ox - a.x, r.oy - a.y, r.oz - a.z};
    const float b1 = dot(tvec, pvec) * inv_det;
    const V3<float> qvec = cross(tvec, e1);
    const float b2 = dot(d, qvec) * inv_det;
    const float t = dot(e2, qvec) * inv_det;
    const bool ok = r_fabs(det) >= 1e-14f && b1 >= 0.0f && b1 <= 1.0f && b2 >= 0.0f && b1 + b2 <= 1.0f &&
                    tmin < t && t < tmax;
    if (ok) {
      t_out = t;
      b1o = b1;
      b2o = b2;
    }
    return ok;
  }
  // sphere: centre a, r^2 b.y
  const V3<float> d{r.dx, r.dy, r.dz};
  const V3<float> oc{a.x - r.ox, a.y - r.oy, a.z - r.oz};
  const float aa = len2(d);
  const float h = dot(d, oc);
  const float cc = len2(oc) - b.y;
  const V3<float> v = oc - d * (h / aa);
  const float disc = aa * (b.y - len2(v));
  if (!(disc >= 0.0f) || aa == 0.0f) return false;
  const float q = h + r_copysign(r_sqrt(disc), h);
  float t0 = cc / q, t1 = q / aa;
  if (q == 0.0f) t0 = t1 = h / aa;
  if (t0 > t1) {
    const float t = t0;
    t0 = t1;
    t1 = t;
  }
  float t = t0;
  if (!(tmin < t && t < tmax)) {
    t = t1;
    if (!(tmin < t && t < tmax)) return false;
  }
  t_out = t;
  b1o = b2o = 0.0f;
  return true;
}

__device__ __forceinline__ bool slot_hit32(const float* __restrict__ slot32, uint32_t k, const Ray32& r, float tmin,
                                           float tmax, float& t_out, float& b1o, float& b2o) {
  return slot_test32(load_slot32(slot32, k), r, tmin, tmax, t_out, b1o, b2o);
}

__device__ __forceinline__ int planes32(const DevScene& S, const Ray32& r, float tmin, float& tmax) {
  int hit = -1;
  for (uint32_t i = 0; i < S.n_planes; ++i) {
    const float4 p = ld_global(S.plane32 + 2 * i), n = ld_global(S.plane32 + 2 * i + 1);
    const float denom = n.x * r.dx + n.y * r.dy + n.z * r.dz;
    if (r_fabs(denom) < 1e-12f) continue;
    const float t = ((p.x - r.ox) * n.x + (p.y - r.oy) * n.y + (p.z - r.oz) * n.z) / denom;
    if (!(tmin < t && t < tmax)) continue;
    tmax = t;
    hit = (int)i;
  }
  return hit;
}

// Inner node record `ref` (< 2^25 records of 32 bytes): a 32-bit byte
// offset from the array base.
__device__ __forceinline__ const float4* node_at(const float4* __restrict__ node2, uint32_t ref) {
  return reinterpret_cast<const float4*>(reinterpret_cast<const char*>(node2) + (ref << 5));
}
static_assert(kNode2Stride * sizeof(float4) == 32, "two-child node records are 32 bytes");
// The half floats of a record word (low and high half).
__device__ __forceinline__ float half_lo(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu));
}
__device__ __forceinline__ float half_hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
// Both children's boxes of a record ({L lo, L hi, R lo, R hi} as 12 halves,
// then the left and right child refs); returns the refs.
struct NodeTest {
  uint32_t l, r;
  bool hl, hr;
  float tl, tr;
};
__device__ __forceinline__ NodeTest node_test(const float4* n, const Ray32& ray, float tmin, float tmax) {
  const uint4 a = ld_global(reinterpret_cast<const uint4*>(n)), b = ld_global(reinterpret_cast<const uint4*>(n + 1));
  NodeTest t;
  t.hl = box32(ray, half_lo(a.x), half_hi(a.x), half_lo(a.y), half_hi(a.y), half_lo(a.z), half_hi(a.z), tmin, tmax,
               t.tl);
  t.hr = box32(ray, half_lo(a.w), half_hi(a.w), half_lo(b.x), half_hi(b.x), half_lo(b.y), half_hi(b.y), tmin, tmax,
               t.tr);
  t.l = b.z;
  t.r = b.w;
  return t;
}

struct Hit32 {
  float t, b1, b2;
  uint32_t prim;  // 0xFFFFFFFF none
  int plane;
};


// "No node": never a valid child reference (a leaf's first slot plus its
// count never exceeds 2^25), and it has the leaf bit, so inner() is false.
constexpr uint32_t kNoRef = 0xFFFFFFFFu;

// Traversal state of one ray, walked as two loops per round (after Aila &
// Laine, "Understanding the Efficiency of Ray Traversal on GPUs", 2009):
// descend inner nodes, parking the first leaf met and continuing from the
// stack, then test the parked leaf. A lane that reaches a leaf early keeps
// descending instead of idling while the rest of its wave is still on inner
// nodes; it leaves the descent at its next leaf or once it has taken
// kDescent steps in the round. The paper ends the descent by a wave-wide vote
// (every lane has a leaf parked), which is ~3.5 % faster on the torus
// workload but makes the order in which a ray meets nodes and leaves depend
// on the other rays of its wave: then which of two hits at equal t a ray
// keeps (a box whose computed entry rounds past a hit on its face is culled
// or not depending on when that hit was found), and its node-visit counts,
// would depend on how paths are grouped into waves. With a rule of the lane's
// own, images and counters are independent of chunking, batching and queue
// order.
constexpr uint32_t kDescent = 24;
// Stack layout: entry i (1-based, i < n) of a walk lives in LDS slot i + 1,
// and a step spills the top to slot n + 1, so slot 1 holds kNoRef whenever
// n is 1 (the step that pushed from an empty stack spilled its kNoRef top
// there) and slot 0 is read only by an empty stack, whose pop ends the walk.
// The slot addresses are then n and n + 1 with no clamping (a walk whose
// inner nodes lie at depth d or less uses slots 0 .. d + 1). The top is
// kNoRef exactly when the stack is empty.
struct Walk {
  uint32_t ref;   // next node to visit (kNoRef: traversal done)
  uint32_t leaf;  // parked leaf (kNoRef: none)
  uint32_t n;     // stack entries: the top in `top`, the rest in LDS slots 2 .. n
  uint32_t top;   // kNoRef when n == 0

  __device__ __forceinline__ bool inner() const { return !(ref & kLeafBit); }
  // Entry n-1, the top after a pop (kNoRef when n == 1; slot 0's contents
  // when n == 0).
  __device__ __forceinline__ uint32_t below(LdsStack s) const { return s.ref[n * kStackStride]; }
  __device__ __forceinline__ uint32_t pop(uint32_t under) {
    const uint32_t r = top;
    top = under;
    n -= n ? 1u : 0u;
    return r;
  }
  // Parks ref if it is a leaf and no leaf is parked, continuing from the stack.
  __device__ __forceinline__ void park(LdsStack s) {
    if (leaf == kNoRef && ref != kNoRef && (ref & kLeafBit)) {
      leaf = ref;
      ref = pop(below(s));
    }
  }
  // One inner node with children l and r, whether their boxes were hit,
  // and whether r is visited first when both are; `under` = below(), read
  // before the node. Branch-free: the top is spilled to its slot on every
  // step (kept only by a push; at an inner node of depth d the stack holds
  // at most d entries, so the slot is at most d + 1), and a child leaf is
  // parked when none is, continuing with the other child or the stack. The
  // child taken is the first when both boxes were hit, else the one hit.
  __device__ __forceinline__ void step(LdsStack s, uint32_t under, uint32_t l, uint32_t r, bool hl, bool hr,
                                       bool r_first) {
    s.ref[(n + 1) * kStackStride] = top;
    const bool both = hl && hr, any = hl || hr;
    const uint32_t first = r_first ? r : l, second = r_first ? l : r;
    const uint32_t c = both ? first : (hl ? l : r);
    const bool park_c = any && leaf == kNoRef && (c & kLeafBit);
    const bool push = both && !park_c;
    const bool popping = !any || (park_c && !both);
    leaf = park_c ? c : leaf;
    ref = popping ? top : (park_c ? second : c);
    top = push ? second : popping ? under : top;
    n = push ? n + 1 : (popping && n) ? n - 1 : n;
  }
  // Leaves the descent: a leaf is parked and `steps` steps were taken.
  __device__ __forceinline__ bool descended(uint32_t steps) const { return leaf != kNoRef && steps >= kDescent; }
};

// A ray's walk in rounds (one descent, then the leaves it parked), so that a
// kernel may hand a lane whose walk has ended a new ray between rounds:
// begin once, then round until it returns true. A ray's rounds depend on the
// ray alone, whatever its lane does before or after it.
struct Trav32 {
  Hit32 h;
  Walk w;
};
// Closest hit: planes, then the root box; false when that already ends the
// walk (h final).
__device__ __forceinline__ bool closest_begin32(const DevScene& S, const Ray32& r, float tmin, float tmax,
                                                LdsStack stk, Counters& c, Trav32& T) {
  T.h.plane = planes32(S, r, tmin, tmax);
  T.h.t = tmax;
  T.h.prim = 0xFFFFFFFFu;
  T.h.b1 = T.h.b2 = 0.0f;
  T.w = Walk{kNoRef, kNoRef, 0, kNoRef};
  if (!S.has_bvh) return false;
  float tn;
  ++c.visits;
  if (!box32(r, S.root_lo[0], S.root_lo[1], S.root_lo[2], S.root_hi[0], S.root_hi[1], S.root_hi[2], tmin, T.h.t, tn))
    return false;
  T.w.ref = S.root_ref;
  T.w.park(stk);
  return true;
}
// One round of the closest-hit walk; true when the walk has ended (h final:
// the hit's slot is then mapped to its primitive id). Called only while the
// walk runs: from a true closest_begin32 until it returns true.
__device__ __forceinline__ bool closest_round32(const DevScene& S, const Ray32& r, float tmin, LdsStack stk,
                                                Counters& c, Trav32& T) {
  Walk& w = T.w;
  Hit32& h = T.h;
  for (uint32_t steps = 1; w.inner(); ++steps) {
    const uint32_t under = w.below(stk);
    const NodeTest nt = node_test(node_at(S.node2, w.ref), r, tmin, h.t);
    c.visits += 2;
    // the nearer child first (by entry distance)
    w.step(stk, under, nt.l, nt.r, nt.hl, nt.hr, nt.tr < nt.tl);
    if (w.descended(steps)) break;
  }
  w.park(stk);  // a leaf popped after a miss
  while (w.leaf != kNoRef) {
    const uint32_t k0 = leaf_first(w.leaf), kn = k0 + leaf_count(w.leaf);
    for (uint32_t k = k0; k < kn; ++k) {
      const Slot32 cur = load_slot32(S.slot32, k);
      ++c.tests;
      float t, b1, b2;
      if (slot_test32(cur, r, tmin, h.t, t, b1, b2)) {
        h.t = t;
        h.b1 = b1;
        h.b2 = b2;
        h.prim = k;  // the slot until the walk ends
      }
    }
    w.leaf = kNoRef;
    w.park(stk);
  }
  if (w.ref != kNoRef) return false;
  if (h.prim != 0xFFFFFFFFu) h.prim = ld_global(S.slot_prim + h.prim);  // slot -> primitive id, once per ray
  return true;
}

__device__ inline Hit32 closest32(const DevScene& S, bool use_bvh, const Ray32& r, float tmin, float tmax,
                                  LdsStack stk, Counters& c) {
  if (!use_bvh) {
    Hit32 h;
    h.plane = planes32(S, r, tmin, tmax);
    h.t = tmax;
    h.prim = 0xFFFFFFFFu;
    h.b1 = h.b2 = 0.0f;
    for (uint32_t i = 0; i < S.n_prims; ++i) {
      ++c.tests;
      float t, b1, b2;
      if (slot_hit32(S.slot32, ld_global(S.prim_slot + i), r, tmin, h.t, t, b1, b2)) {
        h.t = t;
        h.b1 = b1;
        h.b2 = b2;
        h.prim = i;
      }
    }
    return h;
  }
  Trav32 T;
  if (closest_begin32(S, r, tmin, tmax, stk, c, T))
    while (!closest_round32(S, r, tmin, stk, c, T)) {
    }
  return T.h;
}

// Any hit in (tmin, tmax): planes, then the root box; false when that
// already ends the walk (`hit` final).
__device__ __forceinline__ bool any_begin32(const DevScene& S, const Ray32& r, float tmin, float tmax, LdsStack stk,
                                            Counters& c, Walk& w, bool& hit) {
  w = Walk{kNoRef, kNoRef, 0, kNoRef};
  hit = false;
  for (uint32_t i = 0; i < S.n_planes; ++i) {
    const float4 p = ld_global(S.plane32 + 2 * i), n = ld_global(S.plane32 + 2 * i + 1);
    const float denom = n.x * r.dx + n.y * r.dy + n.z * r.dz;
    if (r_fabs(denom) < 1e-12f) continue;
    const float t = ((p.x - r.ox) * n.x + (p.y - r.oy) * n.y + (p.z - r.oz) * n.z) / denom;
    if (tmin < t && t < tmax) {
      hit = true;
      return false;
    }
  }
  if (!S.has_bvh) return false;
  float tn;
  ++c.visits;
  if (!box32(r, S.root_lo[0], S.root_lo[1], S.root_lo[2], S.root_hi[0], S.root_hi[1], S.root_hi[2], tmin, tmax, tn))
    return false;
  w.ref = S.root_ref;
  w.park(stk);
  return true;
}
// One round of the any-hit walk; true when it has ended (`hit` set on a hit).
__device__ __forceinline__ bool any_round32(const DevScene& S, const Ray32& r, float tmin, float tmax, LdsStack stk,
                                            Counters& c, Walk& w, bool& hit) {
  for (uint32_t steps = 1; w.inner(); ++steps) {
    const uint32_t under = w.below(stk);
    const NodeTest nt = node_test(node_at(S.node2, w.ref), r, tmin, tmax);
    c.visits += 2;
    // Any order finds whether some hit lies in (tmin, tmax). Bvh::any
    // visits the left child first (bvh.hpp:106-132); the farther child
    // first (by entry distance) is faster here.
    w.step(stk, under, nt.l, nt.r, nt.hl, nt.hr, nt.tr > nt.tl);
    if (w.descended(steps)) break;
  }
  w.park(stk);  // a leaf popped after a miss
  while (w.leaf != kNoRef) {
    const uint32_t k0 = leaf_first(w.leaf), kn = k0 + leaf_count(w.leaf);
    for (uint32_t k = k0; k < kn; ++k) {
      const Slot32 cur = load_slot32(S.slot32, k);
      ++c.tests;
      float t, b1, b2;
      if (slot_test32(cur, r, tmin, tmax, t, b1, b2)) {
        hit = true;
        w.ref = w.leaf = kNoRef;
        return true;
      }
    }
    w.leaf = kNoRef;
    w.park(stk);
  }
  return w.ref == kNoRef;
}

__device__ inline bool any32(const DevScene& S, bool use_bvh, const Ray32& r, float tmin, float tmax, LdsStack stk,
                             Counters& c) {
  if (!use_bvh) {
    for (uint32_t i = 0; i < S.n_planes; ++i) {
      const float4 p = ld_global(S.plane32 + 2 * i), n = ld_global(S.plane32 + 2 * i + 1);
      const float denom = n.x * r.dx + n.y * r.dy + n.z * r.dz;
      if (r_fabs(denom) < 1e-12f) continue;
      const float t = ((p.x - r.ox) * n.x + (p.y - r.oy) * n.y + (p.z - r.oz) * n.z) / denom;
      if (tmin < t && t < tmax) return true;
    }
    float t, b1, b2;
    for (uint32_t i = 0; i < S.n_prims; ++i) {
      ++c.tests;
      if (slot_hit32(S.slot32, ld_global(S.prim_slot + i), r, tmin, tmax, t, b1, b2)) return true;
    }
    return false;
  }
  Walk w;
  bool hit;
  if (any_begin32(S, r, tmin, tmax, stk, c, w, hit))
    while (!any_round32(S, r, tmin, tmax, stk, c, w, hit)) {
    }
  return hit;
}

}  // namespace rtg
