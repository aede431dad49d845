/* Oracle (test infrastructure): Scene::compile tessellation, procedural
 * meshes and the BVH test scenes, restated from the reference. */
#include <stdlib.h>
#include <string.h>

#include "orc_math.h"

double orc__rng_uniform_range(orc_rng* r, double lo, double hi);

void orc_free(void* p) { free(p); }

/* mesh.cpp:67-88 */
int orc_make_torus(double major, double minor, int segments_u, int segments_v, double** pos, double** nrm,
                   uint32_t* n_vert, uint32_t** faces, uint32_t* n_faces) {
  const int su = segments_u > 3 ? segments_u : 3, sv = segments_v > 3 ? segments_v : 3;
  const size_t nv = (size_t)su * sv;
  double* P = (double*)malloc(nv * 3 * sizeof(double));
  double* N = (double*)malloc(nv * 3 * sizeof(double));
  uint32_t* F = (uint32_t*)malloc(nv * 2 * 3 * sizeof(uint32_t));
  size_t k = 0;
  for (int i = 0; i < su; ++i) {
    const double u = 2 * kPi * i / su;
    for (int j = 0; j < sv; ++j) {
      const double v = 2 * kPi * j / sv;
      const v3 dir = V(cos(u), 0, sin(u));
      const double cv = cos(v);
      const v3 n = vadd(vscale(dir, cv), V(0, sin(v), 0));
      vstore(P + 3 * k, vadd(vscale(dir, major), vscale(n, minor)));
      vstore(N + 3 * k, n);
      ++k;
    }
  }
#define IDX(i, j) ((uint32_t)((((i) % su) * sv) + ((j) % sv)))
  size_t f = 0;
  for (int i = 0; i < su; ++i)
    for (int j = 0; j < sv; ++j) {
      F[f++] = IDX(i, j); F[f++] = IDX(i, j + 1); F[f++] = IDX(i + 1, j);
      F[f++] = IDX(i + 1, j); F[f++] = IDX(i, j + 1); F[f++] = IDX(i + 1, j + 1);
    }
#undef IDX
  *pos = P; *nrm = N; *n_vert = (uint32_t)nv; *faces = F; *n_faces = (uint32_t)(nv * 2);
  return 0;
}

/* mesh.cpp:45-65; the midpoint cache is a hash table (lookup only, so the
 * vertex numbering follows the face loop exactly as with std::map). */
int orc_make_icosphere(int subdivisions, double** pos, double** nrm, uint32_t* n_vert, uint32_t** faces,
                       uint32_t* n_faces) {
  const double t = (1 + sqrt(5.0)) / 2;
  const double init[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                              {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const uint32_t f0[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                              {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  size_t nf = 20, nvtx = 12;
  size_t cap_f = 20, cap_v = 12;
  for (int s = 0; s < subdivisions; ++s) { cap_v += cap_f * 3 / 2; cap_f *= 4; }
  double* P = (double*)malloc(cap_v * 3 * sizeof(double));
  uint32_t* F = (uint32_t*)malloc(cap_f * 3 * sizeof(uint32_t));
  for (int i = 0; i < 12; ++i) vstore(P + 3 * i, vnormalize(vload(init[i])));
  memcpy(F, f0, sizeof f0);
  for (int s = 0; s < subdivisions; ++s) {
    const size_t hcap = nf * 4 + 16;
    uint64_t* keys = (uint64_t*)malloc(hcap * sizeof(uint64_t));
    uint32_t* vals = (uint32_t*)malloc(hcap * sizeof(uint32_t));
    for (size_t i = 0; i < hcap; ++i) keys[i] = ~0ull;
    uint32_t* nxt = (uint32_t*)malloc(nf * 4 * 3 * sizeof(uint32_t));
#define MID(a, b, out)                                                                        \
  do {                                                                                        \
    const uint32_t lo_ = (a) < (b) ? (a) : (b), hi_ = (a) < (b) ? (b) : (a);                  \
    const uint64_t key_ = ((uint64_t)lo_ << 32) | hi_;                                        \
    size_t h_ = (size_t)(orc_splitmix64(key_) % hcap);                                        \
    while (keys[h_] != ~0ull && keys[h_] != key_) h_ = (h_ + 1) % hcap;                       \
    if (keys[h_] == key_) (out) = vals[h_];                                                   \
    else {                                                                                    \
      vstore(P + 3 * nvtx, vnormalize(vadd(vload(P + 3 * (a)), vload(P + 3 * (b)))));        \
      keys[h_] = key_;                                                                        \
      vals[h_] = (uint32_t)nvtx;                                                              \
      (out) = (uint32_t)nvtx++;                                                               \
    }                                                                                         \
  } while (0)
    size_t k = 0;
    for (size_t i = 0; i < nf; ++i) {
      const uint32_t* f = F + 3 * i;
      uint32_t a, b, c;
      MID(f[0], f[1], a);
      MID(f[1], f[2], b);
      MID(f[2], f[0], c);
      const uint32_t tri[4][3] = {{f[0], a, c}, {f[1], b, a}, {f[2], c, b}, {a, b, c}};
      memcpy(nxt + k, tri, sizeof tri);
      k += 12;
    }
#undef MID
    memcpy(F, nxt, k * sizeof(uint32_t));
    nf *= 4;
    free(nxt); free(keys); free(vals);
  }
  double* N = (double*)malloc(nvtx * 3 * sizeof(double));
  memcpy(N, P, nvtx * 3 * sizeof(double));
  *pos = P; *nrm = N; *n_vert = (uint32_t)nvtx; *faces = F; *n_faces = (uint32_t)nf;
  return 0;
}

/* mesh.cpp:20-31 */
static v3 xf_point(const double scale[3], double rot_deg, const double tr[3], v3 p) {
  const v3 s = vmul(p, vload(scale));
  const double a = rot_deg * kPi / 180, c = cos(a), sn = sin(a);
  return vadd(V(c * s.x + sn * s.z, s.y, -sn * s.x + c * s.z), vload(tr));
}
static v3 xf_normal(const double scale[3], double rot_deg, v3 n) {
  const v3 s = V(n.x / scale[0], n.y / scale[1], n.z / scale[2]);
  const double a = rot_deg * kPi / 180, c = cos(a), sn = sin(a);
  return vnormalize_or(V(c * s.x + sn * s.z, s.y, -sn * s.x + c * s.z), n);
}

typedef struct prim_vec { rtg_primitive* p; uint32_t n, cap; uint64_t degenerate; } prim_vec;

static void push_prim(prim_vec* v, const rtg_primitive* p) {
  if (v->n == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 64;
    v->p = (rtg_primitive*)realloc(v->p, (size_t)v->cap * sizeof(rtg_primitive));
  }
  v->p[v->n++] = *p;
}

static void set3(rtg_vec3* d, v3 s) { d->x = s.x; d->y = s.y; d->z = s.z; }

/* scene.cpp:45-49 with shapes.hpp:160-161 (area < 1e-14 => dropped). */
static void add_triangle(prim_vec* v, v3 a, v3 b, v3 c, const v3* n, uint32_t mat) {
  const double area = 0.5 * vlen(vcross(vsub(b, a), vsub(c, a)));
  if (area < 1e-14) { v->degenerate++; return; }
  rtg_primitive p;
  memset(&p, 0, sizeof p);
  p.kind = RTG_PRIM_TRIANGLE;
  p.material = mat;
  set3(&p.v0, a); set3(&p.v1, b); set3(&p.v2, c);
  if (n) {
    p.has_vertex_normals = 1;
    set3(&p.n0, n[0]); set3(&p.n1, n[1]); set3(&p.n2, n[2]);
  }
  push_prim(v, &p);
}

int orc_compile(const orc_object* objs, uint32_t n_objs, const rtg_material* materials, uint32_t n_materials,
                uint32_t n_point_lights, orc_compiled* out) { /* scene.cpp:31-137 */
  memset(out, 0, sizeof *out);
  prim_vec pv = {0};
  rtg_plane* planes = (rtg_plane*)calloc(n_objs ? n_objs : 1, sizeof(rtg_plane));
  uint32_t n_planes = 0;
  for (uint32_t oi = 0; oi < n_objs; ++oi) {
    const orc_object* o = &objs[oi];
    if (o->material >= n_materials) { free(planes); free(pv.p); return -1; }
    switch (o->kind) {
      case ORC_OBJ_SPHERE: {
        rtg_primitive p;
        memset(&p, 0, sizeof p);
        p.kind = RTG_PRIM_SPHERE;
        p.material = o->material;
        set3(&p.v0, vload(o->a));
        p.v1.x = o->scalar;
        push_prim(&pv, &p);
        break;
      }
      case ORC_OBJ_PLANE: {
        rtg_plane* pl = &planes[n_planes++];
        set3(&pl->point, vload(o->a));
        set3(&pl->normal, vnormalize_or(vload(o->b), V(0, 1, 0)));
        pl->material = o->material;
        break;
      }
      case ORC_OBJ_TRIANGLE: add_triangle(&pv, vload(o->a), vload(o->b), vload(o->c), NULL, o->material); break;
      case ORC_OBJ_QUAD: {
        const v3 c = vload(o->a), u = vload(o->b), v = vload(o->c);
        add_triangle(&pv, c, vadd(c, u), vadd(vadd(c, u), v), NULL, o->material);
        add_triangle(&pv, c, vadd(vadd(c, u), v), vadd(c, v), NULL, o->material);
        break;
      }
      case ORC_OBJ_BOX: { /* scene.cpp:68-81 */
        const v3 mn = vload(o->a), mx = vload(o->b);
        const v3 c = vscale(vadd(mn, mx), 0.5);
        const double one[3] = {1, 1, 1}, tr[3] = {c.x, c.y, c.z};
        const v3 h = vscale(vsub(mx, mn), 0.5);
        const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        v3 vv[8];
        for (int i = 0; i < 8; ++i)
          vv[i] = xf_point(one, o->scalar, tr, V(sg[i][0] * h.x, sg[i][1] * h.y, sg[i][2] * h.z));
        const int faces[6][4] = {{4, 5, 6, 7}, {1, 0, 3, 2}, {5, 1, 2, 6}, {0, 4, 7, 3}, {7, 6, 2, 3}, {0, 1, 5, 4}};
        for (int f = 0; f < 6; ++f) {
          add_triangle(&pv, vv[faces[f][0]], vv[faces[f][1]], vv[faces[f][2]], NULL, o->material);
          add_triangle(&pv, vv[faces[f][0]], vv[faces[f][2]], vv[faces[f][3]], NULL, o->material);
        }
        break;
      }
      case ORC_OBJ_MESH: { /* scene.cpp:82-113 */
        double *P, *N;
        uint32_t nv, *F, nf;
        if (o->mesh_source == ORC_MESH_TORUS) {
          const int sv = o->segments_v > 0 ? o->segments_v : (o->segments_u / 2 > 3 ? o->segments_u / 2 : 3);
          orc_make_torus(o->major_radius, o->minor_radius, o->segments_u, sv, &P, &N, &nv, &F, &nf);
        } else {
          orc_make_icosphere(o->subdivisions, &P, &N, &nv, &F, &nf);
        }
        const int smooth = o->smooth != 0;
        for (uint32_t f = 0; f < nf; ++f) {
          v3 pts[3], nn[3];
          for (int k = 0; k < 3; ++k) {
            pts[k] = xf_point(o->scale, o->rotate_y, o->translate, vload(P + 3 * F[3 * f + k]));
            if (smooth) nn[k] = xf_normal(o->scale, o->rotate_y, vload(N + 3 * F[3 * f + k]));
          }
          add_triangle(&pv, pts[0], pts[1], pts[2], smooth ? nn : NULL, o->material);
        }
        free(P); free(N); free(F);
        break;
      }
      default: free(planes); free(pv.p); return -2;
    }
  }
  /* scene.cpp:117-129: point lights first, then emissive primitives. */
  out->lights = (rtg_light_ref*)malloc(sizeof(rtg_light_ref) * (n_point_lights + pv.n + 1));
  for (uint32_t i = 0; i < n_point_lights; ++i) {
    out->lights[out->n_lights].kind = RTG_LIGHT_POINT;
    out->lights[out->n_lights++].index = i;
  }
  out->is_sampled_light = (uint8_t*)calloc(pv.n + 1, 1);
  for (uint32_t i = 0; i < pv.n; ++i) {
    const rtg_material* m = &materials[pv.p[i].material];
    const int emissive = m->kind == RTG_MAT_EMISSIVE &&
                         !(m->c0.x * m->p0 == 0 && m->c0.y * m->p0 == 0 && m->c0.z * m->p0 == 0);
    if (emissive) {
      out->lights[out->n_lights].kind = RTG_LIGHT_EMITTER;
      out->lights[out->n_lights++].index = i;
      out->is_sampled_light[i] = 1;
    }
  }
  double* bounds = (double*)malloc(sizeof(double) * 6 * (pv.n + 1));
  for (uint32_t i = 0; i < pv.n; ++i) orc_prim_bounds(&pv.p[i], bounds + 6 * i);
  out->nodes = (rtg_bvh_node*)malloc(sizeof(rtg_bvh_node) * (2 * (size_t)pv.n + 1));
  out->prim_indices = (uint32_t*)malloc(sizeof(uint32_t) * (pv.n + 1));
  const int rc = orc_bvh_build(bounds, pv.n, 4, out->nodes, &out->n_nodes, out->prim_indices, &out->n_prim_indices,
                               &out->bvh);
  free(bounds);
  out->prims = pv.p;
  out->n_prims = pv.n;
  out->planes = planes;
  out->n_planes = n_planes;
  out->degenerate_skipped = pv.degenerate;
  return rc;
}

void orc_compiled_free(orc_compiled* c) {
  free(c->prims); free(c->is_sampled_light); free(c->planes); free(c->lights); free(c->nodes); free(c->prim_indices);
  memset(c, 0, sizeof *c);
}

/* tests/test_bvh.cpp:11-30 */
uint32_t orc_random_scene_objects(int spheres, int triangles, uint64_t seed, int with_mesh, orc_object* out) {
  orc_rng rng;
  orc_rng_init(&rng, seed, 0xda3e39cb94b95bdbull);
  uint32_t n = 0;
#define RP(s, dst)                                          \
  do {                                                      \
    (dst)[0] = orc__rng_uniform_range(&rng, -(s), (s));     \
    (dst)[1] = orc__rng_uniform_range(&rng, -(s), (s));     \
    (dst)[2] = orc__rng_uniform_range(&rng, -(s), (s));     \
  } while (0)
  for (int i = 0; i < spheres; ++i) {
    orc_object* o = &out[n++];
    memset(o, 0, sizeof *o);
    o->kind = ORC_OBJ_SPHERE;
    RP(10.0, o->a);
    o->scalar = orc__rng_uniform_range(&rng, 0.05, 1.0);
  }
  for (int i = 0; i < triangles; ++i) {
    orc_object* o = &out[n++];
    memset(o, 0, sizeof *o);
    o->kind = ORC_OBJ_TRIANGLE;
    double c[3], r[3];
    RP(10.0, c);
    RP(1.0, r); o->a[0] = c[0] + r[0]; o->a[1] = c[1] + r[1]; o->a[2] = c[2] + r[2];
    RP(1.0, r); o->b[0] = c[0] + r[0]; o->b[1] = c[1] + r[1]; o->b[2] = c[2] + r[2];
    RP(1.0, r); o->c[0] = c[0] + r[0]; o->c[1] = c[1] + r[1]; o->c[2] = c[2] + r[2];
  }
#undef RP
  if (with_mesh) {
    orc_object* o = &out[n++];
    memset(o, 0, sizeof *o);
    o->kind = ORC_OBJ_MESH;
    o->mesh_source = ORC_MESH_ICOSPHERE;
    o->subdivisions = 3;
    o->smooth = 1;
    o->scale[0] = o->scale[1] = o->scale[2] = 3;
  }
  orc_object* fl = &out[n++];
  memset(fl, 0, sizeof *fl);
  fl->kind = ORC_OBJ_PLANE;
  fl->a[1] = -12;
  fl->b[1] = 1;
  return n;
}
