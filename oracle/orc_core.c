/* Oracle (test infrastructure): RNG, BVH build, shapes, scene queries and
 * camera, restated from the reference. See rt_oracle.h. */
#include <stdlib.h>
#include <string.h>

#include "orc_math.h"

/* ===================== RNG: math/rng.hpp:10-51 ===================== */

uint64_t orc_splitmix64(uint64_t x) { /* rng.hpp:10-15 */
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint32_t orc_rng_next_u32(orc_rng* r) { /* rng.hpp:34-40 (PCG32 XSH-RR) */
  const uint64_t old = r->state;
  r->state = old * 6364136223846793005ull + r->inc;
  const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

void orc_rng_init(orc_rng* r, uint64_t seed, uint64_t stream) { /* rng.hpp:22-27 */
  r->state = 0;
  r->inc = (stream << 1u) | 1u;
  orc_rng_next_u32(r);
  r->state += seed;
  orc_rng_next_u32(r);
}

void orc_rng_for_sample(orc_rng* r, uint64_t seed, uint64_t pixel, uint64_t sample) { /* rng.hpp:29-32 */
  const uint64_t h = orc_splitmix64(seed ^ orc_splitmix64(pixel ^ orc_splitmix64(sample)));
  orc_rng_init(r, h, orc_splitmix64(h));
}

double orc_rng_uniform(orc_rng* r) { return orc_rng_next_u32(r) * (1.0 / 4294967296.0); } /* rng.hpp:43 */

static double rng_uniform_range(orc_rng* r, double lo, double hi) { return lo + (hi - lo) * orc_rng_uniform(r); }

static uint32_t rng_below(orc_rng* r, uint32_t n) { return (uint32_t)(orc_rng_uniform(r) * n); } /* rng.hpp:46 */

void orc_rng_stream(uint64_t seed, uint64_t pixel, uint64_t sample, uint32_t* out, int n) {
  orc_rng r;
  orc_rng_for_sample(&r, seed, pixel, sample);
  for (int i = 0; i < n; ++i) out[i] = orc_rng_next_u32(&r);
}

/* ===================== AABB: math/aabb.hpp ===================== */
typedef struct box { v3 lo, hi; } box;

static box box_empty(void) {
  box b = {{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
  return b;
}
static int box_is_empty(box b) { return b.lo.x > b.hi.x || b.lo.y > b.hi.y || b.lo.z > b.hi.z; }
static void box_expand_pt(box* b, v3 p) { /* aabb.hpp:22, vec3.hpp:76-81 (fmin/fmax) */
  b->lo = V(fmin(b->lo.x, p.x), fmin(b->lo.y, p.y), fmin(b->lo.z, p.z));
  b->hi = V(fmax(b->hi.x, p.x), fmax(b->hi.y, p.y), fmax(b->hi.z, p.z));
}
static void box_expand(box* b, box o) {
  b->lo = V(fmin(b->lo.x, o.lo.x), fmin(b->lo.y, o.lo.y), fmin(b->lo.z, o.lo.z));
  b->hi = V(fmax(b->hi.x, o.hi.x), fmax(b->hi.y, o.hi.y), fmax(b->hi.z, o.hi.z));
}
static double box_area(box b) { /* aabb.hpp:28-32 */
  if (box_is_empty(b)) return 0;
  const v3 e = vsub(b.hi, b.lo);
  return 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
}
static int box_longest_axis(box b) { /* aabb.hpp:34-37 */
  const v3 e = box_is_empty(b) ? V(0, 0, 0) : vsub(b.hi, b.lo);
  return (e.x >= e.y && e.x >= e.z) ? 0 : (e.y >= e.z ? 1 : 2);
}
static v3 box_centroid(box b) { return vscale(vadd(b.lo, b.hi), 0.5); }

/* aabb.hpp:57-66: one slab; NaN (parallel ray on a slab plane) = inside. */
static int slab(double lo, double hi, double o, double inv, double* tmin, double* tmax) {
  double t0 = (lo - o) * inv;
  double t1 = (hi - o) * inv;
  if (t0 != t0 || t1 != t1) return 1;
  if (t0 > t1) { double t = t0; t0 = t1; t1 = t; }
  *tmin = t0 > *tmin ? t0 : *tmin;
  *tmax = t1 < *tmax ? t1 : *tmax;
  return *tmin <= *tmax;
}
static int box_hit(const rtg_bvh_node* n, v3 o, v3 inv, double tmin, double tmax) { /* aabb.hpp:43-47 */
  return slab(n->lo.x, n->hi.x, o.x, inv.x, &tmin, &tmax) && slab(n->lo.y, n->hi.y, o.y, inv.y, &tmin, &tmax) &&
         slab(n->lo.z, n->hi.z, o.z, inv.z, &tmin, &tmax);
}

/* ===================== BVH build: src/geometry/bvh.cpp ===================== */
enum { kBins = 16, kMaxDepth = 60 };

typedef struct build_prim { box b; v3 c; uint32_t index; } build_prim;

typedef struct builder {
  build_prim* prims;
  rtg_bvh_node* nodes;
  uint32_t n_nodes;
  orc_bvh_stats* st;
  uint64_t max_leaf;
  int error;
} builder;

static int bin_of(double c, double lo, double scale) { /* bvh.cpp:112-115 */
  const int b = (int)((c - lo) * scale);
  return b < 0 ? 0 : (b > kBins - 1 ? kBins - 1 : b);
}

static void make_leaf(builder* B, uint32_t node, uint64_t begin, uint64_t n) { /* bvh.cpp:117-123 */
  if (n > 0xFFFF) { B->error = 1; return; }
  B->nodes[node].offset = (uint32_t)begin;
  B->nodes[node].count = (uint16_t)n;
  B->st->leaves++;
  if (n > B->st->max_leaf_size) B->st->max_leaf_size = n;
}

/* libstdc++'s std::partition on random-access ranges (the bidirectional
 * two-ended scheme) — the element order it leaves defines prim_indices. */
static uint64_t partition_prims(build_prim* p, uint64_t first, uint64_t last, int ax, double lo, double scale,
                                int split) {
  for (;;) {
    for (;;) {
      if (first == last) return first;
      if (bin_of(vget(p[first].c, ax), lo, scale) <= split) ++first;
      else break;
    }
    --last;
    for (;;) {
      if (first == last) return first;
      if (!(bin_of(vget(p[last].c, ax), lo, scale) <= split)) --last;
      else break;
    }
    build_prim t = p[first];
    p[first] = p[last];
    p[last] = t;
    ++first;
  }
}

static void build_rec(builder* B, uint64_t begin, uint64_t end, uint64_t depth) { /* bvh.cpp:28-109 */
  if (B->error) return;
  const uint32_t node = B->n_nodes++;
  memset(&B->nodes[node], 0, sizeof(rtg_bvh_node));
  box bounds = box_empty(), cb = box_empty();
  for (uint64_t i = begin; i < end; ++i) {
    box_expand(&bounds, B->prims[i].b);
    box_expand_pt(&cb, B->prims[i].c);
  }
  B->nodes[node].lo.x = bounds.lo.x; B->nodes[node].lo.y = bounds.lo.y; B->nodes[node].lo.z = bounds.lo.z;
  B->nodes[node].hi.x = bounds.hi.x; B->nodes[node].hi.y = bounds.hi.y; B->nodes[node].hi.z = bounds.hi.z;
  if (depth > B->st->max_depth) B->st->max_depth = depth;
  const uint64_t n = end - begin;
  if (n <= B->max_leaf || depth >= kMaxDepth) { make_leaf(B, node, begin, n); return; }

  int best_axis = -1, best_split = -1;
  double best_cost = INFINITY;
  const double parent_area = box_area(bounds);
  for (int ax = 0; ax < 3; ++ax) {
    const double lo = vget(cb.lo, ax), hi = vget(cb.hi, ax);
    if (!(hi > lo)) continue;
    box bin_b[kBins];
    uint64_t bin_n[kBins];
    for (int b = 0; b < kBins; ++b) { bin_b[b] = box_empty(); bin_n[b] = 0; }
    const double scale = kBins / (hi - lo);
    for (uint64_t i = begin; i < end; ++i) {
      const int b = bin_of(vget(B->prims[i].c, ax), lo, scale);
      bin_n[b]++;
      box_expand(&bin_b[b], B->prims[i].b);
    }
    double right_area[kBins] = {0};
    uint64_t right_count[kBins] = {0};
    box acc = box_empty();
    uint64_t cnt = 0;
    for (int b = kBins - 1; b > 0; --b) {
      box_expand(&acc, bin_b[b]);
      cnt += bin_n[b];
      right_area[b] = box_area(acc);
      right_count[b] = cnt;
    }
    acc = box_empty();
    cnt = 0;
    for (int b = 0; b < kBins - 1; ++b) {
      box_expand(&acc, bin_b[b]);
      cnt += bin_n[b];
      const uint64_t rc = right_count[b + 1];
      if (cnt == 0 || rc == 0) continue;
      const double denom = parent_area > 1e-30 ? parent_area : 1e-30;
      const double cost = 1.0 + (box_area(acc) * (double)cnt + right_area[b + 1] * (double)rc) / denom;
      if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = b; }
    }
  }

  uint64_t mid;
  if (best_axis < 0) {
    mid = begin + n / 2;
    best_axis = box_longest_axis(cb);
  } else {
    if (best_cost >= (double)n && n <= 16) { make_leaf(B, node, begin, n); return; }
    const double lo = vget(cb.lo, best_axis);
    const double scale = kBins / (vget(cb.hi, best_axis) - lo);
    mid = partition_prims(B->prims, begin, end, best_axis, lo, scale, best_split);
    if (mid == begin || mid == end) mid = begin + n / 2;
  }
  B->nodes[node].axis = (uint8_t)best_axis;
  build_rec(B, begin, mid, depth + 1);
  const uint32_t right = B->n_nodes;
  build_rec(B, mid, end, depth + 1);
  B->nodes[node].offset = right;
  B->nodes[node].count = 0;
}

int orc_bvh_build(const double* bounds, uint32_t n, int max_leaf, rtg_bvh_node* nodes_out, uint32_t* n_nodes,
                  uint32_t* prim_out, uint32_t* n_prim_out, orc_bvh_stats* stats) { /* bvh.cpp:133-163 */
  orc_bvh_stats st;
  memset(&st, 0, sizeof st);
  *n_nodes = 0;
  *n_prim_out = 0;
  build_prim* prims = (build_prim*)malloc(sizeof(build_prim) * (n ? n : 1));
  uint32_t np = 0;
  for (uint32_t i = 0; i < n; ++i) {
    box b = {vload(bounds + 6 * i), vload(bounds + 6 * i + 3)};
    if (box_is_empty(b) || !vis_finite(b.lo) || !vis_finite(b.hi)) continue;
    prims[np].b = b;
    prims[np].c = box_centroid(b);
    prims[np].index = i;
    ++np;
  }
  if (np == 0) {
    free(prims);
    if (stats) *stats = st;
    return 0;
  }
  builder B = {prims, nodes_out, 0, &st, (uint64_t)(max_leaf > 1 ? max_leaf : 1), 0};
  build_rec(&B, 0, np, 0);
  if (B.error) { free(prims); return -1; }
  for (uint32_t i = 0; i < np; ++i) prim_out[i] = prims[i].index;
  free(prims);
  *n_nodes = B.n_nodes;
  *n_prim_out = np;
  box root = {vfrom(nodes_out[0].lo), vfrom(nodes_out[0].hi)};
  double root_area = box_area(root);
  if (!(root_area > 1e-30)) root_area = 1e-30;
  for (uint32_t i = 0; i < B.n_nodes; ++i) {
    box b = {vfrom(nodes_out[i].lo), vfrom(nodes_out[i].hi)};
    const double rel = box_area(b) / root_area;
    st.sah_cost += rel * (nodes_out[i].count > 0 ? (double)nodes_out[i].count : 1.0);
  }
  st.nodes = B.n_nodes;
  if (stats) *stats = st;
  return 0;
}

int orc_bvh_validate(const rtg_bvh_node* nodes, uint32_t n_nodes, const uint32_t* pi, const double* bounds,
                     uint32_t n) { /* bvh.cpp:165-190 */
  const double e = 1e-9;
#define CONTAINS(olo, ohi, ilo, ihi)                                                                      \
  ((ilo).x >= (olo).x - e && (ilo).y >= (olo).y - e && (ilo).z >= (olo).z - e && (ihi).x <= (ohi).x + e && \
   (ihi).y <= (ohi).y + e && (ihi).z <= (ohi).z + e)
  int* seen = (int*)calloc(n ? n : 1, sizeof(int));
  int ok = 1;
  for (uint32_t i = 0; i < n_nodes && ok; ++i) {
    const rtg_bvh_node* nd = &nodes[i];
    if (nd->count > 0) {
      for (uint32_t k = 0; k < nd->count; ++k) {
        const uint32_t p = pi[nd->offset + k];
        v3 lo = vload(bounds + 6 * p), hi = vload(bounds + 6 * p + 3);
        if (!CONTAINS(nd->lo, nd->hi, lo, hi)) { ok = 0; break; }
        seen[p]++;
      }
    } else {
      if (nd->offset >= n_nodes || nd->offset <= i) { ok = 0; break; }
      if (!CONTAINS(nd->lo, nd->hi, nodes[i + 1].lo, nodes[i + 1].hi) ||
          !CONTAINS(nd->lo, nd->hi, nodes[nd->offset].lo, nodes[nd->offset].hi)) { ok = 0; break; }
    }
  }
  for (uint32_t p = 0; p < n && ok; ++p) {
    box b = {vload(bounds + 6 * p), vload(bounds + 6 * p + 3)};
    if (seen[p] != (box_is_empty(b) ? 0 : 1)) ok = 0;
  }
#undef CONTAINS
  free(seen);
  return ok;
}

/* ===================== Shapes: geometry/shapes.hpp ===================== */

void orc_prim_bounds(const rtg_primitive* p, double out[6]) {
  if (p->kind == RTG_PRIM_SPHERE) { /* shapes.hpp:68-71 */
    const double r = fabs(p->v1.x);
    const v3 c = vfrom(p->v0);
    vstore(out, vsub(c, V(r, r, r)));
    vstore(out + 3, vadd(c, V(r, r, r)));
  } else { /* shapes.hpp:152-159 */
    box b = box_empty();
    box_expand_pt(&b, vfrom(p->v0));
    box_expand_pt(&b, vfrom(p->v1));
    box_expand_pt(&b, vfrom(p->v2));
    const v3 pad = V(1e-9, 1e-9, 1e-9);
    vstore(out, vsub(b.lo, pad));
    vstore(out + 3, vadd(b.hi, pad));
  }
}

/* shapes.hpp:25-46 (Ray Tracing Gems ch. 7 stable quadratic). */
static int sphere_hit_t(v3 c, double radius, v3 o, v3 d, double tmin, double tmax, double* t_out) {
  const v3 oc = vsub(c, o);
  const double a = vlen2(d);
  if (a == 0) return 0;
  const double h = vdot(d, oc);
  const double cc = vlen2(oc) - radius * radius;
  const v3 v = vsub(oc, vscale(d, h / a));
  const double disc = a * (radius * radius - vlen2(v));
  if (disc < 0) return 0;
  const double sq = sqrt(disc);
  const double q = h + copysign(sq, h);
  double t0 = cc / q, t1 = q / a;
  if (q == 0) { t0 = t1 = h / a; }
  if (t0 > t1) { double t = t0; t0 = t1; t1 = t; }
  double t = t0;
  if (!(tmin < t && t < tmax)) {
    t = t1;
    if (!(tmin < t && t < tmax)) return 0;
  }
  *t_out = t;
  return 1;
}

/* shapes.hpp:110-127 (Möller–Trumbore, two-sided). */
static int tri_hit_t(const rtg_primitive* p, v3 o, v3 d, double tmin, double tmax, double* t_out, double bary[2]) {
  const v3 v0 = vfrom(p->v0);
  const v3 e1 = vsub(vfrom(p->v1), v0), e2 = vsub(vfrom(p->v2), v0);
  const v3 pvec = vcross(d, e2);
  const double det = vdot(e1, pvec);
  if (fabs(det) < 1e-14) return 0;
  const double inv_det = 1.0 / det;
  const v3 tvec = vsub(o, v0);
  const double b1 = vdot(tvec, pvec) * inv_det;
  if (b1 < 0 || b1 > 1) return 0;
  const v3 qvec = vcross(tvec, e1);
  const double b2 = vdot(d, qvec) * inv_det;
  if (b2 < 0 || b1 + b2 > 1) return 0;
  const double t = vdot(e2, qvec) * inv_det;
  if (!(tmin < t && t < tmax)) return 0;
  *t_out = t;
  bary[0] = b1;
  bary[1] = b2;
  return 1;
}

static int prim_hit_t(const rtg_primitive* p, v3 o, v3 d, double tmin, double tmax, double* t, double e[2]) {
  if (p->kind == RTG_PRIM_SPHERE) return sphere_hit_t(vfrom(p->v0), p->v1.x, o, d, tmin, tmax, t);
  return tri_hit_t(p, o, d, tmin, tmax, t, e);
}

/* hit.hpp:27-31 */
static void set_face_normal(orc_hit* rec, v3 d, v3 outward) {
  rec->front_face = vdot(d, outward) < 0;
  const v3 g = rec->front_face ? outward : vneg(outward);
  vstore(rec->gn, g);
  vstore(rec->normal, g);
}

/* shapes.hpp:49-58 and :129-142 (uv is not stored: no shading code reads it). */
static void prim_finalize(const rtg_primitive* p, v3 o, v3 d, double t, const double bary[2], orc_hit* rec) {
  rec->t = t;
  const v3 pt = vadd(o, vscale(d, t));
  vstore(rec->p, pt);
  if (p->kind == RTG_PRIM_SPHERE) {
    const v3 outward = vdiv(vsub(pt, vfrom(p->v0)), p->v1.x);
    set_face_normal(rec, d, outward);
  } else {
    const v3 v0 = vfrom(p->v0);
    const v3 ng = vnormalize(vcross(vsub(vfrom(p->v1), v0), vsub(vfrom(p->v2), v0)));
    set_face_normal(rec, d, ng);
    if (p->has_vertex_normals) {
      const double w0 = 1 - bary[0] - bary[1];
      const v3 s = vadd(vadd(vscale(vfrom(p->n0), w0), vscale(vfrom(p->n1), bary[0])), vscale(vfrom(p->n2), bary[1]));
      const v3 ns = vnormalize_or(s, ng);
      v3 nrm = rec->front_face ? ns : vneg(ns);
      const v3 g = vload(rec->gn);
      if (vdot(nrm, g) <= 0) nrm = g;
      vstore(rec->normal, nrm);
    }
  }
  rec->material = p->material;
}

/* shapes.hpp:82-97 */
static int plane_intersect(const rtg_plane* pl, v3 o, v3 d, double tmin, double tmax, orc_hit* rec) {
  const v3 n = vfrom(pl->normal);
  const double denom = vdot(n, d);
  if (fabs(denom) < 1e-12) return 0;
  const double t = vdot(vsub(vfrom(pl->point), o), n) / denom;
  if (!(tmin < t && t < tmax)) return 0;
  rec->t = t;
  vstore(rec->p, vadd(o, vscale(d, t)));
  set_face_normal(rec, d, n);
  rec->material = pl->material;
  return 1;
}

/* ===================== Scene queries: src/core/scene.cpp ===================== */

static int intersect_planes(const rtg_scene_desc* s, v3 o, v3 d, double tmin, double* tmax, orc_hit* rec) {
  int hit = 0; /* scene.cpp:152-166 */
  for (uint32_t i = 0; i < s->n_planes; ++i) {
    orc_hit tmp;
    memset(&tmp, 0, sizeof tmp);
    if (plane_intersect(&s->planes[i], o, d, tmin, *tmax, &tmp)) {
      tmp.primitive = i;
      tmp.is_plane = 1;
      *rec = tmp;
      *tmax = tmp.t;
      hit = 1;
    }
  }
  return hit;
}

static int intersect_brute(const rtg_scene_desc* s, v3 o, v3 d, double tmin, double tmax, orc_hit* rec,
                           uint64_t* pt) { /* scene.cpp:200-215 */
  double t_max = tmax;
  const int plane_hit = intersect_planes(s, o, d, tmin, &t_max, rec);
  double best_t = t_max, best_e[2] = {0, 0};
  uint32_t best = 0xFFFFFFFFu;
  for (uint32_t i = 0; i < s->n_prims; ++i) {
    if (pt) ++*pt;
    double th, e[2] = {0, 0};
    if (prim_hit_t(&s->prims[i], o, d, tmin, best_t, &th, e)) {
      best_t = th; best_e[0] = e[0]; best_e[1] = e[1]; best = i;
    }
  }
  if (best != 0xFFFFFFFFu) {
    prim_finalize(&s->prims[best], o, d, best_t, best_e, rec);
    rec->primitive = best;
    rec->is_plane = 0;
    return 1;
  }
  return plane_hit;
}

int orc_intersect(const rtg_scene_desc* s, int use_bvh, const double oo[3], const double dd[3], double tmin,
                  double tmax, orc_hit* rec, uint64_t* nv, uint64_t* pt) { /* scene.cpp:176-198 */
  const v3 o = vload(oo), d = vload(dd);
  memset(rec, 0, sizeof *rec);
  rec->t = INFINITY;
  rec->primitive = 0xFFFFFFFFu;
  rec->front_face = 1;
  if (!use_bvh) return intersect_brute(s, o, d, tmin, tmax, rec, pt);
  double t_max = tmax;
  const int plane_hit = intersect_planes(s, o, d, tmin, &t_max, rec);
  double best_t = t_max, best_e[2] = {0, 0};
  uint32_t best = 0xFFFFFFFFu;
  if (s->n_nodes) { /* Bvh::closest, bvh.hpp:75-104 */
    const v3 inv = V(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
    const int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    uint32_t stack[64];
    int sp = 0;
    uint32_t idx = 0;
    double tm = t_max;
    for (;;) {
      const rtg_bvh_node* n = &s->nodes[idx];
      if (nv) ++*nv;
      if (box_hit(n, o, inv, tmin, tm)) {
        if (n->count > 0) {
          for (uint32_t i = 0; i < n->count; ++i) {
            if (pt) ++*pt;
            const uint32_t pi = s->prim_indices[n->offset + i];
            double th, e[2] = {0, 0};
            double t = tm;
            if (prim_hit_t(&s->prims[pi], o, d, tmin, tm, &th, e)) {
              best_t = th; best_e[0] = e[0]; best_e[1] = e[1]; best = pi;
              t = th;
            }
            if (t < tm) tm = t;
          }
        } else {
          if (neg[n->axis]) { stack[sp++] = idx + 1; idx = n->offset; }
          else { stack[sp++] = n->offset; idx = idx + 1; }
          continue;
        }
      }
      if (sp == 0) break;
      idx = stack[--sp];
    }
  }
  if (best != 0xFFFFFFFFu) {
    prim_finalize(&s->prims[best], o, d, best_t, best_e, rec);
    rec->primitive = best;
    rec->is_plane = 0;
    return 1;
  }
  return plane_hit;
}

int orc_occluded(const rtg_scene_desc* s, int use_bvh, const double oo[3], const double dd[3], double tmax,
                 uint64_t* nv, uint64_t* pt) { /* scene.cpp:217-234 */
  const v3 o = vload(oo), d = vload(dd);
  const double tmin = 1e-6;
  for (uint32_t i = 0; i < s->n_planes; ++i) {
    orc_hit tmp;
    if (plane_intersect(&s->planes[i], o, d, tmin, tmax, &tmp)) return 1;
  }
  double th, e[2];
  if (!use_bvh) {
    for (uint32_t i = 0; i < s->n_prims; ++i) {
      if (pt) ++*pt;
      if (prim_hit_t(&s->prims[i], o, d, tmin, tmax, &th, e)) return 1;
    }
    return 0;
  }
  if (!s->n_nodes) return 0; /* Bvh::any, bvh.hpp:106-132 */
  const v3 inv = V(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  uint32_t stack[64];
  int sp = 0;
  uint32_t idx = 0;
  for (;;) {
    const rtg_bvh_node* n = &s->nodes[idx];
    if (nv) ++*nv;
    if (box_hit(n, o, inv, tmin, tmax)) {
      if (n->count > 0) {
        for (uint32_t i = 0; i < n->count; ++i) {
          if (pt) ++*pt;
          if (prim_hit_t(&s->prims[s->prim_indices[n->offset + i]], o, d, tmin, tmax, &th, e)) return 1;
        }
      } else {
        stack[sp++] = n->offset;
        idx = idx + 1;
        continue;
      }
    }
    if (sp == 0) break;
    idx = stack[--sp];
  }
  return 0;
}

void orc_trace_batch(const rtg_scene_desc* s, int mode, int use_bvh, const double* rays, uint64_t n, rtg_hit* hits,
                     uint64_t* nv, uint64_t* pt) {
  for (uint64_t i = 0; i < n; ++i) {
    const double* r = rays + 8 * i;
    rtg_hit h = {INFINITY, 0xFFFFFFFFu, 0};
    if (mode == 0) {
      orc_hit rec;
      if (orc_intersect(s, use_bvh, r, r + 3, r[6], r[7], &rec, nv, pt)) {
        h.t = rec.t;
        h.prim = rec.primitive;
        h.flags = 1u | (rec.is_plane ? 2u : 0u) | (rec.front_face ? 4u : 0u);
      }
    } else {
      h.flags = orc_occluded(s, use_bvh, r, r + 3, r[7], nv, pt) ? 1u : 0u;
    }
    hits[i] = h;
  }
}

/* ===================== Camera: core/camera.hpp:29-62 ===================== */

void orc_camera_setup(const double lf[3], const double la[3], const double vup[3], double vfov_deg, double aperture,
                      double focus_dist, int width, int height, rtg_camera* out) {
  const double aspect = (double)width / (double)height;
  const v3 origin = vload(lf);
  const v3 view = vsub(vload(la), origin);
  const v3 w = vnormalize_or(vneg(view), V(0, 0, 1));
  v3 up = vnormalize_or(vload(vup), V(0, 1, 0));
  if (fabs(vdot(up, w)) > 0.999999) up = fabs(w.y) < 0.9 ? V(0, 1, 0) : V(0, 0, 1);
  const v3 u = vnormalize(vcross(up, w));
  const v3 v = vcross(w, u);
  double focus;
  if (focus_dist > 0) focus = focus_dist;
  else focus = fmax(vlen(view), 1e-3);
  double fov = vfov_deg < 0.1 ? 0.1 : (vfov_deg > 179 ? 179 : vfov_deg);
  const double theta = fov * kPi / 180;
  const double vp_h = 2 * tan(theta / 2) * focus;
  const double vp_w = vp_h * aspect;
  const v3 vp_u = vscale(u, vp_w);
  const v3 vp_v = vscale(v, -vp_h);
  const v3 du = vdiv(vp_u, width), dv = vdiv(vp_v, height);
  const v3 corner = vsub(vsub(vsub(origin, vscale(w, focus)), vdiv(vp_u, 2)), vdiv(vp_v, 2));
#define PUT(dst, src) (dst).x = (src).x, (dst).y = (src).y, (dst).z = (src).z
  PUT(out->origin, origin); PUT(out->corner, corner); PUT(out->du, du); PUT(out->dv, dv);
  PUT(out->u, u); PUT(out->v, v); PUT(out->w, w);
#undef PUT
  out->lens_radius = aperture / 2;
}

/* sampling.hpp:38-42 */
static void random_in_unit_disk(orc_rng* rng, double* x, double* y) {
  const double r = sqrt(orc_rng_uniform(rng));
  const double phi = 2 * kPi * orc_rng_uniform(rng);
  *x = r * cos(phi);
  *y = r * sin(phi);
}

void orc_camera_generate(const rtg_camera* c, int i, int j, double sx, double sy, orc_rng* rng, double oo[3],
                         double dd[3]) {
  const v3 target = vadd(vadd(vfrom(c->corner), vscale(vfrom(c->du), i + sx)), vscale(vfrom(c->dv), j + sy));
  v3 o = vfrom(c->origin);
  if (c->lens_radius > 0) {
    double dx, dy;
    random_in_unit_disk(rng, &dx, &dy);
    o = vadd(vfrom(c->origin), vscale(vadd(vscale(vfrom(c->u), dx), vscale(vfrom(c->v), dy)), c->lens_radius));
  }
  vstore(oo, o);
  vstore(dd, vnormalize(vsub(target, o)));
}

/* Shared with orc_render.c */
int orc__prim_hit_t(const rtg_primitive* p, const double o[3], const double d[3], double tmin, double tmax,
                    double* t) {
  double e[2];
  return prim_hit_t(p, vload(o), vload(d), tmin, tmax, t, e);
}
uint32_t orc__rng_below(orc_rng* r, uint32_t n) { return rng_below(r, n); }
double orc__rng_uniform_range(orc_rng* r, double lo, double hi) { return rng_uniform_range(r, lo, hi); }

/* AABB::hit with inv = 1/d (aabb.hpp:43-52), for the slab known answers. */
int orc_box_hit(const double lo[3], const double hi[3], const double o[3], const double d[3], double tmin,
                double tmax) {
  rtg_bvh_node n;
  memset(&n, 0, sizeof n);
  n.lo.x = lo[0]; n.lo.y = lo[1]; n.lo.z = lo[2];
  n.hi.x = hi[0]; n.hi.y = hi[1]; n.hi.z = hi[2];
  return box_hit(&n, vload(o), V(1.0 / d[0], 1.0 / d[1], 1.0 / d[2]), tmin, tmax);
}
