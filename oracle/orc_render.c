/* Oracle (test infrastructure): materials, integrators, tile renderer and
 * tone map restated from the reference. See rt_oracle.h. */
#include <stdlib.h>
#include <string.h>

#include "orc_math.h"

int orc__prim_hit_t(const rtg_primitive* p, const double o[3], const double d[3], double tmin, double tmax,
                    double* t);
uint32_t orc__rng_below(orc_rng* r, uint32_t n);

static const double kInvPi = 1.0 / 3.14159265358979323846;

/* ---- sampling.hpp ---- */
static v3 random_unit_vector(orc_rng* rng) { /* sampling.hpp:30-35 */
  const double z = 1 - 2 * orc_rng_uniform(rng);
  const double r = sqrt(fmax(0.0, 1 - z * z));
  const double phi = 2 * kPi * orc_rng_uniform(rng);
  return V(r * cos(phi), r * sin(phi), z);
}
static v3 cosine_hemisphere(orc_rng* rng) { /* sampling.hpp:38-42, 53-57 */
  const double r = sqrt(orc_rng_uniform(rng));
  const double phi = 2 * kPi * orc_rng_uniform(rng);
  const double dx = r * cos(phi), dy = r * sin(phi);
  const double z = sqrt(fmax(0.0, 1 - dx * dx - dy * dy));
  return V(dx, dy, z);
}
static v3 uniform_cone(orc_rng* rng, double cos_max) { /* sampling.hpp:62-67 */
  const double cos_t = 1 - orc_rng_uniform(rng) * (1 - cos_max);
  const double sin_t = sqrt(fmax(0.0, 1 - cos_t * cos_t));
  const double phi = 2 * kPi * orc_rng_uniform(rng);
  return V(sin_t * cos(phi), sin_t * sin(phi), cos_t);
}

/* ---- materials: src/materials/material.cpp ---- */
static v3 texture_sample(const rtg_material* m, v3 p) { /* material.cpp:14-24 */
  const double cs = m->p0;
  if (cs <= 0) return vfrom(m->c0);
  const double k = 1e-4;
  const long long parity = (long long)floor(p.x * cs + k) + (long long)floor(p.y * cs + k) +
                           (long long)floor(p.z * cs + k);
  return (parity & 1) ? vfrom(m->c1) : vfrom(m->c0);
}
static double schlick(double cos_theta, double eta) { /* material.cpp:26-31 */
  double r0 = (1 - eta) / (1 + eta);
  r0 *= r0;
  const double m = 1 - cos_theta;
  return r0 + (1 - r0) * m * m * m * m * m;
}
static v3 emission(const rtg_material* m) { /* material.cpp:76-79 */
  if (m->kind == RTG_MAT_EMISSIVE) return vscale(vfrom(m->c0), m->p0);
  return V(0, 0, 0);
}
static int emits_back_face(const rtg_material* m) { return m->kind == RTG_MAT_EMISSIVE && m->two_sided; }
static v3 emitted(const rtg_material* m, int front) { /* material.cpp:86-90 */
  if (m->kind == RTG_MAT_EMISSIVE && (front || m->two_sided)) return vscale(vfrom(m->c0), m->p0);
  return V(0, 0, 0);
}
static v3 eval_brdf_cos(const rtg_material* m, const orc_hit* rec, v3 wi) { /* material.cpp:94-100 */
  if (m->kind != RTG_MAT_LAMBERTIAN) return V(0, 0, 0);
  const double cos_t = vdot(vload(rec->normal), wi);
  if (cos_t <= 0 || vdot(vload(rec->gn), wi) <= 0) return V(0, 0, 0);
  return vscale(texture_sample(m, vload(rec->p)), cos_t * kInvPi);
}

typedef struct scatter_result { v3 dir, weight; int specular; } scatter_result;

static int scatter(const rtg_material* m, v3 in_dir, const orc_hit* rec, orc_rng* rng,
                   scatter_result* out) { /* material.cpp:33-74 */
  const v3 nrm = vload(rec->normal), gn = vload(rec->gn);
  switch (m->kind) {
    case RTG_MAT_LAMBERTIAN: {
      const onb b = onb_make(nrm);
      out->dir = onb_to_world(&b, cosine_hemisphere(rng));
      if (vdot(out->dir, gn) <= 0) return 0;
      out->weight = texture_sample(m, vload(rec->p));
      out->specular = 0;
      return 1;
    }
    case RTG_MAT_METAL: {
      const v3 unit_in = vnormalize(in_dir);
      v3 d = vreflect(unit_in, nrm);
      if (m->p0 > 0) d = vnormalize(vadd(vnormalize(d), vscale(random_unit_vector(rng), m->p0)));
      if (vdot(d, gn) <= 0) return 0;
      out->dir = d;
      out->weight = vfrom(m->c0);
      out->specular = 1;
      return 1;
    }
    case RTG_MAT_DIELECTRIC: {
      const double eta = rec->front_face ? (1.0 / m->p0) : m->p0;
      const v3 unit_in = vnormalize(in_dir);
      const double cos_i = fmin(vdot(vneg(unit_in), nrm), 1.0);
      v3 refracted;
      const int can_refract = vrefract(unit_in, nrm, eta, &refracted);
      if (!can_refract || orc_rng_uniform(rng) < schlick(cos_i, eta)) out->dir = vreflect(unit_in, nrm);
      else out->dir = vnormalize(refracted);
      out->weight = vfrom(m->c0);
      out->specular = 1;
      return 1;
    }
    default: return 0;
  }
}

/* ---- integrator.cpp ---- */
static v3 background_eval(const rtg_background* bg, v3 dir) { /* scene.cpp:10-14 */
  if (bg->kind == RTG_BG_SOLID) return vscale(vfrom(bg->top), bg->intensity);
  const double t = 0.5 * (vnormalize(dir).y + 1.0);
  return vscale(vadd(vscale(vfrom(bg->bottom), 1 - t), vscale(vfrom(bg->top), t)), bg->intensity);
}

static int prim_intersect_full(const rtg_primitive* p, v3 o, v3 d, double tmin, double tmax, orc_hit* rec);

/* integrator.cpp:14-91 */
static v3 sample_one_light(const rtg_scene_desc* s, const rtg_light_ref* light, const orc_hit* rec,
                           const rtg_material* mat, orc_rng* rng, rtg_stats* st, int use_bvh) {
  const v3 p = vload(rec->p);
  v3 wi;
  double dist = 0;
  v3 rop;
  if (light->kind == RTG_LIGHT_POINT) {
    const rtg_point_light* L = &s->point_lights[light->index];
    const v3 d = vsub(vfrom(L->position), p);
    const double d2 = vlen2(d);
    if (d2 < 1e-16) return V(0, 0, 0);
    dist = sqrt(d2);
    wi = vdiv(d, dist);
    rop = vdiv(vfrom(L->intensity), d2);
  } else {
    const rtg_primitive* prim = &s->prims[light->index];
    const rtg_material* lm = &s->materials[prim->material];
    if (prim->kind == RTG_PRIM_SPHERE) {
      const v3 c = vfrom(prim->v0);
      const double radius = prim->v1.x;
      const v3 oc = vsub(c, p);
      const double d2 = vlen2(oc);
      const double r2 = radius * radius;
      if (d2 <= r2 * (1 + 1e-9)) {
        if (!emits_back_face(lm)) return V(0, 0, 0);
        const v3 n_l = random_unit_vector(rng);
        const v3 d = vsub(vadd(c, vscale(n_l, radius)), p);
        const double dd2 = vlen2(d);
        if (dd2 < 1e-16) return V(0, 0, 0);
        dist = sqrt(dd2);
        wi = vdiv(d, dist);
        const double cos_l = fabs(vdot(n_l, wi));
        if (cos_l < 1e-9) return V(0, 0, 0);
        const double area = 4 * kPi * radius * radius;
        rop = vscale(emission(lm), cos_l * area / dd2);
      } else {
        const double x = r2 / d2;
        const double one_minus_cos = x / (1 + sqrt(1 - x));
        const double cos_max = 1 - one_minus_cos;
        const onb b = onb_make(vdiv(oc, sqrt(d2)));
        wi = onb_to_world(&b, uniform_cone(rng, cos_max));
        orc_hit lh;
        if (!prim_intersect_full(prim, p, wi, 0, INFINITY, &lh)) return V(0, 0, 0);
        dist = lh.t;
        const double pdf = 1 / (2 * kPi * one_minus_cos);
        rop = vdiv(emission(lm), pdf);
      }
    } else {
      const double su = sqrt(orc_rng_uniform(rng)); /* sampling.hpp:70-73 */
      const double b1 = 1 - su, b2 = orc_rng_uniform(rng) * su;
      const v3 v0 = vfrom(prim->v0), v1 = vfrom(prim->v1), v2 = vfrom(prim->v2);
      const v3 q = vadd(vadd(vscale(v0, 1 - b1 - b2), vscale(v1, b1)), vscale(v2, b2));
      const v3 d = vsub(q, p);
      const double d2 = vlen2(d);
      if (d2 < 1e-16) return V(0, 0, 0);
      dist = sqrt(d2);
      wi = vdiv(d, dist);
      const v3 cr = vcross(vsub(v1, v0), vsub(v2, v0));
      const v3 nl = vnormalize(cr);
      const double cos_signed = -vdot(nl, wi);
      if (cos_signed <= 0 && !emits_back_face(lm)) return V(0, 0, 0);
      const double cos_l = fabs(cos_signed);
      if (cos_l < 1e-9) return V(0, 0, 0);
      const double area = 0.5 * vlen(cr);
      const double pdf = d2 / (cos_l * area);
      rop = vdiv(emission(lm), pdf);
    }
  }
  const v3 f = eval_brdf_cos(mat, rec, wi);
  if (vis_zero(f) || vis_zero(rop)) return V(0, 0, 0);
  const v3 gn = vload(rec->gn);
  const v3 origin = offset_ray_origin(p, gn, wi);
  const v3 q = vadd(p, vscale(wi, dist));
  double stmax = vlen(vsub(q, origin)) - 4 * spawn_offset(q);
  if (stmax < 0) stmax = 0; /* std::max(x, 0): returns x unless x < 0 */
  st->shadow_rays++;
  double oo[3], dd[3];
  vstore(oo, origin);
  vstore(dd, wi);
  if (orc_occluded(s, use_bvh, oo, dd, stmax, &st->node_visits, &st->prim_tests)) return V(0, 0, 0);
  return vmul(f, rop);
}

static int prim_intersect_full(const rtg_primitive* p, v3 o, v3 d, double tmin, double tmax, orc_hit* rec) {
  double oo[3], dd[3], t;
  vstore(oo, o);
  vstore(dd, d);
  if (!orc__prim_hit_t(p, oo, dd, tmin, tmax, &t)) return 0;
  /* Sphere::intersect: only t is read by the caller. */
  rec->t = t;
  return 1;
}

static int is_sampled_light(const rtg_scene_desc* s, const orc_hit* rec) { /* integrator.cpp:93-95 */
  return !rec->is_plane && rec->primitive != 0xFFFFFFFFu && s->is_sampled_light[rec->primitive];
}

static int scene_intersect(const rtg_scene_desc* s, int use_bvh, v3 o, v3 d, orc_hit* rec, rtg_stats* st) {
  double oo[3], dd[3];
  vstore(oo, o);
  vstore(dd, d);
  return orc_intersect(s, use_bvh, oo, dd, 1e-6, INFINITY, rec, &st->node_visits, &st->prim_tests);
}

static v3 path_trace(const rtg_scene_desc* s, const rtg_render_settings* cfg, v3 ro, v3 rd, orc_rng* rng,
                     rtg_stats* st) { /* integrator.cpp:97-139 */
  v3 L = V(0, 0, 0), beta = V(1, 1, 1);
  int specular_bounce = 1;
  for (int depth = 0; depth < cfg->max_depth; ++depth) {
    orc_hit rec;
    st->rays++;
    if (!scene_intersect(s, cfg->use_bvh, ro, rd, &rec, st)) {
      L = vadd(L, vmul(beta, background_eval(&s->background, rd)));
      break;
    }
    const rtg_material* mat = &s->materials[rec.material];
    const v3 Le = emitted(mat, rec.front_face);
    if (!vis_zero(Le) && (specular_bounce || !cfg->light_sampling || !is_sampled_light(s, &rec)))
      L = vadd(L, vmul(beta, Le));
    scatter_result sr;
    if (!scatter(mat, rd, &rec, rng, &sr)) break;
    if (cfg->light_sampling && !sr.specular && s->n_lights > 0) {
      const uint32_t n = s->n_lights;
      const rtg_light_ref* light = &s->lights[orc__rng_below(rng, n)];
      const v3 c = sample_one_light(s, light, &rec, mat, rng, st, cfg->use_bvh);
      L = vadd(L, vscale(vmul(beta, c), (double)n));
    }
    beta = vmul(beta, sr.weight);
    specular_bounce = sr.specular;
    if (vis_zero(beta)) break;
    if (depth + 1 >= cfg->rr_min_depth) {
      double q = vmaxc(beta);
      q = q < 0.05 ? 0.05 : (q > 0.95 ? 0.95 : q);
      if (orc_rng_uniform(rng) >= q) { st->rr_terminated++; break; }
      beta = vdiv(beta, q);
    }
    ro = offset_ray_origin(vload(rec.p), vload(rec.gn), sr.dir);
    rd = sr.dir;
  }
  return L;
}

static v3 direct_trace(const rtg_scene_desc* s, const rtg_render_settings* cfg, v3 ro, v3 rd, orc_rng* rng,
                       rtg_stats* st, int depth) { /* integrator.cpp:141-174 */
  st->rays++;
  orc_hit rec;
  if (!scene_intersect(s, cfg->use_bvh, ro, rd, &rec, st)) return background_eval(&s->background, rd);
  const rtg_material* mat = &s->materials[rec.material];
  v3 L = emitted(mat, rec.front_face);
  if (depth >= cfg->max_depth) return L;
  const v3 p = vload(rec.p), gn = vload(rec.gn), nrm = vload(rec.normal);
  if (mat->kind == RTG_MAT_LAMBERTIAN) {
    for (uint32_t i = 0; i < s->n_lights; ++i)
      L = vadd(L, sample_one_light(s, &s->lights[i], &rec, mat, rng, st, cfg->use_bvh));
    return L;
  }
  if (mat->kind == RTG_MAT_METAL) {
    scatter_result sr;
    if (scatter(mat, rd, &rec, rng, &sr)) {
      const v3 c = direct_trace(s, cfg, offset_ray_origin(p, gn, sr.dir), sr.dir, rng, st, depth + 1);
      L = vadd(L, vmul(vfrom(mat->c0), c));
    }
    return L;
  }
  if (mat->kind == RTG_MAT_DIELECTRIC) {
    const v3 in = vnormalize(rd);
    const double eta = rec.front_face ? 1.0 / mat->p0 : mat->p0;
    const double cos_i = fmin(vdot(vneg(in), nrm), 1.0);
    v3 refracted;
    const int can = vrefract(in, nrm, eta, &refracted);
    const double F = can ? schlick(cos_i, eta) : 1.0;
    const v3 rdir = vreflect(in, nrm);
    const v3 c1 = direct_trace(s, cfg, offset_ray_origin(p, gn, rdir), rdir, rng, st, depth + 1);
    L = vadd(L, vmul(vscale(vfrom(mat->c0), F), c1));
    if (can) {
      const v3 tdir = vnormalize(refracted);
      const v3 c2 = direct_trace(s, cfg, offset_ray_origin(p, gn, tdir), tdir, rng, st, depth + 1);
      L = vadd(L, vmul(vscale(vfrom(mat->c0), 1 - F), c2));
    }
    return L;
  }
  return L;
}

void orc_radiance(const rtg_scene_desc* s, const rtg_render_settings* cfg, const double o[3], const double d[3],
                  orc_rng* rng, double out[3], rtg_stats* st) { /* integrator.cpp:184-198 */
  v3 L = V(0, 0, 0);
  switch (cfg->integrator) {
    case RTG_INTEGRATOR_PATH: L = path_trace(s, cfg, vload(o), vload(d), rng, st); break;
    case RTG_INTEGRATOR_DIRECT: L = direct_trace(s, cfg, vload(o), vload(d), rng, st, 0); break;
    default: {
      st->rays++;
      orc_hit rec;
      if (scene_intersect(s, cfg->use_bvh, vload(o), vload(d), &rec, st)) {
        const v3 n = rec.front_face ? vload(rec.normal) : vneg(vload(rec.normal));
        L = vscale(vadd(n, V(1, 1, 1)), 0.5);
      }
    }
  }
  vstore(out, L);
}

int orc_sample_direct_light_at(const rtg_scene_desc* s, const double o[3], const double d[3], uint64_t seed,
                               double out[3]) {
  rtg_stats st;
  memset(&st, 0, sizeof st);
  orc_hit rec;
  if (!orc_intersect(s, 1, o, d, 1e-6, INFINITY, &rec, NULL, NULL)) return 0;
  orc_rng rng;
  orc_rng_init(&rng, seed, 0xda3e39cb94b95bdbull);
  v3 sum = V(0, 0, 0);
  for (uint32_t i = 0; i < s->n_lights; ++i)
    sum = vadd(sum, sample_one_light(s, &s->lights[i], &rec, &s->materials[rec.material], &rng, &st, 1));
  vstore(out, sum);
  return 1;
}

/* renderer.cpp:21-56: one sample of pixel (x, y) with the invalid guard. */
static v3 render_sample(const rtg_scene_desc* s, const rtg_render_settings* cfg, int x, int y, uint32_t smp,
                        rtg_stats* st) {
  const uint64_t pixel = (uint64_t)y * (uint64_t)cfg->width + (uint64_t)x;
  orc_rng rng;
  orc_rng_for_sample(&rng, cfg->seed, pixel, smp);
  const double jx = orc_rng_uniform(&rng);
  const double jy = orc_rng_uniform(&rng);
  double o[3], d[3], L[3];
  orc_camera_generate(&s->camera, x, y, jx, jy, &rng, o, d);
  orc_radiance(s, cfg, o, d, &rng, L, st);
  v3 c = vload(L);
  if (!vis_finite(c) || c.x < 0 || c.y < 0 || c.z < 0) {
    st->invalid_samples++;
    c = V(0, 0, 0);
  }
  st->camera_samples++;
  return c;
}

void orc_render_rows(const rtg_scene_desc* s, const rtg_render_settings* cfg, uint32_t s0, uint32_t s1, int y0,
                     int y1, double* sum, uint32_t* count, rtg_stats* st) {
  for (int y = y0; y < y1; ++y)
    for (int x = 0; x < cfg->width; ++x) {
      const size_t i = (size_t)y * (size_t)cfg->width + (size_t)x;
      for (uint32_t smp = s0; smp < s1; ++smp) {
        const v3 c = render_sample(s, cfg, x, y, smp, st);
        sum[3 * i + 0] += c.x; /* framebuffer.hpp:31-35 */
        sum[3 * i + 1] += c.y;
        sum[3 * i + 2] += c.z;
        count[i]++;
      }
    }
}

void orc_sample_radiance_rows(const rtg_scene_desc* s, const rtg_render_settings* cfg, uint32_t s0, uint32_t s1,
                              int y0, int y1, double* out, rtg_stats* st) {
  const uint32_t k = s1 - s0;
  for (int y = y0; y < y1; ++y)
    for (int x = 0; x < cfg->width; ++x) {
      const size_t i = (size_t)y * (size_t)cfg->width + (size_t)x;
      for (uint32_t smp = s0; smp < s1; ++smp) vstore(out + 3 * (i * k + (smp - s0)), render_sample(s, cfg, x, y, smp, st));
    }
}

/* ---- tonemap.cpp:14-53 ---- */
static double aces(double x) {
  const double a = 2.51, b = 0.03, c = 2.43, d = 0.59, e = 0.14;
  const double v = (x * (a * x + b)) / (x * (c * x + d) + e);
  return v < 0 ? 0 : (v > 1 ? 1 : v);
}
static double clamp01(double v) { return v < 0 ? 0 : (v > 1 ? 1 : v); }

void orc_tone_map(const double in[3], int op, double exposure, double out[3]) {
  v3 c = vscale(vload(in), exposure);
  if (!vis_finite(c)) c = V(0, 0, 0);
  if (op == 0) c = V(clamp01(c.x), clamp01(c.y), clamp01(c.z));
  else if (op == 1) c = V(c.x / (1 + c.x), c.y / (1 + c.y), c.z / (1 + c.z));
  else c = V(aces(c.x), aces(c.y), aces(c.z));
  vstore(out, c);
}

uint8_t orc_linear_to_srgb8(double v) {
  v = clamp01(v);
  const double s = v <= 0.0031308 ? 12.92 * v : 1.055 * pow(v, 1.0 / 2.4) - 0.055;
  return (uint8_t)lround(clamp01(s) * 255.0);
}

void orc_resolve(const double* sum, const uint32_t* count, int w, int h, int op, double exposure, uint8_t* rgb) {
  size_t k = 0;
  for (int i = 0; i < w * h; ++i) {
    const v3 m = count[i] ? vdiv(vload(sum + 3 * i), (double)count[i]) : V(0, 0, 0);
    double in[3], out[3];
    vstore(in, m);
    orc_tone_map(in, op, exposure, out);
    rgb[k++] = orc_linear_to_srgb8(out[0]);
    rgb[k++] = orc_linear_to_srgb8(out[1]);
    rgb[k++] = orc_linear_to_srgb8(out[2]);
  }
}

/* The torus ray set: n rays from Rng(seed) (math/rng.hpp:22-27); origin =
 * c + e * U(-1,1)^3 with c, e the centroid and extent of the BVH root box
 * (a box twice the scene's), direction = random_unit_vector
 * (sampling.hpp:30-35); every component rounded to float, then widened
 * back to double, so FP32 and FP64 traversals see the same rays. */
void orc_ray_set(uint64_t seed, uint64_t n, const double lo[3], const double hi[3], double* rays) {
  orc_rng rng;
  orc_rng_init(&rng, seed, 0xda3e39cb94b95bdbull);
  const v3 c = vscale(vadd(vload(lo), vload(hi)), 0.5), e = vsub(vload(hi), vload(lo));
  for (uint64_t i = 0; i < n; ++i) {
    double* r = rays + 8 * i;
    const double ux = -1 + 2 * orc_rng_uniform(&rng);
    const double uy = -1 + 2 * orc_rng_uniform(&rng);
    const double uz = -1 + 2 * orc_rng_uniform(&rng);
    const v3 d = random_unit_vector(&rng);
    r[0] = (double)(float)(c.x + e.x * ux);
    r[1] = (double)(float)(c.y + e.y * uy);
    r[2] = (double)(float)(c.z + e.z * uz);
    r[3] = (double)(float)d.x;
    r[4] = (double)(float)d.y;
    r[5] = (double)(float)d.z;
    r[6] = 1e-6;
    r[7] = INFINITY;
  }
}
