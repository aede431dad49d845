/* Internal to the oracle (test infrastructure): FP64 vector helpers written
 * in the reference's evaluation order (math/vec3.hpp, math/ray.hpp,
 * math/sampling.hpp). Compile with -ffp-contract=off. */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <math.h>

#include "rt_oracle.h"

typedef struct v3 { double x, y, z; } v3;

static const double kPi = 3.14159265358979323846;

static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vload(const double* p) { return V(p[0], p[1], p[2]); }
static inline v3 vfrom(rtg_vec3 a) { return V(a.x, a.y, a.z); }
static inline void vstore(double* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
/* vec3.hpp:68-69: s * a and a * s both evaluate a.c * s. */
static inline v3 vscale(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
/* vec3.hpp:48,70: a / s is a * (1.0 / s). */
static inline v3 vdiv(v3 a, double s) { return vscale(a, 1.0 / s); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vlen2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double vlen(v3 a) { return sqrt(vlen2(a)); }
static inline int vis_zero(v3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }
static inline int vis_finite(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static inline double vmaxc(v3 a) { return a.x > a.y ? (a.x > a.z ? a.x : a.z) : (a.y > a.z ? a.y : a.z); }
static inline double vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* vec3.hpp:85-91 */
static inline v3 vnormalize_or(v3 v, v3 fallback) {
  const double len2 = vlen2(v);
  if (!(len2 > 1e-9 * 1e-9) || !isfinite(len2)) return fallback;
  return vdiv(v, sqrt(len2));
}
static inline v3 vnormalize(v3 v) { return vnormalize_or(v, V(0, 0, 0)); }
/* vec3.hpp:94: v - 2 * dot(v, n) * n */
static inline v3 vreflect(v3 v, v3 n) { return vsub(v, vscale(n, 2 * vdot(v, n))); }
/* vec3.hpp:99-106 */
static inline int vrefract(v3 uv, v3 n, double eta, v3* out) {
  const double cos_i = fmin(vdot(vneg(uv), n), 1.0);
  const double sin2_t = eta * eta * (1.0 - cos_i * cos_i);
  if (sin2_t > 1.0) return 0;
  const double cos_t = sqrt(1.0 - sin2_t);
  *out = vadd(vscale(uv, eta), vscale(n, eta * cos_i - cos_t));
  return 1;
}

/* ray.hpp:42-58 */
static inline double spawn_offset(v3 p) {
  double m = fabs(p.x);
  if (fabs(p.y) > m) m = fabs(p.y);
  if (fabs(p.z) > m) m = fabs(p.z);
  if (1.0 > m) m = 1.0;
  return 1e-7 * m * 16;
}
static inline v3 offset_ray_origin(v3 p, v3 gn, v3 dir) {
  const double d = spawn_offset(p);
  return vdot(dir, gn) > 0 ? vadd(p, vscale(gn, d)) : vsub(p, vscale(gn, d));
}

/* sampling.hpp:13-25 (Duff et al. 2017 ONB) */
typedef struct onb { v3 u, v, w; } onb;
static inline onb onb_make(v3 n) {
  onb b;
  b.w = n;
  const double sign = copysign(1.0, n.z);
  const double a = -1.0 / (sign + n.z);
  const double bb = n.x * n.y * a;
  b.u = V(1.0 + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  b.v = V(bb, sign + n.y * n.y * a, -n.y);
  return b;
}
static inline v3 onb_to_world(const onb* b, v3 a) {
  return vadd(vadd(vscale(b->u, a.x), vscale(b->v, a.y)), vscale(b->w, a.z));
}

#endif
