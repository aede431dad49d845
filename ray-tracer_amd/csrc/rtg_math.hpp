// rtg_math.hpp — device vector math for the path-tracing kernels.
//
// One template serves both precision modes. For T = double every operation
// follows the reference's evaluation order (math/vec3.hpp, math/ray.hpp,
// math/sampling.hpp) and the FP64 translation unit is built with
// -ffp-contract=off, so each double operation rounds exactly as on the CPU.
// For T = float the same code is built with contraction on and fast
// reciprocal / square root.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtg {

template <class T>
struct V3 {
  T x, y, z;
};

template <class T>
__device__ __forceinline__ V3<T> v3(T x, T y, T z) {
  return V3<T>{x, y, z};
}
template <class T>
__device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
__device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
__device__ __forceinline__ V3<T> operator-(V3<T> a) {
  return {-a.x, -a.y, -a.z};
}
// Component-wise product (colours, albedo filtering).
template <class T>
__device__ __forceinline__ V3<T> operator*(V3<T> a, V3<T> b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
// a * s and s * a both evaluate a.c * s (vec3.hpp:68-69).
template <class T>
__device__ __forceinline__ V3<T> operator*(V3<T> a, T s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class T>
__device__ __forceinline__ T dot(V3<T> a, V3<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
__device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
__device__ __forceinline__ T len2(V3<T> a) {
  return a.x * a.x + a.y * a.y + a.z * a.z;
}
template <class T>
__device__ __forceinline__ bool is_zero(V3<T> a) {
  return a.x == 0 && a.y == 0 && a.z == 0;
}
template <class T>
__device__ __forceinline__ T max_comp(V3<T> a) {
  return a.x > a.y ? (a.x > a.z ? a.x : a.z) : (a.y > a.z ? a.y : a.z);
}
template <class T>
__device__ __forceinline__ T get(V3<T> a, int i) {
  return i == 0 ? a.x : (i == 1 ? a.y : a.z);
}

// ---- precision-dependent primitives ----
__device__ __forceinline__ double r_sqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float r_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double r_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float r_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double r_fabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float r_fabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double r_fmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float r_fmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double r_fmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float r_fmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double r_copysign(double a, double b) { return __builtin_copysign(a, b); }
__device__ __forceinline__ float r_copysign(float a, float b) { return __builtin_copysignf(a, b); }
__device__ __forceinline__ bool r_isfinite(double a) { return __builtin_isfinite(a); }
__device__ __forceinline__ bool r_isfinite(float a) { return __builtin_isfinite(a); }
__device__ __forceinline__ double r_floor(double a) { return __builtin_floor(a); }
__device__ __forceinline__ float r_floor(float a) { return __builtin_floorf(a); }
__device__ __forceinline__ void r_sincos(double a, double* s, double* c) {
  *s = sin(a);
  *c = cos(a);
}
__device__ __forceinline__ void r_sincos(float a, float* s, float* c) { __sincosf(a, s, c); }

template <class T>
__device__ __forceinline__ V3<T> div(V3<T> a, T s) {  // vec3.hpp:48,70: a * (1 / s)
  return a * r_rcp(s);
}
template <class T>
__device__ __forceinline__ bool is_finite(V3<T> a) {
  return r_isfinite(a.x) && r_isfinite(a.y) && r_isfinite(a.z);
}

// vec3.hpp:85-91. Doubles divide exactly as the reference (v * (1/sqrt)).
__device__ __forceinline__ V3<double> normalize_or(V3<double> v, V3<double> fb) {
  const double l2 = len2(v);
  if (!(l2 > 1e-9 * 1e-9) || !r_isfinite(l2)) return fb;
  return v * (1.0 / r_sqrt(l2));
}
__device__ __forceinline__ V3<float> normalize_or(V3<float> v, V3<float> fb) {
  const float l2 = len2(v);
  if (!(l2 > 1e-18f) || !r_isfinite(l2)) return fb;
  return v * __builtin_amdgcn_rsqf(l2);
}
template <class T>
__device__ __forceinline__ V3<T> normalize(V3<T> v) {
  return normalize_or(v, V3<T>{0, 0, 0});
}

// vec3.hpp:94
template <class T>
__device__ __forceinline__ V3<T> reflect(V3<T> v, V3<T> n) {
  return v - n * (T(2) * dot(v, n));
}
// vec3.hpp:99-106
template <class T>
__device__ __forceinline__ bool refract(V3<T> uv, V3<T> n, T eta, V3<T>* out) {
  const T cos_i = r_fmin(dot(-uv, n), T(1));
  const T sin2_t = eta * eta * (T(1) - cos_i * cos_i);
  if (sin2_t > T(1)) return false;
  const T cos_t = r_sqrt(T(1) - sin2_t);
  *out = uv * eta + n * (eta * cos_i - cos_t);
  return true;
}

// ray.hpp:42-58
template <class T>
__device__ __forceinline__ T spawn_offset(V3<T> p) {
  T m = r_fabs(p.x);
  if (r_fabs(p.y) > m) m = r_fabs(p.y);
  if (r_fabs(p.z) > m) m = r_fabs(p.z);
  if (T(1) > m) m = T(1);
  return T(1e-7) * m * T(16);
}
template <class T>
__device__ __forceinline__ V3<T> offset_ray_origin(V3<T> p, V3<T> gn, V3<T> dir) {
  const T d = spawn_offset(p);
  return dot(dir, gn) > 0 ? p + gn * d : p - gn * d;
}

// sampling.hpp:13-25 (Duff et al. 2017)
template <class T>
struct Onb {
  V3<T> u, v, w;
};
template <class T>
__device__ __forceinline__ Onb<T> make_onb(V3<T> n) {
  Onb<T> b;
  b.w = n;
  const T sign = r_copysign(T(1), n.z);
  const T a = T(-1) / (sign + n.z);
  const T bb = n.x * n.y * a;
  b.u = V3<T>{T(1) + sign * n.x * n.x * a, sign * bb, -sign * n.x};
  b.v = V3<T>{bb, sign + n.y * n.y * a, -n.y};
  return b;
}
template <class T>
__device__ __forceinline__ V3<T> to_world(const Onb<T>& b, V3<T> a) {
  return b.u * a.x + b.v * a.y + b.w * a.z;
}

template <class T>
struct Consts;
template <>
struct Consts<double> {
  static constexpr double pi = 3.14159265358979323846;
  static constexpr double inv_pi = 1.0 / 3.14159265358979323846;
};
template <>
struct Consts<float> {
  static constexpr float pi = 3.14159265358979323846f;
  static constexpr float inv_pi = (float)(1.0 / 3.14159265358979323846);
};

}  // namespace rtg
