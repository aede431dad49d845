// vec.hpp — host FP64 vector helpers. Each operation is written in the
// evaluation order of the reference's Vec3 (include/raytracer/math/vec3.hpp)
// and the host code is built with -ffp-contract=off, so compiled scenes and
// BVHs come out bit-identical to the reference's.
#pragma once
#include <cmath>

#include "rtg.h"

namespace rtm {

struct D3 {
  double x = 0, y = 0, z = 0;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator-(D3 a) { return {-a.x, -a.y, -a.z}; }
inline D3 operator*(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator*(double s, D3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator/(D3 a, double s) { return a * (1.0 / s); }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double length2(D3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double length(D3 a) { return std::sqrt(length2(a)); }
inline bool finite3(D3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
inline D3 normalize_or(D3 v, D3 fallback) {
  const double l2 = length2(v);
  if (!(l2 > 1e-9 * 1e-9) || !std::isfinite(l2)) return fallback;
  return v / std::sqrt(l2);
}
inline D3 normalize(D3 v) { return normalize_or(v, D3{}); }
inline D3 vmin(D3 a, D3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline D3 vmax(D3 a, D3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline rtg_vec3 to_rtg(D3 a) { return {a.x, a.y, a.z}; }
inline D3 from_rtg(rtg_vec3 a) { return {a.x, a.y, a.z}; }

constexpr double kPi = 3.14159265358979323846;

}  // namespace rtm
