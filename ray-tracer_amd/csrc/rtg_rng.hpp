// rtg_rng.hpp — random numbers for the two precision modes.
//
// Pcg32: the reference's per-sample stream (math/rng.hpp:10-51), seeded by
//   splitmix64 of (seed, pixel, sample). Used by the FP64 mode so the device
//   consumes exactly the draws the CPU renderer consumes.
// Hash4: stateless counter-based generator for the FP32 mode. Every draw is
//   a pure function of (seed, pixel, sample, bounce, dimension), so a path's
//   numbers do not depend on queue order, chunking or which kernel runs it.
//   The mixing function is the pcg4d hash of Jarzynski & Olano, "Hash
//   Functions for GPU Rendering", JCGT 9(3), 2020.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtg {

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Pcg32 {
  uint64_t state, inc;

  __device__ __forceinline__ uint32_t next_u32() {
    const uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
  }
  __device__ __forceinline__ void init(uint64_t seed, uint64_t stream) {
    state = 0;
    inc = (stream << 1u) | 1u;
    next_u32();
    state += seed;
    next_u32();
  }
  __device__ __forceinline__ static Pcg32 for_sample(uint64_t seed, uint64_t pixel, uint64_t sample) {
    const uint64_t h = splitmix64(seed ^ splitmix64(pixel ^ splitmix64(sample)));
    Pcg32 r;
    r.init(h, splitmix64(h));
    return r;
  }
  __device__ __forceinline__ double uniform() { return next_u32() * (1.0 / 4294967296.0); }
};

// pcg4d (Jarzynski & Olano 2020).
__device__ __forceinline__ uint4 pcg4d(uint4 v) {
  v.x = v.x * 1664525u + 1013904223u;
  v.y = v.y * 1664525u + 1013904223u;
  v.z = v.z * 1664525u + 1013904223u;
  v.w = v.w * 1664525u + 1013904223u;
  v.x += v.y * v.w;
  v.y += v.z * v.x;
  v.z += v.x * v.y;
  v.w += v.y * v.z;
  v.x ^= v.x >> 16u;
  v.y ^= v.y >> 16u;
  v.z ^= v.z >> 16u;
  v.w ^= v.w >> 16u;
  v.x += v.y * v.w;
  v.y += v.z * v.x;
  v.z += v.x * v.y;
  v.w += v.y * v.z;
  return v;
}

// Key of one (seed, pixel, sample) path.
__device__ __forceinline__ uint2 path_key(uint64_t seed, uint32_t pixel, uint32_t sample) {
  const uint64_t h = splitmix64(seed ^ splitmix64(((uint64_t)pixel << 32) ^ (uint64_t)sample));
  return make_uint2((uint32_t)h, (uint32_t)(h >> 32));
}

// Four 32-bit draws for (key, bounce, group).
__device__ __forceinline__ uint4 draw4(uint2 key, uint32_t bounce, uint32_t group) {
  return pcg4d(make_uint4(key.x, key.y, bounce, group));
}

// [0, 1) with 24 bits, exactly representable in float.
__device__ __forceinline__ float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

}  // namespace rtg
