// rtg_f32.hip — FP32 instantiation of the path-tracing kernels (the
// production path: float arithmetic, counter-hash RNG, LDS traversal stack).
#include <algorithm>

#include "rtg_launch.hpp"

namespace rtg {

namespace {
inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

template <int STACK>
void chunk_impl(const DevScene& S, const PathBuffers<float>& B, const ChunkParams& P, hipStream_t s,
                const LaunchHooks* h) {
  const uint32_t g = grid_of(P.n_paths);
  if (P.integrator != RTG_INTEGRATOR_PATH) {
    RTG_LAUNCH(h, RTG_K_DIRECT, (k_direct<float, STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P);
  } else {
    RTG_LAUNCH(h, RTG_K_GENERATE, k_generate<float>, dim3(grid_of(std::max<uint64_t>(P.n_paths, (uint64_t)(P.max_depth + 1) * kSubQueues))), dim3(kBlock), 0, s, S, B, P);
    const bool nee = P.light_sampling && S.n_lights > 0;
    for (int b = 0; b < P.max_depth; ++b) {
      RTG_LAUNCH(h, RTG_K_EXTEND, (k_extend<float, STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      RTG_LAUNCH(h, RTG_K_SHADE, k_shade<float>, dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      if (nee) RTG_LAUNCH(h, RTG_K_SHADOW, (k_shadow<float, STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P, b);
    }
  }
  RTG_LAUNCH(h, RTG_K_ACCUMULATE, k_accumulate<float>, dim3(grid_of(P.n_pix)), dim3(kBlock), 0, s, B, P);
}

template <int STACK>
void trace_impl(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                unsigned long long* stats, hipStream_t s, const LaunchHooks* h) {
  RTG_LAUNCH(h, RTG_K_TRACE, (k_trace<float, STACK>), dim3(grid_of(n)), dim3(kBlock), 0, s, S, rays, n, mode, use_bvh,
             hits, stats);
}
}  // namespace

void run_chunk(const DevScene& S, const PathBuffers<float>& B, const ChunkParams& P, uint32_t stack, hipStream_t s,
               const LaunchHooks* h) {
  if (stack <= kStackClasses[0]) chunk_impl<20>(S, B, P, s, h);
  else if (stack <= kStackClasses[1]) chunk_impl<32>(S, B, P, s, h);
  else chunk_impl<64>(S, B, P, s, h);
}

void run_trace32(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                 unsigned long long* stats, uint32_t stack, hipStream_t s, const LaunchHooks* h) {
  if (stack <= kStackClasses[0]) trace_impl<20>(S, rays, n, mode, use_bvh, hits, stats, s, h);
  else if (stack <= kStackClasses[1]) trace_impl<32>(S, rays, n, mode, use_bvh, hits, stats, s, h);
  else trace_impl<64>(S, rays, n, mode, use_bvh, hits, stats, s, h);
}

}  // namespace rtg
