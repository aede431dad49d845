// rtg_f64.hip — FP64 instantiation of the path-tracing kernels (the exact
// mode: the reference's algorithm, layouts and PCG32 stream in double).
// Built with -ffp-contract=off so no multiply-add is fused: every operation
// rounds as in the reference's x86-64 build.
#include <algorithm>

#include "rtg_launch.hpp"

namespace rtg {

namespace {
inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// tonemap.cpp:14-53
__device__ double clamp01(double v) { return v < 0 ? 0 : (v > 1 ? 1 : v); }
__device__ double aces(double x) {
  const double a = 2.51, b = 0.03, c = 2.43, d = 0.59, e = 0.14;
  const double v = (x * (a * x + b)) / (x * (c * x + d) + e);
  return v < 0 ? 0 : (v > 1 ? 1 : v);
}
__device__ uint8_t srgb8(double v) {
  v = clamp01(v);
  const double s = v <= 0.0031308 ? 12.92 * v : 1.055 * pow(v, 1.0 / 2.4) - 0.055;
  return (uint8_t)llround(clamp01(s) * 255.0);
}

__global__ void k_resolve(const double* sum, const uint32_t* count, int n, int op, double exposure, uint8_t* rgb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c[3] = {0, 0, 0};
  if (count[i]) {
    const double inv = 1.0 / (double)count[i];
    for (int k = 0; k < 3; ++k) c[k] = sum[3 * (size_t)i + k] * inv;
  }
  for (int k = 0; k < 3; ++k) c[k] = c[k] * exposure;
  if (!(isfinite(c[0]) && isfinite(c[1]) && isfinite(c[2]))) c[0] = c[1] = c[2] = 0;
  for (int k = 0; k < 3; ++k) {
    double v = c[k];
    if (op == 0) v = clamp01(v);
    else if (op == 1) v = v / (1 + v);
    else v = aces(v);
    rgb[3 * (size_t)i + k] = srgb8(v);
  }
}
}  // namespace

namespace {
template <int STACK>
void chunk_impl(const DevScene& S, const PathBuffers<double>& B, const ChunkParams& P, hipStream_t s,
                const SideStream* side, const LaunchHooks* h, hipEvent_t acc_wait, hipEvent_t acc_done) {
  const uint32_t g = grid_of(P.n_paths);
  if (P.integrator != RTG_INTEGRATOR_PATH) {
    RTG_LAUNCH(h, RTG_K_DIRECT, (k_direct64<STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P);
  } else {
    RTG_LAUNCH(h, RTG_K_GENERATE, k_generate<double>, dim3(grid_of(generate_threads(P))), dim3(kBlock), 0, s, S, B, P);
    const bool nee = P.light_sampling && S.n_lights > 0;
    // With a side stream, the any-hit kernel of bounce b runs beside the
    // closest-hit kernel of bounce b + 1 only: it adds to the radiance and
    // reads the one shadow-queue buffer, so the shade kernel of bounce b + 1
    // waits for it and the additions keep the reference's order.
    bool pending = false;
    for (int b = 0; b < P.max_depth; ++b) {
      RTG_LAUNCH(h, RTG_K_EXTEND, (k_extend<double, STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      if (pending) hipStreamWaitEvent(s, side->shadowed[0], 0);
      pending = false;
      RTG_LAUNCH(h, RTG_K_SHADE, k_shade<double>, dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      if (!nee) continue;
      if (side) {
        hipEventRecord(side->shaded, s);
        hipStreamWaitEvent(side->stream, side->shaded, 0);
        RTG_LAUNCH(h, RTG_K_SHADOW, (k_shadow<double, STACK>), dim3(g), dim3(kBlock), 0, side->stream, S, B, P, b);
        hipEventRecord(side->shadowed[0], side->stream);
        pending = true;
      } else {
        RTG_LAUNCH(h, RTG_K_SHADOW, (k_shadow<double, STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      }
    }
    if (pending) hipStreamWaitEvent(s, side->shadowed[0], 0);
  }
  if (acc_wait) hipStreamWaitEvent(s, acc_wait, 0);
  RTG_LAUNCH(h, RTG_K_ACCUMULATE, k_accumulate<double>, dim3(accumulate_blocks<double>(P)), dim3(kBlock), 0, s, B, P);
  if (acc_done) hipEventRecord(acc_done, s);
}

template <int STACK>
void trace_impl(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                unsigned long long* stats, hipStream_t s, const LaunchHooks* h) {
  RTG_LAUNCH(h, RTG_K_TRACE, (k_trace<double, STACK>), dim3(grid_of(n)), dim3(kBlock), 0, s, S, rays, n, mode,
             use_bvh, hits, stats);
}
}  // namespace

// The reference's traversal pushes one sibling per inner node on the path,
// so its stack, kept in LDS, needs the BVH depth in entries: the FP32 stack
// classes serve it.
void run_chunk(const DevScene& S, const PathBuffers<double>& B, const ChunkParams& P, uint32_t stack, hipStream_t s,
               const SideStream* side, const LaunchHooks* h, hipEvent_t acc_wait, hipEvent_t acc_done) {
  switch (stack_class_of(stack)) {
    case 0: chunk_impl<19>(S, B, P, s, side, h, acc_wait, acc_done); break;
    case 1: chunk_impl<31>(S, B, P, s, side, h, acc_wait, acc_done); break;
    default: chunk_impl<63>(S, B, P, s, side, h, acc_wait, acc_done); break;
  }
}

void run_trace64(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                 unsigned long long* stats, uint32_t stack, hipStream_t s, const LaunchHooks* h) {
  switch (stack_class_of(stack)) {
    case 0: trace_impl<19>(S, rays, n, mode, use_bvh, hits, stats, s, h); break;
    case 1: trace_impl<31>(S, rays, n, mode, use_bvh, hits, stats, s, h); break;
    default: trace_impl<63>(S, rays, n, mode, use_bvh, hits, stats, s, h); break;
  }
}

void run_resolve(const double* sum, const uint32_t* count, int w, int h, int op, double exposure, uint8_t* rgb,
                 hipStream_t s) {
  const int n = w * h;
  hipLaunchKernelGGL(k_resolve, dim3((n + 255) / 256), dim3(256), 0, s, sum, count, n, op, exposure, rgb);
}

}  // namespace rtg
