// rtg_f32.hip — FP32 instantiation of the path-tracing kernels (the
// production path: float arithmetic, counter-hash RNG, LDS traversal stack).
#include <algorithm>
#include <cstdlib>

#include "rtg_launch.hpp"

namespace rtg {

namespace {
inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// Workgroups of `kernel` resident on the whole device at once.
template <class K>
uint32_t resident_blocks(K kernel) {
  int dev = 0, cus = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kBlock, 0) != hipSuccess || n < 1) n = 1;
  return (uint32_t)n * (uint32_t)cus;
}

template <int STACK>
void chunk_impl(const DevScene& S, const PathBuffers<float>& B, const ChunkParams& P, hipStream_t s,
                const SideStream* side, const LaunchHooks* h, hipEvent_t acc_wait, hipEvent_t acc_done) {
  const uint32_t g = grid_of(P.n_paths);
  if (P.integrator == RTG_INTEGRATOR_DIRECT) {
    // The weighted ray queue: generation 0 in one kernel, then one extend /
    // shade / shadow round per depth (rounds past the last non-empty
    // generation exit at once), the depth-first recomputation (which returns
    // at once unless a generation overflowed) and the radiance.
    DirectMem M;
    direct_layout(static_cast<char*>(P.frames), P.frame_stride, P.dq_entries, S.n_lights, P.max_depth,
                  P.dq_forced != 0, &M);
    // Generations >= 1 run grid-stride over their queue with at most one
    // device-full of workgroups: a generation's size is known only on the
    // device, and a grid over every slot cost ~31 us per launch even when
    // the generation was empty (8 Mi-sample chunks).
    static const uint32_t r_ext = resident_blocks(k_dextend<STACK>), r_shade = resident_blocks(k_dshade<float>),
                          r_shadow = resident_blocks(k_dshadow<STACK>), r_dfs = resident_blocks(k_ddfs<STACK>);
    if (!M.forced) {
      const uint32_t slots = std::max<uint32_t>(M.cap, (uint32_t)kSubQueues * M.sub_gen);
      RTG_LAUNCH(h, RTG_K_GENERATE, k_dclear<float>, dim3(grid_of(dclear_threads())), dim3(kBlock), 0, s, M);
      RTG_LAUNCH(h, RTG_K_EXTEND, (k_dfirst<STACK>), dim3(g), dim3(kBlock), 0, s, S, P, M);
      for (int b = 1; b <= P.max_depth; ++b) {
        RTG_LAUNCH(h, RTG_K_EXTEND, (k_dextend<STACK>), dim3(std::min(grid_of(slots), r_ext)), dim3(kBlock), 0, s, S,
                   P, M, b);
        RTG_LAUNCH(h, RTG_K_SHADE, k_dshade<float>, dim3(std::min(grid_of(slots), r_shade)), dim3(kBlock), 0, s, S, P,
                   M, b);
        if (S.n_lights && b < P.max_depth)
          RTG_LAUNCH(h, RTG_K_SHADOW, (k_dshadow<STACK>),
                     dim3(std::min(grid_of((uint64_t)kSubQueues * M.sub_lit), r_shadow)), dim3(kBlock), 0, s, S, P, M,
                     b);
      }
    }
    M.dfs_threads = std::min(M.dfs_threads, r_dfs * (uint32_t)kBlock);  // fewer threads, each with more samples
    RTG_LAUNCH(h, RTG_K_DIRECT, (k_ddfs<STACK>), dim3(grid_of(M.dfs_threads)), dim3(kBlock), 0, s, S, P, M);
    RTG_LAUNCH(h, RTG_K_ACCUMULATE, k_dfinish<float>, dim3(grid_of(dfinish_threads(P))), dim3(kBlock), 0, s, B, P, M);
  } else if (P.integrator == RTG_INTEGRATOR_NORMALS) {
    RTG_LAUNCH(h, RTG_K_DIRECT, (k_direct32<STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P);
  } else {
    RTG_LAUNCH(h, RTG_K_GENERATE, k_generate<float>, dim3(grid_of(generate_threads(P))), dim3(kBlock), 0, s, S, B, P);
    const bool nee = P.light_sampling && S.n_lights > 0;
    // With a side stream, the any-hit kernels run there in bounce order,
    // beside the closest-hit and shade kernels of the following bounces: the
    // any-hit kernel of bounce b reads buffer b & 1 of the shadow queue and
    // adds to the shadow radiance, which nothing else touches until the
    // accumulation. The shade kernel of bounce b rewrites buffer b & 1, so it
    // waits for the any-hit kernel of bounce b - 2; the accumulation waits
    // for the last. Every path's additions keep their order.
    bool pending[2] = {false, false};
    // The traversal kernels are persistent (at most one device-full of
    // workgroups, idle lanes refilled between rounds of their walks) unless
    // RTG_PERSISTENT=0 (one ray per thread; the same results).
    static const uint32_t r_ext = resident_blocks(k_extend_dyn<STACK>), r_sh = resident_blocks(k_shadow_dyn<STACK>);
    const char* pe = std::getenv("RTG_PERSISTENT");
    const bool persistent = !(pe && pe[0] == '0');
    const uint32_t g_ext = std::min(g, r_ext), g_sh = std::min(g, r_sh);
    for (int b = 0; b < P.max_depth; ++b) {
      if (persistent)
        RTG_LAUNCH(h, RTG_K_EXTEND, (k_extend_dyn<STACK>), dim3(g_ext), dim3(kBlock), 0, s, S, B, P, b);
      else
        RTG_LAUNCH(h, RTG_K_EXTEND, (k_extend<float, STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      if (pending[b & 1]) hipStreamWaitEvent(s, side->shadowed[b & 1], 0);
      pending[b & 1] = false;
      RTG_LAUNCH(h, RTG_K_SHADE, k_shade<float>, dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      if (!nee) continue;
      if (side) {
        hipEventRecord(side->shaded, s);
        hipStreamWaitEvent(side->stream, side->shaded, 0);
        if (persistent)
          RTG_LAUNCH(h, RTG_K_SHADOW, (k_shadow_dyn<STACK>), dim3(g_sh), dim3(kBlock), 0, side->stream, S, B, P, b);
        else
          RTG_LAUNCH(h, RTG_K_SHADOW, (k_shadow<float, STACK>), dim3(g), dim3(kBlock), 0, side->stream, S, B, P, b);
        hipEventRecord(side->shadowed[b & 1], side->stream);
        pending[b & 1] = true;
      } else if (persistent) {
        RTG_LAUNCH(h, RTG_K_SHADOW, (k_shadow_dyn<STACK>), dim3(g_sh), dim3(kBlock), 0, s, S, B, P, b);
      } else {
        RTG_LAUNCH(h, RTG_K_SHADOW, (k_shadow<float, STACK>), dim3(g), dim3(kBlock), 0, s, S, B, P, b);
      }
    }
    for (int p = 0; p < 2; ++p)
      if (pending[p]) hipStreamWaitEvent(s, side->shadowed[p], 0);
  }
  if (acc_wait) hipStreamWaitEvent(s, acc_wait, 0);
  RTG_LAUNCH(h, RTG_K_ACCUMULATE, k_accumulate<float>, dim3(accumulate_blocks<float>(P)), dim3(kBlock), 0, s, B, P);
  if (acc_done) hipEventRecord(acc_done, s);
}

template <int STACK>
void batch_impl(const BatchJob* jobs, const PathBuffers<float>& B, const BatchPlans& L, int depth, const BatchGrids& g,
                hipStream_t s, const SideStream* side, const LaunchHooks* h) {
  RTG_LAUNCH(h, RTG_K_GENERATE, k_batch_generate<float>, dim3(g.generate), dim3(kBlock), 0, s, jobs, B, L);
  RTG_LAUNCH(h, RTG_K_PLAN, k_batch_plan<float>, dim3(1), dim3(kMaxBatchJobs), 0, s, jobs, L, -1);
  // As in a single render, the any-hit kernels run on the side stream; the
  // shade and planning kernels of bounce b (which rewrite buffer b & 1 of
  // the shadow queue and of the shadow plan) wait for the any-hit kernel of
  // bounce b - 2.
  bool pending[2] = {false, false};
  for (int b = 0; b < depth; ++b) {
    RTG_LAUNCH(h, RTG_K_EXTEND, (k_wide_extend_dyn<STACK>), dim3(g.trace), dim3(kBlock), 0, s, jobs, B, L, b);
    if (pending[b & 1]) hipStreamWaitEvent(s, side->shadowed[b & 1], 0);
    pending[b & 1] = false;
    RTG_LAUNCH(h, RTG_K_SHADE, k_batch_shade<float>, dim3(g.shade), dim3(kBlock), 0, s, jobs, B, L, b);
    RTG_LAUNCH(h, RTG_K_PLAN, k_batch_plan<float>, dim3(1), dim3(kMaxBatchJobs), 0, s, jobs, L, b);
    if (side) {
      hipEventRecord(side->shaded, s);
      hipStreamWaitEvent(side->stream, side->shaded, 0);
      RTG_LAUNCH(h, RTG_K_SHADOW, (k_wide_shadow_dyn<STACK>), dim3(g.trace), dim3(kBlock), 0, side->stream, jobs, B,
                 L, b);
      hipEventRecord(side->shadowed[b & 1], side->stream);
      pending[b & 1] = true;
    } else {
      RTG_LAUNCH(h, RTG_K_SHADOW, (k_wide_shadow_dyn<STACK>), dim3(g.trace), dim3(kBlock), 0, s, jobs, B, L, b);
    }
  }
  for (int p = 0; p < 2; ++p)
    if (pending[p]) hipStreamWaitEvent(s, side->shadowed[p], 0);
  RTG_LAUNCH(h, RTG_K_ACCUMULATE, k_batch_accumulate<float>, dim3(g.accumulate), dim3(kBlock), 0, s, jobs, B, L);
}

// A wide batch pass: the grids cover every path of the batch (path_blocks
// virtual blocks; generate and accumulate blocks as planned on the host).
template <int STACK>
void wide_impl(const BatchJob* jobs, const PathBuffers<float>& B, const BatchPlans& L, int depth, const BatchGrids& g,
               hipStream_t s, const SideStream* side, const LaunchHooks* h, hipEvent_t acc_wait, hipEvent_t acc_done) {
  RTG_LAUNCH(h, RTG_K_GENERATE, k_wide_generate<float>, dim3(g.generate), dim3(kBlock), 0, s, jobs, B, L);
  RTG_LAUNCH(h, RTG_K_PLAN, k_batch_plan<float>, dim3(1), dim3(kMaxBatchJobs), 0, s, jobs, L, -1);
  // Traversal as in a single render: persistent workgroups, idle lanes refilled.
  static const uint32_t r_ext = resident_blocks(k_wide_extend_dyn<STACK>),
                        r_sh = resident_blocks(k_wide_shadow_dyn<STACK>);
  const uint32_t g_ext = std::min(g.trace, r_ext), g_sh = std::min(g.trace, r_sh);
  bool pending[2] = {false, false};
  for (int b = 0; b < depth; ++b) {
    RTG_LAUNCH(h, RTG_K_EXTEND, (k_wide_extend_dyn<STACK>), dim3(g_ext), dim3(kBlock), 0, s, jobs, B, L, b);
    if (pending[b & 1]) hipStreamWaitEvent(s, side->shadowed[b & 1], 0);
    pending[b & 1] = false;
    RTG_LAUNCH(h, RTG_K_SHADE, k_wide_shade<float>, dim3(g.shade), dim3(kBlock), 0, s, jobs, B, L, b);
    RTG_LAUNCH(h, RTG_K_PLAN, k_batch_plan<float>, dim3(1), dim3(kMaxBatchJobs), 0, s, jobs, L, b);
    if (side) {
      hipEventRecord(side->shaded, s);
      hipStreamWaitEvent(side->stream, side->shaded, 0);
      RTG_LAUNCH(h, RTG_K_SHADOW, (k_wide_shadow_dyn<STACK>), dim3(g_sh), dim3(kBlock), 0, side->stream, jobs, B, L,
                 b);
      hipEventRecord(side->shadowed[b & 1], side->stream);
      pending[b & 1] = true;
    } else {
      RTG_LAUNCH(h, RTG_K_SHADOW, (k_wide_shadow_dyn<STACK>), dim3(g_sh), dim3(kBlock), 0, s, jobs, B, L, b);
    }
  }
  for (int p = 0; p < 2; ++p)
    if (pending[p]) hipStreamWaitEvent(s, side->shadowed[p], 0);
  if (acc_wait) hipStreamWaitEvent(s, acc_wait, 0);
  RTG_LAUNCH(h, RTG_K_ACCUMULATE, k_wide_accumulate<float>, dim3(g.accumulate), dim3(kBlock), 0, s, jobs, B, L);
  if (acc_done) hipEventRecord(acc_done, s);
}

template <int STACK>
uint32_t occupancy_trace() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wide_extend_dyn<STACK>, kBlock, 0) != hipSuccess) n = 1;
  int m = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&m, k_wide_shadow_dyn<STACK>, kBlock, 0) != hipSuccess) m = 1;
  return (uint32_t)std::max(1, std::min(n, m));
}

template <int STACK>
void trace_impl(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                unsigned long long* stats, hipStream_t s, const LaunchHooks* h) {
  RTG_LAUNCH(h, RTG_K_TRACE, (k_trace<float, STACK>), dim3(grid_of(n)), dim3(kBlock), 0, s, S, rays, n, mode, use_bvh,
             hits, stats);
}
}  // namespace

void run_chunk(const DevScene& S, const PathBuffers<float>& B, const ChunkParams& P, uint32_t stack, hipStream_t s,
               const SideStream* side, const LaunchHooks* h, hipEvent_t acc_wait, hipEvent_t acc_done) {
  switch (stack_class_of(stack)) {
    case 0: chunk_impl<19>(S, B, P, s, side, h, acc_wait, acc_done); break;
    case 1: chunk_impl<31>(S, B, P, s, side, h, acc_wait, acc_done); break;
    default: chunk_impl<63>(S, B, P, s, side, h, acc_wait, acc_done); break;
  }
}

void run_batch(const BatchJob* jobs, const PathBuffers<float>& B, const BatchPlans& L, int depth, uint32_t stack,
               const BatchGrids& g, hipStream_t s, const SideStream* side, const LaunchHooks* h) {
  switch (stack_class_of(stack)) {
    case 0: batch_impl<19>(jobs, B, L, depth, g, s, side, h); break;
    case 1: batch_impl<31>(jobs, B, L, depth, g, s, side, h); break;
    default: batch_impl<63>(jobs, B, L, depth, g, s, side, h); break;
  }
}

void run_wide(const BatchJob* jobs, const PathBuffers<float>& B, const BatchPlans& L, int depth, uint32_t stack,
              const BatchGrids& g, hipStream_t s, const SideStream* side, const LaunchHooks* h, hipEvent_t acc_wait,
              hipEvent_t acc_done) {
  switch (stack_class_of(stack)) {
    case 0: wide_impl<19>(jobs, B, L, depth, g, s, side, h, acc_wait, acc_done); break;
    case 1: wide_impl<31>(jobs, B, L, depth, g, s, side, h, acc_wait, acc_done); break;
    default: wide_impl<63>(jobs, B, L, depth, g, s, side, h, acc_wait, acc_done); break;
  }
}

void batch_occupancy(uint32_t stack, uint32_t* trace, uint32_t* shade, uint32_t* other) {
  switch (stack_class_of(stack)) {
    case 0: *trace = occupancy_trace<19>(); break;
    case 1: *trace = occupancy_trace<31>(); break;
    default: *trace = occupancy_trace<63>(); break;
  }
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batch_shade<float>, kBlock, 0) != hipSuccess) n = 1;
  *shade = (uint32_t)std::max(1, n);
  int m = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&m, k_batch_generate<float>, kBlock, 0) != hipSuccess) m = 1;
  int a = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_batch_accumulate<float>, kBlock, 0) != hipSuccess) a = 1;
  *other = (uint32_t)std::max(1, std::min(m, a));
}

void run_trace32(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                 unsigned long long* stats, uint32_t stack, hipStream_t s, const LaunchHooks* h) {
  switch (stack_class_of(stack)) {
    case 0: trace_impl<19>(S, rays, n, mode, use_bvh, hits, stats, s, h); break;
    case 1: trace_impl<31>(S, rays, n, mode, use_bvh, hits, stats, s, h); break;
    default: trace_impl<63>(S, rays, n, mode, use_bvh, hits, stats, s, h); break;
  }
}

}  // namespace rtg
