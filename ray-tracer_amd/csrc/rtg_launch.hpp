// rtg_launch.hpp — host-side launchers implemented by the two kernel
// translation units (rtg_f32.hip, rtg_f64.hip) and called by rtg_api.hip.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "rtg_direct.hpp"
#include "rtg_kernels.hpp"

namespace rtg {

// LDS stack classes of the FP32 traversal kernels (slots per lane). A
// walk whose inner nodes lie at depth d or less uses slots 0 .. d + 1, so a
// scene whose BVH depth + 1 is `stack` (inner nodes at depth stack - 2 or
// less) fits a class of `stack` slots. 19 slots keep a traversal workgroup
// under 10 KB of LDS: 16 workgroups (8 waves per SIMD) fit a CU.
constexpr uint32_t kStackClasses[3] = {19, 31, 63};
inline int stack_class_of(uint32_t stack) {
  const uint32_t need = stack > 1 ? stack : 1;
  return need <= kStackClasses[0] ? 0 : need <= kStackClasses[1] ? 1 : 2;
}

// Optional per-launch timing (profiling): `events` hands out a pair of
// events for a launch of RTG_K_* kind `kind`, which the kernel's dispatch
// records at its start and end (hipExtLaunchKernel). The pair measures the
// kernel alone and adds no packets to the stream, unlike event records
// around the launch (measured: ~2.5 % of a torus render, 25 % of a 256x256
// Cornell render). Null events (creation failed): the launch is not timed.
struct LaunchHooks {
  void* user;
  void (*events)(void* user, int kind, hipEvent_t* start, hipEvent_t* stop);
};
#define RTG_LAUNCH(hooks, kind, kernel, grid, block, shmem, stream, ...)                          \
  do {                                                                                            \
    if (hooks) {                                                                                  \
      hipEvent_t rtg_start_ = nullptr, rtg_stop_ = nullptr;                                       \
      (hooks)->events((hooks)->user, (kind), &rtg_start_, &rtg_stop_);                            \
      hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, rtg_start_, rtg_stop_, 0u, __VA_ARGS__); \
    } else {                                                                                      \
      hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                        \
    }                                                                                             \
  } while (0)

// A second stream for work of a pass that may overlap the main stream's:
// the any-hit kernels run there, in bounce order, beside the closest-hit and
// shade kernels of the following bounces (FP32: they read their own shadow
// queue buffer and add to the shadow radiance only). The shade kernel of
// bounce b waits for the any-hit kernel of bounce b - 2, which read the same
// buffer. Null: everything on the main stream.
struct SideStream {
  hipStream_t stream;
  hipEvent_t shaded;       // main stream: shade of the current bounce done
  hipEvent_t shadowed[2];  // side stream: any-hit of the last bounce of each parity done
};

// Runs one chunk of a render pass: generate, max_depth x (extend, shade,
// shadow), accumulate; FP32 direct: the weighted ray queue (rtg_direct.hpp);
// FP64 direct and normals: the per-sample kernel. The
// accumulation waits for `acc_wait` (if not null: another lane's latest
// accumulation) and records `acc_done` (if not null).
void run_chunk(const DevScene& S, const PathBuffers<float>& B, const ChunkParams& P, uint32_t stack, hipStream_t s,
               const SideStream* side, const LaunchHooks* hooks, hipEvent_t acc_wait = nullptr,
               hipEvent_t acc_done = nullptr);
void run_chunk(const DevScene& S, const PathBuffers<double>& B, const ChunkParams& P, uint32_t stack, hipStream_t s,
               const SideStream* side, const LaunchHooks* hooks, hipEvent_t acc_wait = nullptr,
               hipEvent_t acc_done = nullptr);

// Grid sizes of one batched pass (persistent workgroups; see k_batch_*).
struct BatchGrids {
  uint32_t generate, trace, shade, accumulate;
};
// One batched pass (FP32 path integrator): generate, max_depth x (extend,
// shade, plan, shadow), accumulate over the jobs of `jobs` (device table).
void run_batch(const BatchJob* jobs, const PathBuffers<float>& B, const BatchPlans& L, int depth, uint32_t stack,
               const BatchGrids& g, hipStream_t s, const SideStream* side, const LaunchHooks* hooks);
// One wide batch pass (grids covering every path of the batch, one virtual
// block per workgroup); the accumulation waits for `acc_wait` and records
// `acc_done` when not null.
void run_wide(const BatchJob* jobs, const PathBuffers<float>& B, const BatchPlans& L, int depth, uint32_t stack,
              const BatchGrids& g, hipStream_t s, const SideStream* side, const LaunchHooks* hooks,
              hipEvent_t acc_wait = nullptr, hipEvent_t acc_done = nullptr);
// Resident workgroups per CU of the batch kernels (for the persistent grids).
void batch_occupancy(uint32_t stack, uint32_t* trace, uint32_t* shade, uint32_t* other);

void run_trace32(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                 unsigned long long* stats, uint32_t stack, hipStream_t s, const LaunchHooks* hooks);
void run_trace64(const DevScene& S, const double* rays, uint64_t n, int mode, int use_bvh, rtg_hit* hits,
                 unsigned long long* stats, uint32_t stack, hipStream_t s, const LaunchHooks* hooks);

// Tone map + sRGB8 (tonemap.cpp:14-53), FP64 like the reference.
void run_resolve(const double* sum, const uint32_t* count, int w, int h, int op, double exposure, uint8_t* rgb,
                 hipStream_t s);

}  // namespace rtg
