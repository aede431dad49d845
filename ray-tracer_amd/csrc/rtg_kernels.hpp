// rtg_kernels.hpp — the wavefront path tracer and batch trace kernels,
// templated on the precision mode. Included by rtg_f32.hip and rtg_f64.hip,
// each of which instantiates one mode (the FP64 unit is compiled with
// -ffp-contract=off so it rounds like the reference CPU build).
//
// One render pass is processed in chunks of paths; path id
//   pid = (pixel - p0) * k + (sample - s0)
// so a chunk covers whole pixels with all k samples of the pass. Per chunk:
//   generate                       camera rays (renderer.cpp:31-37)
//   for bounce = 0 .. max_depth-1  (integrator.cpp:101-138)
//     extend    closest hit of the active queue
//     shade     emission, scatter, light choice -> shadow queue, Russian
//               roulette, next ray -> next queue (compacted: one atomic per
//               wavefront, ballot + mbcnt for the lane offsets)
//     shadow    any-hit of the shadow queue; unoccluded -> L += contribution
//   accumulate per-pixel sums in sample order (framebuffer.hpp:31-35)
// FP32 direct runs as a weighted ray queue (rtg_direct.hpp). FP64 direct
// (which must draw the reference's PCG32 stream in its depth-first order) and
// normals run one kernel per sample (direct's recursion a depth-first walk
// over an explicit frame stack).
#pragma once
#include "rtg_shade.hpp"

namespace rtg {

constexpr int kBlock = 128;

// kStRedone: direct-integrator chunks recomputed depth first (their queue filled).
enum StatIdx {
  kStCamera = 0, kStRays, kStShadow, kStRR, kStInvalid, kStVisits, kStTests, kStSVisits, kStSTests, kStRedone, kStCount
};

template <class T>
struct RayArrays {
  T *ox, *oy, *oz, *dx, *dy, *dz;  // the ray
  T *bx, *by, *bz;                 // throughput
  uint8_t* spec;                   // specular_bounce
  uint64_t *rs, *ri;               // PCG32 state / increment (FP64 mode; FP32: null)
};
template <class T>
struct PathBuffers {
  // Rays of bounce b with their paths' throughput and specular flag in
  // ray[b & 1], by queue slot (bounce 0: rays by path id; throughput 1 and
  // the flag set for every path, not stored), and the closest hits of bounce
  // b by the same slot: the closest-hit and shade kernels read and write
  // them with the queue's own (coalesced) order, and the shade kernel reads
  // them without waiting for the path id.
  RayArrays<T> ray[2];
  T *lx, *ly, *lz;                 // radiance (by path id)
  T *ht, *hb1, *hb2;               // hit t, barycentrics (by queue slot)
  uint32_t* hprim;                 // hit: primitive, kHitPlane | plane, or kHitNone (by queue slot)
  uint32_t* q[2];                  // active queues (kSubQueues sub-queues of sub_cap each)
  uint32_t* counts;                // sub-queue sizes (x kCountStride): [b*kSubQueues + s] active, [(kMaxBounce+b)*kSubQueues + s] shadow
  uint32_t* sq_pid;                // shadow queue (same sub-queue layout)
  T *sox, *soy, *soz, *sdx, *sdy, *sdz, *stm, *scx, *scy, *scz;
  // FP32: the shadow queue has two buffers (bounce parity; the second at
  // sq_stride) and shadow-ray radiance is summed apart from the path's own
  // (slx..: added to L when the path is accumulated), so the any-hit kernel
  // of bounce b can run beside the shade kernel of bounce b + 1. FP64 keeps
  // one buffer (sq_stride 0) and adds to L in the reference's order.
  T *slx, *sly, *slz;
  uint32_t sq_stride;
  unsigned long long* stats;       // kStatSlots slots of kStatStride counters
  uint32_t capacity;               // paths per chunk
  uint32_t sub_cap;                // entries per sub-queue
};
constexpr int kMaxBounce = 1024;
// Queue appends and statistics are spread over independent counters (one
// per sub-queue / slot, chosen by block index) so that no single address
// takes an atomic from every block of a launch.
constexpr int kSubQueues = 32;
// Queue counters sit kCountStride words (64 bytes) apart. The sub-queue
// counters take an atomic from every workgroup that appends; packed into one
// cache line they serialise there (measured: the direct integrator's shade
// kernel 19.9 -> 14.3 ms per sphere-field render with them apart; 16, 32,
// 64 and 1024 words measure the same).
#ifndef RTG_COUNT_STRIDE
#define RTG_COUNT_STRIDE 16
#endif
constexpr uint32_t kCountStride = RTG_COUNT_STRIDE;
constexpr int kStatSlots = 256;
constexpr int kStatStride = 16;  // counters per slot (128-byte lines)
__host__ __device__ constexpr uint32_t sub_queue_cap(uint32_t capacity) {
  return ((capacity + kBlock - 1) / kBlock + kSubQueues - 1) / kSubQueues * kBlock;
}

// One job of a pass: a range of pixels x a range of samples of one render.
// A single render's chunk is one job at path id 0; a batch packs many jobs
// into one set of path arrays, each with its own path-id range, queue
// storage, queue counters and statistics slots.
struct ChunkParams {
  uint32_t width, height;
  uint32_t p0, k, s0;  // pixels [p0, p0 + n_pix), samples [s0, s0 + k)
  uint32_t n_pix, n_paths;
  int32_t max_depth, rr_min_depth, integrator, light_sampling, use_bvh;
  int32_t sample_major;  // path id layout: 0 pixel-major (samples of a pixel adjacent), 1 sample-major
  uint64_t seed;
  double* acc_sum;      // width*height*3
  uint32_t* acc_count;  // width*height
  uint32_t pid_base;       // first path id of the job
  uint32_t q_base;         // first queue slot of the job
  uint32_t sub_cap;        // entries per sub-queue
  uint32_t shadow_counts;  // offset of the shadow-queue counters in `counts`
  uint32_t stat_slots;     // statistics slots of the job
  uint32_t zero_stats;     // generate clears the job's statistics slots
  uint32_t* counts;        // queue sizes (x kCountStride): [b*kSubQueues + s] active, [shadow_counts + b*kSubQueues + s] shadow
  unsigned long long* stats;
  void* frames;            // direct integrator: FP64 recursion frames (kFrameWords x max_depth per path), FP32 queues
  uint32_t frame_stride;   // paths per frame field (the chunk capacity)
  uint32_t dq_entries;     // FP32 direct: ray slots of one generation
  uint32_t dq_forced;      // FP32 direct: depth first from the start
};

// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Compacted appends of the whole block to two queues at once (the active
// queue and the shadow queue): wave ballots, an LDS total per block, one
// global atomic per queue per block on the block's own sub-queue counter.
// The block's shadow entries are grouped by `cls` (the light they go to:
// class 0 first), so the waves that trace them next hold rays to one light
// from nearby points rather than a mix of directions. Must be called by
// every thread of the block. Returns -1 or the slot.
struct AppendSlots {
  int64_t a, b;
};
__device__ __forceinline__ AppendSlots block_append2(uint32_t* counts_a, uint32_t* counts_b, uint32_t sub_cap,
                                                     bool pa, bool pb, uint32_t cls, uint32_t lb) {
  __shared__ uint32_t s_wave[3][kBlock / 64];
  __shared__ uint32_t s_base[2];
  const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const unsigned long long ma = __ballot(pa), mb0 = __ballot(pb && cls == 0), mb1 = __ballot(pb && cls != 0);
  if (lane == 0) {
    s_wave[0][wave] = (uint32_t)__popcll(ma);
    s_wave[1][wave] = (uint32_t)__popcll(mb0);
    s_wave[2][wave] = (uint32_t)__popcll(mb1);
  }
  __syncthreads();
  const uint32_t sub = lb % kSubQueues;
  if (threadIdx.x < 2) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < kBlock / 64; ++w)
      tot += threadIdx.x == 0 ? s_wave[0][w] : s_wave[1][w] + s_wave[2][w];
    uint32_t* c = (threadIdx.x == 0 ? counts_a : counts_b) + (size_t)sub * kCountStride;
    s_base[threadIdx.x] = tot ? atomicAdd(c, tot) : 0u;
  }
  __syncthreads();
  uint32_t oa = s_base[0], ob0 = s_base[1], ob1 = s_base[1];
  for (uint32_t w = 0; w < kBlock / 64; ++w) {
    ob1 += s_wave[1][w];  // every class-0 entry of the block comes first
    if (w < wave) {
      oa += s_wave[0][w];
      ob0 += s_wave[1][w];
      ob1 += s_wave[2][w];
    }
  }
  AppendSlots r;
  r.a = pa ? (int64_t)sub * sub_cap + oa + lanes_below(ma) : -1;
  r.b = pb ? (int64_t)sub * sub_cap + (cls == 0 ? ob0 + lanes_below(mb0) : ob1 + lanes_below(mb1)) : -1;
  __syncthreads();  // s_wave / s_base are reused by the block's next job
  return r;
}

// Maps dense index i of a queue to its storage slot: the block reads the
// kSubQueues counts once into LDS as a prefix sum. Returns false past the end.
struct QueueView {
  uint32_t total;
  bool past_end;  // the block has no entries: it returns before any barrier
};
// Every wave sums the kSubQueues sub-queue sizes itself (one load per lane
// and a shuffle scan), so both waves of a block agree on the total without
// LDS or a barrier, and blocks past the end of the queue (the grid covers
// every path, the queue shrinks with each bounce) exit at once. Other blocks
// publish the exclusive prefix sums in pre[] for queue_slot. A sub-queue
// holds at most `cap` entries (appends past it are dropped by the producer).
__device__ __forceinline__ QueueView load_queue(const uint32_t* counts, uint32_t* pre, uint32_t lb,
                                                uint32_t cap = ~0u, uint32_t stride = 1) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t c = lane < (uint32_t)kSubQueues ? counts[lane * stride] : 0u;
  c = c < cap ? c : cap;
  uint32_t x = c;
  for (int off = 1; off < kSubQueues; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  const uint32_t total = __shfl(x, kSubQueues - 1, 64);
  if (lb * (uint32_t)kBlock >= total) return QueueView{total, true};
  if (threadIdx.x < (uint32_t)kSubQueues) pre[threadIdx.x] = x - c;
  if (threadIdx.x == 0) pre[kSubQueues] = total;
  __syncthreads();
  return QueueView{total, false};
}
__device__ __forceinline__ uint32_t queue_slot(const uint32_t* pre, uint32_t sub_cap, uint32_t i) {
  int lo = 0, hi = kSubQueues - 1;  // last k with pre[k] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return (uint32_t)lo * sub_cap + (i - pre[lo]);
}

// Block-level reduction of per-thread counters into one of `slots` slots
// (the host sums the slots).
__device__ __forceinline__ void flush_stats(unsigned long long* g, const uint32_t* v, int n, uint32_t slots,
                                            uint32_t lb) {
  __shared__ unsigned long long sh[kStCount];
  if (threadIdx.x < kStCount) sh[threadIdx.x] = 0;
  __syncthreads();
  for (int i = 0; i < n; ++i)
    if (v[i]) atomicAdd(&sh[i], (unsigned long long)v[i]);
  __syncthreads();
  if (threadIdx.x < (unsigned)n && sh[threadIdx.x])
    atomicAdd(&g[(lb % slots) * kStatStride + threadIdx.x], sh[threadIdx.x]);
  __syncthreads();  // sh is reused by the block's next job
}
__device__ __forceinline__ void flush_stats(const ChunkParams& P, const uint32_t* v, uint32_t lb) {
  flush_stats(P.stats, v, kStCount, P.stat_slots, lb);
}

template <class T>
struct ModeOps;

template <>
struct ModeOps<double> {
  using R = PcgDraws;
  __device__ static Hit64 closest(const DevScene& S, bool bvh, V3<double> o, V3<double> d, LdsStack stk,
                                 Counters& c) {
    return closest64(S, bvh, o, d, 1e-6, __builtin_inf(), stk, c);
  }
  __device__ static bool occluded(const DevScene& S, bool bvh, V3<double> o, V3<double> d, double tmax,
                                  LdsStack stk, Counters& c) {
    return any64(S, bvh, o, d, 1e-6, tmax, stk, c);
  }
  __device__ static void finalize(const DevScene& S, V3<double> o, V3<double> d, double t, double b1, double b2,
                                  uint32_t prim, int plane, HitRec<double>& rec) {
    finalize64(S, o, d, t, b1, b2, prim, plane, rec);
  }
  __device__ static bool light(const DevScene& S, uint32_t li, const HitRec<double>& rec, const Mat<double>& m, R& rng,
                               ShadowReq<double>& req) {
    return sample_light64(S, li, rec, m, rng, req);
  }
  __device__ static void camera(const DevScene& S, int x, int y, double sx, double sy, R& rng, V3<double>& o,
                                V3<double>& d) {
    camera64(S, x, y, sx, sy, rng, o, d);
  }
};

template <>
struct ModeOps<float> {
  using R = HashDraws;
  __device__ static Hit32 closest(const DevScene& S, bool bvh, V3<float> o, V3<float> d, LdsStack stk, Counters& c) {
    const Ray32 r = make_ray32(S, o.x, o.y, o.z, d.x, d.y, d.z);
    return closest32(S, bvh, r, 1e-6f, __builtin_inff(), stk, c);
  }
  __device__ static bool occluded(const DevScene& S, bool bvh, V3<float> o, V3<float> d, float tmax, LdsStack stk,
                                  Counters& c) {
    const Ray32 r = make_ray32(S, o.x, o.y, o.z, d.x, d.y, d.z);
    return any32(S, bvh, r, 1e-6f, tmax, stk, c);
  }
  __device__ static void finalize(const DevScene& S, V3<float> o, V3<float> d, float t, float b1, float b2,
                                  uint32_t prim, int plane, HitRec<float>& rec) {
    finalize32(S, o, d, t, b1, b2, prim, plane, rec);
  }
  __device__ static bool light(const DevScene& S, uint32_t li, const HitRec<float>& rec, const Mat<float>& m, R& rng,
                               ShadowReq<float>& req) {
    return sample_light32(S, li, rec, m, rng, req);
  }
  __device__ static void camera(const DevScene& S, int x, int y, float sx, float sy, R& rng, V3<float>& o,
                                V3<float>& d) {
    camera32(S, x, y, sx, sy, rng, o, d);
  }
};

// Draw source of one path at one bounce.
template <class T>
struct DrawSrc;
// FP64: the path's PCG32 stream, its state kept with the ray by queue slot
// (bounce b at slot k of ray[b & 1]); a path that lives on hands it to the
// next bounce's slot. FP32: a counter hash, nothing kept.
template <>
struct DrawSrc<double> {
  struct Next {
    uint64_t state = 0, inc = 0;
  };
  __device__ static PcgDraws make(const PathBuffers<double>& B, const ChunkParams&, int b, uint32_t k, uint32_t,
                                  uint32_t) {
    const RayArrays<double>& R = B.ray[b & 1];
    PcgDraws d;
    d.pcg.state = R.rs[k];
    d.pcg.inc = R.ri[k];
    return d;
  }
  __device__ static Next next(const PcgDraws& d) { return Next{d.pcg.state, d.pcg.inc}; }
  __device__ static void store(const PathBuffers<double>& B, int b, uint32_t k, const Next& n) {
    const RayArrays<double>& R = B.ray[b & 1];
    R.rs[k] = n.state;
    R.ri[k] = n.inc;
  }
};
template <>
struct DrawSrc<float> {
  struct Next {};
  __device__ static HashDraws make(const PathBuffers<float>&, const ChunkParams& P, int b, uint32_t, uint32_t pixel,
                                   uint32_t sample) {
    return HashDraws(path_key(P.seed, pixel, sample), (uint32_t)b);
  }
  __device__ static Next next(const HashDraws&) { return Next{}; }
  __device__ static void store(const PathBuffers<float>&, int, uint32_t, const Next&) {}
};

static_assert(kBlock == (int)kStackStride, "traversal stacks are laid out for kBlock threads");
__device__ __forceinline__ LdsStack lds_stack(uint32_t* ref) { return LdsStack{ref + threadIdx.x}; }

// ---------------------------------------------------------------------------
// Path id -> (pixel, sample). Pixel-major keeps the samples of one pixel
// in adjacent lanes (their rays start together: the most coherent
// traversal); sample-major puts neighbouring pixels of one sample together
// (coalesced accumulation). Results do not depend on the layout.
__device__ __forceinline__ uint32_t pid_pixel(const ChunkParams& P, uint32_t pid) {
  const uint32_t l = pid - P.pid_base;
  return P.p0 + (P.sample_major ? l % P.n_pix : l / P.k);
}
__device__ __forceinline__ uint32_t pid_sample(const ChunkParams& P, uint32_t pid) {
  const uint32_t l = pid - P.pid_base;
  return P.s0 + (P.sample_major ? l / P.n_pix : l % P.k);
}
__device__ __forceinline__ uint32_t pid_of(const ChunkParams& P, uint32_t j, uint32_t s) {
  return P.pid_base + (P.sample_major ? s * P.n_pix + j : j * P.k + s);
}

// Threads a job's generate step needs: one per path, and enough to clear
// the job's queue counters (and statistics slots when asked).
__host__ __device__ inline uint32_t generate_threads(const ChunkParams& P) {
  uint32_t n = P.n_paths;
  const uint32_t nz = (uint32_t)(P.max_depth + 1) * kSubQueues;
  if (nz > n) n = nz;
  if (P.zero_stats && P.stat_slots * kStatStride > n) n = P.stat_slots * kStatStride;
  return n;
}

// generate: one camera ray per path (renderer.cpp:31-37). The bounce-0
// queue is the identity and is not stored. lb = the job's block index.
template <class T>
__device__ __forceinline__ void generate_body(const DevScene& S, const PathBuffers<T>& B, const ChunkParams& P,
                                              uint32_t lb) {
  const uint32_t l = lb * kBlock + threadIdx.x;
  const uint32_t nz = (uint32_t)(P.max_depth + 1) * kSubQueues;
  if (l < nz) {
    P.counts[(size_t)l * kCountStride] = l == 0 ? P.n_paths : 0u;
    P.counts[(size_t)(P.shadow_counts + l) * kCountStride] = 0u;
  }
  if (P.zero_stats && l < P.stat_slots * kStatStride) P.stats[l] = 0;
  if (l >= P.n_paths) return;
  const uint32_t pid = P.pid_base + l;
  const uint32_t pixel = pid_pixel(P, pid);
  const uint32_t sample = pid_sample(P, pid);
  const int x = (int)(pixel % P.width), y = (int)(pixel / P.width);
  V3<T> o, d;
  if constexpr (sizeof(T) == 8) {
    PcgDraws rng;
    rng.pcg = Pcg32::for_sample(P.seed, pixel, sample);
    const double jx = rng.next();
    const double jy = rng.next();
    ModeOps<T>::camera(S, x, y, jx, jy, rng, o, d);
    B.ray[0].rs[pid] = rng.pcg.state;
    B.ray[0].ri[pid] = rng.pcg.inc;
  } else {
    HashDraws rng(path_key(P.seed, pixel, sample), 0xFFFFFFFFu);
    const float jx = rng.next();
    const float jy = rng.next();
    ModeOps<T>::camera(S, x, y, jx, jy, rng, o, d);
  }
  const RayArrays<T>& R = B.ray[0];
  R.ox[pid] = o.x; R.oy[pid] = o.y; R.oz[pid] = o.z;
  R.dx[pid] = d.x; R.dy[pid] = d.y; R.dz[pid] = d.z;
  B.lx[pid] = 0; B.ly[pid] = 0; B.lz[pid] = 0;
  if constexpr (sizeof(T) == 4) {
    B.slx[pid] = 0; B.sly[pid] = 0; B.slz[pid] = 0;
  }
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_generate(DevScene S, PathBuffers<T> B, ChunkParams P) {
  generate_body<T>(S, B, P, blockIdx.x);
}

// Queue slot of active entry i at bounce b (its path id at bounce 0, whose
// queue is the identity and is not stored), and the entry's path id.
__device__ __forceinline__ uint32_t active_slot(const ChunkParams& P, const uint32_t* pre, int b, uint32_t i) {
  return b == 0 ? P.pid_base + i : P.q_base + queue_slot(pre, P.sub_cap, i);
}
template <class T>
__device__ __forceinline__ uint32_t slot_pid(const PathBuffers<T>& B, int b, uint32_t slot) {
  return b == 0 ? slot : B.q[b & 1][slot];
}

// A closest hit for the shade kernel in one word (primitive ids stay below
// 2^25), which reads the distance of hits only and the barycentrics of
// primitive hits only: a miss stores its word, a plane hit its word and
// distance.
constexpr uint32_t kHitNone = 0xFFFFFFFFu, kHitPlane = 0x80000000u;
template <class T, class H>
__device__ __forceinline__ void store_hit(const PathBuffers<T>& B, uint32_t k, const H& h) {
  B.hprim[k] = h.prim != 0xFFFFFFFFu ? h.prim : h.plane >= 0 ? kHitPlane | (uint32_t)h.plane : kHitNone;
  if (h.prim != 0xFFFFFFFFu || h.plane >= 0) B.ht[k] = h.t;
  if (h.prim != 0xFFFFFFFFu) {
    B.hb1[k] = h.b1;
    B.hb2[k] = h.b2;
  }
}

// extend: closest hit for the active queue of bounce b.
template <class T, int STACK>
__device__ __forceinline__ void extend_body(const DevScene& S, const PathBuffers<T>& B, const ChunkParams& P, int b,
                                            uint32_t lb, uint32_t* sref, uint32_t* pre) {
  const QueueView qv = load_queue(P.counts + (size_t)b * kSubQueues * kCountStride, pre, lb, ~0u, kCountStride);
  if (qv.past_end) return;
  const uint32_t i = lb * kBlock + threadIdx.x;
  uint32_t st[kStCount] = {0};
  if (i < qv.total) {
    const uint32_t k = active_slot(P, pre, b, i);
    const RayArrays<T>& R = B.ray[b & 1];
    const V3<T> o{R.ox[k], R.oy[k], R.oz[k]}, d{R.dx[k], R.dy[k], R.dz[k]};
    Counters c{0, 0};
    const auto h = ModeOps<T>::closest(S, P.use_bvh != 0, o, d, lds_stack(sref), c);
    store_hit(B, k, h);
    st[kStRays] = 1;
    st[kStVisits] = c.visits;
    st[kStTests] = c.tests;
  }
  flush_stats(P, st, lb);
}
template <class T, int STACK>
__global__ __launch_bounds__(kBlock) void k_extend(DevScene S, PathBuffers<T> B, ChunkParams P, int b) {
  __shared__ uint32_t sref[(STACK > 0 ? STACK : 1) * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  extend_body<T, STACK>(S, B, P, b, blockIdx.x, sref, pre);
}

// shade: one bounce of path_trace (integrator.cpp:103-137) after the hit.
template <class T>
__device__ __forceinline__ void shade_body(const DevScene& S, const PathBuffers<T>& B, const ChunkParams& P, int b,
                                           uint32_t lb, uint32_t* pre) {
  const QueueView qv = load_queue(P.counts + (size_t)b * kSubQueues * kCountStride, pre, lb, ~0u, kCountStride);
  if (qv.past_end) return;
  const uint32_t i = lb * kBlock + threadIdx.x;
  uint32_t st[kStCount] = {0};
  bool alive = false, shadow = false;
  uint32_t pid = 0, shadow_cls = 0;
  ShadowReq<T> req;
  V3<T> nro, ndir, nbeta;  // the next ray, throughput and draws of a path that lives on
  bool nspec = false;
  typename DrawSrc<T>::Next ndraw{};
  if (i < qv.total) {
    const uint32_t k = active_slot(P, pre, b, i);
    const RayArrays<T>& R = B.ray[b & 1];
    const V3<T> d{R.dx[k], R.dy[k], R.dz[k]};
    const uint32_t word = B.hprim[k];
    const uint32_t prim = (word & kHitPlane) ? 0xFFFFFFFFu : word;
    const int plane = word == kHitNone ? -1 : (word & kHitPlane) ? (int)(word & ~kHitPlane) : -1;
    pid = slot_pid(B, b, k);
    const uint32_t pixel = pid_pixel(P, pid), sample = pid_sample(P, pid);
    V3<T> beta{T(1), T(1), T(1)};
    if (b > 0) beta = V3<T>{R.bx[k], R.by[k], R.bz[k]};
    // The path's radiance is read and written only by a bounce that adds
    // to it (a miss, or emission counted here); most bounces add nothing.
    if (prim == 0xFFFFFFFFu && plane < 0) {
      V3<T> L{B.lx[pid], B.ly[pid], B.lz[pid]};
      L = L + beta * background<T>(S, d);
      B.lx[pid] = L.x; B.ly[pid] = L.y; B.lz[pid] = L.z;
    } else {  // the origin and the specular flag are read by hits only
      const V3<T> o{R.ox[k], R.oy[k], R.oz[k]};
      const bool spec = b == 0 || R.spec[k] != 0;
      auto rng = DrawSrc<T>::make(B, P, b, k, pixel, sample);
      HitRec<T> rec;
      ModeOps<T>::finalize(S, o, d, B.ht[k], B.hb1[k], B.hb2[k], prim, plane, rec);
      const Mat<T> mat = load_mat<T>(S, rec.material);
      const V3<T> Le = emitted(mat, rec.front_face);
      const bool sampled = !rec.is_plane && S.sampled[rec.prim] != 0;
      if (!is_zero(Le) && (spec || !P.light_sampling || !sampled)) {
        V3<T> L{B.lx[pid], B.ly[pid], B.lz[pid]};
        L = L + beta * Le;
        B.lx[pid] = L.x; B.ly[pid] = L.y; B.lz[pid] = L.z;
      }
      Scatter<T> sr;
      if (scatter(mat, d, rec, rng, sr)) {
        if (P.light_sampling && !sr.specular && S.n_lights > 0) {
          const uint32_t nl = S.n_lights;
          const uint32_t li = rng.below(nl);
          if (ModeOps<T>::light(S, li, rec, mat, rng, req)) {
            shadow = true;
            shadow_cls = li & 1u;
            req.f_rop = (beta * req.f_rop) * T(nl);
          }
        }
        beta = beta * sr.weight;
        bool live = !is_zero(beta);
        if (live && b + 1 >= P.rr_min_depth) {
          T q = max_comp(beta);
          q = q < T(0.05) ? T(0.05) : (q > T(0.95) ? T(0.95) : q);
          if (rng.next() >= q) {
            st[kStRR] = 1;
            live = false;
          } else {
            beta = beta * (T(1) / q);
          }
        }
        if (live && b + 1 < P.max_depth) {
          nro = offset_ray_origin(rec.p, rec.gn, sr.dir);
          ndir = sr.dir;
          nbeta = beta;
          nspec = sr.specular;
          ndraw = DrawSrc<T>::next(rng);
          alive = true;
        }
      }
    }
  }
  // Converged: compacted appends to the next active queue and the shadow queue.
  const AppendSlots a = block_append2(P.counts + (size_t)(b + 1) * kSubQueues * kCountStride,
                                      P.counts + (size_t)(P.shadow_counts + b * kSubQueues) * kCountStride, P.sub_cap,
                                      alive, shadow,
                                      shadow_cls, lb);
  if (alive) {
    const uint32_t k = P.q_base + a.a;
    const RayArrays<T>& R = B.ray[(b + 1) & 1];
    B.q[(b + 1) & 1][k] = pid;
    R.ox[k] = nro.x; R.oy[k] = nro.y; R.oz[k] = nro.z;
    R.dx[k] = ndir.x; R.dy[k] = ndir.y; R.dz[k] = ndir.z;
    R.bx[k] = nbeta.x; R.by[k] = nbeta.y; R.bz[k] = nbeta.z;
    R.spec[k] = nspec ? 1 : 0;
    DrawSrc<T>::store(B, b + 1, k, ndraw);
  }
  if (shadow) {
    const int64_t si = P.q_base + a.b + (int64_t)(b & 1) * B.sq_stride;
    B.sq_pid[si] = pid;
    B.sox[si] = req.o.x; B.soy[si] = req.o.y; B.soz[si] = req.o.z;
    B.sdx[si] = req.d.x; B.sdy[si] = req.d.y; B.sdz[si] = req.d.z;
    B.stm[si] = req.tmax;
    B.scx[si] = req.f_rop.x; B.scy[si] = req.f_rop.y; B.scz[si] = req.f_rop.z;
    st[kStShadow] = 1;
  }
  flush_stats(P, st, lb);
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_shade(DevScene S, PathBuffers<T> B, ChunkParams P, int b) {
  __shared__ uint32_t pre[kSubQueues + 1];
  shade_body<T>(S, B, P, b, blockIdx.x, pre);
}

// shadow: any-hit of the shadow queue of bounce b (Scene::occluded).
template <class T, int STACK>
__device__ __forceinline__ void shadow_body(const DevScene& S, const PathBuffers<T>& B, const ChunkParams& P, int b,
                                            uint32_t lb, uint32_t* sref, uint32_t* pre) {
  const QueueView qv =
      load_queue(P.counts + (size_t)(P.shadow_counts + b * kSubQueues) * kCountStride, pre, lb, ~0u, kCountStride);
  if (qv.past_end) return;
  const uint32_t i = lb * kBlock + threadIdx.x;
  uint32_t st[kStCount] = {0};
  if (i < qv.total) {
    const uint32_t k = P.q_base + queue_slot(pre, P.sub_cap, i) + (uint32_t)(b & 1) * B.sq_stride;
    const uint32_t pid = B.sq_pid[k];
    const V3<T> o{B.sox[k], B.soy[k], B.soz[k]}, d{B.sdx[k], B.sdy[k], B.sdz[k]};
    Counters c{0, 0};
    const bool occ = ModeOps<T>::occluded(S, P.use_bvh != 0, o, d, B.stm[k], lds_stack(sref), c);
    if (!occ) {
      if constexpr (sizeof(T) == 4) {
        B.slx[pid] = B.slx[pid] + B.scx[k];
        B.sly[pid] = B.sly[pid] + B.scy[k];
        B.slz[pid] = B.slz[pid] + B.scz[k];
      } else {
        B.lx[pid] = B.lx[pid] + B.scx[k];
        B.ly[pid] = B.ly[pid] + B.scy[k];
        B.lz[pid] = B.lz[pid] + B.scz[k];
      }
    }
    st[kStSVisits] = c.visits;
    st[kStSTests] = c.tests;
  }
  flush_stats(P, st, lb);
}
template <class T, int STACK>
__global__ __launch_bounds__(kBlock) void k_shadow(DevScene S, PathBuffers<T> B, ChunkParams P, int b) {
  __shared__ uint32_t sref[(STACK > 0 ? STACK : 1) * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  shadow_body<T, STACK>(S, B, P, b, blockIdx.x, sref, pre);
}

// Persistent FP32 walks of a bounce's closest-hit and any-hit queues. The
// grid holds at most what the device runs at once; each workgroup takes a
// contiguous range of the queue (whole workgroups of rays), and between
// rounds of its walk (rtg_trace.hpp) a wave hands its idle lanes the range's
// next rays — after Aila & Laine 2009, replacing terminated rays — so lanes
// whose rays end early keep working while the longest walks of their wave
// finish. A ray's walk and counters depend on the ray alone: results are
// bitwise those of one ray per thread (k_extend / k_shadow).
constexpr uint32_t kRefill = 32;  // a wave takes new rays once half its lanes are idle
// A workgroup's contiguous range of the queue (whole workgroups of rays).
// (Interleaving the workgroups' 128-ray chunks instead, so that they all run
// out together, measured the same.)
struct Range {
  uint32_t r0, r1;
};
__device__ __forceinline__ Range block_range(uint32_t total) {
  const uint32_t per = (uint32_t)(((uint64_t)total + (uint64_t)gridDim.x * kBlock - 1) / ((uint64_t)gridDim.x * kBlock)) * kBlock;
  const uint64_t r0 = (uint64_t)blockIdx.x * per;
  const uint32_t a = r0 < total ? (uint32_t)r0 : total;
  return Range{a, total - a < per ? total : a + per};
}
// Next ray indices for the idle lanes of a converged wave: true while the
// workgroup's range lasts; `i` the lane's queue index (>= r1 when none).
__device__ __forceinline__ bool refill(uint32_t* cursor, const Range& rg, unsigned long long idle, bool& more,
                                       uint32_t& i) {
  const uint32_t n_idle = (uint32_t)__popcll(idle);
  if (!more || (n_idle < kRefill && n_idle != 64u)) return false;
  uint32_t base = 0;
  if ((threadIdx.x & 63u) == 0) base = atomicAdd(cursor, n_idle);
  base = __shfl(base, 0, 64);
  if (base >= rg.r1) {
    more = false;
    return false;
  }
  i = base + lanes_below(idle);
  return true;
}

// Closest hits of queue indices [rg.r0, rg.r1) of bounce b (`cursor`, in
// LDS, starts at rg.r0), idle lanes refilled between rounds; adds rays,
// node visits and primitive tests to st.
template <int STACK>
__device__ __forceinline__ void extend_range(const DevScene& S, const PathBuffers<float>& B, const ChunkParams& P,
                                             int b, uint32_t* sref, const uint32_t* pre, uint32_t* cursor,
                                             const Range& rg, uint32_t* st) {
  const LdsStack stk = lds_stack(sref);
  Counters c{0, 0};
  Trav32 T;
  T.w = Walk{kNoRef, kNoRef, 0, kNoRef};
  Ray32 r{};
  uint32_t k = 0;  // the ray's queue slot
  bool active = false, more = true;
  auto finish = [&]() { store_hit(B, k, T.h); };
  for (;;) {
    uint32_t i = 0;
    if (refill(cursor, rg, __ballot(!active), more, i) && !active && i < rg.r1) {
      k = active_slot(P, pre, b, i);
      const RayArrays<float>& R = B.ray[b & 1];
      r = make_ray32(S, R.ox[k], R.oy[k], R.oz[k], R.dx[k], R.dy[k], R.dz[k]);
      ++st[kStRays];
      if (P.use_bvh) {
        active = closest_begin32(S, r, 1e-6f, __builtin_inff(), stk, c, T);
      } else {  // brute force (Scene::intersect_brute): one pass over the primitives
        T.h = closest32(S, false, r, 1e-6f, __builtin_inff(), stk, c);
        active = false;
      }
      if (!active) finish();
    }
    if (!__ballot(active)) {
      if (!more) break;
      continue;
    }
    if (active && closest_round32(S, r, 1e-6f, stk, c, T)) {
      finish();
      active = false;
    }
  }
  st[kStVisits] += c.visits;
  st[kStTests] += c.tests;
}

// Any hits of shadow-queue indices [rg.r0, rg.r1) of bounce b, as above;
// an unoccluded request adds its contribution to its path's shadow radiance.
template <int STACK>
__device__ __forceinline__ void shadow_range(const DevScene& S, const PathBuffers<float>& B, const ChunkParams& P,
                                             int b, uint32_t* sref, const uint32_t* pre, uint32_t* cursor,
                                             const Range& rg, uint32_t* st) {
  const LdsStack stk = lds_stack(sref);
  Counters c{0, 0};
  Walk w{kNoRef, kNoRef, 0, kNoRef};
  Ray32 r{};
  uint32_t k = 0;
  float tmax = 0.0f;
  bool active = false, more = true, hit = false;
  auto finish = [&]() {
    if (!hit) {
      const uint32_t pid = B.sq_pid[k];
      B.slx[pid] = B.slx[pid] + B.scx[k];
      B.sly[pid] = B.sly[pid] + B.scy[k];
      B.slz[pid] = B.slz[pid] + B.scz[k];
    }
  };
  for (;;) {
    uint32_t i = 0;
    if (refill(cursor, rg, __ballot(!active), more, i) && !active && i < rg.r1) {
      k = P.q_base + queue_slot(pre, P.sub_cap, i) + (uint32_t)(b & 1) * B.sq_stride;
      r = make_ray32(S, B.sox[k], B.soy[k], B.soz[k], B.sdx[k], B.sdy[k], B.sdz[k]);
      tmax = B.stm[k];
      if (P.use_bvh) {
        active = any_begin32(S, r, 1e-6f, tmax, stk, c, w, hit);
      } else {
        hit = any32(S, false, r, 1e-6f, tmax, stk, c);
        active = false;
      }
      if (!active) finish();
    }
    if (!__ballot(active)) {
      if (!more) break;
      continue;
    }
    if (active && any_round32(S, r, 1e-6f, tmax, stk, c, w, hit)) {
      finish();
      active = false;
    }
  }
  st[kStSVisits] += c.visits;
  st[kStSTests] += c.tests;
}

// Held to 64 VGPRs (8 waves per SIMD; 70 unconstrained, 7 waves): nine
// spill to scratch, outside the node and triangle loops, and the kernel
// alone runs 4 % faster on the torus workload (19.2 vs 20.0 ms per render).
template <int STACK>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_extend_dyn(DevScene S,
                                                                                                  PathBuffers<float> B,
                                                                                                  ChunkParams P, int b) {
  __shared__ uint32_t sref[STACK * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  __shared__ uint32_t cursor;
  const QueueView qv = load_queue(P.counts + (size_t)b * kSubQueues * kCountStride, pre, blockIdx.x, ~0u, kCountStride);
  if (qv.past_end) return;
  const Range rg = block_range(qv.total);
  if (threadIdx.x == 0) cursor = rg.r0;
  __syncthreads();
  uint32_t st[kStCount] = {0};
  extend_range<STACK>(S, B, P, b, sref, pre, &cursor, rg, st);
  flush_stats(P, st, blockIdx.x);
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_shadow_dyn(DevScene S, PathBuffers<float> B, ChunkParams P, int b) {
  __shared__ uint32_t sref[STACK * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  __shared__ uint32_t cursor;
  const QueueView qv = load_queue(P.counts + (size_t)(P.shadow_counts + b * kSubQueues) * kCountStride, pre,
                                  blockIdx.x, ~0u, kCountStride);
  if (qv.past_end) return;
  const Range rg = block_range(qv.total);
  if (threadIdx.x == 0) cursor = rg.r0;
  __syncthreads();
  uint32_t st[kStCount] = {0};
  shadow_range<STACK>(S, B, P, b, sref, pre, &cursor, rg, st);
  flush_stats(P, st, blockIdx.x);
}

// accumulate: the invalid-sample guard and Framebuffer::add in sample order
// (renderer.cpp:39-45, framebuffer.hpp:31-35).
//
// With pixel-major path ids a pixel's k samples are contiguous, so one
// thread per pixel would read with a stride of k values. Instead a block
// takes a run of whole pixels (at most kAccTile<T> samples), loads it into
// LDS with coalesced reads (one pad slot per pixel keeps the per-pixel reads
// free of bank conflicts) and each thread adds one pixel's samples in
// sample order from LDS. Passes wider than the tile, and sample-major ids,
// read global memory directly.
template <class T>
constexpr uint32_t kAccTile = 4096 / sizeof(T);
template <class T>
constexpr uint32_t kAccLds = 3 * 2 * kAccTile<T>;  // values of LDS per block (x, y, z with pads)

template <class T>
__host__ __device__ inline uint32_t acc_pixels_per_block(const ChunkParams& P) {
  return (P.sample_major || P.k > kAccTile<T>) ? (uint32_t)kBlock : kAccTile<T> / P.k;
}
template <class T>
__host__ __device__ inline uint32_t accumulate_blocks(const ChunkParams& P) {
  const uint32_t ppb = acc_pixels_per_block<T>(P);
  return (P.n_pix + ppb - 1) / ppb;
}

template <class T>
__device__ __forceinline__ void add_pixel(const ChunkParams& P, uint32_t pixel, const T* x, const T* y, const T* z,
                                          const T* xs, const T* ys, const T* zs, uint32_t stride, uint32_t* st) {
  double sx = P.acc_sum[3 * (size_t)pixel], sy = P.acc_sum[3 * (size_t)pixel + 1],
         sz = P.acc_sum[3 * (size_t)pixel + 2];
  for (uint32_t s = 0; s < P.k; ++s) {
    T a = x[s * stride], b = y[s * stride], c = z[s * stride];
    if (xs) {  // FP32 path integrator: plus the shadow-ray radiance
      a = a + xs[s * stride];
      b = b + ys[s * stride];
      c = c + zs[s * stride];
    }
    if (!(r_isfinite(a) && r_isfinite(b) && r_isfinite(c)) || a < 0 || b < 0 || c < 0) {
      ++st[kStInvalid];
      a = b = c = 0;
    }
    sx += (double)a;
    sy += (double)b;
    sz += (double)c;
  }
  P.acc_sum[3 * (size_t)pixel] = sx;
  P.acc_sum[3 * (size_t)pixel + 1] = sy;
  P.acc_sum[3 * (size_t)pixel + 2] = sz;
  P.acc_count[pixel] += P.k;
  st[kStCamera] += P.k;
}

template <class T>
__device__ __forceinline__ void accumulate_body(const PathBuffers<T>& B, const ChunkParams& P, uint32_t lb, T* tile) {
  uint32_t st[kStCount] = {0};
  const bool ls = sizeof(T) == 4 && P.integrator == RTG_INTEGRATOR_PATH;  // shadow radiance kept apart
  if (P.sample_major || P.k > kAccTile<T>) {
    const uint32_t j = lb * kBlock + threadIdx.x;
    if (j < P.n_pix) {
      const uint32_t p0 = pid_of(P, j, 0), stride = P.sample_major ? P.n_pix : 1u;
      add_pixel<T>(P, P.p0 + j, B.lx + p0, B.ly + p0, B.lz + p0, ls ? B.slx + p0 : nullptr, ls ? B.sly + p0 : nullptr,
                   ls ? B.slz + p0 : nullptr, stride, st);
    }
  } else {
    const uint32_t ppb = kAccTile<T> / P.k, j0 = lb * ppb;
    const uint32_t np = P.n_pix - j0 < ppb ? P.n_pix - j0 : ppb;
    const uint32_t n = np * P.k, base = P.pid_base + j0 * P.k;
    T* tx = tile;
    T* ty = tile + 2 * kAccTile<T>;
    T* tz = tile + 4 * kAccTile<T>;
    for (uint32_t e = threadIdx.x; e < n; e += kBlock) {
      const uint32_t at = e + e / P.k;  // pixel e / k at (k + 1) * (e / k)
      tx[at] = ls ? B.lx[base + e] + B.slx[base + e] : B.lx[base + e];
      ty[at] = ls ? B.ly[base + e] + B.sly[base + e] : B.ly[base + e];
      tz[at] = ls ? B.lz[base + e] + B.slz[base + e] : B.lz[base + e];
    }
    __syncthreads();
    for (uint32_t jj = threadIdx.x; jj < np; jj += kBlock) {
      const uint32_t at = jj * (P.k + 1);
      add_pixel<T>(P, P.p0 + j0 + jj, tx + at, ty + at, tz + at, nullptr, nullptr, nullptr, 1u, st);
    }
    __syncthreads();  // the tile is reused by the block's next job
  }
  flush_stats(P, st, lb);
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_accumulate(PathBuffers<T> B, ChunkParams P) {
  __shared__ T tile[kAccLds<T>];
  accumulate_body<T>(B, P, blockIdx.x, tile);
}

// ---------------------------------------------------------------------------
// direct / normals integrators: one sample per thread (integrator.cpp:141-198);
// FP64 direct and normals in either precision (FP32 direct: rtg_direct.hpp).
//
// direct_trace recurses: a metal hit adds albedo x (its reflection's
// radiance), a dielectric hit tint x F x (reflection) + tint x (1 - F) x
// (refraction), nested as the reference nests them, so FP64 sums round as
// the CPU's. The recursion is a depth-first walk over frames. The frame being
// worked on lives in registers; the frames of the enclosing calls wait in
// HBM, one field per row of `frame_stride` lanes (the paths of a chunk side
// by side, so a wave's pushes and pops coalesce), max_depth frames per path:
// any depth the scene format allows (1000) is followed.
constexpr int kFrameWords = 16;  // L, w1, w2, o2, d2 (3 each) and phase | second << 2
enum : int { kFrL = 0, kFrW1 = 3, kFrW2 = 6, kFrO2 = 9, kFrD2 = 12, kFrFlags = 15 };

template <class T>
struct FrameStack {
  T* base;
  uint32_t stride, lane;
  __device__ T& at(int frame, int field) const {
    return base[((size_t)frame * kFrameWords + (size_t)field) * stride + lane];
  }
  __device__ void put(int frame, int field, const V3<T>& v) const {
    at(frame, field) = v.x;
    at(frame, field + 1) = v.y;
    at(frame, field + 2) = v.z;
  }
  __device__ V3<T> get(int frame, int field) const {
    return V3<T>{at(frame, field), at(frame, field + 1), at(frame, field + 2)};
  }
};

template <class T, int STACK>
__device__ V3<T> direct_sample(const DevScene& S, const ChunkParams& P, uint32_t lane, V3<T> o, V3<T> d,
                               typename ModeOps<T>::R& rng, LdsStack stk, uint32_t* st) {
  const FrameStack<T> fs{static_cast<T*>(P.frames), P.frame_stride, lane};
  // Two modes. Tracing: (a, b) is the ray of the frame at depth `top`.
  // Returning: a is the radiance of a finished call, handed to the frame
  // at `top` (its caller; none when top < 0). Only these few values are
  // carried from step to step; a frame's weights and pending second ray wait
  // in HBM.
  bool tracing = true;
  int top = 0;
  V3<T> a = o, b = d;
  while (true) {
    if (!tracing) {
      if (top < 0) return a;
      const int flags = (int)fs.at(top, kFrFlags);
      V3<T> L = fs.get(top, kFrL);
      if ((flags & 3) == 1) {
        L = L + fs.get(top, kFrW1) * a;
        if (flags & 4) {  // then the refracted ray
          fs.put(top, kFrL, L);
          fs.at(top, kFrFlags) = T(2);
          a = fs.get(top, kFrO2);
          b = fs.get(top, kFrD2);
          ++top;
          tracing = true;
          continue;
        }
      } else {
        L = L + fs.get(top, kFrW2) * a;
      }
      a = L;
      --top;
      continue;
    }
    const V3<T> ro = a, rd = b;
    Counters c{0, 0};
    ++st[kStRays];
    const auto h = ModeOps<T>::closest(S, P.use_bvh != 0, ro, rd, stk, c);
    st[kStVisits] += c.visits;
    st[kStTests] += c.tests;
    tracing = false;  // unless a child ray follows
    if (h.prim == 0xFFFFFFFFu && h.plane < 0) {
      // Normals: a miss is black (integrator.cpp:190); direct: the sky.
      a = P.integrator == RTG_INTEGRATOR_NORMALS ? V3<T>{0, 0, 0} : background<T>(S, rd);
      --top;
      continue;
    }
    HitRec<T> rec;
    ModeOps<T>::finalize(S, ro, rd, h.t, h.b1, h.b2, h.prim, h.plane, rec);
    const Mat<T> mat = load_mat<T>(S, rec.material);
    V3<T> L = emitted(mat, rec.front_face);
    if (P.integrator == RTG_INTEGRATOR_NORMALS) {
      const V3<T> nn = rec.front_face ? rec.normal : -rec.normal;
      a = (nn + V3<T>{1, 1, 1}) * T(0.5);
      --top;
      continue;
    }
    if (top < P.max_depth) {
      if (mat.kind == RTG_MAT_LAMBERTIAN) {
        for (uint32_t li = 0; li < S.n_lights; ++li) {
          ShadowReq<T> req;
          if (ModeOps<T>::light(S, li, rec, mat, rng, req)) {
            ++st[kStShadow];
            Counters c2{0, 0};
            if (!ModeOps<T>::occluded(S, P.use_bvh != 0, req.o, req.d, req.tmax, stk, c2)) L = L + req.f_rop;
            st[kStSVisits] += c2.visits;
            st[kStSTests] += c2.tests;
          }
        }
      } else if (mat.kind == RTG_MAT_METAL) {
        Scatter<T> sr;
        if (scatter(mat, rd, rec, rng, sr)) {
          fs.put(top, kFrL, L);
          fs.put(top, kFrW1, mat.c0);
          fs.at(top, kFrFlags) = T(1);
          a = offset_ray_origin(rec.p, rec.gn, sr.dir);
          b = sr.dir;
          ++top;
          tracing = true;
          continue;
        }
      } else if (mat.kind == RTG_MAT_DIELECTRIC) {
        const V3<T> in = normalize(rd);
        const T eta = rec.front_face ? T(1) / mat.p0 : mat.p0;
        const T cos_i = r_fmin(dot(-in, rec.normal), T(1));
        V3<T> refracted;
        const bool can = refract(in, rec.normal, eta, &refracted);
        const T F = can ? schlick(cos_i, eta) : T(1);
        const V3<T> rdir = reflect(in, rec.normal);
        fs.put(top, kFrL, L);
        fs.put(top, kFrW1, mat.c0 * F);
        fs.at(top, kFrFlags) = T(can ? 5 : 1);
        if (can) {
          const V3<T> tdir = normalize(refracted);
          fs.put(top, kFrW2, mat.c0 * (T(1) - F));
          fs.put(top, kFrO2, offset_ray_origin(rec.p, rec.gn, tdir));
          fs.put(top, kFrD2, tdir);
        }
        a = offset_ray_origin(rec.p, rec.gn, rdir);
        b = rdir;
        ++top;
        tracing = true;
        continue;
      }
    }
    a = L;
    --top;
  }
}

template <class T, int STACK>
__device__ __forceinline__ void direct_body(const DevScene& S, const PathBuffers<T>& B, const ChunkParams& P,
                                            uint32_t* sref) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t pid = P.pid_base + l;
  uint32_t st[kStCount] = {0};
  if (l < P.n_paths) {
    const uint32_t pixel = pid_pixel(P, pid), sample = pid_sample(P, pid);
    const int x = (int)(pixel % P.width), y = (int)(pixel / P.width);
    V3<T> o, d, L;
    if constexpr (sizeof(T) == 8) {
      PcgDraws rng;
      rng.pcg = Pcg32::for_sample(P.seed, pixel, sample);
      const double jx = rng.next();
      const double jy = rng.next();
      ModeOps<T>::camera(S, x, y, jx, jy, rng, o, d);
      L = direct_sample<T, STACK>(S, P, l, o, d, rng, lds_stack(sref), st);
    } else {
      HashDraws rng(path_key(P.seed, pixel, sample), 0xFFFFFFFFu);
      const float jx = rng.next();
      const float jy = rng.next();
      ModeOps<T>::camera(S, x, y, jx, jy, rng, o, d);
      rng.bounce = 0;
      rng.idx = 0;
      L = direct_sample<T, STACK>(S, P, l, o, d, rng, lds_stack(sref), st);
    }
    B.lx[pid] = L.x;
    B.ly[pid] = L.y;
    B.lz[pid] = L.z;
  }
  flush_stats(P, st, blockIdx.x);
}
// FP32 at 4 waves per SIMD (128 VGPRs, a 60-byte spill; 134 VGPRs and 3
// waves otherwise); FP64 as the compiler allocates it (217 VGPRs, 2 waves).
template <int STACK>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(STACK <= 31 ? 4 : 1))) void k_direct32(DevScene S,
                                                                                        PathBuffers<float> B,
                                                                                        ChunkParams P) {
  __shared__ uint32_t sref[(STACK > 0 ? STACK : 1) * kBlock];
  direct_body<float, STACK>(S, B, P, sref);
}
template <int STACK>
__global__ __launch_bounds__(kBlock) void k_direct64(DevScene S, PathBuffers<double> B, ChunkParams P) {
  __shared__ uint32_t sref[(STACK > 0 ? STACK : 1) * kBlock];
  direct_body<double, STACK>(S, B, P, sref);
}

// ---------------------------------------------------------------------------
// Batch ray queries (rtg_trace): rays {o, d, tmin, tmax} as doubles.
template <class T, int STACK>
__global__ __launch_bounds__(kBlock) void k_trace(DevScene S, const double* __restrict__ rays, uint64_t n, int mode,
                                                  int use_bvh, rtg_hit* __restrict__ hits,
                                                  unsigned long long* stats) {
  __shared__ uint32_t sref[(STACK > 0 ? STACK : 1) * kBlock];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t st[kStCount] = {0};
  if (i < n) {
    const double* r = rays + 8 * i;
    rtg_hit out;
    out.t = __builtin_inf();
    out.prim = 0xFFFFFFFFu;
    out.flags = 0;
    Counters c{0, 0};
    const LdsStack stk = lds_stack(sref);
    if constexpr (sizeof(T) == 8) {
      const V3<double> o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
      if (mode == RTG_TRACE_CLOSEST) {
        const Hit64 h = closest64(S, use_bvh != 0, o, d, r[6], r[7], stk, c);
        if (h.prim != 0xFFFFFFFFu || h.plane >= 0) {
          HitRec<double> rec;
          finalize64(S, o, d, h.t, h.b1, h.b2, h.prim, h.plane, rec);
          out.t = h.t;
          out.prim = h.prim != 0xFFFFFFFFu ? h.prim : (uint32_t)h.plane;
          out.flags = 1u | (h.prim == 0xFFFFFFFFu ? 2u : 0u) | (rec.front_face ? 4u : 0u);
        }
      } else {
        out.flags = any64(S, use_bvh != 0, o, d, 1e-6, r[7], stk, c) ? 1u : 0u;
      }
    } else {
      const Ray32 ray = make_ray32(S, (float)r[0], (float)r[1], (float)r[2], (float)r[3], (float)r[4], (float)r[5]);
      if (mode == RTG_TRACE_CLOSEST) {
        const Hit32 h = closest32(S, use_bvh != 0, ray, (float)r[6], (float)r[7], stk, c);
        if (h.prim != 0xFFFFFFFFu || h.plane >= 0) {
          HitRec<float> rec;
          finalize32(S, V3<float>{ray.ox, ray.oy, ray.oz}, V3<float>{ray.dx, ray.dy, ray.dz}, h.t, h.b1, h.b2,
                     h.prim, h.plane, rec);
          out.t = h.t;
          out.prim = h.prim != 0xFFFFFFFFu ? h.prim : (uint32_t)h.plane;
          out.flags = 1u | (h.prim == 0xFFFFFFFFu ? 2u : 0u) | (rec.front_face ? 4u : 0u);
        }
      } else {
        out.flags = any32(S, use_bvh != 0, ray, 1e-6f, (float)r[7], stk, c) ? 1u : 0u;
      }
    }
    hits[i] = out;
    st[kStRays] = 1;
    st[kStVisits] = c.visits;
    st[kStTests] = c.tests;
  }
  flush_stats(stats, st, kStCount, kStatSlots, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Batches: many jobs (different scenes, resolutions, sample ranges, depths)
// share one set of path arrays and one launch per step. A plan holds, per
// job, the first virtual block of the job (prefix sums, n_jobs + 1 entries).
// Workgroups are persistent and walk the virtual blocks round-robin (the
// traversal kernels walk contiguous shares job by job: k_wide_*_dyn), so
// each virtual block belongs to one job: the job's scene and settings are
// uniform across the workgroup and sit in scalar registers.
struct BatchJob {
  DevScene S;
  ChunkParams P;
};
constexpr int kMaxBatchJobs = 256;
constexpr int kBatchDepth = 64;                                 // max_depth limit of batched jobs
constexpr uint32_t kBatchShadowCounts = (kBatchDepth + 1) * kSubQueues;
constexpr uint32_t kBatchCounts = 2 * kBatchShadowCounts;       // counters per job
constexpr uint32_t kBatchStatSlots = 32;

struct BatchPlans {
  const uint32_t* n_jobs;  // the batch's job count, written by the host
  const uint32_t* gen;     // written by the host
  const uint32_t* acc;     // written by the host
  uint32_t* act;           // active queue of the next extend/shade (k_batch_plan)
  uint32_t* shd;           // shadow queues of the shadow steps (k_batch_plan), one plan per bounce parity
};
// The job count is read from device memory, so a captured sequence of batch
// launches (a graph) replays for any number of jobs.
__device__ __forceinline__ uint32_t jobs_of(const BatchPlans& L) { return __builtin_amdgcn_readfirstlane(*L.n_jobs); }

__device__ __forceinline__ uint32_t blocks_of(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// Job of virtual block vb: the last j with plan[j] <= vb (jobs without
// blocks repeat their successor's start and are skipped). vb is uniform
// across the workgroup, so the search runs on scalar loads of the plan
// (read-only during the launch; it stays in the scalar cache) and the
// workgroup keeps its LDS for traversal stacks.
__device__ __forceinline__ uint32_t job_of(const uint32_t* __restrict__ plan, uint32_t n_jobs, uint32_t vb) {
  uint32_t lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (plan[mid] <= vb) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// Runs f(job, local block) over every virtual block of a plan.
template <class F>
__device__ __forceinline__ void batch_loop(const uint32_t* __restrict__ plan, uint32_t n_jobs, F&& f) {
  const uint32_t total = __builtin_amdgcn_readfirstlane(plan[n_jobs]);
  for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const uint32_t j = job_of(plan, n_jobs, vb);
    f(j, vb - __builtin_amdgcn_readfirstlane(plan[j]));
  }
}

// Plan of the active queue of bounce b + 1 and of the shadow queue of
// bounce b (b = -1: the camera rays of bounce 0 only). One workgroup.
template <class T>
__global__ __launch_bounds__(kMaxBatchJobs) void k_batch_plan(const BatchJob* __restrict__ jobs, BatchPlans L, int b) {
  __shared__ uint32_t sa[kMaxBatchJobs], ss[kMaxBatchJobs];
  const uint32_t j = threadIdx.x;
  const uint32_t n_jobs = jobs_of(L);
  uint32_t na = 0, ns = 0;
  if (j < n_jobs) {
    const ChunkParams& P = jobs[j].P;
    if (b + 1 < P.max_depth)
      for (int s = 0; s < kSubQueues; ++s) na += P.counts[((size_t)(b + 1) * kSubQueues + s) * kCountStride];
    if (b >= 0 && b < P.max_depth)
      for (int s = 0; s < kSubQueues; ++s) ns += P.counts[((size_t)P.shadow_counts + b * kSubQueues + s) * kCountStride];
  }
  sa[j] = blocks_of(na);
  ss[j] = blocks_of(ns);
  __syncthreads();
  for (uint32_t off = 1; off < kMaxBatchJobs; off <<= 1) {  // inclusive scan
    const uint32_t a = j >= off ? sa[j - off] : 0u, c = j >= off ? ss[j - off] : 0u;
    __syncthreads();
    sa[j] += a;
    ss[j] += c;
    __syncthreads();
  }
  uint32_t* shd = L.shd + (b & 1) * (kMaxBatchJobs + 1);
  if (j < n_jobs) {
    L.act[j + 1] = sa[j];
    shd[j + 1] = ss[j];
  }
  if (j == 0) {
    L.act[0] = 0;
    shd[0] = 0;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_generate(const BatchJob* __restrict__ jobs, PathBuffers<T> B,
                                                           BatchPlans L) {
  batch_loop(L.gen, jobs_of(L), [&](uint32_t j, uint32_t lb) { generate_body<T>(jobs[j].S, B, jobs[j].P, lb); });
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_shade(const BatchJob* __restrict__ jobs, PathBuffers<T> B,
                                                        BatchPlans L, int b) {
  __shared__ uint32_t pre[kSubQueues + 1];
  batch_loop(L.act, jobs_of(L), [&](uint32_t j, uint32_t lb) { shade_body<T>(jobs[j].S, B, jobs[j].P, b, lb, pre); });
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_accumulate(const BatchJob* __restrict__ jobs, PathBuffers<T> B,
                                                             BatchPlans L) {
  __shared__ T tile[kAccLds<T>];
  batch_loop(L.acc, jobs_of(L), [&](uint32_t j, uint32_t lb) { accumulate_body<T>(B, jobs[j].P, lb, tile); });
}

// Wide batches: one virtual block per workgroup over a grid that covers
// every path of the batch (workgroups past the plan's end exit at once), for
// batches large enough that launching such grids costs nothing next to the
// work. Without the persistent loop a workgroup holds no state across
// blocks, so these kernels keep the registers and occupancy of the
// single-render kernels.
template <class F>
__device__ __forceinline__ void batch_block(const uint32_t* __restrict__ plan, uint32_t n_jobs, F&& f) {
  const uint32_t vb = blockIdx.x;
  if (vb >= __builtin_amdgcn_readfirstlane(plan[n_jobs])) return;
  const uint32_t j = job_of(plan, n_jobs, vb);
  f(j, vb - __builtin_amdgcn_readfirstlane(plan[j]));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_wide_generate(const BatchJob* __restrict__ jobs, PathBuffers<T> B,
                                                          BatchPlans L) {
  batch_block(L.gen, jobs_of(L), [&](uint32_t j, uint32_t lb) { generate_body<T>(jobs[j].S, B, jobs[j].P, lb); });
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_wide_shade(const BatchJob* __restrict__ jobs, PathBuffers<T> B,
                                                       BatchPlans L, int b) {
  __shared__ uint32_t pre[kSubQueues + 1];
  batch_block(L.act, jobs_of(L), [&](uint32_t j, uint32_t lb) { shade_body<T>(jobs[j].S, B, jobs[j].P, b, lb, pre); });
}

// Wide batch passes with the persistent traversal of a single render: a
// workgroup takes a contiguous share of the plan's virtual blocks and walks
// it job by job (a job's scene and settings stay uniform across the
// workgroup); within a job its idle lanes are refilled between rounds, and
// at the job's end the workgroup drains before taking the next job.
template <class F>
__device__ __forceinline__ void wide_segments(const uint32_t* __restrict__ plan, uint32_t n_jobs, F&& f) {
  const uint32_t total = __builtin_amdgcn_readfirstlane(plan[n_jobs]);
  const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
  uint32_t vb = blockIdx.x * per;
  const uint32_t vend = vb + per < total ? vb + per : total;
  while (vb < vend) {
    const uint32_t j = job_of(plan, n_jobs, vb);
    const uint32_t jb0 = __builtin_amdgcn_readfirstlane(plan[j]), jb1 = __builtin_amdgcn_readfirstlane(plan[j + 1]);
    const uint32_t seg = jb1 < vend ? jb1 : vend;
    f(j, vb - jb0, seg - jb0);  // the job's virtual blocks [a, b)
    vb = seg;
  }
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_wide_extend_dyn(const BatchJob* __restrict__ jobs, PathBuffers<float> B,
                                                            BatchPlans L, int b) {
  __shared__ uint32_t sref[STACK * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  __shared__ uint32_t cursor;
  wide_segments(L.act, jobs_of(L), [&](uint32_t j, uint32_t a, uint32_t e) {
    const DevScene& S = jobs[j].S;
    const ChunkParams& P = jobs[j].P;
    const QueueView qv = load_queue(P.counts + (size_t)b * kSubQueues * kCountStride, pre, a, ~0u, kCountStride);
    if (qv.past_end) return;
    const Range rg{a * kBlock, e * kBlock < qv.total ? e * kBlock : qv.total};
    if (threadIdx.x == 0) cursor = rg.r0;
    __syncthreads();
    uint32_t st[kStCount] = {0};
    extend_range<STACK>(S, B, P, b, sref, pre, &cursor, rg, st);
    flush_stats(P, st, a);  // ends with a barrier: pre and cursor are free for the next job
  });
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_wide_shadow_dyn(const BatchJob* __restrict__ jobs, PathBuffers<float> B,
                                                            BatchPlans L, int b) {
  __shared__ uint32_t sref[STACK * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  __shared__ uint32_t cursor;
  wide_segments(L.shd + (b & 1) * (kMaxBatchJobs + 1), jobs_of(L), [&](uint32_t j, uint32_t a, uint32_t e) {
    const DevScene& S = jobs[j].S;
    const ChunkParams& P = jobs[j].P;
    const QueueView qv = load_queue(P.counts + (size_t)(P.shadow_counts + b * kSubQueues) * kCountStride, pre, a,
                                    ~0u, kCountStride);
    if (qv.past_end) return;
    const Range rg{a * kBlock, e * kBlock < qv.total ? e * kBlock : qv.total};
    if (threadIdx.x == 0) cursor = rg.r0;
    __syncthreads();
    uint32_t st[kStCount] = {0};
    shadow_range<STACK>(S, B, P, b, sref, pre, &cursor, rg, st);
    flush_stats(P, st, a);
  });
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_wide_accumulate(const BatchJob* __restrict__ jobs, PathBuffers<T> B,
                                                            BatchPlans L) {
  __shared__ T tile[kAccLds<T>];
  batch_block(L.acc, jobs_of(L), [&](uint32_t j, uint32_t lb) { accumulate_body<T>(B, jobs[j].P, lb, tile); });
}

}  // namespace rtg
