// rtg_api.hip — the C ABI of include/rtg.h: device contexts, scene upload
// (flattening the host-built BVH into the FP32 two-child layout), device
// accumulation buffers, the render-pass driver and batch ray queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "rtg_launch.hpp"

using namespace rtg;

// Device scratch of rtg_render_pass_static, the entry point that allocates
// nothing: the statistics slots and one chunk's path state live in this
// module-scope array. Every device the library runs on holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (static_scratch_of), never by
// the array's name, which on the host is not a device address.
constexpr uint64_t kStaticScratchBytes = RTG_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
thread_local std::string g_error;

// Address of `scratch` on each device, looked up once per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

// With `device` current: its copy of the scratch array.
hipError_t static_scratch_of(int device, char** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<char*>(p);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

// Zero fill (a kernel, so a caller may capture it in a graph).
__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Path-state slots of a context.
constexpr int kBatchBuffers = 2;
constexpr int kBatchBuffers2 = 5;  // batch groups of the second lane
constexpr int kLaneBuffers = 3;    // FP32 chunks of the second lane
constexpr int kLane64Buffers = 4;  // FP64 chunks of the second lane
// Batches: every job's queue storage is rounded up to kSubQueues sub-queues
// of whole workgroups, at most kSubQueues * kBlock entries more than its
// paths.
constexpr uint64_t kBatchQueueSlack = (uint64_t)kMaxBatchJobs * kSubQueues * kBlock;

struct Status {
  rtg_status code;
  std::string msg;
};
}  // namespace

struct rtg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  SideStream side{};  // default priority, like the main stream
  std::string err;
  unsigned long long* stats = nullptr;
  unsigned long long* stats_host = nullptr;  // pinned copy of `stats`, written at the end of a call
  // Statistics slots of a batch's single-render jobs (one kStatSlots region
  // per job), and their pinned copy.
  unsigned long long* job_stats = nullptr;
  unsigned long long* job_stats_host = nullptr;
  uint32_t job_stats_cap = 0;
  int n_cu = 0;
  // Path state, grown on demand: FP32 chunks, FP64 chunks, FP32 batches.
  void* path_mem[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  uint64_t path_bytes[6] = {0, 0, 0, 0, 0, 0};
  PathBuffers<float> pb32{};
  // A second lane for FP32 path-integrator passes: chunks alternate between
  // (stream, side, pb32) and (stream2, side2, pb32b), so two chunks' kernels
  // interleave and fill each other's tails.
  hipStream_t stream2 = nullptr;
  SideStream side2{};
  PathBuffers<float> pb32b{};
  PathBuffers<double> pb64b{};
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t acc_done[2] = {nullptr, nullptr};  // each lane's latest accumulation
  PathBuffers<double> pb64{};
  PathBuffers<float> pbb{}, pbb2{};  // batch groups, one set per lane
  uint32_t cap[6] = {0, 0, 0, 0, 0, 0};
  // Batches: device job table and plans per lane; pinned host staging of
  // the job tables and plans of every group of a call, and of their
  // statistics slots.
  BatchJob* batch_jobs[2] = {nullptr, nullptr};
  uint32_t* batch_plans[2] = {nullptr, nullptr};  // job count, then 5 x (kMaxBatchJobs + 1): gen, acc, act, shd x 2
  void* batch_stage = nullptr;
  size_t batch_stage_cap = 0;  // groups
  unsigned long long* batch_stats_host = nullptr;
  size_t batch_stats_cap = 0;  // segments
  uint32_t batch_occ[3][3] = {};    // per stack class: trace, shade, other workgroups per CU
  // Captured batch passes, one per (max depth, stack class): the launches of
  // a pass replay as one graph. Their arguments point into the batch
  // buffers, so they are dropped whenever those are reallocated.
  struct BatchGraph {
    int depth, cls;
    bool side;  // captured with the any-hit kernels on the side stream
    hipGraphExec_t exec;
  };
  std::vector<BatchGraph> batch_graphs;
  uint8_t* rgb_dev = nullptr;       // rtg_resolve staging
  size_t rgb_cap = 0;
  // Profiling: an event pair per launch, resolved when the call completes.
  bool profiling = false;
  uint32_t profile_kinds = 0;  // RTG_K_* kinds whose launches are timed
  std::vector<hipEvent_t> events;
  size_t events_used = 0;
  std::vector<int> event_kind;  // kind of each begin event
  rtg_profile profile{};
  LaunchHooks hooks{};
  char* scratch_dev = nullptr;  // this device's copy of `scratch` (rtg_render_pass_static)
  rtg_tuning tuning{};          // scheduling knobs (zero: built-in default)
  void* frame_mem = nullptr;    // direct integrator: recursion frames of a chunk
  uint64_t frame_bytes = 0;
};

void drop_batch_graphs(rtg_context* ctx) {
  for (auto& g : ctx->batch_graphs) hipGraphExecDestroy(g.exec);
  ctx->batch_graphs.clear();
}

// Device arrays of a scene, shared by the handles made from it with other
// cameras (rtg_scene_with_camera); freed with the last of them.
struct DeviceArrays {
  int device = 0;
  std::vector<void*> ptrs;
  ~DeviceArrays() {
    if (ptrs.empty()) return;
    hipSetDevice(device);
    for (void* p : ptrs) hipFree(p);
  }
};

struct rtg_scene {
  rtg_context* ctx = nullptr;
  DevScene dev{};
  std::shared_ptr<DeviceArrays> arrays;
  uint64_t bytes = 0;
  uint32_t stack = 0;  // BVH depth + 1
};

struct rtg_accum {
  rtg_context* ctx = nullptr;
  int32_t w = 0, h = 0;
  double* sum = nullptr;
  uint32_t* count = nullptr;
};

struct rtg_buffer {
  rtg_context* ctx = nullptr;
  void* ptr = nullptr;
  uint64_t bytes = 0;
};

namespace {

rtg_status fail(rtg_context* ctx, rtg_status code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  g_error = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                              \
  do {                                                                                                  \
    const hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                               \
      return fail(ctx, e_ == hipErrorOutOfMemory ? RTG_ERR_OUT_OF_MEMORY : RTG_ERR_HIP,                 \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                                   \
  } while (0)

// Outward rounding of the double boxes to float: the float box always
// contains the double box.
float round_down(double v) {
  float f = (float)v;
  if ((double)f > v) f = std::nextafter(f, -INFINITY);
  return f;
}
float round_up(double v) {
  float f = (float)v;
  if ((double)f < v) f = std::nextafter(f, INFINITY);
  return f;
}
float bits_f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
float4 f4(float x, float y, float z, float w) { return float4{x, y, z, w}; }

struct NodeFrame {
  float origin[3];
  int exp2;
};
// The frame of the FP32 node records: box coordinates are stored as half
// floats of (b - origin) * 2^-e, origin the centre of the BVH's root box
// (a float) and e the power of two that brings the root box within
// +-2^14 of it (clamped to [-60, 20]; a box past the half range becomes
// infinite, which still bounds it). A record's precision then follows the
// scene's own extent, not where the scene sits in world space.
NodeFrame node_frame(const rtg_scene_desc* d) {
  NodeFrame f{{0.0f, 0.0f, 0.0f}, 0};
  if (!d->n_nodes) return f;
  const rtg_bvh_node& r = d->nodes[0];
  const double lo[3] = {r.lo.x, r.lo.y, r.lo.z}, hi[3] = {r.hi.x, r.hi.y, r.hi.z};
  double half = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double c = 0.5 * lo[a] + 0.5 * hi[a];
    f.origin[a] = std::isfinite(c) ? (float)c : 0.0f;
    half = std::max({half, std::fabs(lo[a] - f.origin[a]), std::fabs(hi[a] - f.origin[a])});
  }
  int e = 0;
  if (half > 0.0 && std::isfinite(half)) e = (int)std::ceil(std::log2(half / 16384.0));
  f.exp2 = std::min(20, std::max(-60, e));
  return f;
}
// A coordinate in the frame as the nearest half float at or below (lo) or
// at or above (hi) it, so the half box contains the double box: v - origin
// is split exactly into s + err (Knuth's two-sum), both scaled by 2^-e
// (exact), and the nearest half of s is stepped outward when it lies inside
// s + err (|err| is below half an ulp of s, far below a half step).
uint32_t half_bits(_Float16 h) { return __builtin_bit_cast(uint16_t, h); }
uint32_t frame_half(const NodeFrame& f, int axis, double v, bool up) {
  const double a = v, b = -(double)f.origin[axis];
  const double s = a + b, bb = s - a, err = (a - (s - bb)) + (b - bb);
  const double x = std::ldexp(s, -f.exp2), xe = std::ldexp(err, -f.exp2);
  const _Float16 h = (_Float16)x;
  const double hv = (double)h;
  uint32_t bits = half_bits(h);
  if (!up && (hv > x || (hv == x && xe < 0.0)))  // the next half down
    bits = bits == 0u ? 0x8001u : (bits & 0x8000u) ? bits + 1u : bits - 1u;
  if (up && (hv < x || (hv == x && xe > 0.0)))  // the next half up
    bits = bits == 0x8000u ? 0x0001u : (bits & 0x8000u) ? bits - 1u : bits + 1u;
  return bits;
}
float half_value(uint32_t b) { return (float)__builtin_bit_cast(_Float16, (uint16_t)b); }

// The camera as generate() reads it: FP64 for the exact mode, FP32 origin,
// corner - origin, du, dv, u, v and lens radius for the fast mode.
void set_camera(DevScene& S, const rtg_camera& c) {
  S.cam = c;
  const double cam[19] = {c.origin.x, c.origin.y, c.origin.z, c.corner.x - c.origin.x, c.corner.y - c.origin.y,
                          c.corner.z - c.origin.z, c.du.x, c.du.y, c.du.z, c.dv.x, c.dv.y, c.dv.z,
                          c.u.x, c.u.y, c.u.z, c.v.x, c.v.y, c.v.z, c.lens_radius};
  for (int i = 0; i < 19; ++i) S.cam32[i] = (float)cam[i];
}

struct Upload {
  rtg_scene* s;
  rtg_context* ctx;
  template <class T>
  rtg_status put(const T* host, size_t n, const T** dev) {
    if (n == 0) {
      *dev = nullptr;
      return RTG_OK;
    }
    void* p = nullptr;
    HIP_TRY(ctx, hipMalloc(&p, n * sizeof(T)));
    s->arrays->ptrs.push_back(p);
    s->bytes += n * sizeof(T);
    HIP_TRY(ctx, hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
    *dev = static_cast<const T*>(p);
    return RTG_OK;
  }
};

rtg_status validate_desc(rtg_context* ctx, const rtg_scene_desc* d) {
  auto bad = [&](const std::string& m) { return fail(ctx, RTG_ERR_INVALID_ARG, "scene: " + m); };
  if ((d->n_prims && !d->prims) || (d->n_planes && !d->planes) || (d->n_materials && !d->materials) ||
      (d->n_lights && !d->lights) || (d->n_nodes && !d->nodes) || (d->n_prim_indices && !d->prim_indices) ||
      (d->n_point_lights && !d->point_lights) || (d->n_prims && !d->is_sampled_light))
    return bad("null array with non-zero count");
  for (uint32_t i = 0; i < d->n_prims; ++i) {
    if (d->prims[i].material >= d->n_materials) return bad("primitive material out of range");
    if (d->prims[i].kind > RTG_PRIM_TRIANGLE) return bad("unknown primitive kind");
  }
  for (uint32_t i = 0; i < d->n_planes; ++i)
    if (d->planes[i].material >= d->n_materials) return bad("plane material out of range");
  for (uint32_t i = 0; i < d->n_lights; ++i) {
    const rtg_light_ref& l = d->lights[i];
    if (l.kind == RTG_LIGHT_POINT ? l.index >= d->n_point_lights : l.index >= d->n_prims)
      return bad("light index out of range");
  }
  for (uint32_t i = 0; i < d->n_prim_indices; ++i)
    if (d->prim_indices[i] >= d->n_prims) return bad("prim_indices out of range");
  for (uint32_t i = 0; i < d->n_nodes; ++i) {
    const rtg_bvh_node& n = d->nodes[i];
    if (n.count > 0) {
      if ((uint64_t)n.offset + n.count > d->n_prim_indices) return bad("leaf range out of prim_indices");
    } else {
      if (n.offset <= i || n.offset >= d->n_nodes || i + 1 >= d->n_nodes) return bad("interior child out of range");
      if (n.axis > 2) return bad("split axis out of range");
    }
  }
  return RTG_OK;
}

void profile_events(void* user, int kind, hipEvent_t* start, hipEvent_t* stop) {
  auto* ctx = static_cast<rtg_context*>(user);
  if (!(ctx->profile_kinds & (1u << kind))) return;
  while (ctx->events.size() < ctx->events_used + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    ctx->events.push_back(e);
  }
  if (ctx->event_kind.size() < ctx->events_used + 1) ctx->event_kind.resize(ctx->events_used + 1);
  ctx->event_kind[ctx->events_used] = kind;
  *start = ctx->events[ctx->events_used++];
  *stop = ctx->events[ctx->events_used++];
}

// Adds every recorded start/stop pair to the totals. Called when the profile
// is read and when many events are pending, not after every call, so that
// the host does not read event times while the next pass could be running.
void profile_flush_events(rtg_context* ctx) {
  for (size_t i = 0; i + 1 < ctx->events_used; i += 2) {
    float ms = 0;
    hipEventSynchronize(ctx->events[i + 1]);
    if (hipEventElapsedTime(&ms, ctx->events[i], ctx->events[i + 1]) == hipSuccess) {
      const int k = ctx->event_kind[i];
      ctx->profile.ms[k] += ms;
      ctx->profile.launches[k]++;
    }
  }
  ctx->events_used = 0;
}

// After the stream has drained: a call's counters (its launch times stay
// pending until profile_flush_events).
void profile_resolve(rtg_context* ctx, const unsigned long long* h) {
  if (!ctx->profiling) return;
  constexpr size_t kPendingEvents = 1u << 13;
  if (ctx->events_used >= kPendingEvents) profile_flush_events(ctx);
  ctx->profile.closest_rays += h[kStRays];
  ctx->profile.closest_visits += h[kStVisits];
  ctx->profile.closest_tests += h[kStTests];
  ctx->profile.shadow_rays += h[kStShadow];
  ctx->profile.shadow_visits += h[kStSVisits];
  ctx->profile.shadow_tests += h[kStSTests];
}

// The side stream for overlapping a pass's any-hit kernels, unless
// RTG_OVERLAP=0 (everything on the main stream).
const SideStream* side_of(rtg_context* ctx, bool lane2 = false) {
  const char* e = std::getenv("RTG_OVERLAP");
  return (e && e[0] == '0') ? nullptr : lane2 ? &ctx->side2 : &ctx->side;
}

const LaunchHooks* hooks_of(rtg_context* ctx) {
  if (!ctx->profiling) return nullptr;
  ctx->hooks.user = ctx;
  ctx->hooks.events = profile_events;
  return &ctx->hooks;
}

// The FP32 two-child records of the reference BVH's interior nodes, in its
// DFS order (node i's record is the i-th interior node's): both child boxes
// rounded outward from the double boxes, so each float box contains its
// double box, and the two child refs (leaf: first slot + count; interior:
// record index) and the split axis. max_depth = the BVH's depth.
bool build_node_records(const rtg_scene_desc* d, const NodeFrame& fr, std::vector<float4>& node2,
                        uint32_t& max_depth, std::string& err) {
  std::vector<int32_t> n2_index(d->n_nodes, -1);
  uint32_t n_interior = 0;
  max_depth = 0;
  std::vector<std::pair<uint32_t, uint32_t>> st;  // (node, depth)
  if (d->n_nodes) st.push_back({0, 0});
  while (!st.empty()) {
    auto [i, dep] = st.back();
    st.pop_back();
    max_depth = std::max(max_depth, dep);
    const rtg_bvh_node& n = d->nodes[i];
    if (n.count == 0) {
      if (n.offset >= d->n_nodes || i + 1 >= d->n_nodes) {
        err = "scene: BVH child index out of range";
        return false;
      }
      st.push_back({n.offset, dep + 1});
      st.push_back({i + 1, dep + 1});
    } else if (n.count > kMaxLeafCount) {
      err = "scene: BVH leaf with more than 64 primitives";
      return false;
    }
    if (dep > 63) {
      err = "scene: BVH deeper than 63 levels";
      return false;
    }
  }
  for (uint32_t i = 0; i < d->n_nodes; ++i)
    if (d->nodes[i].count == 0) n2_index[i] = (int32_t)n_interior++;
  if (n_interior >= kMaxSlots) {  // node records are addressed by 32-bit byte offsets
    err = "scene: more than 2^25 interior BVH nodes";
    return false;
  }
  auto child_ref = [&](uint32_t c) -> uint32_t {
    const rtg_bvh_node& n = d->nodes[c];
    return n.count > 0 ? leaf_ref(n.offset, n.count) : (uint32_t)n2_index[c];
  };
  node2.assign((size_t)n_interior * kNode2Stride, f4(0, 0, 0, 0));
  for (uint32_t i = 0; i < d->n_nodes; ++i) {
    const rtg_bvh_node& n = d->nodes[i];
    if (n.count > 0) continue;
    const rtg_bvh_node& L = d->nodes[i + 1];
    const rtg_bvh_node& R = d->nodes[n.offset];
    float4* o = &node2[(size_t)n2_index[i] * kNode2Stride];
    uint32_t h[12];
    const rtg_bvh_node* child[2] = {&L, &R};
    for (int c = 0; c < 2; ++c) {
      const double lo[3] = {child[c]->lo.x, child[c]->lo.y, child[c]->lo.z};
      const double hi[3] = {child[c]->hi.x, child[c]->hi.y, child[c]->hi.z};
      for (int a = 0; a < 3; ++a) {
        h[6 * c + a] = frame_half(fr, a, lo[a], false);
        h[6 * c + 3 + a] = frame_half(fr, a, hi[a], true);
      }
    }
    o[0] = f4(bits_f(h[0] | h[1] << 16), bits_f(h[2] | h[3] << 16), bits_f(h[4] | h[5] << 16),
              bits_f(h[6] | h[7] << 16));
    o[1] = f4(bits_f(h[8] | h[9] << 16), bits_f(h[10] | h[11] << 16), bits_f(child_ref(i + 1)),
              bits_f(child_ref(n.offset)));
  }
  return true;
}

}  // namespace

extern "C" {

int rtg_abi_version(void) { return RTG_ABI_VERSION; }

rtg_status rtg_set_profiling(rtg_context* ctx, int32_t on) {
  if (!ctx) return fail(nullptr, RTG_ERR_INVALID_ARG, "rtg_set_profiling: NULL context");
  return rtg_set_profiling_kinds(ctx, on ? (1u << RTG_K_COUNT) - 1u : 0u);
}

rtg_status rtg_set_profiling_kinds(rtg_context* ctx, uint32_t kinds) {
  if (!ctx) return fail(nullptr, RTG_ERR_INVALID_ARG, "rtg_set_profiling_kinds: NULL context");
  if (kinds >> RTG_K_COUNT) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_set_profiling_kinds: unknown kernel kind");
  ctx->profiling = kinds != 0;
  ctx->profile_kinds = kinds;
  ctx->profile = rtg_profile{};
  ctx->events_used = 0;
  return RTG_OK;
}

rtg_status rtg_get_profile(rtg_context* ctx, rtg_profile* out) {
  if (!ctx || !out) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_get_profile: NULL argument");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  profile_flush_events(ctx);
  *out = ctx->profile;
  return RTG_OK;
}

const char* rtg_last_error(const rtg_context* ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

rtg_status rtg_open(int device, rtg_context** out) {
  if (!out) return fail(nullptr, RTG_ERR_INVALID_ARG, "rtg_open: out is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(nullptr, RTG_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(nullptr, RTG_ERR_INVALID_ARG, "rtg_open: device out of range");
  auto* c = new rtg_context;
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side.shaded, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side.shadowed[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side.shadowed[1], hipEventDisableTiming);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side2.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side2.shaded, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side2.shadowed[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side2.shadowed[1], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->acc_done[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->acc_done[1], hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&c->stats, sizeof(unsigned long long) * kStatSlots * kStatStride);
  if (e == hipSuccess) e = hipHostMalloc(&c->stats_host, sizeof(unsigned long long) * kStatSlots * kStatStride);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = static_scratch_of(device, &c->scratch_dev);
  if (e != hipSuccess) {
    const std::string m = std::string("rtg_open: ") + hipGetErrorString(e);
    delete c;
    return fail(nullptr, RTG_ERR_HIP, m);
  }
  *out = c;
  return RTG_OK;
}

rtg_status rtg_close(rtg_context* ctx) {
  if (!ctx) return RTG_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  drop_batch_graphs(ctx);
  for (void* p : ctx->path_mem)
    if (p) hipFree(p);
  if (ctx->frame_mem) hipFree(ctx->frame_mem);
  if (ctx->stats) hipFree(ctx->stats);
  if (ctx->stats_host) hipHostFree(ctx->stats_host);
  if (ctx->job_stats) hipFree(ctx->job_stats);
  if (ctx->job_stats_host) hipHostFree(ctx->job_stats_host);
  for (int l = 0; l < 2; ++l) {
    if (ctx->batch_jobs[l]) hipFree(ctx->batch_jobs[l]);
    if (ctx->batch_plans[l]) hipFree(ctx->batch_plans[l]);
  }
  if (ctx->batch_stage) hipHostFree(ctx->batch_stage);
  if (ctx->batch_stats_host) hipHostFree(ctx->batch_stats_host);
  if (ctx->rgb_dev) hipFree(ctx->rgb_dev);
  if (ctx->ev0) hipEventDestroy(ctx->ev0);
  if (ctx->ev1) hipEventDestroy(ctx->ev1);
  if (ctx->side.stream) {
    hipStreamSynchronize(ctx->side.stream);
    hipStreamDestroy(ctx->side.stream);
  }
  if (ctx->side.shaded) hipEventDestroy(ctx->side.shaded);
  for (hipEvent_t e : ctx->side.shadowed)
    if (e) hipEventDestroy(e);
  for (hipStream_t st : {ctx->stream2, ctx->side2.stream})
    if (st) {
      hipStreamSynchronize(st);
      hipStreamDestroy(st);
    }
  for (hipEvent_t e : {ctx->side2.shaded, ctx->side2.shadowed[0], ctx->side2.shadowed[1], ctx->fork, ctx->join,
                       ctx->acc_done[0], ctx->acc_done[1]})
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : ctx->events) hipEventDestroy(e);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return RTG_OK;
}

rtg_status rtg_device_info(rtg_context* ctx, char* name, size_t name_len, int* mps, uint64_t* hbm) {
  if (!ctx) return fail(nullptr, RTG_ERR_INVALID_ARG, "rtg_device_info: NULL context");
  hipDeviceProp_t p;
  HIP_TRY(ctx, hipGetDeviceProperties(&p, ctx->device));
  if (name && name_len) {
    std::strncpy(name, p.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (mps) *mps = p.multiProcessorCount;
  if (hbm) *hbm = p.totalGlobalMem;
  return RTG_OK;
}

rtg_status rtg_synchronize(rtg_context* ctx) {
  if (!ctx) return fail(nullptr, RTG_ERR_INVALID_ARG, "NULL context");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return RTG_OK;
}

rtg_status rtg_scene_node_records(const rtg_scene_desc* desc, float* out, uint64_t cap, uint64_t* n) {
  if (!desc || !n) return fail(nullptr, RTG_ERR_INVALID_ARG, "rtg_scene_node_records: NULL argument");
  if (desc->n_nodes && !desc->nodes) return fail(nullptr, RTG_ERR_INVALID_ARG, "scene: null array with non-zero count");
  std::vector<float4> node2;
  uint32_t depth = 0;
  std::string err;
  if (!build_node_records(desc, node_frame(desc), node2, depth, err)) return fail(nullptr, RTG_ERR_UNSUPPORTED, err);
  *n = node2.size() / kNode2Stride;
  const uint64_t m = std::min<uint64_t>(cap, *n);
  for (uint64_t r = 0; out && r < m; ++r) {  // 16 floats per record: the 12 half planes widened, refs, axis, 0
    uint32_t w[8];
    std::memcpy(w, &node2[r * kNode2Stride], sizeof w);
    float* o = out + 16 * r;
    for (int k = 0; k < 12; ++k) o[k] = half_value((w[k / 2] >> (16 * (k & 1))) & 0xFFFFu);
    o[12] = bits_f(w[6]);
    o[13] = bits_f(w[7]);
    o[14] = 0.0f;
    o[15] = 0.0f;
  }
  // the split axis of the reference node each record stands for
  if (out && m) {
    uint64_t r = 0;
    for (uint32_t i = 0; i < desc->n_nodes && r < m; ++i)
      if (desc->nodes[i].count == 0) out[16 * r++ + 14] = bits_f(desc->nodes[i].axis);
  }
  return RTG_OK;
}

rtg_status rtg_scene_node_frame(const rtg_scene_desc* desc, float origin[3], int32_t* exp2) {
  if (!desc || !origin || !exp2) return fail(nullptr, RTG_ERR_INVALID_ARG, "rtg_scene_node_frame: NULL argument");
  if (desc->n_nodes && !desc->nodes) return fail(nullptr, RTG_ERR_INVALID_ARG, "scene: null array with non-zero count");
  const NodeFrame f = node_frame(desc);
  for (int a = 0; a < 3; ++a) origin[a] = f.origin[a];
  *exp2 = f.exp2;
  return RTG_OK;
}

// ---------------------------------------------------------------------------
rtg_status rtg_scene_create(rtg_context* ctx, const rtg_scene_desc* d, rtg_scene** out) {
  if (!ctx || !d || !out) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_scene_create: NULL argument");
  if (rtg_status st = validate_desc(ctx, d)) return st;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  auto* s = new rtg_scene;
  s->ctx = ctx;
  s->arrays = std::make_shared<DeviceArrays>();
  s->arrays->device = ctx->device;
  Upload up{s, ctx};
  DevScene& S = s->dev;
  auto cleanup = [&](rtg_status st) {
    delete s;
    return st;
  };
  const uint32_t np = d->n_prims;

  // --- BVH depth and the FP32 two-child node layout ---
  std::vector<float4> node2;
  uint32_t max_depth = 0;
  const NodeFrame frame = node_frame(d);
  {
    std::string err;
    if (!build_node_records(d, frame, node2, max_depth, err)) return cleanup(fail(ctx, RTG_ERR_UNSUPPORTED, err));
  }
  if (np > kMaxSlots) return cleanup(fail(ctx, RTG_ERR_UNSUPPORTED, "scene: more than 2^25 primitives"));
  s->stack = max_depth + 1;

  // --- slots: leaf order first, then primitives outside the BVH ---
  std::vector<uint32_t> prim_slot(np, 0xFFFFFFFFu);
  std::vector<uint32_t> slot_prim;
  slot_prim.reserve(np);
  for (uint32_t k = 0; k < d->n_prim_indices; ++k) {
    prim_slot[d->prim_indices[k]] = (uint32_t)slot_prim.size();
    slot_prim.push_back(d->prim_indices[k]);
  }
  for (uint32_t i = 0; i < np; ++i)
    if (prim_slot[i] == 0xFFFFFFFFu) {
      prim_slot[i] = (uint32_t)slot_prim.size();
      slot_prim.push_back(i);
    }
  std::vector<float> slot32((size_t)slot_prim.size() * kSlot32Floats, 0.0f);
  std::vector<float4> shade32((size_t)np * kShade32Stride);
  for (uint32_t k = 0; k < slot_prim.size(); ++k) {
    const uint32_t i = slot_prim[k];
    const rtg_primitive& p = d->prims[i];
    float* o = &slot32[(size_t)k * kSlot32Floats];
    if (p.kind == RTG_PRIM_SPHERE) {
      const double r = p.v1.x;
      const float c[9] = {(float)p.v0.x, (float)p.v0.y, (float)p.v0.z, bits_f(kSlotSphere), (float)(r * r), 0, 0, 0, 0};
      std::copy(c, c + 9, o);
    } else {
      const float c[9] = {(float)p.v0.x,          (float)p.v0.y,          (float)p.v0.z,
                          (float)(p.v1.x - p.v0.x), (float)(p.v1.y - p.v0.y), (float)(p.v1.z - p.v0.z),
                          (float)(p.v2.x - p.v0.x), (float)(p.v2.y - p.v0.y), (float)(p.v2.z - p.v0.z)};
      std::copy(c, c + 9, o);
      // the sphere marker is a NaN payload; a NaN edge component (it never
      // hits) is written as the default NaN
      if (std::isnan(o[3])) o[3] = std::numeric_limits<float>::quiet_NaN();
    }
  }
  for (uint32_t i = 0; i < np; ++i) {
    const rtg_primitive& p = d->prims[i];
    float4* o = &shade32[(size_t)i * kShade32Stride];
    if (p.kind == RTG_PRIM_SPHERE) {
      o[0] = f4((float)p.v0.x, (float)p.v0.y, (float)p.v0.z, bits_f(p.material | kShadeSphere));
      o[1] = f4((float)p.v1.x, (float)(1.0 / p.v1.x), 0, 0);
      o[2] = o[3] = f4(0, 0, 0, 0);
    } else {
      const double e1[3] = {p.v1.x - p.v0.x, p.v1.y - p.v0.y, p.v1.z - p.v0.z};
      const double e2[3] = {p.v2.x - p.v0.x, p.v2.y - p.v0.y, p.v2.z - p.v0.z};
      double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      const double l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
      const double inv = (l2 > 1e-18 && std::isfinite(l2)) ? 1.0 / std::sqrt(l2) : 0.0;
      const uint32_t w = p.material | (p.has_vertex_normals ? kShadeSmooth : 0u);
      o[0] = f4((float)(n[0] * inv), (float)(n[1] * inv), (float)(n[2] * inv), bits_f(w));
      o[1] = f4((float)p.n0.x, (float)p.n0.y, (float)p.n0.z, (float)p.n1.x);
      o[2] = f4((float)p.n1.y, (float)p.n1.z, (float)p.n2.x, (float)p.n2.y);
      o[3] = f4((float)p.n2.z, 0, 0, 0);
    }
  }
  std::vector<float4> mat32((size_t)d->n_materials * 2);
  for (uint32_t i = 0; i < d->n_materials; ++i) {
    const rtg_material& m = d->materials[i];
    const bool textured = m.kind == RTG_MAT_LAMBERTIAN && (float)m.p0 > 0.0f;  // checker (material.cpp:14-24)
    mat32[2 * i] = f4((float)m.c0.x, (float)m.c0.y, (float)m.c0.z,
                      bits_f(m.kind | (m.two_sided ? 256u : 0u) | (textured ? 512u : 0u)));
    mat32[2 * i + 1] = f4((float)m.c1.x, (float)m.c1.y, (float)m.c1.z, (float)m.p0);
  }
  std::vector<float4> plane32((size_t)d->n_planes * 2);
  for (uint32_t i = 0; i < d->n_planes; ++i) {
    const rtg_plane& p = d->planes[i];
    plane32[2 * i] = f4((float)p.point.x, (float)p.point.y, (float)p.point.z, bits_f(p.material));
    plane32[2 * i + 1] = f4((float)p.normal.x, (float)p.normal.y, (float)p.normal.z, 0);
  }
  std::vector<float4> light32((size_t)d->n_lights * kLight32Stride, f4(0, 0, 0, 0));
  for (uint32_t i = 0; i < d->n_lights; ++i) {
    const rtg_light_ref& l = d->lights[i];
    float4* o = &light32[(size_t)i * kLight32Stride];
    if (l.kind == RTG_LIGHT_POINT) {
      const rtg_point_light& pl = d->point_lights[l.index];
      o[0] = f4((float)pl.position.x, (float)pl.position.y, (float)pl.position.z, bits_f(kL32Point));
      o[1] = f4((float)pl.intensity.x, (float)pl.intensity.y, (float)pl.intensity.z, 0);
      continue;
    }
    const rtg_primitive& p = d->prims[l.index];
    const rtg_material& m = d->materials[p.material];
    const double em[3] = {m.c0.x * m.p0, m.c0.y * m.p0, m.c0.z * m.p0};
    const float ts = m.two_sided ? 1.0f : 0.0f;
    if (p.kind == RTG_PRIM_SPHERE) {
      const double r = p.v1.x;
      o[0] = f4((float)p.v0.x, (float)p.v0.y, (float)p.v0.z, bits_f(kL32Sphere));
      o[1] = f4((float)r, (float)(r * r), (float)(4 * M_PI * r * r), ts);
      o[2] = f4((float)em[0], (float)em[1], (float)em[2], 0);
    } else {
      const double e1[3] = {p.v1.x - p.v0.x, p.v1.y - p.v0.y, p.v1.z - p.v0.z};
      const double e2[3] = {p.v2.x - p.v0.x, p.v2.y - p.v0.y, p.v2.z - p.v0.z};
      const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      const double inv = len > 1e-9 ? 1.0 / len : 0.0;
      o[0] = f4((float)p.v0.x, (float)p.v0.y, (float)p.v0.z, bits_f(kL32Triangle));
      o[1] = f4((float)e1[0], (float)e1[1], (float)e1[2], (float)(0.5 * len));
      o[2] = f4((float)e2[0], (float)e2[1], (float)e2[2], ts);
      o[3] = f4((float)em[0], (float)em[1], (float)em[2], 0);
      o[4] = f4((float)(n[0] * inv), (float)(n[1] * inv), (float)(n[2] * inv), 0);
    }
  }

  // --- uploads ---
  rtg_status st = RTG_OK;
#define PUT(host, n, dst)                                  \
  if ((st = up.put((host), (n), &(dst))) != RTG_OK) return cleanup(st)
  PUT(node2.data(), node2.size(), S.node2);
  PUT(slot32.data(), slot32.size(), S.slot32);
  PUT(slot_prim.data(), slot_prim.size(), S.slot_prim);
  PUT(shade32.data(), shade32.size(), S.shade32);
  PUT(prim_slot.data(), prim_slot.size(), S.prim_slot);
  PUT(mat32.data(), mat32.size(), S.mat32);
  PUT(plane32.data(), plane32.size(), S.plane32);
  PUT(light32.data(), light32.size(), S.light32);
  PUT(d->nodes, (size_t)d->n_nodes, S.nodes);
  PUT(d->prim_indices, (size_t)d->n_prim_indices, S.prim_indices);
  PUT(d->prims, (size_t)np, S.prims);
  PUT(d->is_sampled_light, (size_t)np, S.sampled);
  PUT(d->materials, (size_t)d->n_materials, S.materials);
  PUT(d->planes, (size_t)d->n_planes, S.planes);
  PUT(d->lights, (size_t)d->n_lights, S.lights);
  PUT(d->point_lights, (size_t)d->n_point_lights, S.point_lights);
#undef PUT
  S.has_bvh = d->n_nodes > 0;
  S.n_nodes = d->n_nodes;
  S.n_prims = np;
  S.n_slots = (uint32_t)slot_prim.size();
  S.n_planes = d->n_planes;
  S.n_lights = d->n_lights;
  S.n_materials = d->n_materials;
  S.stack_depth = s->stack;
  if (d->n_nodes) {
    const rtg_bvh_node& r = d->nodes[0];
    S.root_ref = r.count > 0 ? leaf_ref(r.offset, r.count) : 0u;
    const double lo[3] = {r.lo.x, r.lo.y, r.lo.z}, hi[3] = {r.hi.x, r.hi.y, r.hi.z};
    for (int a = 0; a < 3; ++a) {  // in the node frame, as the records' boxes
      S.root_lo[a] = half_value(frame_half(frame, a, lo[a], false));
      S.root_hi[a] = half_value(frame_half(frame, a, hi[a], true));
    }
  }
  for (int a = 0; a < 3; ++a) S.frame_o[a] = frame.origin[a];
  S.frame_s = std::ldexp(1.0f, frame.exp2);
  set_camera(S, d->camera);
  S.bg = d->background;
  const rtg_background& b = d->background;
  const double bg[7] = {b.top.x, b.top.y, b.top.z, b.bottom.x, b.bottom.y, b.bottom.z, b.intensity};
  for (int i = 0; i < 7; ++i) S.bg32[i] = (float)bg[i];
  S.bg_solid = b.kind == RTG_BG_SOLID;
  *out = s;
  return RTG_OK;
}

rtg_status rtg_scene_with_camera(rtg_context* ctx, const rtg_scene* base, const rtg_camera* camera, rtg_scene** out) {
  if (!ctx || !base || !camera || !out) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_scene_with_camera: NULL argument");
  if (base->ctx->device != ctx->device)
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_scene_with_camera: the scene lives on another device");
  auto* s = new rtg_scene(*base);
  s->ctx = ctx;
  set_camera(s->dev, *camera);
  *out = s;
  return RTG_OK;
}

rtg_status rtg_scene_release(rtg_scene* s) {
  if (!s) return RTG_OK;
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  delete s;
  return RTG_OK;
}

uint64_t rtg_scene_bytes(const rtg_scene* s) { return s ? s->bytes : 0; }

uint64_t rtg_context_bytes(const rtg_context* ctx) {
  if (!ctx) return 0;
  uint64_t b = 0;
  for (uint64_t x : ctx->path_bytes) b += x;
  b += sizeof(unsigned long long) * kStatSlots * kStatStride;                   // stats
  b += sizeof(unsigned long long) * kStatSlots * kStatStride * ctx->job_stats_cap;  // single-render job stats
  for (int l = 0; l < 2; ++l)
    if (ctx->batch_jobs[l]) b += sizeof(BatchJob) * kMaxBatchJobs + sizeof(uint32_t) * (1 + 5 * (kMaxBatchJobs + 1));
  b += ctx->rgb_cap;
  b += ctx->frame_bytes;
  return b;
}

// ---------------------------------------------------------------------------
rtg_status rtg_accum_create(rtg_context* ctx, int32_t w, int32_t h, rtg_accum** out) {
  if (!ctx || !out || w < 1 || h < 1 || w > 16384 || h > 16384)
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_accum_create: bad arguments");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  auto* a = new rtg_accum;
  a->ctx = ctx;
  a->w = w;
  a->h = h;
  const size_t n = (size_t)w * h;
  hipError_t e = hipMalloc(&a->sum, n * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->count, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemsetAsync(a->sum, 0, n * 3 * sizeof(double), ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->count, 0, n * sizeof(uint32_t), ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    if (a->sum) hipFree(a->sum);
    if (a->count) hipFree(a->count);
    delete a;
    return fail(ctx, e == hipErrorOutOfMemory ? RTG_ERR_OUT_OF_MEMORY : RTG_ERR_HIP,
                std::string("rtg_accum_create: ") + hipGetErrorString(e));
  }
  *out = a;
  return RTG_OK;
}

rtg_status rtg_accum_release(rtg_accum* a) {
  if (!a) return RTG_OK;
  hipSetDevice(a->ctx->device);
  hipStreamSynchronize(a->ctx->stream);
  hipFree(a->sum);
  hipFree(a->count);
  delete a;
  return RTG_OK;
}

rtg_status rtg_accum_clear(rtg_context* ctx, rtg_accum* a) {
  if (!ctx || !a) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_accum_clear: NULL argument");
  const size_t n = (size_t)a->w * a->h;
  HIP_TRY(ctx, hipMemsetAsync(a->sum, 0, n * 3 * sizeof(double), ctx->stream));
  HIP_TRY(ctx, hipMemsetAsync(a->count, 0, n * sizeof(uint32_t), ctx->stream));
  return RTG_OK;
}

rtg_status rtg_accum_read(rtg_context* ctx, const rtg_accum* a, double* sum, uint32_t* count) {
  if (!ctx || !a) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_accum_read: NULL argument");
  const size_t n = (size_t)a->w * a->h;
  if (sum) HIP_TRY(ctx, hipMemcpyAsync(sum, a->sum, n * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (count) HIP_TRY(ctx, hipMemcpyAsync(count, a->count, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return RTG_OK;
}

uint32_t rtg_accum_min_samples(rtg_context* ctx, const rtg_accum* a) {
  if (!ctx || !a) return 0;
  std::vector<uint32_t> c((size_t)a->w * a->h);
  if (rtg_accum_read(ctx, a, nullptr, c.data()) != RTG_OK) return 0;
  return *std::min_element(c.begin(), c.end());
}

rtg_status rtg_resolve(rtg_context* ctx, const rtg_accum* a, int32_t tone_map, double exposure, uint8_t* rgb) {
  if (!ctx || !a || !rgb || tone_map < 0 || tone_map > 2)
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_resolve: bad arguments");
  const size_t n = (size_t)a->w * a->h * 3;
  if (ctx->rgb_cap < n) {  // device staging of the 8-bit image, kept per context
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->rgb_dev) HIP_TRY(ctx, hipFree(ctx->rgb_dev));
    ctx->rgb_dev = nullptr;
    ctx->rgb_cap = 0;
    HIP_TRY(ctx, hipMalloc((void**)&ctx->rgb_dev, n));
    ctx->rgb_cap = n;
  }
  run_resolve(a->sum, a->count, a->w, a->h, tone_map, exposure, ctx->rgb_dev, ctx->stream);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(rgb, ctx->rgb_dev, n, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return RTG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
namespace {

// Path state for `cap` paths laid out from `base` (the path arrays,
// sub-queues and queue counters, and for batches the statistics slots).
// Every array has c entries: c = room for all sub-queues of one chunk, or
// (batches) of every job of a batch. Returns the bytes the layout takes;
// with base == nullptr only the size is computed and B is left alone.
template <class T>
uint64_t lay_out_paths(char* base, uint32_t cap, bool batch, unsigned long long* stats, PathBuffers<T>* B) {
  const uint32_t sub_cap = sub_queue_cap(cap);
  const uint64_t c = batch ? (uint64_t)cap + kBatchQueueSlack : (uint64_t)sub_cap * kSubQueues;  // >= cap
  auto al = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
  // two buffers of 6 ray + 3 throughput arrays of T and a specular flag;
  // 3 L + 3 hit + 10 shadow arrays of T; FP64: two buffers of PCG32 state
  // and increment (u64); 4 u32 (2 queues, the hit word, sq_pid); counts.
  // FP32: a second shadow queue (10 T + sq_pid) and 3 shadow-radiance arrays.
  constexpr bool f32 = sizeof(T) == 4;
  const uint64_t sq = f32 ? 2 * c : c;  // shadow-queue entries
  const uint64_t n_counts =
      (batch ? (uint64_t)kMaxBatchJobs * kBatchCounts : 2ull * kMaxBounce * kSubQueues) * kCountStride;
  const uint64_t n_stats = batch ? (uint64_t)kMaxBatchJobs * kBatchStatSlots * kStatStride : 0;
  const uint64_t bytes = al(c * sizeof(T)) * (24 + (f32 ? 3 : 0)) + al(sq * sizeof(T)) * 10 + al(c * 8) * (f32 ? 0 : 4) +
                         al(c * 4) * 3 + al(sq * 4) + al(c) * 2 + al(sizeof(uint32_t) * n_counts) + al(8 * n_stats);
  if (!base) return bytes;
  char* p = base;
  auto take = [&](uint64_t b) {
    char* r = p;
    p += al(b);
    return r;
  };
  T** tarr[] = {&B->ray[0].ox, &B->ray[0].oy, &B->ray[0].oz, &B->ray[0].dx, &B->ray[0].dy, &B->ray[0].dz,
                &B->ray[0].bx, &B->ray[0].by, &B->ray[0].bz, &B->ray[1].ox, &B->ray[1].oy, &B->ray[1].oz,
                &B->ray[1].dx, &B->ray[1].dy, &B->ray[1].dz, &B->ray[1].bx, &B->ray[1].by, &B->ray[1].bz,
                &B->lx,        &B->ly,        &B->lz,        &B->ht,        &B->hb1,       &B->hb2};
  for (T** a : tarr) *a = reinterpret_cast<T*>(take(c * sizeof(T)));
  T** sarr[] = {&B->sox, &B->soy, &B->soz, &B->sdx, &B->sdy, &B->sdz, &B->stm, &B->scx, &B->scy, &B->scz};
  for (T** a : sarr) *a = reinterpret_cast<T*>(take(sq * sizeof(T)));
  B->slx = B->sly = B->slz = nullptr;
  if (f32) {
    B->slx = reinterpret_cast<T*>(take(c * sizeof(T)));
    B->sly = reinterpret_cast<T*>(take(c * sizeof(T)));
    B->slz = reinterpret_cast<T*>(take(c * sizeof(T)));
  }
  B->sq_stride = f32 ? (uint32_t)c : 0u;
  for (RayArrays<T>& R : B->ray) {  // FP64: the PCG32 stream of the path at each slot
    R.rs = f32 ? nullptr : reinterpret_cast<uint64_t*>(take(c * 8));
    R.ri = f32 ? nullptr : reinterpret_cast<uint64_t*>(take(c * 8));
  }
  B->q[0] = reinterpret_cast<uint32_t*>(take(c * 4));
  B->q[1] = reinterpret_cast<uint32_t*>(take(c * 4));
  B->hprim = reinterpret_cast<uint32_t*>(take(c * 4));
  B->sq_pid = reinterpret_cast<uint32_t*>(take(sq * 4));
  B->ray[0].spec = reinterpret_cast<uint8_t*>(take(c));
  B->ray[1].spec = reinterpret_cast<uint8_t*>(take(c));
  B->counts = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * n_counts));
  B->stats = batch ? reinterpret_cast<unsigned long long*>(take(8 * n_stats)) : stats;
  B->capacity = cap;
  B->sub_cap = sub_cap;
  return bytes;
}

// Grows the context's path state `which` to room for `cap` paths.
template <class T>
rtg_status ensure_paths(rtg_context* ctx, int which, uint32_t cap, PathBuffers<T>& B) {
  if (ctx->cap[which] >= cap) return RTG_OK;
  const bool batch = which == kBatchBuffers || which == kBatchBuffers2;
  if (which == kBatchBuffers) drop_batch_graphs(ctx);  // graphs hold the first lane's buffers
  if (ctx->path_mem[which]) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream2));
    HIP_TRY(ctx, hipFree(ctx->path_mem[which]));
    ctx->path_mem[which] = nullptr;
    ctx->cap[which] = 0;
  }
  const uint64_t bytes = lay_out_paths<T>(nullptr, cap, batch, ctx->stats, nullptr);
  void* mem = nullptr;
  HIP_TRY(ctx, hipMalloc(&mem, bytes));
  lay_out_paths<T>(static_cast<char*>(mem), cap, batch, ctx->stats, &B);
  ctx->path_mem[which] = mem;
  ctx->path_bytes[which] = bytes;
  ctx->cap[which] = cap;
  return RTG_OK;
}

// Paths per device chunk and per batch group. Large chunks keep the later
// bounces' queues long enough to fill 256 CUs (measured on the torus
// workload: 1 Mi paths 1.56, 4 Mi 2.56, 16 Mi 2.99 G rays/s); 32 Mi paths of
// FP32 state are ~4.6 GB. RTG_CHUNK_PATHS overrides the context's tuning.
constexpr uint32_t kDefaultChunkPaths = 1u << 25;
uint32_t chunk_paths_of(const rtg_context* ctx) {
  if (const char* e = std::getenv("RTG_CHUNK_PATHS")) {
    const long v = std::atol(e);
    if (v >= 1024 && v <= (1L << 28)) return (uint32_t)v;
  }
  return ctx->tuning.chunk_paths ? ctx->tuning.chunk_paths : kDefaultChunkPaths;
}

int32_t pid_layout_default() {
  if (const char* e = std::getenv("RTG_SAMPLE_MAJOR")) return std::atoi(e) != 0;
  return 0;
}

void add_counters(rtg_stats* out, const unsigned long long* h) {
  out->camera_samples += h[kStCamera];
  out->rays += h[kStRays];
  out->shadow_rays += h[kStShadow];
  out->rr_terminated += h[kStRR];
  out->invalid_samples += h[kStInvalid];
  out->node_visits += h[kStVisits] + h[kStSVisits];
  out->prim_tests += h[kStTests] + h[kStSTests];
  out->redone_chunks += h[kStRedone];
}

// Enqueues the copy of the statistics slots into pinned memory (before the
// call's final event, so one wait covers both).
rtg_status stage_stats(rtg_context* ctx) {
  HIP_TRY(ctx, hipMemcpyAsync(ctx->stats_host, ctx->stats, sizeof(unsigned long long) * kStatSlots * kStatStride,
                              hipMemcpyDeviceToHost, ctx->stream));
  return RTG_OK;
}

// After the final event: sums the staged slots.
void read_stats(rtg_context* ctx, rtg_stats* out, float ms) {
  const unsigned long long* slots = ctx->stats_host;
  unsigned long long h[kStCount] = {0};
  for (int sl = 0; sl < kStatSlots; ++sl)
    for (int k = 0; k < kStCount; ++k) h[k] += slots[(size_t)sl * kStatStride + k];
  profile_resolve(ctx, h);
  if (!out) return;
  *out = rtg_stats{};
  add_counters(out, h);
  out->elapsed_ms = ms;
}

rtg_status check_settings(rtg_context* ctx, const char* fn_name, const rtg_render_settings* cfg,
                          const rtg_accum* acc, uint32_t s0, uint32_t s1) {
  auto fn = [&]() { return std::string(fn_name); };
  if (cfg->width != acc->w || cfg->height != acc->h)
    return fail(ctx, RTG_ERR_INVALID_ARG, fn() + ": settings size differs from the accumulation buffer");
  if (s1 <= s0) return fail(ctx, RTG_ERR_INVALID_ARG, fn() + ": empty sample range");
  if (cfg->max_depth < 1 || cfg->max_depth >= kMaxBounce)
    return fail(ctx, RTG_ERR_INVALID_ARG, fn() + ": max_depth must be in [1, 1023]");
  if (cfg->integrator < 0 || cfg->integrator > 2 || cfg->precision < 0 || cfg->precision > 1)
    return fail(ctx, RTG_ERR_INVALID_ARG, fn() + ": bad integrator or precision");
  return RTG_OK;
}

// The per-render fields of a job's parameters.
ChunkParams render_params(const rtg_render_settings* cfg, const rtg_accum* acc) {
  ChunkParams P{};
  P.width = (uint32_t)acc->w;
  P.height = (uint32_t)acc->h;
  P.max_depth = cfg->max_depth;
  P.rr_min_depth = cfg->rr_min_depth;
  P.integrator = cfg->integrator;
  P.light_sampling = cfg->light_sampling;
  P.use_bvh = cfg->use_bvh;
  P.sample_major = pid_layout_default();
  P.seed = cfg->seed;
  P.acc_sum = acc->sum;
  P.acc_count = acc->count;
  return P;
}

// Chunk capacity: the configured maximum, or less when the work is smaller
// (rounded up to a power of two so buffers are reused across passes).
uint32_t capacity_for(const rtg_context* ctx, uint64_t paths) {
  const uint32_t lim = chunk_paths_of(ctx);
  uint32_t cap = 1024;
  while (cap < paths && cap < lim) cap <<= 1;
  return std::min(cap, lim);
}

// Direct integrator memory of a chunk: FP64 keeps one sample's recursion
// frames per path (max_depth frames); FP32 runs the weighted ray queue
// (rtg_direct.hpp). None for the other integrators, whose walks never
// descend.
bool direct_forced(const rtg_render_settings* cfg, uint32_t n_lights) {
  const char* e = std::getenv("RTG_DIRECT");
  if (e && std::strcmp(e, "dfs") == 0) return true;
  return cfg->max_depth > kDirectWaveDepth || n_lights > kDirectWaveLights;
}
// Ray slots of one FP32 direct generation: twice the chunk's samples, and
// at least two per thread of a block in every sub-queue (a block appends to
// one sub-queue, so small chunks would otherwise overflow at the first
// generation). RTG_DIRECT_GROWTH sets the factor alone (tests use tiny
// queues); a generation that outgrows its slots makes its chunk run depth
// first, with the same result.
uint32_t direct_entries(uint32_t cap) {
  double f = 2.0, lo = 2.0 * kBlock * kSubQueues;
  if (const char* e = std::getenv("RTG_DIRECT_GROWTH")) {
    f = std::atof(e);
    lo = kSubQueues;
  }
  return (uint32_t)std::clamp(f * cap, lo, (double)(1u << 30));
}
uint64_t chunk_extra_bytes(const rtg_render_settings* cfg, uint32_t n_lights, uint32_t cap) {
  if (cfg->integrator != RTG_INTEGRATOR_DIRECT) return 0;
  if (cfg->precision == RTG_PRECISION_F64) return (uint64_t)cfg->max_depth * kFrameWords * 8 * cap;
  return direct_layout(nullptr, cap, direct_entries(cap), n_lights, cfg->max_depth, direct_forced(cfg, n_lights),
                       nullptr);
}
void set_direct(ChunkParams& P, const rtg_render_settings* cfg, uint32_t n_lights, uint32_t cap) {
  P.frame_stride = cap;
  P.dq_entries = direct_entries(cap);
  P.dq_forced = direct_forced(cfg, n_lights) ? 1u : 0u;
}

// Chunks of a direct render hold at most this much integrator memory: 4 Mi
// FP64 paths at depth 12, 64 Ki at depth 1000.
constexpr uint64_t kFrameBudget = 4ull << 30;
uint32_t frame_capacity(const rtg_render_settings* cfg, uint32_t n_lights, uint32_t cap) {
  while (cap > 1024 && chunk_extra_bytes(cfg, n_lights, cap) > kFrameBudget) cap >>= 1;
  return cap;
}

rtg_status ensure_frames(rtg_context* ctx, uint64_t bytes) {
  if (ctx->frame_bytes >= bytes) return RTG_OK;
  if (ctx->frame_mem) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream2));
    HIP_TRY(ctx, hipFree(ctx->frame_mem));
    ctx->frame_mem = nullptr;
    ctx->frame_bytes = 0;
  }
  HIP_TRY(ctx, hipMalloc(&ctx->frame_mem, bytes));
  ctx->frame_bytes = bytes;
  return RTG_OK;
}

// The largest chunk (a power of two, at least one workgroup of paths) whose
// path state and `extra(cap)` more bytes fit `bytes`; 0 if none does.
template <class Extra>
uint32_t static_capacity(bool f64, uint64_t bytes, Extra extra) {
  for (uint32_t cap = 1u << 25; cap >= (uint32_t)kBlock; cap >>= 1) {
    const uint64_t paths = f64 ? lay_out_paths<double>(nullptr, cap, false, nullptr, nullptr)
                               : lay_out_paths<float>(nullptr, cap, false, nullptr, nullptr);
    const uint64_t need = ((paths + 255) & ~uint64_t(255)) + extra(cap);
    if (need <= bytes) return cap;
  }
  return 0;
}

// Chunks of a path-integrator pass on two lanes (passes of at least
// kTwoLanePaths paths, or of one full chunk if chunks are smaller) unless
// RTG_LANES=1 or the context is tuned to one lane. The direct and normals
// kernels, one launch per chunk, fill the GPU alone and lose ~7 % to the
// split (measured).
constexpr uint64_t kTwoLanePaths = 1ull << 22;
bool two_lanes_enabled(const rtg_context* ctx) {
  const char* e = std::getenv("RTG_LANES");
  if (e && e[0] == '1') return false;
  return ctx->tuning.lanes != 1;
}

// One pass (samples [s0, s1) of pixels [px0, px1)) of one render, and the
// statistics slots (kStatSlots) its kernels add to.
struct PassItem {
  const rtg_scene* scene;
  const rtg_render_settings* cfg;
  rtg_accum* acc;
  uint32_t s0, s1, px0, px1;
  unsigned long long* stats;
};

// Runs `items` in order with no host synchronisation between chunks or
// items (the caller clears the statistics slots before and reads them after).
// Each item is cut into chunks of whole pixels x a sample sub-range;
// path-integrator chunks alternate between two lanes (streams, path
// buffers), so a chunk's kernels run beside the other lane's, across item
// boundaries too (a small pass fills the GPU poorly alone: its bounces'
// launches are short and end in tails). A chunk's accumulation waits for the
// other lane's latest one and each lane runs its chunks in order, so every
// pixel's additions stay in sample order: the result is bitwise that of the
// items one after another. With `cancel`, at most two chunks are in flight:
// before each chunk the lane's previous one has finished and the flag is
// read. Returns the items completed in *done_out.
rtg_status run_items(rtg_context* ctx, const PassItem* items, uint32_t n, const volatile int32_t* cancel,
                     uint32_t* done_out) {
  const uint64_t chunk = chunk_paths_of(ctx);
  const uint64_t split_last = std::min<uint64_t>(kTwoLanePaths, chunk);
  const bool lanes = two_lanes_enabled(ctx);
  auto lanes_of = [&](uint32_t i) { return items[i].cfg->integrator == RTG_INTEGRATOR_PATH && lanes; };
  auto paths_of = [&](uint32_t i) { return (uint64_t)(items[i].s1 - items[i].s0) * (items[i].px1 - items[i].px0); };
  // An item of at least `split` paths is cut into (at least) two chunks, one
  // per lane. An item followed by another shares the GPU with it anyway, so
  // only the last is cut from kTwoLanePaths on (measured on the torus pass
  // schedule: cutting the 16- or 4-spp passes too, 5.20 / 5.14 vs 5.24 G
  // rays/s).
  auto cap_of = [&](uint32_t i) {
    const uint64_t paths = paths_of(i);
    return frame_capacity(items[i].cfg, items[i].scene->dev.n_lights,
                          capacity_for(ctx, lanes_of(i) && paths >= (i + 1 < n ? chunk : split_last)
                                                ? (paths + 1) / 2
                                                : paths));
  };
  uint32_t cap32 = 0, cap64 = 0;
  uint64_t frames = 0;
  bool two = false;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t& c = items[i].cfg->precision == RTG_PRECISION_F64 ? cap64 : cap32;
    c = std::max(c, cap_of(i));
    frames = std::max(frames, chunk_extra_bytes(items[i].cfg, items[i].scene->dev.n_lights, cap_of(i)));
    two = two || (lanes_of(i) && (n > 1 || paths_of(i) >= split_last));
  }
  if (frames)
    if (rtg_status st = ensure_frames(ctx, frames)) return st;
  if (cap32) {
    if (rtg_status st = ensure_paths(ctx, 0, cap32, ctx->pb32)) return st;
    if (two)
      if (rtg_status st = ensure_paths(ctx, kLaneBuffers, cap32, ctx->pb32b)) return st;
  }
  if (cap64) {
    if (rtg_status st = ensure_paths(ctx, 1, cap64, ctx->pb64)) return st;
    if (two)
      if (rtg_status st = ensure_paths(ctx, kLane64Buffers, cap64, ctx->pb64b)) return st;
  }
  if (two) {  // the second lane starts after the main stream's work so far
    HIP_TRY(ctx, hipEventRecord(ctx->fork, ctx->stream));
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream2, ctx->fork, 0));
  }
  bool used[2] = {false, false};
  uint32_t lane = 0, done = 0;
  bool cancelled = false;
  for (uint32_t i = 0; i < n && !cancelled; ++i) {
    const PassItem& it = items[i];
    const bool f64 = it.cfg->precision == RTG_PRECISION_F64;
    const bool alt = two && lanes_of(i);
    const uint32_t k_total = it.s1 - it.s0, cap = cap_of(i);
    // Sample sub-ranges outermost keep every pixel's additions in sample order.
    for (uint32_t sa = it.s0; sa < it.s1 && !cancelled;) {
      const uint32_t k = std::min(k_total - (sa - it.s0), cap);
      const uint32_t pix_per_chunk = std::max(1u, cap / k);
      for (uint32_t p0 = it.px0; p0 < it.px1; p0 += pix_per_chunk) {
        if (!alt) lane = 0;
        if (cancel) {
          if (used[lane]) HIP_TRY(ctx, hipEventSynchronize(ctx->acc_done[lane]));
          if (*cancel) {
            cancelled = true;
            break;
          }
        }
        const bool lane2 = lane == 1;
        const PathBuffers<float>& B32 = lane2 ? ctx->pb32b : ctx->pb32;
        const PathBuffers<double>& B64 = lane2 ? ctx->pb64b : ctx->pb64;
        ChunkParams P = render_params(it.cfg, it.acc);
        P.p0 = p0;
        P.k = k;
        P.s0 = sa;
        P.n_pix = std::min(pix_per_chunk, it.px1 - p0);
        P.n_paths = P.n_pix * k;
        P.sub_cap = f64 ? B64.sub_cap : B32.sub_cap;
        P.shadow_counts = kMaxBounce * kSubQueues;
        P.stat_slots = kStatSlots;
        P.counts = f64 ? B64.counts : B32.counts;
        P.stats = it.stats;
        P.frames = ctx->frame_mem;
        set_direct(P, it.cfg, it.scene->dev.n_lights, cap);
        const hipStream_t st = lane2 ? ctx->stream2 : ctx->stream;
        const hipEvent_t wait = used[lane ^ 1] ? ctx->acc_done[lane ^ 1] : nullptr;
        const hipEvent_t rec = (two || cancel) ? ctx->acc_done[lane] : nullptr;
        if (f64)
          run_chunk(it.scene->dev, B64, P, it.scene->stack, st, side_of(ctx, lane2), hooks_of(ctx), wait, rec);
        else
          run_chunk(it.scene->dev, B32, P, it.scene->stack, st, side_of(ctx, lane2), hooks_of(ctx), wait, rec);
        HIP_TRY(ctx, hipGetLastError());
        used[lane] = true;
        if (alt) lane ^= 1;
      }
      sa += k;
    }
    if (!cancelled) ++done;
  }
  if (used[1]) {  // the main stream continues after the second lane's work
    HIP_TRY(ctx, hipEventRecord(ctx->join, ctx->stream2));
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->join, 0));
  }
  if (done_out) *done_out = done;
  return cancelled ? RTG_ERR_CANCELLED : RTG_OK;
}

// Unobserved passes of one call merged into one item unless RTG_MERGE_PASSES=0.
bool merge_passes_enabled() {
  const char* e = std::getenv("RTG_MERGE_PASSES");
  return !(e && e[0] == '0');
}

// Passes [bounds[i], bounds[i + 1]) of pixels [px0, px1), i < n_pass, of one
// render (run_items: passes side by side, bitwise the passes one by one).
rtg_status render_ranges(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* cfg,
                         const uint32_t* bounds, uint32_t n_pass, uint32_t px0, uint32_t px1, rtg_accum* acc,
                         const volatile int32_t* cancel, uint32_t* passes_done, rtg_stats* stats, const char* fn) {
  // Without a cancel flag nobody sees the passes apart: they run as one item
  // of all their samples, cut into full chunks (every pixel still adds its
  // samples in sample order, so the sums and counters are bitwise those of
  // the passes one by one). Small early passes no longer leave the GPU half
  // idle in their late bounces. With a cancel flag the passes stay items of
  // their own, so a cancelled call reports the passes it completed.
  const bool merge = cancel == nullptr && merge_passes_enabled();
  std::vector<PassItem> items;
  if (merge) items.push_back(PassItem{scene, cfg, acc, bounds[0], bounds[n_pass], px0, px1, ctx->stats});
  else
    for (uint32_t i = 0; i < n_pass; ++i)
      items.push_back(PassItem{scene, cfg, acc, bounds[i], bounds[i + 1], px0, px1, ctx->stats});
  HIP_TRY(ctx, hipMemsetAsync(ctx->stats, 0, sizeof(unsigned long long) * kStatSlots * kStatStride, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  uint32_t done = 0;
  const rtg_status rs = run_items(ctx, items.data(), (uint32_t)items.size(), cancel, &done);
  if (merge && rs == RTG_OK) done = n_pass;
  if (rs != RTG_OK && rs != RTG_ERR_CANCELLED) return rs;
  if (rtg_status st = stage_stats(ctx)) return st;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  read_stats(ctx, stats, ms);
  if (passes_done) *passes_done = done;
  if (rs == RTG_ERR_CANCELLED) return fail(ctx, RTG_ERR_CANCELLED, std::string(fn) + ": cancelled");
  return RTG_OK;
}

rtg_status render_range(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* cfg, uint32_t s0,
                        uint32_t s1, uint32_t px0, uint32_t px1, rtg_accum* acc, const volatile int32_t* cancel,
                        rtg_stats* stats, const char* fn) {
  const uint32_t bounds[2] = {s0, s1};
  return render_ranges(ctx, scene, cfg, bounds, 1, px0, px1, acc, cancel, nullptr, stats, fn);
}

// ---------------------------------------------------------------------------
// Batches
struct Segment {
  uint32_t job, p0, n_pix, s0, k;
};

// A group of at least this many paths runs as a wide batch pass (grids over
// every path, direct launches: the launches cost nothing next to the work,
// and the kernels keep the single-render kernels' registers); smaller
// groups run the persistent batch kernels, whose fixed grids replay from a
// captured graph.
constexpr uint64_t kDefaultWidePaths = 1ull << 22;
uint64_t wide_paths_of(const rtg_context* ctx) {
  return ctx->tuning.wide_paths ? ctx->tuning.wide_paths : kDefaultWidePaths;
}

bool batchable(const rtg_job& j) {
  return j.settings.precision == RTG_PRECISION_F32 && j.settings.integrator == RTG_INTEGRATOR_PATH &&
         j.settings.max_depth <= kBatchDepth;
}

// Batch passes replay as captured graphs unless RTG_BATCH_GRAPHS=0 (direct
// launches; also used while per-launch profiling is on).
bool batch_graphs_enabled() {
  const char* e = std::getenv("RTG_BATCH_GRAPHS");
  return !(e && e[0] == '0');
}

constexpr size_t kPlanWords = 1 + 5 * (kMaxBatchJobs + 1);
// Host staging of one group: its job table, then the job count and the
// generate / accumulate plans.
constexpr size_t kStageBytes = sizeof(BatchJob) * kMaxBatchJobs + sizeof(uint32_t) * (1 + 2 * (kMaxBatchJobs + 1));
constexpr size_t kJobStatWords = (size_t)kBatchStatSlots * kStatStride;

rtg_status ensure_batch_tables(rtg_context* ctx, size_t groups, size_t segments) {
  for (int l = 0; l < 2; ++l)
    if (!ctx->batch_jobs[l]) {
      HIP_TRY(ctx, hipMalloc(&ctx->batch_jobs[l], sizeof(BatchJob) * kMaxBatchJobs));
      HIP_TRY(ctx, hipMalloc(&ctx->batch_plans[l], sizeof(uint32_t) * kPlanWords));
    }
  // Previous calls have completed (each ends with a synchronisation), so
  // the pinned buffers are free to be replaced.
  if (ctx->batch_stage_cap < groups) {
    if (ctx->batch_stage) HIP_TRY(ctx, hipHostFree(ctx->batch_stage));
    ctx->batch_stage = nullptr;
    ctx->batch_stage_cap = 0;
    HIP_TRY(ctx, hipHostMalloc(&ctx->batch_stage, kStageBytes * groups));
    ctx->batch_stage_cap = groups;
  }
  if (ctx->batch_stats_cap < segments) {
    if (ctx->batch_stats_host) HIP_TRY(ctx, hipHostFree(ctx->batch_stats_host));
    ctx->batch_stats_host = nullptr;
    ctx->batch_stats_cap = 0;
    HIP_TRY(ctx, hipHostMalloc(&ctx->batch_stats_host, sizeof(unsigned long long) * kJobStatWords * segments));
    ctx->batch_stats_cap = segments;
  }
  return RTG_OK;
}

int stack_class(uint32_t stack) { return stack_class_of(stack); }

// One lane of batch groups: a main stream, its side stream, path buffers
// and device job tables.
struct BatchLane {
  hipStream_t stream;
  const SideStream* side;
  const PathBuffers<float>* B;
  BatchJob* jobs;
  uint32_t* plans;
  uint32_t cap;
};

// Enqueues one batched pass over segments seg[0..n) (n <= kMaxBatchJobs,
// total paths <= the lane's capacity) on `lane`, staging its tables in
// `stage` and copying its statistics slots to `stats_out` (pinned; read
// after the call's final synchronisation). The accumulation waits for
// `acc_wait` and records `acc_done` (if not null): the other lane's groups
// may write the same pixels in an earlier sample range.
rtg_status run_group(rtg_context* ctx, const rtg_job* jobs, const Segment* seg, uint32_t n, const BatchLane& lane,
                     bool first_lane, void* stage, unsigned long long* stats_out, hipEvent_t acc_wait,
                     hipEvent_t acc_done) {
  auto* hj = static_cast<BatchJob*>(stage);
  auto* hn = reinterpret_cast<uint32_t*>(hj + kMaxBatchJobs);
  uint32_t* hgen = hn + 1;
  uint32_t* hacc = hgen + (kMaxBatchJobs + 1);
  hn[0] = n;
  const PathBuffers<float>& B = *lane.B;
  uint64_t pid = 0, q = 0;
  uint32_t gen_blocks = 0, acc_blocks = 0, path_blocks = 0, stack = 1;
  int depth = 1;
  for (uint32_t i = 0; i < n; ++i) {
    const rtg_job& J = jobs[seg[i].job];
    ChunkParams P = render_params(&J.settings, J.acc);
    P.p0 = seg[i].p0;
    P.k = seg[i].k;
    P.s0 = seg[i].s0;
    P.n_pix = seg[i].n_pix;
    P.n_paths = P.n_pix * P.k;
    P.pid_base = (uint32_t)pid;
    P.q_base = (uint32_t)q;
    P.sub_cap = sub_queue_cap(P.n_paths);
    P.shadow_counts = kBatchShadowCounts;
    P.stat_slots = kBatchStatSlots;
    P.zero_stats = 1;
    P.counts = B.counts + (size_t)i * kBatchCounts * kCountStride;
    P.stats = B.stats + (size_t)i * kJobStatWords;
    pid += P.n_paths;
    q += (uint64_t)P.sub_cap * kSubQueues;
    hj[i].S = J.scene->dev;
    hj[i].P = P;
    hgen[i] = gen_blocks;
    gen_blocks += (generate_threads(P) + kBlock - 1) / kBlock;
    hacc[i] = acc_blocks;
    acc_blocks += accumulate_blocks<float>(P);
    path_blocks += (P.n_paths + kBlock - 1) / kBlock;
    depth = std::max(depth, P.max_depth);
    stack = std::max(stack, J.scene->stack);
  }
  hgen[n] = gen_blocks;
  hacc[n] = acc_blocks;
  if (pid > lane.cap || q > (uint64_t)lane.cap + kBatchQueueSlack)
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_render_batch: internal error: group exceeds the batch buffers");
  HIP_TRY(ctx, hipMemcpyAsync(lane.jobs, hj, sizeof(BatchJob) * n, hipMemcpyHostToDevice, lane.stream));
  HIP_TRY(ctx, hipMemcpyAsync(lane.plans, hn, sizeof(uint32_t) * (1 + 2 * (kMaxBatchJobs + 1)), hipMemcpyHostToDevice,
                              lane.stream));
  constexpr uint32_t kP = kMaxBatchJobs + 1;
  uint32_t* plans = lane.plans;
  const BatchPlans L{plans, plans + 1, plans + 1 + kP, plans + 1 + 2 * kP, plans + 1 + 3 * kP};
  if (pid >= wide_paths_of(ctx)) {
    BatchGrids g;
    g.generate = std::max(1u, gen_blocks);
    g.trace = g.shade = std::max(1u, path_blocks);
    g.accumulate = std::max(1u, acc_blocks);
    run_wide(lane.jobs, B, L, depth, stack, g, lane.stream, lane.side, hooks_of(ctx), acc_wait, acc_done);
    HIP_TRY(ctx, hipGetLastError());
  } else {
    const int cls = stack_class(stack);
    uint32_t* occ = ctx->batch_occ[cls];
    if (!occ[0]) batch_occupancy(stack, &occ[0], &occ[1], &occ[2]);
    const uint32_t cu = (uint32_t)std::max(1, ctx->n_cu);
    // Persistent grids: resident workgroups walk the virtual blocks. A direct
    // launch trims them to the work; a graph keeps them fixed (its launches
    // replay with the arguments they were captured with) and the surplus
    // workgroups find no block and exit. Graphs hold the first lane's buffers.
    const bool graph = first_lane && !ctx->profiling && batch_graphs_enabled();
    BatchGrids g;
    g.generate = std::max(1u, graph ? occ[2] * cu : std::min(gen_blocks, occ[2] * cu));
    g.trace = std::max(1u, graph ? occ[0] * cu : std::min(path_blocks, occ[0] * cu));
    g.shade = std::max(1u, graph ? occ[1] * cu : std::min(path_blocks, occ[1] * cu));
    g.accumulate = std::max(1u, graph ? occ[2] * cu : std::min(acc_blocks, occ[2] * cu));
    if (acc_wait) HIP_TRY(ctx, hipStreamWaitEvent(lane.stream, acc_wait, 0));
    if (graph) {
      hipGraphExec_t exec = nullptr;
      for (const auto& e : ctx->batch_graphs)
        if (e.depth == depth && e.cls == cls && e.side == (lane.side != nullptr)) exec = e.exec;
      if (!exec) {
        hipGraph_t gr = nullptr;
        HIP_TRY(ctx, hipStreamBeginCapture(lane.stream, hipStreamCaptureModeThreadLocal));
        run_batch(lane.jobs, B, L, depth, stack, g, lane.stream, lane.side, nullptr);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(lane.stream, &gr);
        hipError_t ie = hipSuccess;
        if (le == hipSuccess && ce == hipSuccess) ie = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        if (gr) hipGraphDestroy(gr);
        for (hipError_t e : {le, ce, ie})
          if (e != hipSuccess)
            return fail(ctx, RTG_ERR_HIP,
                        std::string("rtg_render_batch: capturing a batch pass: ") + hipGetErrorString(e));
        ctx->batch_graphs.push_back({depth, cls, lane.side != nullptr, exec});
      }
      HIP_TRY(ctx, hipGraphLaunch(exec, lane.stream));
    } else {
      run_batch(lane.jobs, B, L, depth, stack, g, lane.stream, lane.side, hooks_of(ctx));
      HIP_TRY(ctx, hipGetLastError());
    }
    if (acc_done) HIP_TRY(ctx, hipEventRecord(acc_done, lane.stream));
  }
  HIP_TRY(ctx, hipMemcpyAsync(stats_out, B.stats, sizeof(unsigned long long) * kJobStatWords * n,
                              hipMemcpyDeviceToHost, lane.stream));
  return RTG_OK;
}

}  // namespace

extern "C" {

rtg_status rtg_render_pass(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* cfg, uint32_t s0,
                           uint32_t s1, rtg_accum* acc, const volatile int32_t* cancel, rtg_stats* stats) {
  if (!ctx || !scene || !cfg || !acc) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_render_pass: NULL argument");
  if (rtg_status st = check_settings(ctx, "rtg_render_pass", cfg, acc, s0, s1)) return st;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return render_range(ctx, scene, cfg, s0, s1, 0, (uint32_t)acc->w * (uint32_t)acc->h, acc, cancel, stats,
                      "rtg_render_pass");
}

rtg_status rtg_render_passes(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* cfg,
                             const uint32_t* bounds, uint32_t n_passes, rtg_accum* acc, const volatile int32_t* cancel,
                             uint32_t* passes_done, rtg_stats* stats) {
  if (passes_done) *passes_done = 0;
  if (!ctx || !scene || !cfg || !acc || (n_passes && !bounds))
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_render_passes: NULL argument");
  if (n_passes == 0) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_render_passes: no passes");
  for (uint32_t i = 0; i < n_passes; ++i)
    if (rtg_status st = check_settings(ctx, "rtg_render_passes", cfg, acc, bounds[i], bounds[i + 1])) return st;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return render_ranges(ctx, scene, cfg, bounds, n_passes, 0, (uint32_t)acc->w * (uint32_t)acc->h, acc, cancel,
                       passes_done, stats, "rtg_render_passes");
}

rtg_status rtg_render_pass_static(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* cfg,
                                  uint32_t s0, uint32_t s1, rtg_accum* acc, const volatile int32_t* cancel,
                                  rtg_stats* stats) {
  static const char* const name = "rtg_render_pass_static";
  auto fn = [&]() { return std::string(name); };  // built only for an error message
  if (!ctx || !scene || !cfg || !acc) return fail(ctx, RTG_ERR_INVALID_ARG, fn() + ": NULL argument");
  if (rtg_status st = check_settings(ctx, name, cfg, acc, s0, s1)) return st;
  if (!ctx->scratch_dev) return fail(ctx, RTG_ERR_HIP, fn() + ": the device scratch array was not set up");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const bool f64 = cfg->precision == RTG_PRECISION_F64;
  const uint32_t n_pix = (uint32_t)acc->w * (uint32_t)acc->h;
  const uint64_t paths = (uint64_t)(s1 - s0) * n_pix;
  // The scratch holds the statistics slots, then one chunk's path state.
  const uint32_t n_slots = kStatSlots * kStatStride;
  const uint64_t slot_bytes = ((uint64_t)n_slots * 8 + 255) & ~uint64_t(255);
  unsigned long long* slots = reinterpret_cast<unsigned long long*>(ctx->scratch_dev);
  char* base = ctx->scratch_dev + slot_bytes;
  const uint32_t n_lights = scene->dev.n_lights;
  auto extra = [&](uint32_t c) { return chunk_extra_bytes(cfg, n_lights, c); };
  const uint32_t cap = std::min(capacity_for(ctx, paths), static_capacity(f64, kStaticScratchBytes - slot_bytes, extra));
  if (cap == 0)
    return fail(ctx, RTG_ERR_UNSUPPORTED,
                fn() + ": the scratch array holds no chunk of 128 paths at this max_depth (" +
                    std::to_string(cfg->max_depth) + ")");
  PathBuffers<float> B32{};
  PathBuffers<double> B64{};
  const uint64_t need = f64 ? lay_out_paths<double>(base, cap, false, slots, &B64)
                            : lay_out_paths<float>(base, cap, false, slots, &B32);
  const uint64_t extra_bytes = extra(cap);
  if (slot_bytes + ((need + 255) & ~uint64_t(255)) + extra_bytes > kStaticScratchBytes)
    return fail(ctx, RTG_ERR_OUT_OF_MEMORY, fn() + ": chunk state exceeds the scratch array");
  char* frames = base + ((need + 255) & ~uint64_t(255));
  hipLaunchKernelGGL(k_zero_u64, dim3((n_slots + 127) / 128), dim3(128), 0, ctx->stream, slots, n_slots);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  // One lane: chunks of whole pixels x a sample sub-range, in sample order
  // per pixel (the chunking of rtg_render_pass, bitwise the same result).
  bool cancelled = false, used = false;
  for (uint32_t sa = s0; sa < s1 && !cancelled;) {
    const uint32_t k = std::min(s1 - sa, cap);
    const uint32_t pix_per_chunk = std::max(1u, cap / k);
    for (uint32_t p0 = 0; p0 < n_pix; p0 += pix_per_chunk) {
      if (cancel) {
        if (used) HIP_TRY(ctx, hipEventSynchronize(ctx->acc_done[0]));
        if (*cancel) {
          cancelled = true;
          break;
        }
      }
      ChunkParams P = render_params(cfg, acc);
      P.p0 = p0;
      P.k = k;
      P.s0 = sa;
      P.n_pix = std::min(pix_per_chunk, n_pix - p0);
      P.n_paths = P.n_pix * k;
      P.sub_cap = f64 ? B64.sub_cap : B32.sub_cap;
      P.shadow_counts = kMaxBounce * kSubQueues;
      P.stat_slots = kStatSlots;
      P.counts = f64 ? B64.counts : B32.counts;
      P.stats = slots;
      P.frames = extra_bytes ? frames : nullptr;
      set_direct(P, cfg, n_lights, cap);
      const hipEvent_t rec = cancel ? ctx->acc_done[0] : nullptr;
      if (f64)
        run_chunk(scene->dev, B64, P, scene->stack, ctx->stream, side_of(ctx), nullptr, nullptr, rec);
      else
        run_chunk(scene->dev, B32, P, scene->stack, ctx->stream, side_of(ctx), nullptr, nullptr, rec);
      HIP_TRY(ctx, hipGetLastError());
      used = true;
    }
    sa += k;
  }
  if (!stats && !cancel) return RTG_OK;  // enqueued on the context's stream
  HIP_TRY(ctx, hipMemcpyAsync(ctx->stats_host, slots, sizeof(unsigned long long) * n_slots, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  read_stats(ctx, stats, ms);
  if (cancelled) return fail(ctx, RTG_ERR_CANCELLED, fn() + ": cancelled");
  return RTG_OK;
}

rtg_status rtg_render_batch(rtg_context* ctx, const rtg_job* jobs, uint32_t n_jobs, rtg_stats* stats) {
  if (!ctx || (n_jobs && !jobs)) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_render_batch: NULL argument");
  const std::string fn = "rtg_render_batch";
  std::vector<uint32_t> p0(n_jobs), p1(n_jobs);
  for (uint32_t i = 0; i < n_jobs; ++i) {
    const rtg_job& J = jobs[i];
    const std::string fi = fn + ": job " + std::to_string(i);
    if (!J.scene || !J.acc) return fail(ctx, RTG_ERR_INVALID_ARG, fi + ": NULL scene or accumulation buffer");
    if (rtg_status st = check_settings(ctx, fi.c_str(), &J.settings, J.acc, J.s0, J.s1)) return st;
    const uint32_t n_pix = (uint32_t)J.acc->w * (uint32_t)J.acc->h;
    p0[i] = J.p1 == 0 ? 0 : J.p0;
    p1[i] = J.p1 == 0 ? n_pix : J.p1;
    if (p0[i] >= p1[i] || p1[i] > n_pix) return fail(ctx, RTG_ERR_INVALID_ARG, fi + ": bad pixel range");
  }
  for (uint32_t i = 0; i < n_jobs; ++i)
    for (uint32_t j = i + 1; j < n_jobs; ++j)
      if (jobs[i].acc == jobs[j].acc && p0[i] < p1[j] && p0[j] < p1[i])
        return fail(ctx, RTG_ERR_INVALID_ARG,
                    fn + ": jobs " + std::to_string(i) + " and " + std::to_string(j) +
                        " write overlapping pixels of one accumulation buffer");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<rtg_stats> acc_stats(n_jobs, rtg_stats{});
  // Jobs outside the batch kernels' scope (FP64, direct / normals, deeper
  // than kBatchDepth) through the single-render kernels: one after another
  // (run_items, their chunks alternating between the lanes), each with its
  // own statistics slots. Every FP32 path-integrator job, whatever its
  // size, shares the batch launches. Jobs write different accumulation
  // buffers or disjoint pixels, so their order does not matter.
  uint64_t batch_paths = 0;
  std::vector<uint8_t> in_batch(n_jobs, 0);
  std::vector<uint32_t> solo;
  for (uint32_t i = 0; i < n_jobs; ++i) {
    if (batchable(jobs[i])) {
      in_batch[i] = 1;
      batch_paths += (uint64_t)(jobs[i].s1 - jobs[i].s0) * (p1[i] - p0[i]);
    } else {
      solo.push_back(i);
    }
  }
  const size_t region = (size_t)kStatSlots * kStatStride;
  if (!solo.empty()) {
    const uint32_t ns = (uint32_t)solo.size();
    if (ctx->job_stats_cap < ns) {
      if (ctx->job_stats) HIP_TRY(ctx, hipFree(ctx->job_stats));
      if (ctx->job_stats_host) HIP_TRY(ctx, hipHostFree(ctx->job_stats_host));
      ctx->job_stats = nullptr;
      ctx->job_stats_host = nullptr;
      ctx->job_stats_cap = 0;
      HIP_TRY(ctx, hipMalloc(&ctx->job_stats, sizeof(unsigned long long) * region * ns));
      HIP_TRY(ctx, hipHostMalloc(&ctx->job_stats_host, sizeof(unsigned long long) * region * ns));
      ctx->job_stats_cap = ns;
    }
    std::vector<PassItem> items(ns);
    for (uint32_t k = 0; k < ns; ++k) {
      const rtg_job& J = jobs[solo[k]];
      items[k] = PassItem{J.scene, &J.settings, J.acc, J.s0, J.s1, p0[solo[k]], p1[solo[k]], ctx->job_stats + k * region};
    }
    HIP_TRY(ctx, hipMemsetAsync(ctx->job_stats, 0, sizeof(unsigned long long) * region * ns, ctx->stream));
    if (rtg_status st = run_items(ctx, items.data(), ns, nullptr, nullptr)) return st;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->job_stats_host, ctx->job_stats, sizeof(unsigned long long) * region * ns,
                                hipMemcpyDeviceToHost, ctx->stream));
  }
  // Segments of at most `cap` paths (whole pixels x a sample sub-range),
  // packed in job order into groups of at most `cap` paths and
  // kMaxBatchJobs segments.
  std::vector<Segment> segs;
  std::vector<std::pair<size_t, size_t>> groups;
  const uint32_t cap = batch_paths ? capacity_for(ctx, batch_paths) : 0;
  if (batch_paths) {
    for (uint32_t i = 0; i < n_jobs; ++i) {
      if (!in_batch[i]) continue;
      const uint32_t k_total = jobs[i].s1 - jobs[i].s0;
      for (uint32_t sa = jobs[i].s0; sa < jobs[i].s1;) {
        const uint32_t k = std::min(k_total - (sa - jobs[i].s0), cap);
        const uint32_t per = std::max(1u, cap / k);
        for (uint32_t p = p0[i]; p < p1[i]; p += per) segs.push_back(Segment{i, p, std::min(per, p1[i] - p), sa, k});
        sa += k;
      }
    }
    for (size_t a = 0; a < segs.size();) {
      size_t b = a;
      uint64_t paths = 0;
      while (b < segs.size() && b - a < (size_t)kMaxBatchJobs && paths + (uint64_t)segs[b].n_pix * segs[b].k <= cap) {
        paths += (uint64_t)segs[b].n_pix * segs[b].k;
        ++b;
      }
      groups.emplace_back(a, b);
      a = b;
    }
    // Groups alternate between two lanes (streams and path buffers), so one
    // group's kernels fill the other's tails; a group's accumulation waits
    // for the other lane's latest, so every pixel adds its sample ranges in
    // order.
    const bool two = groups.size() > 1 && two_lanes_enabled(ctx);
    if (rtg_status st = ensure_paths(ctx, kBatchBuffers, cap, ctx->pbb)) return st;
    if (two)
      if (rtg_status st = ensure_paths(ctx, kBatchBuffers2, cap, ctx->pbb2)) return st;
    if (rtg_status st = ensure_batch_tables(ctx, groups.size(), segs.size())) return st;
    const BatchLane lanes[2] = {
        {ctx->stream, side_of(ctx), &ctx->pbb, ctx->batch_jobs[0], ctx->batch_plans[0], ctx->cap[kBatchBuffers]},
        {ctx->stream2, side_of(ctx, true), &ctx->pbb2, ctx->batch_jobs[1], ctx->batch_plans[1],
         ctx->cap[kBatchBuffers2]}};
    if (two) {
      HIP_TRY(ctx, hipEventRecord(ctx->fork, ctx->stream));
      HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream2, ctx->fork, 0));
    }
    bool used[2] = {false, false};
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const int l = two ? (int)(gi & 1) : 0;
      const hipEvent_t wait = (two && used[l ^ 1]) ? ctx->acc_done[l ^ 1] : nullptr;
      const hipEvent_t rec = two ? ctx->acc_done[l] : nullptr;
      char* stage = static_cast<char*>(ctx->batch_stage) + gi * kStageBytes;
      unsigned long long* out = ctx->batch_stats_host + groups[gi].first * kJobStatWords;
      if (rtg_status st = run_group(ctx, jobs, segs.data() + groups[gi].first,
                                    (uint32_t)(groups[gi].second - groups[gi].first), lanes[l], l == 0, stage, out,
                                    wait, rec))
        return st;
      used[l] = true;
    }
    if (used[1]) {
      HIP_TRY(ctx, hipEventRecord(ctx->join, ctx->stream2));
      HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->join, 0));
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  unsigned long long tot[kStCount] = {0};
  for (size_t k = 0; k < solo.size(); ++k) {
    unsigned long long h[kStCount] = {0};
    for (int sl = 0; sl < kStatSlots; ++sl)
      for (int c = 0; c < kStCount; ++c) h[c] += ctx->job_stats_host[k * region + (size_t)sl * kStatStride + c];
    add_counters(&acc_stats[solo[k]], h);
    for (int c = 0; c < kStCount; ++c) tot[c] += h[c];
  }
  for (size_t k = 0; k < segs.size(); ++k) {
    unsigned long long h[kStCount] = {0};
    for (uint32_t sl = 0; sl < kBatchStatSlots; ++sl)
      for (int c = 0; c < kStCount; ++c) h[c] += ctx->batch_stats_host[(k * kBatchStatSlots + sl) * kStatStride + c];
    add_counters(&acc_stats[segs[k].job], h);
    for (int c = 0; c < kStCount; ++c) tot[c] += h[c];
  }
  profile_resolve(ctx, tot);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats)
    for (uint32_t i = 0; i < n_jobs; ++i) {
      stats[i] = acc_stats[i];
      stats[i].elapsed_ms = ms;
    }
  return RTG_OK;
}

// ---------------------------------------------------------------------------
// Scheduling knobs and the autotuner.
rtg_status rtg_get_tuning(rtg_context* ctx, rtg_tuning* out) {
  if (!ctx || !out) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_get_tuning: NULL argument");
  *out = rtg_tuning{};
  out->chunk_paths = chunk_paths_of(ctx);
  out->wide_paths = (uint32_t)std::min<uint64_t>(wide_paths_of(ctx), 0xFFFFFFFFu);
  out->lanes = two_lanes_enabled(ctx) ? 2 : 1;
  return RTG_OK;
}

rtg_status rtg_set_tuning(rtg_context* ctx, const rtg_tuning* t) {
  if (!ctx || !t) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_set_tuning: NULL argument");
  if (t->chunk_paths && (t->chunk_paths < 1024 || t->chunk_paths > (1u << 28)))
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_set_tuning: chunk_paths must be in [1024, 2^28]");
  if (t->lanes > 2) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_set_tuning: lanes must be 1 or 2");
  ctx->tuning = *t;
  ctx->tuning.pad = 0;
  return RTG_OK;
}

}  // extern "C"

namespace {

// One cache line: {"device":"<name>","cus":N,"key":"<key>","chunk_paths":N,
// "wide_paths":N,"lanes":N,"default_ms":X,"best_ms":X}.
std::string tune_line(const std::string& dev, int cus, const std::string& key, const rtg_tune_result& r) {
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "{\"device\":\"%s\",\"cus\":%d,\"key\":\"%s\",\"chunk_paths\":%u,\"wide_paths\":%u,\"lanes\":%u,"
                "\"default_ms\":%.3f,\"best_ms\":%.3f}",
                dev.c_str(), cus, key.c_str(), r.best.chunk_paths, r.best.wide_paths, r.best.lanes, r.default_ms,
                r.best_ms);
  return buf;
}

bool tune_field(const std::string& line, const char* name, double* v) {
  const std::string tag = std::string("\"") + name + "\":";
  const size_t at = line.find(tag);
  if (at == std::string::npos) return false;
  char* end = nullptr;
  *v = std::strtod(line.c_str() + at + tag.size(), &end);
  return end != line.c_str() + at + tag.size();
}

// The cached result for (device, cus, key), if any.
bool tune_lookup(const char* path, const std::string& dev, int cus, const std::string& key, rtg_tune_result* r) {
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  const std::string want = "{\"device\":\"" + dev + "\",\"cus\":" + std::to_string(cus) + ",\"key\":\"" + key + "\",";
  char buf[1024];
  bool found = false;
  while (std::fgets(buf, sizeof buf, f)) {
    const std::string line(buf);
    if (line.compare(0, want.size(), want) != 0) continue;
    double c, w, l, d, b;
    if (tune_field(line, "chunk_paths", &c) && tune_field(line, "wide_paths", &w) && tune_field(line, "lanes", &l) &&
        tune_field(line, "default_ms", &d) && tune_field(line, "best_ms", &b)) {
      *r = rtg_tune_result{};
      r->best.chunk_paths = (uint32_t)c;
      r->best.wide_paths = (uint32_t)w;
      r->best.lanes = (uint32_t)l;
      r->default_ms = d;
      r->best_ms = b;
      found = true;  // the last entry wins
    }
  }
  std::fclose(f);
  return found;
}

}  // namespace

extern "C" {

rtg_status rtg_autotune(rtg_context* ctx, const rtg_job* jobs, const uint32_t* call_jobs, uint32_t n_calls,
                        const char* cache_path, const char* key, rtg_tune_result* out) {
  if (!ctx || !jobs || !call_jobs || !n_calls || !out) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_autotune: NULL argument");
  const std::string k = key ? key : "";
  if (k.find_first_of("\"\\\n") != std::string::npos)
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_autotune: the key must not hold quotes, backslashes or newlines");
  hipDeviceProp_t prop;
  HIP_TRY(ctx, hipGetDeviceProperties(&prop, ctx->device));
  const std::string dev = prop.gcnArchName;
  if (cache_path && tune_lookup(cache_path, dev, prop.multiProcessorCount, k, out)) {
    if (rtg_status st = rtg_set_tuning(ctx, &out->best)) return st;
    return RTG_OK;
  }
  std::vector<rtg_accum*> accs;
  uint64_t total = 0;
  for (uint32_t c = 0; c < n_calls; ++c) total += call_jobs[c];
  for (uint64_t i = 0; i < total; ++i)
    if (std::find(accs.begin(), accs.end(), jobs[i].acc) == accs.end()) accs.push_back(jobs[i].acc);
  // Wall time of the whole workload under `t`.
  auto time_with = [&](const rtg_tuning& t, double* ms) -> rtg_status {
    if (rtg_status st = rtg_set_tuning(ctx, &t)) return st;
    for (rtg_accum* a : accs)
      if (rtg_status st = rtg_accum_clear(ctx, a)) return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t at = 0;
    for (uint32_t c = 0; c < n_calls; ++c) {
      if (rtg_status st = rtg_render_batch(ctx, jobs + at, call_jobs[c], nullptr)) return st;
      at += call_jobs[c];
    }
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RTG_OK;
  };
  const rtg_tuning saved = ctx->tuning;
  rtg_tuning best{};
  double best_ms = 0, warm = 0;
  uint32_t trials = 0;
  rtg_status st = time_with(best, &warm);  // grows the buffers; not compared
  if (st == RTG_OK) st = time_with(best, &best_ms);
  const double default_ms = best_ms;
  // Coordinate descent: each knob in turn over its candidates, the others
  // held at the best found so far.
  const uint32_t chunks[] = {1u << 23, 1u << 24, 1u << 25, 1u << 26};
  const uint32_t wides[] = {1u << 20, 1u << 22, 1u << 24, 0xFFFFFFFFu};
  const uint32_t lanes[] = {1, 2};
  for (int knob = 0; knob < 3 && st == RTG_OK; ++knob) {
    const uint32_t* cand = knob == 0 ? chunks : knob == 1 ? wides : lanes;
    const int n = knob == 2 ? 2 : 4;
    for (int c = 0; c < n && st == RTG_OK; ++c) {
      rtg_tuning t = best;
      (knob == 0 ? t.chunk_paths : knob == 1 ? t.wide_paths : t.lanes) = cand[c];
      const uint32_t effective = knob == 0 ? (best.chunk_paths ? best.chunk_paths : kDefaultChunkPaths)
                                 : knob == 1 ? (best.wide_paths ? best.wide_paths : (uint32_t)kDefaultWidePaths)
                                             : (best.lanes ? best.lanes : 2u);
      if (cand[c] == effective) continue;  // the incumbent, already timed
      double ms = 0;
      st = time_with(t, &ms);
      ++trials;
      if (st == RTG_OK && ms < best_ms) {
        best = t;
        best_ms = ms;
      }
    }
  }
  for (rtg_accum* a : accs) rtg_accum_clear(ctx, a);
  if (st != RTG_OK) {
    ctx->tuning = saved;
    return st;
  }
  // Re-time the default and the winner side by side for the report.
  double d2 = 0, b2 = 0;
  if ((st = time_with(rtg_tuning{}, &d2)) != RTG_OK || (st = time_with(best, &b2)) != RTG_OK) {
    ctx->tuning = saved;
    return st;
  }
  for (rtg_accum* a : accs) rtg_accum_clear(ctx, a);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *out = rtg_tune_result{};
  rtg_set_tuning(ctx, &best);
  rtg_get_tuning(ctx, &out->best);
  out->default_ms = std::min(default_ms, d2);
  out->best_ms = std::min(best_ms, b2);
  out->trials = trials;
  if (cache_path) {
    if (FILE* f = std::fopen(cache_path, "a")) {
      std::fprintf(f, "%s\n", tune_line(dev, prop.multiProcessorCount, k, *out).c_str());
      std::fclose(f);
    }
  }
  return RTG_OK;
}

// ---------------------------------------------------------------------------
rtg_status rtg_buffer_create(rtg_context* ctx, uint64_t bytes, rtg_buffer** out) {
  if (!ctx || !out || bytes == 0) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_buffer_create: bad arguments");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  void* p = nullptr;
  HIP_TRY(ctx, hipMalloc(&p, bytes));
  auto* b = new rtg_buffer;
  b->ctx = ctx;
  b->ptr = p;
  b->bytes = bytes;
  *out = b;
  return RTG_OK;
}

rtg_status rtg_buffer_release(rtg_buffer* b) {
  if (!b) return RTG_OK;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  hipFree(b->ptr);
  delete b;
  return RTG_OK;
}

rtg_status rtg_buffer_write(rtg_context* ctx, rtg_buffer* b, const void* src, uint64_t bytes) {
  if (!ctx || !b || !src || bytes > b->bytes) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_buffer_write: bad arguments");
  HIP_TRY(ctx, hipMemcpyAsync(b->ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return RTG_OK;
}

rtg_status rtg_buffer_read(rtg_context* ctx, const rtg_buffer* b, void* dst, uint64_t bytes) {
  if (!ctx || !b || !dst || bytes > b->bytes) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_buffer_read: bad arguments");
  HIP_TRY(ctx, hipMemcpyAsync(dst, b->ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return RTG_OK;
}

rtg_status rtg_trace(rtg_context* ctx, const rtg_scene* scene, const rtg_buffer* rays, uint64_t n,
                     const rtg_trace_opts* opts, rtg_buffer* hits, rtg_stats* stats) {
  if (!ctx || !scene || !rays || !hits || !opts) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_trace: NULL argument");
  if (n * 64 > rays->bytes || n * sizeof(rtg_hit) > hits->bytes)
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_trace: buffers too small for n rays");
  if (opts->mode < 0 || opts->mode > 1 || opts->precision < 0 || opts->precision > 1)
    return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_trace: bad mode or precision");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemsetAsync(ctx->stats, 0, sizeof(unsigned long long) * kStatSlots * kStatStride, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (n) {
    if (opts->precision == RTG_PRECISION_F64)
      run_trace64(scene->dev, static_cast<const double*>(rays->ptr), n, opts->mode, opts->use_bvh,
                  static_cast<rtg_hit*>(hits->ptr), ctx->stats, scene->stack, ctx->stream, hooks_of(ctx));
    else
      run_trace32(scene->dev, static_cast<const double*>(rays->ptr), n, opts->mode, opts->use_bvh,
                  static_cast<rtg_hit*>(hits->ptr), ctx->stats, scene->stack, ctx->stream, hooks_of(ctx));
    HIP_TRY(ctx, hipGetLastError());
  }
  if (rtg_status st = stage_stats(ctx)) return st;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  read_stats(ctx, stats, ms);
  return RTG_OK;
}

rtg_status rtg_trace_host(rtg_context* ctx, const rtg_scene* scene, const double* rays, uint64_t n,
                          const rtg_trace_opts* opts, rtg_hit* hits, rtg_stats* stats) {
  if (!ctx || !scene || (n && (!rays || !hits))) return fail(ctx, RTG_ERR_INVALID_ARG, "rtg_trace_host: NULL argument");
  if (n == 0) {
    if (stats) *stats = rtg_stats{};
    return RTG_OK;
  }
  rtg_buffer *rb = nullptr, *hb = nullptr;
  rtg_status st = rtg_buffer_create(ctx, n * 64, &rb);
  if (st == RTG_OK) st = rtg_buffer_create(ctx, n * sizeof(rtg_hit), &hb);
  if (st == RTG_OK) st = rtg_buffer_write(ctx, rb, rays, n * 64);
  if (st == RTG_OK) st = rtg_trace(ctx, scene, rb, n, opts, hb, stats);
  if (st == RTG_OK) st = rtg_buffer_read(ctx, hb, hits, n * sizeof(rtg_hit));
  rtg_buffer_release(rb);
  rtg_buffer_release(hb);
  return st;
}

}  // extern "C"
