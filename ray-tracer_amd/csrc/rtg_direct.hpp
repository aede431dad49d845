// rtg_direct.hpp — the direct integrator (integrator.cpp:141-174) in FP32 as
// a weighted ray queue.
//
// direct_trace is a tree walk. A ray's radiance is its hit's emission (the
// sky on a miss), plus at a diffuse hit one estimate per light, plus at a
// metal hit albedo x the radiance of its reflection, plus at a dielectric hit
// tint x F x (reflection) + tint x (1 - F) x (refraction), down to max_depth.
// Expanded, a sample's radiance is the sum over the tree's nodes of (the
// product of the weights from the camera ray down to the node) x (the node's
// own terms). It is evaluated here breadth first, like the path integrator's
// wavefront: generation g holds the rays at depth g, each with its weight.
// Generation 0 — the camera rays, coherent, one per sample — runs as one
// kernel that also traces its shadow rays (k_dfirst); per later generation
//   extend  closest hit of every ray                                (k_dextend)
//   shade   the node's emission / sky term; at a diffuse hit one shadow
//           request per light (the hit joins the lit queue); the children
//           (weight x albedo, or the two Fresnel-weighted rays) appended to
//           generation g + 1                                       (k_dshade)
//   shadow  per lit hit, any-hit of its requests; its own term plus the
//           unoccluded lights' terms go to its sample               (k_dshadow)
//
// A sample's terms come from several threads when its tree branches, in no
// fixed order, so they are summed as 64-bit fixed-point integers with 32
// fraction bits, whose sum does not depend on the order of the additions (the
// accumulation of SPFP: Le Grand, Goetz & Walker, Comput. Phys. Commun. 184
// (2013) 374-380). A generation-0 node is its sample's only one, so its
// thread stores the sums; later generations add atomically. Every random draw
// is keyed on (seed, pixel, sample, tree node, draw index), a tree node being
// a hash of the branch choices from the camera ray. A sample's radiance is
// thus a function of the sample alone: bitwise independent of chunking, queue
// order and of which kernel computed which term.
//
// A generation can outgrow its queue (a dielectric hit has two children).
// Then the chunk is marked, and after the last generation k_ddfs recomputes
// every sample of the chunk depth first — the same terms from the same draws
// into the same fixed-point sums, so the same bits — in the queues' memory.
// Scenes deeper than kDirectWaveDepth, or with more than kDirectWaveLights
// lights, run depth first from the start.
//
// Terms are clamped to +-2^26 before conversion (so thousands of them cannot
// overflow a sum); a non-finite term makes the sample invalid, as the
// reference's NaN/Inf check at render_tile does for the whole sample.
#pragma once
#include "rtg_kernels.hpp"

namespace rtg {

constexpr int kDirectWaveDepth = 64;
constexpr uint32_t kDirectWaveLights = 8;
constexpr uint32_t kDfsWords = 11;  // a pending ray of the depth-first walk: o, d, w, node, depth
constexpr uint32_t kDfsMinThreads = 16384;
constexpr uint32_t kCountWords = 2 * (kDirectWaveDepth + 2) * kSubQueues;


struct DirectMem {
  uint32_t* flag;             // [0] != 0: a generation overflowed its queue
  unsigned long long* stats;  // statistics slots of the wavefront kernels (kept only without overflow)
  unsigned long long* sum;    // fixed-point sums, channel-major: sum[c * cap + l]
  uint32_t* bad;              // a sample met a non-finite term
  float4 *e0[2], *e1[2], *e2[2];  // rays of generation g & 1: {o, node}, {d, sample}, {weight, 0}
  float4* hit[2];                 // {t, b1, b2, prim}
  int32_t* plane[2];
  float4* lit;           // lit hits of one generation: {own term, sample}
  float4 *s0, *s1, *s2;  // their shadow requests, n_light per lit slot: {o, tmax}, {d, cast ? 1 : 0}, {term, 0}
  float* frames;         // k_ddfs: pending rays, field-major (kDfsWords x max_depth rows of dfs_threads)
  uint32_t* cnt;         // queue sizes: generation g's sub-queue s at (g * kSubQueues + s) * cstride, lit
                         // hits after kDirectWaveDepth + 2 generations
  uint32_t cstride;
  uint32_t cap;          // samples of a chunk (generation 0 uses slots [0, cap))
  uint32_t sub_gen;      // ray slots per sub-queue of generations >= 1
  uint32_t sub_lit;      // lit-hit slots per sub-queue
  uint32_t n_light;      // shadow requests per lit hit (the scene's lights)
  uint32_t dfs_threads;
  uint32_t forced;       // 1: depth first from the start
};

// Lays out the direct integrator's memory of a chunk of `cap` samples at
// `base` (or only sizes it when base is null); returns the bytes used.
// `entries`: ray slots of one generation >= 1 (rounded up to a multiple of kSubQueues).
__host__ __device__ inline uint64_t direct_layout(char* base, uint32_t cap, uint32_t entries, uint32_t n_lights,
                                                  int max_depth, bool forced, DirectMem* M) {
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) -> char* {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) & ~uint64_t(255);
    return p;
  };
  DirectMem m{};
  m.cap = cap;
  m.forced = forced ? 1u : 0u;
  m.flag = reinterpret_cast<uint32_t*>(take(256));
  m.stats = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * kStatSlots * kStatStride));
  m.sum = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 3 * (uint64_t)cap));
  m.bad = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * (uint64_t)cap));
  m.cstride = kCountStride;
  m.cnt = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * kCountWords * kCountStride));
  const uint64_t frame_row = (uint64_t)kDfsWords * (uint64_t)max_depth * sizeof(float);  // per thread
  const uint64_t rest = off;
  uint64_t wave = 0;
  if (!forced) {
    m.sub_gen = entries > (uint32_t)kSubQueues ? (entries + kSubQueues - 1) / kSubQueues : 1u;
    const uint64_t slots = (uint64_t)cap > (uint64_t)kSubQueues * m.sub_gen ? cap : (uint64_t)kSubQueues * m.sub_gen;
    // A generation's shade launch has at most ceil(slots / kBlock) blocks;
    // block b appends at most kBlock lit hits to sub-queue b % kSubQueues: a
    // lit sub-queue never overflows.
    const uint64_t blocks = (slots + kBlock - 1) / kBlock;
    m.sub_lit = (uint32_t)(((blocks + kSubQueues - 1) / kSubQueues) * kBlock);
    m.n_light = n_lights;
    for (int g = 0; g < 2; ++g) {
      m.e0[g] = reinterpret_cast<float4*>(take(16 * slots));
      m.e1[g] = reinterpret_cast<float4*>(take(16 * slots));
      m.e2[g] = reinterpret_cast<float4*>(take(16 * slots));
      m.hit[g] = reinterpret_cast<float4*>(take(16 * slots));
      m.plane[g] = reinterpret_cast<int32_t*>(take(4 * slots));
    }
    const uint64_t lit = (uint64_t)kSubQueues * m.sub_lit, sh = lit * n_lights;
    m.lit = reinterpret_cast<float4*>(take(16 * lit));
    m.s0 = reinterpret_cast<float4*>(take(16 * sh));
    m.s1 = reinterpret_cast<float4*>(take(16 * sh));
    m.s2 = reinterpret_cast<float4*>(take(16 * sh));
    wave = off - rest;
  }
  // The depth-first walk reuses the queues' memory (it runs after them).
  uint64_t t = forced ? cap : (wave / frame_row) / kBlock * kBlock;
  const uint64_t t_min = cap < kDfsMinThreads ? cap : kDfsMinThreads;
  if (t < t_min) t = t_min;
  if (t > cap) t = cap;
  m.dfs_threads = (uint32_t)t;
  m.frames = base ? reinterpret_cast<float*>(base + rest) : nullptr;
  const uint64_t frames = (t * frame_row + 255) & ~uint64_t(255);
  off = rest + (frames > wave ? frames : wave);
  if (M) *M = m;
  return off;
}

// ---------------------------------------------------------------------------
// Tree nodes: the camera ray is node 0; child c of node n is a hash of both
// (0xFFFFFFFF is the camera's own draws).
__device__ __forceinline__ uint32_t child_node(uint32_t n, uint32_t c) {
  uint32_t h = n * 0x9E3779B1u + (c + 1u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return h == 0xFFFFFFFFu ? 0xFFFFFFFEu : h;
}
// Draws of a node: group 0 scatters (metal), group li + 1 samples light li.
__device__ __forceinline__ HashDraws node_draws(uint2 key, uint32_t node, uint32_t group) {
  HashDraws r(key, node);
  r.idx = group * 4u;
  return r;
}

// Fixed point: 32 fraction bits in a signed 64-bit integer.
constexpr float kFixMax = 67108864.0f;  // 2^26
__device__ __forceinline__ long long to_fixed(float v) {
  const float a = v > kFixMax ? kFixMax : (v < -kFixMax ? -kFixMax : v);
  return (long long)__builtin_rint((double)a * 4294967296.0);
}
__device__ __forceinline__ float from_fixed(unsigned long long s) { return (float)((double)(long long)s * 0x1p-32); }

// A node's own term and children (integrator.cpp:141-174), shared by the
// wavefront and the depth-first walk so both compute the same floats.
struct DirectNode {
  V3<float> term;  // emission or sky, already weighted
  uint32_t n_child;
  V3<float> o[2], d[2], w[2];
  uint32_t node[2];
  bool lights;  // a diffuse hit below max_depth: every light gets an estimate
};
__device__ __forceinline__ DirectNode direct_node(const DevScene& S, uint2 key, uint32_t node, int depth, int max_depth,
                                                  V3<float> rd, V3<float> w, bool miss, const HitRec<float>& rec,
                                                  const Mat<float>& mat) {
  DirectNode r;
  r.n_child = 0;
  r.lights = false;
  if (miss) {
    r.term = w * background<float>(S, rd);
    return r;
  }
  r.term = w * emitted(mat, rec.front_face);
  if (depth >= max_depth) return r;
  if (mat.kind == RTG_MAT_LAMBERTIAN) {
    r.lights = S.n_lights > 0;
  } else if (mat.kind == RTG_MAT_METAL) {
    HashDraws rng = node_draws(key, node, 0);
    Scatter<float> sr;
    if (scatter(mat, rd, rec, rng, sr)) {
      r.n_child = 1;
      r.o[0] = offset_ray_origin(rec.p, rec.gn, sr.dir);
      r.d[0] = sr.dir;
      r.w[0] = w * mat.c0;
      r.node[0] = child_node(node, 0);
    }
  } else if (mat.kind == RTG_MAT_DIELECTRIC) {
    const V3<float> in = normalize(rd);
    const float eta = rec.front_face ? 1.0f / mat.p0 : mat.p0;
    const float cos_i = r_fmin(dot(-in, rec.normal), 1.0f);
    V3<float> refracted;
    const bool can = refract(in, rec.normal, eta, &refracted);
    const float F = can ? schlick(cos_i, eta) : 1.0f;
    const V3<float> rdir = reflect(in, rec.normal);
    r.n_child = can ? 2 : 1;
    r.o[0] = offset_ray_origin(rec.p, rec.gn, rdir);
    r.d[0] = rdir;
    r.w[0] = w * (mat.c0 * F);
    r.node[0] = child_node(node, 0);
    if (can) {
      const V3<float> tdir = normalize(refracted);
      r.o[1] = offset_ray_origin(rec.p, rec.gn, tdir);
      r.d[1] = tdir;
      r.w[1] = w * (mat.c0 * (1.0f - F));
      r.node[1] = child_node(node, 1);
    }
  }
  return r;
}

// Light li's estimate at a diffuse hit (integrator.cpp:14-91 up to the
// shadow ray), weighted; false when no shadow ray is cast.
__device__ __forceinline__ bool direct_light(const DevScene& S, uint2 key, uint32_t node, uint32_t li,
                                             const HitRec<float>& rec, const Mat<float>& mat, V3<float> w,
                                             ShadowReq<float>& req) {
  HashDraws rng = node_draws(key, node, li + 1u);
  if (!sample_light32(S, li, rec, mat, rng, req)) return false;
  req.f_rop = w * req.f_rop;
  return true;
}

__device__ __forceinline__ bool finite3(V3<float> v) {
  return r_isfinite(v.x) && r_isfinite(v.y) && r_isfinite(v.z);
}
// A node's terms in fixed point; `bad` once one is not finite (it is then
// left out: the sample is invalid anyway).
struct FixSum {
  long long x, y, z;
  bool bad;
  __device__ __forceinline__ void add(V3<float> v) {
    if (!finite3(v)) {
      bad = true;
      return;
    }
    x += to_fixed(v.x);
    y += to_fixed(v.y);
    z += to_fixed(v.z);
  }
};
// Hands a node's terms to sample l: generation 0 stores (its node is the
// sample's only one), later generations add atomically (a sample's nodes of
// one generation may sit in any threads).
__device__ __forceinline__ void put_terms(const DirectMem& M, uint32_t l, const FixSum& f, int g) {
  unsigned long long* sx = M.sum + l;
  unsigned long long* sy = M.sum + M.cap + l;
  unsigned long long* sz = M.sum + 2 * (size_t)M.cap + l;
  if (g == 0) {
    *sx = (unsigned long long)f.x;
    *sy = (unsigned long long)f.y;
    *sz = (unsigned long long)f.z;
    M.bad[l] = f.bad ? 1u : 0u;
    return;
  }
  if (f.x) atomicAdd(sx, (unsigned long long)f.x);
  if (f.y) atomicAdd(sy, (unsigned long long)f.y);
  if (f.z) atomicAdd(sz, (unsigned long long)f.z);
  if (f.bad) atomicOr(M.bad + l, 1u);
}

__device__ __forceinline__ uint32_t* gen_count(const DirectMem& M, int g) {
  return M.cnt + (size_t)g * kSubQueues * M.cstride;
}
__device__ __forceinline__ uint32_t* lit_count(const DirectMem& M, int g) {
  return M.cnt + (size_t)(kDirectWaveDepth + 2 + g) * kSubQueues * M.cstride;
}

// Reserves na slots in counter ca and nb in cb per thread of the block (one
// atomic per counter per block; every thread must call it); returns the
// thread's first offsets.
struct Reserve2 {
  uint32_t a, b;
};
__device__ __forceinline__ Reserve2 block_reserve2(uint32_t* ca, uint32_t na, uint32_t* cb, uint32_t nb) {
  __shared__ uint32_t s_tot[2][kBlock / 64];
  __shared__ uint32_t s_base[2];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x / 64;
  uint32_t xa = na, xb = nb;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t ya = __shfl_up(xa, off, 64), yb = __shfl_up(xb, off, 64);
    if (lane >= (uint32_t)off) {
      xa += ya;
      xb += yb;
    }
  }
  if (lane == 63) {
    s_tot[0][wave] = xa;
    s_tot[1][wave] = xb;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < kBlock / 64; ++w) tot += s_tot[threadIdx.x][w];
    s_base[threadIdx.x] = tot ? atomicAdd(threadIdx.x ? cb : ca, tot) : 0u;
  }
  __syncthreads();
  uint32_t ba = s_base[0], bb = s_base[1];
  for (uint32_t w = 0; w < wave; ++w) {
    ba += s_tot[0][w];
    bb += s_tot[1][w];
  }
  __syncthreads();
  return Reserve2{ba + xa - na, bb + xb - nb};
}

// Reserves n consecutive slots per lane of a converged wave in sub-queue
// counter `c` (one atomic per wave); returns the lane's first offset.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* c, uint32_t n) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t x = n;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  const uint32_t total = __shfl(x, 63, 64);
  uint32_t base = 0;
  if (lane == 63 && total) base = atomicAdd(c, total);
  base = __shfl(base, 63, 64);
  return base + x - n;
}

__device__ __forceinline__ uint2 sample_key(const ChunkParams& P, uint32_t l) {
  const uint32_t pid = P.pid_base + l;
  return path_key(P.seed, pid_pixel(P, pid), pid_sample(P, pid));
}

// ---------------------------------------------------------------------------
// Clears the queue counters, the wavefront's statistics slots and the
// overflow flag of a chunk.
template <class T>  // float (a template so that every unit may include this file)
__global__ __launch_bounds__(kBlock) void k_dclear(DirectMem M) {
  const uint32_t l = blockIdx.x * kBlock + threadIdx.x;
  if (l < kCountWords) M.cnt[(size_t)l * M.cstride] = 0u;
  if (l < (uint32_t)(kStatSlots * kStatStride)) M.stats[l] = 0;
  if (l == 0) M.flag[0] = 0;
}
__host__ __device__ inline uint32_t dclear_threads() {
  return kCountWords > (uint32_t)(kStatSlots * kStatStride) ? kCountWords : (uint32_t)(kStatSlots * kStatStride);
}

// Generation 0 in one kernel: the camera ray (renderer.cpp:31-37), its
// closest hit, the node's terms with every light's shadow ray traced in the
// same thread, and the children appended to generation 1. A generation-0
// node is its sample's only one: its sums are stored. Camera rays of
// neighbouring pixels, and their shadow rays to one light, are coherent, so
// one thread per sample fits here; the queue takes over where trees branch
// (torus workload, direct: 27.9 ms per render with generation 0 as separate
// extend / shade / shadow kernels, 22.8 ms with this kernel).
template <int STACK>
__global__ __launch_bounds__(kBlock) void k_dfirst(DevScene S, ChunkParams P, DirectMem M) {
  __shared__ uint32_t sref[STACK * kBlock];
  const uint32_t l = blockIdx.x * kBlock + threadIdx.x;
  const bool live = l < P.n_paths;
  uint32_t st[kStCount] = {0};
  DirectNode n;
  n.n_child = 0;
  n.lights = false;
  HitRec<float> rec;
  Mat<float> mat;
  uint2 key = make_uint2(0, 0);
  const V3<float> w{1.0f, 1.0f, 1.0f};
  if (live) {
    const uint32_t pid = P.pid_base + l;
    const uint32_t pixel = pid_pixel(P, pid), sample = pid_sample(P, pid);
    key = path_key(P.seed, pixel, sample);
    V3<float> o, d;
    {
      HashDraws rng(key, 0xFFFFFFFFu);
      const float jx = rng.next();
      const float jy = rng.next();
      camera32(S, (int)(pixel % P.width), (int)(pixel / P.width), jx, jy, rng, o, d);
    }
    Counters c{0, 0};
    const Hit32 h = closest32(S, P.use_bvh != 0, make_ray32(S, o.x, o.y, o.z, d.x, d.y, d.z), 1e-6f, __builtin_inff(),
                              lds_stack(sref), c);
    st[kStRays] = 1;
    st[kStVisits] = c.visits;
    st[kStTests] = c.tests;
    const bool miss = h.prim == 0xFFFFFFFFu && h.plane < 0;
    if (!miss) {
      finalize32(S, o, d, h.t, h.b1, h.b2, h.prim, h.plane, rec);
      mat = load_mat32(S, rec.material);
    }
    n = direct_node(S, key, 0u, 0, P.max_depth, d, w, miss, rec, mat);
  }
  // The children first (the append is block-wide), so only the lights'
  // inputs stay live across the shadow walks.
  const uint32_t sub = blockIdx.x % kSubQueues;
  uint32_t* c1 = gen_count(M, 1) + (size_t)sub * M.cstride;
  const uint32_t cpos = block_reserve2(c1, n.n_child, c1, 0u).a;
#pragma unroll
  for (uint32_t j = 0; j < 2; ++j) {
    if (j >= n.n_child) continue;
    const uint32_t at = cpos + j;
    if (at >= M.sub_gen) {
      atomicOr(M.flag, 1u);
      continue;
    }
    const uint32_t k = sub * M.sub_gen + at;
    M.e0[1][k] = make_float4(n.o[j].x, n.o[j].y, n.o[j].z, __uint_as_float(n.node[j]));
    M.e1[1][k] = make_float4(n.d[j].x, n.d[j].y, n.d[j].z, __uint_as_float(l));
    M.e2[1][k] = make_float4(n.w[j].x, n.w[j].y, n.w[j].z, 0.0f);
  }
  if (live) {
    FixSum f{0, 0, 0, false};
    f.add(n.term);
    if (n.lights)
      for (uint32_t li = 0; li < S.n_lights; ++li) {
        ShadowReq<float> req;
        if (!direct_light(S, key, 0u, li, rec, mat, w, req)) continue;
        Counters c{0, 0};
        if (!any32(S, P.use_bvh != 0, make_ray32(S, req.o.x, req.o.y, req.o.z, req.d.x, req.d.y, req.d.z), 1e-6f,
                   req.tmax, lds_stack(sref), c))
          f.add(req.f_rop);
        ++st[kStShadow];
        st[kStSVisits] += c.visits;
        st[kStSTests] += c.tests;
      }
    put_terms(M, l, f, 0);
  }
  flush_stats(M.stats, st, kStCount, kStatSlots, blockIdx.x);
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_dextend(DevScene S, ChunkParams P, DirectMem M, int g) {
  __shared__ uint32_t sref[STACK * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  const QueueView qv = load_queue(gen_count(M, g), pre, blockIdx.x, M.sub_gen, M.cstride);
  if (qv.past_end) return;
  uint32_t st[kStCount] = {0};
  for (uint32_t vb = blockIdx.x; vb * kBlock < qv.total; vb += gridDim.x) {
    const uint32_t i = vb * kBlock + threadIdx.x;
    if (i >= qv.total) continue;
    const uint32_t k = queue_slot(pre, M.sub_gen, i);
    const float4 a = M.e0[g & 1][k], b = M.e1[g & 1][k];
    Counters c{0, 0};
    const Hit32 h = closest32(S, P.use_bvh != 0, make_ray32(S, a.x, a.y, a.z, b.x, b.y, b.z), 1e-6f, __builtin_inff(),
                              lds_stack(sref), c);
    M.hit[g & 1][k] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.prim));
    M.plane[g & 1][k] = h.plane;
    ++st[kStRays];
    st[kStVisits] += c.visits;
    st[kStTests] += c.tests;
  }
  flush_stats(M.stats, st, kStCount, kStatSlots, blockIdx.x);
}

template <class T>  // float (a template so that every unit may include this file)
__global__ __launch_bounds__(kBlock) void k_dshade(DevScene S, ChunkParams P, DirectMem M, int g) {
  __shared__ uint32_t pre[kSubQueues + 1];
  const QueueView qv = load_queue(gen_count(M, g), pre, blockIdx.x, M.sub_gen, M.cstride);
  if (qv.past_end) return;
  for (uint32_t vb = blockIdx.x; vb * kBlock < qv.total; vb += gridDim.x) {  // uniform in the block
    const uint32_t i = vb * kBlock + threadIdx.x;
    const bool live = i < qv.total;
    DirectNode n;
    n.n_child = 0;
    n.lights = false;
    HitRec<float> rec;
    Mat<float> mat;
    uint32_t l = 0, node = 0;
    uint2 key = make_uint2(0, 0);
    V3<float> w{0, 0, 0};
    if (live) {
      const uint32_t k = queue_slot(pre, M.sub_gen, i);
      const float4 a = M.e0[g & 1][k], b = M.e1[g & 1][k], c = M.e2[g & 1][k], h = M.hit[g & 1][k];
      const int plane = M.plane[g & 1][k];
      const uint32_t prim = __float_as_uint(h.w);
      const V3<float> o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
      node = __float_as_uint(a.w);
      l = __float_as_uint(b.w);
      w = V3<float>{c.x, c.y, c.z};
      key = sample_key(P, l);
      const bool miss = prim == 0xFFFFFFFFu && plane < 0;
      if (!miss) {
        finalize32(S, o, d, h.x, h.y, h.z, prim, plane, rec);
        mat = load_mat32(S, rec.material);
      }
      n = direct_node(S, key, node, g, P.max_depth, d, w, miss, rec, mat);
      if (!n.lights && (!is_zero(n.term) || !finite3(n.term))) {  // its sample's terms of this node
        FixSum f{0, 0, 0, false};
        f.add(n.term);
        put_terms(M, l, f, g);
      }
    }
    const uint32_t sub = vb % kSubQueues;
    // Children to generation g + 1; a full sub-queue marks the chunk.
    const Reserve2 rv = block_reserve2(gen_count(M, g + 1) + sub * M.cstride, n.n_child,
                                       lit_count(M, g) + sub * M.cstride, n.lights ? 1u : 0u);
    const uint32_t cpos = rv.a, lpos = rv.b;
#pragma unroll
    for (uint32_t j = 0; j < 2; ++j) {  // constant indices: the children stay in registers
      if (j >= n.n_child) continue;
      const uint32_t at = cpos + j;
      if (at >= M.sub_gen) {
        atomicOr(M.flag, 1u);
        continue;
      }
      const uint32_t k = sub * M.sub_gen + at;
      const int ng = (g + 1) & 1;
      M.e0[ng][k] = make_float4(n.o[j].x, n.o[j].y, n.o[j].z, __uint_as_float(n.node[j]));
      M.e1[ng][k] = make_float4(n.d[j].x, n.d[j].y, n.d[j].z, __uint_as_float(l));
      M.e2[ng][k] = make_float4(n.w[j].x, n.w[j].y, n.w[j].z, 0.0f);
    }
    // A lit hit: its own term and one request per light (the light's own
    // draws; `cast` 0 when the estimate is zero and no shadow ray is cast).
    if (n.lights) {
      if (lpos >= M.sub_lit) {  // cannot happen (sizing); kept as a guard
        atomicOr(M.flag, 1u);
      } else {
        const uint32_t k = sub * M.sub_lit + lpos;
        M.lit[k] = make_float4(n.term.x, n.term.y, n.term.z, __uint_as_float(l));
        for (uint32_t li = 0; li < M.n_light; ++li) {
          ShadowReq<float> req;
          const size_t r = (size_t)k * M.n_light + li;
          if (direct_light(S, key, node, li, rec, mat, w, req)) {
            M.s0[r] = make_float4(req.o.x, req.o.y, req.o.z, req.tmax);
            M.s1[r] = make_float4(req.d.x, req.d.y, req.d.z, 1.0f);
            M.s2[r] = make_float4(req.f_rop.x, req.f_rop.y, req.f_rop.z, 0.0f);
          } else {
            M.s1[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          }
        }
      }
    }
  }
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_dshadow(DevScene S, ChunkParams P, DirectMem M, int g) {
  __shared__ uint32_t sref[STACK * kBlock];
  __shared__ uint32_t pre[kSubQueues + 1];
  const QueueView qv = load_queue(lit_count(M, g), pre, blockIdx.x, M.sub_lit, M.cstride);
  if (qv.past_end) return;
  uint32_t st[kStCount] = {0};
  for (uint32_t vb = blockIdx.x; vb * kBlock < qv.total; vb += gridDim.x) {
    const uint32_t i = vb * kBlock + threadIdx.x;
    if (i >= qv.total) continue;
    const uint32_t k = queue_slot(pre, M.sub_lit, i);
    // Which lights are visible first (only a mask lives across the walks),
    // then the terms.
    uint32_t vis = 0;
    for (uint32_t li = 0; li < M.n_light; ++li) {
      const size_t r = (size_t)k * M.n_light + li;
      const float4 b = M.s1[r];
      if (b.w == 0.0f) continue;
      const float4 a = M.s0[r];
      Counters c{0, 0};
      if (!any32(S, P.use_bvh != 0, make_ray32(S, a.x, a.y, a.z, b.x, b.y, b.z), 1e-6f, a.w, lds_stack(sref), c))
        vis |= 1u << li;
      ++st[kStShadow];
      st[kStSVisits] += c.visits;
      st[kStSTests] += c.tests;
    }
    const float4 e = M.lit[k];
    FixSum f{0, 0, 0, false};
    f.add(V3<float>{e.x, e.y, e.z});
    for (uint32_t li = 0; vis >> li; ++li)
      if ((vis >> li) & 1u) {
        const float4 t = M.s2[(size_t)k * M.n_light + li];
        f.add(V3<float>{t.x, t.y, t.z});
      }
    put_terms(M, __float_as_uint(e.w), f, g);
  }
  flush_stats(M.stats, st, kStCount, kStatSlots, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Depth first: one thread walks a sample's tree, the second child of each
// dielectric split waiting in `frames`. Same terms, draws and fixed-point
// sums as the wavefront; it runs when the chunk overflowed (or is forced),
// else it returns at once.
template <int STACK>
__global__ __launch_bounds__(kBlock) void k_ddfs(DevScene S, ChunkParams P, DirectMem M) {
  __shared__ uint32_t sref[STACK * kBlock];
  if (!M.forced && M.flag[0] == 0) return;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  uint32_t st[kStCount] = {0};
  const uint32_t T = M.dfs_threads;
  auto fr = [&](int f, int field) -> float& {
    return M.frames[((size_t)f * kDfsWords + (size_t)field) * T + t];
  };
  for (uint32_t l = t; l < P.n_paths; l += T) {
    const uint32_t pid = P.pid_base + l;
    const uint32_t pixel = pid_pixel(P, pid), sample = pid_sample(P, pid);
    const uint2 key = path_key(P.seed, pixel, sample);
    V3<float> o, d, w{1.0f, 1.0f, 1.0f};
    {
      HashDraws rng(key, 0xFFFFFFFFu);
      const float jx = rng.next();
      const float jy = rng.next();
      camera32(S, (int)(pixel % P.width), (int)(pixel / P.width), jx, jy, rng, o, d);
    }
    uint32_t node = 0;
    int depth = 0, top = 0;
    FixSum f{0, 0, 0, false};
    while (true) {
      ++st[kStRays];
      Counters c{0, 0};
      const Hit32 h = closest32(S, P.use_bvh != 0, make_ray32(S, o.x, o.y, o.z, d.x, d.y, d.z), 1e-6f,
                                __builtin_inff(), lds_stack(sref), c);
      st[kStVisits] += c.visits;
      st[kStTests] += c.tests;
      const bool miss = h.prim == 0xFFFFFFFFu && h.plane < 0;
      HitRec<float> rec;
      Mat<float> mat;
      if (!miss) {
        finalize32(S, o, d, h.t, h.b1, h.b2, h.prim, h.plane, rec);
        mat = load_mat32(S, rec.material);
      }
      const DirectNode n = direct_node(S, key, node, depth, P.max_depth, d, w, miss, rec, mat);
      f.add(n.term);
      if (n.lights)
        for (uint32_t li = 0; li < S.n_lights; ++li) {
          ShadowReq<float> req;
          if (!direct_light(S, key, node, li, rec, mat, w, req)) continue;
          ++st[kStShadow];
          Counters c2{0, 0};
          if (!any32(S, P.use_bvh != 0, make_ray32(S, req.o.x, req.o.y, req.o.z, req.d.x, req.d.y, req.d.z), 1e-6f,
                     req.tmax, lds_stack(sref), c2))
            f.add(req.f_rop);
          st[kStSVisits] += c2.visits;
          st[kStSTests] += c2.tests;
        }
      if (n.n_child == 2) {  // the refracted ray waits
        fr(top, 0) = n.o[1].x; fr(top, 1) = n.o[1].y; fr(top, 2) = n.o[1].z;
        fr(top, 3) = n.d[1].x; fr(top, 4) = n.d[1].y; fr(top, 5) = n.d[1].z;
        fr(top, 6) = n.w[1].x; fr(top, 7) = n.w[1].y; fr(top, 8) = n.w[1].z;
        fr(top, 9) = __uint_as_float(n.node[1]);
        fr(top, 10) = __int_as_float(depth + 1);
        ++top;
      }
      if (n.n_child) {
        o = n.o[0];
        d = n.d[0];
        w = n.w[0];
        node = n.node[0];
        ++depth;
        continue;
      }
      if (top == 0) break;
      --top;
      o = V3<float>{fr(top, 0), fr(top, 1), fr(top, 2)};
      d = V3<float>{fr(top, 3), fr(top, 4), fr(top, 5)};
      w = V3<float>{fr(top, 6), fr(top, 7), fr(top, 8)};
      node = __float_as_uint(fr(top, 9));
      depth = __float_as_int(fr(top, 10));
    }
    put_terms(M, l, f, 0);  // the whole sample: stored
  }
  flush_stats(P, st, blockIdx.x);
}

// The sums as the samples' radiance (NaN for a sample that met a non-finite
// term, so the accumulation counts it invalid); the wavefront's statistics
// join the job's unless the chunk was recomputed depth first.
template <class T>  // float (a template so that every unit may include this file)
__global__ __launch_bounds__(kBlock) void k_dfinish(PathBuffers<T> B, ChunkParams P, DirectMem M) {
  const uint32_t l = blockIdx.x * kBlock + threadIdx.x;
  const bool redone = M.forced || M.flag[0] != 0;
  if (l < P.n_paths) {
    const uint32_t pid = P.pid_base + l;
    const bool bad = M.bad[l] != 0;
    B.lx[pid] = bad ? __builtin_nanf("") : from_fixed(M.sum[l]);
    B.ly[pid] = bad ? __builtin_nanf("") : from_fixed(M.sum[M.cap + l]);
    B.lz[pid] = bad ? __builtin_nanf("") : from_fixed(M.sum[2 * (size_t)M.cap + l]);
  }
  if (l < (uint32_t)(kStatSlots * kStatStride)) {
    const uint32_t slot = l / kStatStride, c = l % kStatStride;
    if (!redone) {
      const unsigned long long v = M.stats[l];
      if (v) atomicAdd(P.stats + (size_t)(slot % P.stat_slots) * kStatStride + c, v);
    } else if (!M.forced && l == 0) {
      atomicAdd(P.stats + kStRedone, 1ull);
    }
  }
}
__host__ __device__ inline uint32_t dfinish_threads(const ChunkParams& P) {
  return P.n_paths > (uint32_t)(kStatSlots * kStatStride) ? P.n_paths : (uint32_t)(kStatSlots * kStatStride);
}

}  // namespace rtg
