/*
 * rtg.h — C ABI of the MI355X path-tracing back end (librtg.so).
 *
 * The library renders and traces a compiled scene on one gfx950 GPU. The
 * host side (JSON scene load, Scene::compile, binned-SAH BVH build) stays in
 * C++ and hands the compiled scene over as plain arrays, field for field as
 * the reference's `Scene` holds them (include/raytracer/core/scene.hpp:33-56).
 * No C++ exception crosses this boundary: every entry point returns an
 * rtg_status and rtg_last_error() gives the message.
 *
 * Threading: one context per host thread. Calls on one context are
 * serialised on that context's HIP stream; contexts on the same device run
 * concurrently on their own streams.
 */
#ifndef RTG_H
#define RTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTG_ABI_VERSION 3

typedef int32_t rtg_status;
enum {
  RTG_OK = 0,
  RTG_ERR_INVALID_ARG = -1,   /* bad argument or inconsistent scene arrays   */
  RTG_ERR_OUT_OF_MEMORY = -2, /* device allocation failed / admission refused */
  RTG_ERR_HIP = -3,           /* a HIP runtime call failed                    */
  RTG_ERR_CANCELLED = -4,     /* the cancel flag was raised during the call  */
  RTG_ERR_UNSUPPORTED = -5,   /* feature not available on this device/build  */
  RTG_ERR_NO_DEVICE = -6      /* no HIP device                               */
};

/* ------------------------------------------------------------------------ */
/* Compiled scene, as the reference's Scene::compile produces it (all FP64). */

typedef struct rtg_vec3 { double x, y, z; } rtg_vec3;

/* geometry/bvh.hpp:20-26 (BvhNode): DFS order, left child = index + 1,
 * interior `offset` = right child, leaf `offset` = first prim_indices slot. */
typedef struct rtg_bvh_node {
  rtg_vec3 lo, hi;
  uint32_t offset;
  uint16_t count; /* 0 = interior */
  uint8_t axis;
  uint8_t pad;
} rtg_bvh_node;

enum { RTG_PRIM_SPHERE = 0, RTG_PRIM_TRIANGLE = 1 };
/* core/scene.hpp:17 (Primitive = variant<Sphere, Triangle>),
 * geometry/shapes.hpp:11-14 (Sphere), :101-105 (Triangle).
 * Sphere: v0 = centre, v1.x = radius. Triangle: v0, v1, v2 and, when
 * has_vertex_normals, n0, n1, n2. */
typedef struct rtg_primitive {
  uint32_t kind;
  uint32_t material;
  uint32_t has_vertex_normals;
  uint32_t pad;
  rtg_vec3 v0, v1, v2;
  rtg_vec3 n0, n1, n2;
} rtg_primitive;

/* geometry/shapes.hpp:77-80 (Plane): unit normal, tested outside the BVH. */
typedef struct rtg_plane {
  rtg_vec3 point, normal;
  uint32_t material, pad;
} rtg_plane;

enum { RTG_MAT_LAMBERTIAN = 0, RTG_MAT_METAL = 1, RTG_MAT_DIELECTRIC = 2, RTG_MAT_EMISSIVE = 3 };
/* materials/material.hpp:14-56. c0: Lambertian albedo a | Metal albedo |
 * Dielectric tint | Emissive colour; c1: Lambertian albedo2 (checker b);
 * p0: checker_scale | roughness | ior | intensity. */
typedef struct rtg_material {
  uint32_t kind;
  uint32_t two_sided;
  rtg_vec3 c0, c1;
  double p0;
} rtg_material;

/* core/scene.hpp:19-29 (PointLight, LightRef). */
typedef struct rtg_point_light { rtg_vec3 position, intensity; } rtg_point_light;
enum { RTG_LIGHT_POINT = 0, RTG_LIGHT_EMITTER = 1 };
typedef struct rtg_light_ref { uint32_t kind, index; } rtg_light_ref;

/* core/camera.hpp:29-51 after setup(): what generate() reads. */
typedef struct rtg_camera {
  rtg_vec3 origin, corner, du, dv, u, v, w;
  double lens_radius;
} rtg_camera;

enum { RTG_BG_GRADIENT = 0, RTG_BG_SOLID = 1 };
/* core/scene_desc.hpp:35-42 (Background). */
typedef struct rtg_background {
  uint32_t kind, pad;
  rtg_vec3 top, bottom;
  double intensity;
} rtg_background;

typedef struct rtg_scene_desc {
  const rtg_primitive* prims;
  uint32_t n_prims;
  const uint8_t* is_sampled_light; /* n_prims entries */
  const rtg_plane* planes;
  uint32_t n_planes;
  const rtg_material* materials;
  uint32_t n_materials;
  const rtg_point_light* point_lights;
  uint32_t n_point_lights;
  const rtg_light_ref* lights;
  uint32_t n_lights;
  const rtg_bvh_node* nodes;
  uint32_t n_nodes;
  const uint32_t* prim_indices;
  uint32_t n_prim_indices;
  rtg_camera camera;
  rtg_background background;
} rtg_scene_desc;

/* ------------------------------------------------------------------------ */
/* Render settings: the per-sample part of RenderSettings
 * (core/scene_desc.hpp:20-35) plus the device precision mode. */

enum { RTG_INTEGRATOR_PATH = 0, RTG_INTEGRATOR_DIRECT = 1, RTG_INTEGRATOR_NORMALS = 2 };
/* RTG_PRECISION_F32: float arithmetic, stateless counter-hash RNG keyed on
 *   (seed, pixel, sample, bounce, dimension) — the production path.
 * RTG_PRECISION_F64: double arithmetic and the reference's own per-sample
 *   PCG32 stream (math/rng.hpp:29-32) — reproduces the CPU renderer's
 *   samples up to the last-bit differences of device sin/cos. */
enum { RTG_PRECISION_F32 = 0, RTG_PRECISION_F64 = 1 };

typedef struct rtg_render_settings {
  int32_t width, height;
  int32_t max_depth;
  int32_t rr_min_depth;
  int32_t integrator;
  int32_t light_sampling;
  int32_t use_bvh;
  int32_t precision;
  uint64_t seed;
} rtg_render_settings;

/* PathStats (rendering/integrator.hpp:11-16), TraversalCounters
 * (geometry/bvh.hpp:34-37) with the same counting rules. */
typedef struct rtg_stats {
  uint64_t camera_samples;
  uint64_t rays;          /* closest-hit rays (camera + bounces) */
  uint64_t shadow_rays;   /* any-hit visibility rays actually cast */
  uint64_t rr_terminated;
  uint64_t invalid_samples;
  uint64_t node_visits;   /* BVH node box tests */
  uint64_t prim_tests;    /* primitive intersection tests */
  double elapsed_ms;      /* device time of the call (HIP events) */
  uint64_t redone_chunks; /* FP32 direct: chunks whose ray queue filled, recomputed
                             depth first (same result; a measure of wasted work) */
} rtg_stats;

/* ------------------------------------------------------------------------ */
typedef struct rtg_context rtg_context;
typedef struct rtg_scene rtg_scene;
typedef struct rtg_accum rtg_accum;
typedef struct rtg_buffer rtg_buffer;

/* Opens `device` (HIP ordinal) and creates a stream. */
rtg_status rtg_open(int device, rtg_context** out);
rtg_status rtg_close(rtg_context* ctx);
/* Message of the last failed call on `ctx` (ctx-owned; NULL ctx = global). */
const char* rtg_last_error(const rtg_context* ctx);
int rtg_abi_version(void);
/* Device name and size (multiprocessors, HBM bytes). */
rtg_status rtg_device_info(rtg_context* ctx, char* name, size_t name_len, int* multiprocessors,
                           uint64_t* hbm_bytes);

/* Uploads a compiled scene: flattens the BVH into 32-byte two-child records
 * (half-float bounds in the scene's node frame, rounded outward) and the
 * primitives into leaf order.
 * Replaces the scene argument of Renderer (rendering/renderer.hpp:52). */
rtg_status rtg_scene_create(rtg_context* ctx, const rtg_scene_desc* desc, rtg_scene** out);
/* The FP32 node records rtg_scene_create builds from `desc`, computed on
 * the host (no device is touched), 16 floats per interior node of the
 * reference BVH, in its DFS order: left child box lo.xyz, hi.xyz, right
 * child box lo.xyz, hi.xyz — the record's half floats, widened exactly, in
 * the node frame of rtg_scene_node_frame (world = origin + v * 2^exp2, each
 * box rounded outward from the double box, so it contains it) — then the
 * left and right child refs, the split axis of the reference node (the
 * device record does not hold it) as u32 bit patterns, and a zero. Writes
 * min(cap, *n) records to `out` (may be NULL) and the record count to *n.
 * The layout replaces BvhNode (include/raytracer/geometry/bvh.hpp:20-26) on
 * the device; this entry point exposes it to tools and tests. */
rtg_status rtg_scene_node_records(const rtg_scene_desc* desc, float* out, uint64_t cap, uint64_t* n);
/* The node frame of `desc`'s FP32 node records: the centre of the BVH's
 * root box (as floats) and the power of two the records' coordinates are
 * scaled by. No counterpart in the reference (its nodes hold doubles);
 * for tools and tests that read the records. */
rtg_status rtg_scene_node_frame(const rtg_scene_desc* desc, float origin[3], int32_t* exp2);
rtg_status rtg_scene_release(rtg_scene* scene);
/* A handle on `base`'s device arrays with another camera — the same scene
 * rendered at another image size needs the camera set up for that size
 * (core/camera.hpp:29-51). No geometry is copied; the arrays are freed when
 * the last handle sharing them is released. */
rtg_status rtg_scene_with_camera(rtg_context* ctx, const rtg_scene* base, const rtg_camera* camera,
                                 rtg_scene** out);
/* Device bytes held by the scene (admission by HBM footprint). */
uint64_t rtg_scene_bytes(const rtg_scene* scene);
/* Device bytes the context holds now: path state and queues of every lane
 * (grown on demand to the largest chunk or batch group seen), statistics
 * slots, batch tables, resolve staging. With rtg_scene_bytes and the
 * accumulation buffers, the HBM footprint of a render service. */
uint64_t rtg_context_bytes(const rtg_context* ctx);

/* Accumulation buffer: per-pixel FP64 sum + u32 count (the reference
 * Framebuffer, rendering/framebuffer.hpp:16-50), resident on the device. */
rtg_status rtg_accum_create(rtg_context* ctx, int32_t width, int32_t height, rtg_accum** out);
rtg_status rtg_accum_release(rtg_accum* acc);
rtg_status rtg_accum_clear(rtg_context* ctx, rtg_accum* acc);
/* Copies the buffer to host memory: sum = width*height*3 doubles (RGB,
 * row-major, top row first), count = width*height u32. Either may be NULL. */
rtg_status rtg_accum_read(rtg_context* ctx, const rtg_accum* acc, double* sum, uint32_t* count);
uint32_t rtg_accum_min_samples(rtg_context* ctx, const rtg_accum* acc);

/* Renders samples [s0, s1) of every pixel into `acc` — one progressive pass
 * of Renderer::render (rendering/renderer.cpp:58-105) with render_tile's
 * per-sample loop (:21-56). `cancel`, if non-NULL, is polled between device
 * chunks; a raised flag returns RTG_ERR_CANCELLED with whole chunks kept. */
rtg_status rtg_render_pass(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* settings,
                           uint32_t s0, uint32_t s1, rtg_accum* acc, const volatile int32_t* cancel,
                           rtg_stats* stats);

/* Several progressive passes in one call: passes [bounds[i], bounds[i+1])
 * for i < n_passes (bounds increasing), the loop of Renderer::render
 * (rendering/renderer.cpp:58-105) when no on_pass callback needs the
 * framebuffer between passes. Bitwise the result of calling
 * rtg_render_pass for each pass in order (every pixel's additions stay in
 * sample order), but without a host synchronisation between passes: the
 * chunks of consecutive passes run side by side on the context's two lanes.
 * `stats` totals the passes. `cancel`, if non-NULL, is read before each
 * chunk with at most two chunks in flight; a raised flag returns
 * RTG_ERR_CANCELLED with whole chunks kept. *passes_done (may be NULL): the
 * passes completed. */
rtg_status rtg_render_passes(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* settings,
                             const uint32_t* bounds, uint32_t n_passes, rtg_accum* acc,
                             const volatile int32_t* cancel, uint32_t* passes_done, rtg_stats* stats);

/* rtg_render_pass without device allocation. This entry point has no
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The path state and statistics slots live
 * in a fixed device array of RTG_STATIC_SCRATCH_BYTES inside librtg.so, one
 * copy per device, held until the process ends; the call makes no
 * hipMalloc / hipFree and clears what must start at zero with a kernel.
 * A pass of any size is accepted: it is rendered on one lane in chunks of at
 * most 256 Ki paths (the path state that fits the array, in either
 * precision), each pixel's samples added in order, so the accumulation
 * buffer and the counters are bitwise those of rtg_render_pass. Every call on one
 * device shares that device's array: calls that overlap on one device (from
 * contexts on different host threads) race and must not be issued.
 * With stats == NULL and cancel == NULL the call returns once its work is
 * enqueued on the context's stream (rtg_synchronize or a later call on the
 * context orders after it); otherwise it waits for the pass, as
 * rtg_render_pass does. Arguments and errors as rtg_render_pass. */
#define RTG_STATIC_SCRATCH_BYTES (64ull << 20)
rtg_status rtg_render_pass_static(rtg_context* ctx, const rtg_scene* scene, const rtg_render_settings* settings,
                                  uint32_t s0, uint32_t s1, rtg_accum* acc, const volatile int32_t* cancel,
                                  rtg_stats* stats);

/* Tone map + sRGB encode on the device (rendering/tonemap.cpp:14-53):
 * writes width*height*3 bytes. tone_map: 0 clamp, 1 reinhard, 2 aces. */
rtg_status rtg_resolve(rtg_context* ctx, const rtg_accum* acc, int32_t tone_map, double exposure, uint8_t* rgb);

/* Batched rendering: many jobs — each one pass (samples [s0, s1) of pixels
 * [p0, p1)) of one render with its own scene, settings and accumulation
 * buffer — rendered together. FP32 path-integrator jobs with max_depth <= 64
 * share the path arrays and every kernel launch (planned on the device, so a
 * batch of small renders fills the GPU like one large one; groups of at
 * least rtg_tuning.wide_paths paths run as wide passes); FP64, direct,
 * normals and deeper jobs run through the single-render path in the same
 * call. Every job's result is bitwise the result of rendering it alone with
 * rtg_render_pass.
 * Jobs that write the same accumulation buffer must cover disjoint pixel
 * ranges. stats (n_jobs entries, may be NULL) receives each job's counters;
 * elapsed_ms is the device time of the whole call. The reference renders one
 * scene per Renderer (renderer.cpp:58-105); this has no counterpart there. */
typedef struct rtg_job {
  const rtg_scene* scene;
  rtg_accum* acc;
  rtg_render_settings settings; /* width/height = the accumulation buffer's */
  uint32_t s0, s1;              /* samples [s0, s1) */
  uint32_t p0, p1;              /* pixels [p0, p1), row-major; p1 = 0: all */
} rtg_job;
rtg_status rtg_render_batch(rtg_context* ctx, const rtg_job* jobs, uint32_t n_jobs, rtg_stats* stats);

/* ------------------------------------------------------------------------ */
/* Scheduling knobs of a context. None of them changes a result (images and
 * counters are bitwise independent of how passes are cut into chunks and
 * groups and spread over lanes); they only change speed. Zero fields mean
 * the built-in default. The reference has no counterpart: its tile size and
 * thread count (RenderSettings, core/scene_desc.hpp:31-32) play this role
 * for the CPU. */
typedef struct rtg_tuning {
  uint32_t chunk_paths; /* paths per chunk and per batch group (queue capacity), 1024..2^28; default 32 Mi */
  uint32_t wide_paths;  /* batch groups of at least this many paths run as wide passes; default 4 Mi */
  uint32_t lanes;       /* 1 or 2 lanes for chunks and batch groups; default 2 */
  uint32_t pad;
} rtg_tuning;
/* The values in effect (defaults filled in). */
rtg_status rtg_get_tuning(rtg_context* ctx, rtg_tuning* out);
rtg_status rtg_set_tuning(rtg_context* ctx, const rtg_tuning* tuning);

typedef struct rtg_tune_result {
  rtg_tuning best;
  double default_ms; /* the workload with the built-in defaults */
  double best_ms;    /* the workload with `best` */
  uint32_t trials;   /* configurations timed; 0 = taken from the cache */
  uint32_t pad;
} rtg_tune_result;
/* The autotuner: times a workload — n_calls rtg_render_batch calls, call i
 * rendering the next call_jobs[i] entries of `jobs` — under candidate
 * tunings (coordinate descent over chunk_paths, wide_paths and lanes,
 * starting from the defaults), keeps the fastest in the context and
 * returns it with both timings. The jobs' accumulation buffers are
 * overwritten and left cleared. With cache_path, results are cached per
 * device (name, compute units) and `key` (the caller's name for the
 * workload class; no quotes) in a JSON-lines file: a cached entry is used
 * without timing, a new result is appended. */
rtg_status rtg_autotune(rtg_context* ctx, const rtg_job* jobs, const uint32_t* call_jobs, uint32_t n_calls,
                        const char* cache_path, const char* key, rtg_tune_result* out);

/* ------------------------------------------------------------------------ */
/* Batch ray queries: Scene::intersect / occluded / intersect_brute
 * (src/core/scene.cpp:176-234) over n rays.
 * Ray record: 8 doubles {ox, oy, oz, dx, dy, dz, tmin, tmax}.
 * Hit record (rtg_hit): t, primitive index (or plane index) and flags. */
typedef struct rtg_hit {
  double t;
  uint32_t prim;  /* 0xFFFFFFFF = miss */
  uint32_t flags; /* bit0 hit, bit1 is_plane, bit2 front_face */
} rtg_hit;

enum { RTG_TRACE_CLOSEST = 0, RTG_TRACE_ANY = 1 };
typedef struct rtg_trace_opts {
  int32_t mode;      /* RTG_TRACE_CLOSEST | RTG_TRACE_ANY */
  int32_t use_bvh;   /* 0 = brute force over all primitives */
  int32_t precision; /* RTG_PRECISION_F32 | RTG_PRECISION_F64 */
  int32_t pad;
} rtg_trace_opts;

/* Device buffers for resident inputs/outputs. */
rtg_status rtg_buffer_create(rtg_context* ctx, uint64_t bytes, rtg_buffer** out);
rtg_status rtg_buffer_release(rtg_buffer* buf);
rtg_status rtg_buffer_write(rtg_context* ctx, rtg_buffer* buf, const void* src, uint64_t bytes);
rtg_status rtg_buffer_read(rtg_context* ctx, const rtg_buffer* buf, void* dst, uint64_t bytes);

/* Traces n rays held in `rays` (n*8 doubles) into `hits` (n rtg_hit). For
 * RTG_TRACE_ANY, hits[i].flags bit0 = occluded in (tmin, tmax). */
rtg_status rtg_trace(rtg_context* ctx, const rtg_scene* scene, const rtg_buffer* rays, uint64_t n,
                     const rtg_trace_opts* opts, rtg_buffer* hits, rtg_stats* stats);
/* Same, with host arrays (copies in and out; PCIe time included). */
rtg_status rtg_trace_host(rtg_context* ctx, const rtg_scene* scene, const double* rays, uint64_t n,
                          const rtg_trace_opts* opts, rtg_hit* hits, rtg_stats* stats);

rtg_status rtg_synchronize(rtg_context* ctx);

/* ------------------------------------------------------------------------ */
/* Per-kernel device timing of the render passes and traces issued since
 * profiling was switched on (a pair of HIP events per launch, recorded by
 * the kernel's dispatch at its start and end: hipExtLaunchKernel; batch
 * passes launch directly instead of replaying graphs while profiling is
 * on). kind: 0 generate, 1 extend (closest hit), 2 shade,
 * 3 shadow (any hit), 4 accumulate, 5 direct/normals sample, 6 trace,
 * 7 batch planning. */
enum { RTG_K_GENERATE = 0, RTG_K_EXTEND, RTG_K_SHADE, RTG_K_SHADOW, RTG_K_ACCUMULATE, RTG_K_DIRECT, RTG_K_TRACE,
       RTG_K_PLAN, RTG_K_COUNT };
typedef struct rtg_profile {
  uint64_t launches[RTG_K_COUNT];
  double ms[RTG_K_COUNT];
  /* work done by the closest-hit kernels (extend, trace, direct) and by the
   * any-hit kernel (shadow): rays, node boxes tested, primitive tests */
  uint64_t closest_rays, closest_visits, closest_tests;
  uint64_t shadow_rays, shadow_visits, shadow_tests;
} rtg_profile;
rtg_status rtg_set_profiling(rtg_context* ctx, int32_t on); /* on: every kind; clears the totals */
/* Times only the launches of the kinds in `kinds` (bit 1 << RTG_K_*; 0 =
 * off; clears the totals). The work totals are kept either way. Fewer
 * timed launches perturb a pass less: bench.py times the closest-hit
 * launches only. No reference counterpart (a measurement hook). */
rtg_status rtg_set_profiling_kinds(rtg_context* ctx, uint32_t kinds);
rtg_status rtg_get_profile(rtg_context* ctx, rtg_profile* out);

#ifdef __cplusplus
}
#endif
#endif /* RTG_H */
