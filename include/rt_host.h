/*
 * rt_host.h — C ABI of the host library (librthost.so): JSON scenes,
 * Scene::compile and the GPU renderer behind one plain-C surface.
 *
 * The rt_* entry points are the reference's own foreign-function interface
 * for this path — the WebAssembly front end apps/rt_wasm.cpp:114-184
 * (rt_validate, rt_message, rt_threads, rt_running, rt_cancel, rt_start) —
 * with the JavaScript hooks Module.onFrame / Module.onDone replaced by C
 * callbacks (rt_wasm.cpp:67-108). The rth_* entry points expose compiled
 * scenes to tools and tests.
 */
#ifndef RT_HOST_H
#define RT_HOST_H

#include <stdint.h>

#include "rtg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rt_wasm.cpp:114-121: "" if the scene parses and validates, otherwise the
 * first problem found (string owned by the library). */
const char* rt_validate(const char* json);
/* rt_wasm.cpp:123-124: reason for the last failed rt_start. */
const char* rt_message(void);
/* rt_wasm.cpp:126-127: renders in flight per job (one device stream). */
int rt_threads(void);
/* rt_wasm.cpp:129-135 */
int rt_running(void);
/* rt_wasm.cpp:137-143: requests a stop; a final frame and on_done(cancelled)
 * follow. Returns immediately. */
void rt_cancel(void);

typedef struct rt_frame_info { /* RenderProgress, rt_wasm.cpp:77-84 */
  int32_t pass;
  uint32_t spp_done;
  int32_t spp_target;
  int32_t width, height;
  int32_t finished, cancelled;
  int32_t pad;
  double elapsed_ms;
  uint64_t rays;
} rt_frame_info;

typedef struct rt_done_info { /* RenderStatsSnapshot, rt_wasm.cpp:88-96 */
  int32_t cancelled;
  uint32_t spp_done;
  double elapsed_ms;
  uint64_t rays, shadow_rays, invalid_samples;
  double mrays_per_sec;
  int32_t threads;
  int32_t pad;
} rt_done_info;

/* rgba: width*height*4 bytes, valid for the duration of the call. */
typedef void (*rt_frame_fn)(void* user, int32_t job, const uint8_t* rgba, const rt_frame_info* info);
typedef void (*rt_done_fn)(void* user, int32_t job, const rt_done_info* info);

/* rt_wasm.cpp:145-182: starts render `job` on the GPU; max_pass caps the
 * samples per progressive pass (0 = 32). Frames and the final done record
 * arrive on the render thread. Returns 0, or 1 if the scene was rejected or
 * a render is still running (reason in rt_message()). */
int rt_start(int32_t job, const char* json, int32_t max_pass, rt_frame_fn on_frame, rt_done_fn on_done, void* user);
/* Blocks until the current render (if any) has finished. */
void rt_wait(void);

/* ---------------------------------------------------------------------- */
typedef struct rth_scene rth_scene;

/* CLI-style overrides (apps/rt_cli.cpp:119-135); negative = keep. */
typedef struct rth_overrides {
  int32_t width, height, spp, max_depth, rr_min_depth, integrator, light_sampling, use_bvh;
  int64_t seed;
  double exposure; /* <= 0 = keep */
} rth_overrides;

/* parse + overrides + validate + Scene::compile. 0 on success, else 1 with
 * rth_last_error() set (the reference's SceneError text). */
int rth_scene_load(const char* json, const rth_overrides* ov, rth_scene** out);
void rth_scene_free(rth_scene* s);
const char* rth_last_error(void);
/* View of the compiled arrays (owned by the handle). */
const rtg_scene_desc* rth_scene_desc(const rth_scene* s);
void rth_scene_settings(const rth_scene* s, int32_t precision, rtg_render_settings* out);
/* {triangles, spheres, planes, degenerate_skipped, lights, bvh nodes,
 *  bvh leaves, bvh depth, spp, tone_map} */
void rth_scene_info(const rth_scene* s, uint64_t out[10]);
double rth_scene_bvh_build_ms(const rth_scene* s);
/* scene_to_json of the parsed scene, compact; returns length or -needed. */
int64_t rth_dump_scene(const char* json, char* out, uint64_t cap);

typedef struct rth_render_stats {
  uint64_t camera_samples, rays, shadow_rays, rr_terminated, invalid_samples, node_visits, prim_tests, passes;
  double elapsed_ms, device_ms;
  uint32_t spp_done;
  int32_t cancelled;
} rth_render_stats;

/* Renders the scene's full spp with GpuRenderer (progressive passes capped
 * at max_pass, or one pass). Outputs may be NULL: sum (w*h*3 doubles),
 * count (w*h), rgb (w*h*3, tone mapped). */
int rth_render(rth_scene* s, int32_t device, int32_t precision, int32_t progressive, int32_t max_pass, double* sum,
               uint32_t* count, uint8_t* rgb, rth_render_stats* stats);

/* ---------------------------------------------------------------------- */
/* The served render loop: render requests over scenes resident on one GPU,
 * scheduled in ticks of batched passes (rtg_render_batch). Previews render
 * all samples of their pass in the next tick; finals advance their
 * progressive passes (1, 1, 2, 4, ... capped at max_pass, renderer.cpp:66-69)
 * in slices sized per tick (see rts_create); requests are admitted first come
 * first served while the HBM footprint — resident scenes, the path state and
 * queues of the server's context, the admitted requests' accumulation
 * buffers — fits budget_bytes. It has no
 * counterpart in the reference, whose editor backend starts one `rt` process
 * per job (web/backend/jobs.go:130-259). Errors: 1 + rth_last_error(). */
typedef struct rts_server rts_server;

typedef struct rts_request {
  int32_t scene; /* index returned by rts_add_scene */
  int32_t width, height, spp, max_depth, max_pass;
  int32_t final_kind; /* 0 preview, 1 final */
  int32_t pad;
  uint64_t seed;
  double arrival_ms; /* from the start of the run; non-decreasing */
} rts_request;

typedef struct rts_result {
  double arrival_ms, start_ms, first_frame_ms, done_ms; /* from the start of the run */
  uint64_t rays, shadow_rays;
  uint32_t frames, spp_done;
  uint64_t image_hash; /* FNV-1a over 64-bit words: every final sum (double
                          bits), then every count widened to 64 bits */
} rts_result;

typedef struct rts_stats {
  double wall_ms;
  uint64_t rays; /* closest-hit + shadow rays */
  uint32_t ticks, max_active;
  uint64_t peak_accum_bytes;
  double gpu_ms;                 /* wall time inside render calls */
  uint64_t fixed_bytes;          /* resident scenes + path state and queues, at the end */
  uint64_t peak_footprint_bytes; /* the largest footprint admission saw */
} rts_stats;

/* budget_bytes 0: 64 GiB; tick_paths: final-render paths per tick, 0 = tuned
 * (each tick aimed at 7 ms with previews in it, 16 ms without, from the
 * scenes' measured cost per path). */
int rts_create(int32_t device, uint64_t budget_bytes, uint32_t tick_paths, rts_server** out);
/* Replaces the footprint budget (0: 64 GiB) for the next rts_run. */
int rts_set_budget(rts_server* s, uint64_t budget_bytes);
/* Uploads a compiled scene (kept resident); its render settings supply the
 * tone map, exposure, roulette depth and light sampling of every request. */
int rts_add_scene(rts_server* s, const rth_scene* scene, int32_t* index);
/* Replays n requests in real time: served = 1 for the tick scheduler, 0 for
 * the plain loop (one request at a time through rtg_render_pass). Blocks
 * until every request is done. */
int rts_run(rts_server* s, const rts_request* reqs, uint32_t n, int32_t served, rts_result* results,
            rts_stats* stats);
void rts_free(rts_server* s);

#ifdef __cplusplus
}
#endif
#endif /* RT_HOST_H */
