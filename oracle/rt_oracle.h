/*
 * rt_oracle.h — the oracle: a plain-C, FP64 restatement of the reference
 * path tracer's hot path (BVH build, traversal, shapes, materials, path /
 * direct / normals integrators, tile renderer, tone map) and of the
 * Scene::compile tessellation it consumes.
 *
 * THIS IS TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only as the checker; the
 * product (librtg.so, librthost.so, rt) never links or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference checkout). Arithmetic is written in the reference's
 * evaluation order and built with -ffp-contract=off, so results are
 * bit-identical to the reference build on the same libm.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#include "../include/rtg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: math/rng.hpp:10-51 ---- */
typedef struct orc_rng { uint64_t state, inc; } orc_rng;
uint64_t orc_splitmix64(uint64_t x);
void orc_rng_init(orc_rng* r, uint64_t seed, uint64_t stream);
void orc_rng_for_sample(orc_rng* r, uint64_t seed, uint64_t pixel, uint64_t sample);
uint32_t orc_rng_next_u32(orc_rng* r);
double orc_rng_uniform(orc_rng* r);
/* Fills out[n] with successive next_u32() of for_sample(seed, pixel, sample). */
void orc_rng_stream(uint64_t seed, uint64_t pixel, uint64_t sample, uint32_t* out, int n);

/* ---- BVH: src/geometry/bvh.cpp ---- */
typedef struct orc_bvh_stats {
  uint64_t nodes, leaves, max_depth, max_leaf_size;
  double sah_cost;
} orc_bvh_stats;

/* bounds: n * 6 doubles (lo xyz, hi xyz). nodes_out needs room for 2n-1
 * nodes, prim_out for n indices. Returns 0, or -1 if a leaf exceeds 0xFFFF. */
int orc_bvh_build(const double* bounds, uint32_t n, int max_leaf, rtg_bvh_node* nodes_out, uint32_t* n_nodes,
                  uint32_t* prim_out, uint32_t* n_prim_out, orc_bvh_stats* stats);
int orc_bvh_validate(const rtg_bvh_node* nodes, uint32_t n_nodes, const uint32_t* prim_indices,
                     const double* bounds, uint32_t n);
void orc_prim_bounds(const rtg_primitive* p, double out[6]);
int orc_box_hit(const double lo[3], const double hi[3], const double o[3], const double d[3], double tmin,
                double tmax);

/* ---- Scene queries: src/core/scene.cpp:139-234 ---- */
typedef struct orc_hit {
  double t;
  double p[3];
  double normal[3];
  double gn[3];
  int32_t front_face;
  uint32_t material;
  uint32_t primitive;
  int32_t is_plane;
} orc_hit;

int orc_intersect(const rtg_scene_desc* s, int use_bvh, const double o[3], const double d[3], double tmin,
                  double tmax, orc_hit* rec, uint64_t* node_visits, uint64_t* prim_tests);
int orc_occluded(const rtg_scene_desc* s, int use_bvh, const double o[3], const double d[3], double tmax,
                 uint64_t* node_visits, uint64_t* prim_tests);
/* Batch: rays n*8 doubles {o, d, tmin, tmax}; hits in the rtg_hit layout.
 * mode 0 closest (tmin/tmax from the ray), 1 any-hit (occluded uses its own
 * tmin = 1e-6, as Scene::occluded). */
void orc_trace_batch(const rtg_scene_desc* s, int mode, int use_bvh, const double* rays, uint64_t n,
                     rtg_hit* hits, uint64_t* node_visits, uint64_t* prim_tests);

/* ---- Camera: core/camera.hpp:29-62 ---- */
void orc_camera_setup(const double look_from[3], const double look_at[3], const double vup[3], double vfov_deg,
                      double aperture, double focus_dist, int width, int height, rtg_camera* out);
void orc_camera_generate(const rtg_camera* c, int i, int j, double sx, double sy, orc_rng* rng, double o[3],
                         double d[3]);

/* ---- Integrators + renderer: src/rendering/{integrator,renderer}.cpp ---- */
void orc_radiance(const rtg_scene_desc* s, const rtg_render_settings* cfg, const double o[3], const double d[3],
                  orc_rng* rng, double L[3], rtg_stats* st);
/* sample_direct_light (integrator.cpp:205-209) at the closest hit of a ray. */
int orc_sample_direct_light_at(const rtg_scene_desc* s, const double o[3], const double d[3], uint64_t seed,
                               double L[3]);
/* render_tile over rows [y0, y1) for samples [s0, s1): adds into sum
 * (w*h*3) and count (w*h) exactly as Framebuffer::add in sample order. */
void orc_render_rows(const rtg_scene_desc* s, const rtg_render_settings* cfg, uint32_t s0, uint32_t s1, int y0,
                     int y1, double* sum, uint32_t* count, rtg_stats* st);
/* The same, but writes each sample's radiance (after the invalid-sample
 * guard) to out[((y*w + x) * (s1-s0) + (s-s0)) * 3 + c]. */
void orc_sample_radiance_rows(const rtg_scene_desc* s, const rtg_render_settings* cfg, uint32_t s0, uint32_t s1,
                              int y0, int y1, double* out, rtg_stats* st);

/* Seeded ray set over twice the BVH root box (see orc_render.c). */
void orc_ray_set(uint64_t seed, uint64_t n, const double lo[3], const double hi[3], double* rays);

/* ---- Tone map: src/rendering/tonemap.cpp ---- */
void orc_tone_map(const double in[3], int op, double exposure, double out[3]);
uint8_t orc_linear_to_srgb8(double v);
void orc_resolve(const double* sum, const uint32_t* count, int w, int h, int op, double exposure, uint8_t* rgb);

/* ---- Scene compile: src/core/scene.cpp:31-137, src/geometry/mesh.cpp ---- */
enum { ORC_OBJ_SPHERE = 0, ORC_OBJ_PLANE, ORC_OBJ_TRIANGLE, ORC_OBJ_QUAD, ORC_OBJ_BOX, ORC_OBJ_MESH };
enum { ORC_MESH_ICOSPHERE = 0, ORC_MESH_TORUS = 1 };
typedef struct orc_object {
  int32_t kind;
  uint32_t material;
  double a[3], b[3], c[3]; /* sphere: centre; plane: point, normal; triangle: v0 v1 v2;
                              quad: corner u v; box: min max */
  double scalar;           /* sphere radius; box rotate_y */
  /* mesh */
  int32_t mesh_source, subdivisions, segments_u, segments_v, smooth, pad;
  double major_radius, minor_radius, rotate_y;
  double scale[3], translate[3];
} orc_object;

typedef struct orc_compiled {
  rtg_primitive* prims;
  uint32_t n_prims;
  uint8_t* is_sampled_light;
  rtg_plane* planes;
  uint32_t n_planes;
  rtg_light_ref* lights;
  uint32_t n_lights;
  rtg_bvh_node* nodes;
  uint32_t n_nodes;
  uint32_t* prim_indices;
  uint32_t n_prim_indices;
  uint64_t degenerate_skipped;
  orc_bvh_stats bvh;
} orc_compiled;

/* materials[] is needed for the emissive light list. Returns 0 on success. */
int orc_compile(const orc_object* objs, uint32_t n_objs, const rtg_material* materials, uint32_t n_materials,
                uint32_t n_point_lights, orc_compiled* out);
void orc_compiled_free(orc_compiled* c);

/* Procedural meshes (mesh.cpp:45-88): positions/normals as n*3 doubles,
 * faces n*3 u32; caller frees with orc_free. */
int orc_make_torus(double major, double minor, int su, int sv, double** pos, double** nrm, uint32_t* n_vert,
                   uint32_t** faces, uint32_t* n_faces);
int orc_make_icosphere(int subdivisions, double** pos, double** nrm, uint32_t* n_vert, uint32_t** faces,
                       uint32_t* n_faces);
void orc_free(void* p);

/* tests/test_bvh.cpp:11-30 random_scene, for BVH ≡ brute-force checks. */
uint32_t orc_random_scene_objects(int spheres, int triangles, uint64_t seed, int with_mesh, orc_object* out);

#ifdef __cplusplus
}
#endif
#endif
