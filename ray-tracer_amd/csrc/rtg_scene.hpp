// rtg_scene.hpp — device-resident scene layout shared by the host upload
// code (rtg_api.cpp) and the kernels.
//
// FP32 traversal data (the fast path):
//   node2  — one 32-byte record per INTERIOR node of the reference BVH, in
//            the reference's DFS order: the boxes of both children as half
//            floats in the scene's node frame ((b - frame_o) / frame_s,
//            rounded outward from the double boxes) and two child refs.
//            Testing both children from one record halves the dependent
//            loads per level compared with the reference's one-box nodes;
//            half floats halve the bytes each level moves through the
//            vector memory pipeline, which bounds the traversal kernels.
//   slot32 — primitives in leaf order (slot k = prim_indices[k]), nine
//            floats (36 bytes): triangle {v0} {e1} {e2}; sphere {centre}
//            {kSlotSphere, r^2, 0} {0}. slot_prim maps a slot to its
//            primitive id (read once per ray, for the closest hit).
// FP64 traversal data (the exact mode): the reference's own node and
//   primitive arrays, uploaded unchanged.
// Shading data is indexed by primitive id.
#pragma once
#include <stdint.h>

#include "../../include/rtg.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
struct float4 {
  float x, y, z, w;
};
#endif

namespace rtg {

// Child reference: bit31 = leaf; interior: node2 index;
// leaf: bits 30..25 = count-1 (1..64 primitives), bits 24..0 = first slot.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeafCount = 64;
constexpr uint32_t kMaxSlots = 1u << 25;
__host__ __device__ inline uint32_t leaf_ref(uint32_t first, uint32_t count) {
  return kLeafBit | ((count - 1) << 25) | first;
}
__host__ __device__ inline uint32_t leaf_first(uint32_t ref) { return ref & 0x1FFFFFFu; }
__host__ __device__ inline uint32_t leaf_count(uint32_t ref) { return ((ref >> 25) & 63u) + 1; }

// A sphere slot's fourth word: a quiet NaN whose payload no triangle edge
// carries (the host writes a triangle's NaN edge component as the default
// NaN, which tests the same: it never hits).
constexpr uint32_t kSlotSphere = 0x7FC5F1E3u;

// shade32 flags in the material word
constexpr uint32_t kShadeSmooth = 1u << 30;
constexpr uint32_t kShadeSphere = 1u << 31;

enum : uint32_t { kL32Point = 0, kL32Sphere = 1, kL32Triangle = 2 };
constexpr int kLight32Stride = 5;  // float4 per light
constexpr int kShade32Stride = 4;  // float4 per primitive
constexpr int kSlot32Floats = 9;   // floats per slot
constexpr int kNode2Stride = 2;    // float4 per interior node

struct DevScene {
  // FP32 traversal
  const float4* node2;
  const float* slot32;
  const uint32_t* slot_prim;  // slot -> primitive id
  uint32_t root_ref;  // child ref of the root (interior: 0)
  uint32_t has_bvh;
  float root_lo[3], root_hi[3];  // the root box in the node frame (half values)
  float frame_o[3], frame_s;     // node frame: world = frame_o + frame_s * frame coordinate
  uint32_t stack_depth;  // BVH depth + 1
  uint32_t n_slots;
  // FP64 (reference layout)
  const rtg_bvh_node* nodes;
  const uint32_t* prim_indices;
  const rtg_primitive* prims;
  uint32_t n_nodes;
  uint32_t n_prims;
  // shading
  const float4* shade32;
  const uint32_t* prim_slot;
  const uint8_t* sampled;
  const rtg_material* materials;
  const float4* mat32;  // 2 per material: {c0, kind | two_sided<<8 | checker<<9} {c1, p0}
  const rtg_plane* planes;
  const float4* plane32;  // 2 per plane: {point, material} {normal, 0}
  const rtg_light_ref* lights;
  const float4* light32;
  const rtg_point_light* point_lights;
  uint32_t n_planes, n_lights, n_materials, pad0;
  rtg_camera cam;
  rtg_background bg;
  // FP32 camera: origin, corner - origin, du, dv, u, v; lens radius
  float cam32[19];
  float bg32[7];  // top.xyz, bottom.xyz, intensity
  uint32_t bg_solid;
};

}  // namespace rtg
