// prt_device.h — device-side data layout shared by the host builder and the
// HIP kernels of libprt.
//
// Layout in HBM (all arrays 256-byte aligned):
//   nodes   : DevNode[n_nodes], 128 B each — a 4-wide BVH node with the child
//             boxes as six float4 (lo x/y/z, hi x/y/z over the 4 children) so
//             one node is eight 16-byte loads;
//   tris    : float4[3*n_tris], 48 B per triangle in leaf order:
//             (v0.xyz, entity), (v1.xyz, face), (v2.xyz, unused) — the float32
//             vertices of the reference's Obj storage (ObjLoader.h:39-75);
//   tri_nrm : float[3*n_tris] f32 face normal (ObjLoader.cpp:577-700);
//   tri_alb : float[3*n_tris] per-face albedo or null;
//   entities, materials, lights: small tables read through the scalar cache.
#pragma once
#include <stdint.h>

#ifndef PRT_MAX_LEAF
#define PRT_MAX_LEAF 8
#endif
// a leaf child stores count-1 in bits 27..30 and is negative after ~
static_assert(PRT_MAX_LEAF >= 1 && PRT_MAX_LEAF <= 16, "leaf size must fit the 4-bit count of a child code");
// LDS traversal-stack entries of the per-scene kernels (12: five 256-thread
// blocks of a traversal kernel fit one CU's LDS)
#ifndef PRT_STACK_LDS
#define PRT_STACK_LDS 12
#endif
// the BVH's top levels (at most 1 + 4 + 16 nodes) are numbered first and
// copied into LDS by every traversal block
#ifndef PRT_TOP_LEVELS
#define PRT_TOP_LEVELS 3
#endif
#ifndef PRT_TOP_NODES
#define PRT_TOP_NODES 21  // 1 + 4 + 16
#endif
// smallest LDS stack depth any kernel variant uses (sizes the overflow area)
#define PRT_STACK_LDS_MIN 8
#define PRT_STACK_TOTAL 128
// traversal-stack overflow areas per thread (closest-hit and shadow kernels
// run concurrently); each holds the entries beyond the smallest LDS stack
// that the scene's BVH depth can need (ovf_words_for)
#define PRT_OVF_REGIONS 2
inline int ovf_words_for(int bvh_depth) {
  const int entries = 3 * bvh_depth + 1 - PRT_STACK_LDS_MIN;  // deepest stack of a 4-wide BVH
  return 2 * (entries > 1 ? entries : 1);
}

struct __attribute__((aligned(16))) DevNode {
  float lo_x[4], lo_y[4], lo_z[4];
  float hi_x[4], hi_y[4], hi_z[4];
  // >= 0: internal node index; < 0: leaf, ~child = first | (count-1) << 27
  int32_t child[4];
  int32_t pad[4];
};
static_assert(sizeof(DevNode) == 128, "node record must be 128 bytes");

static inline int32_t prt_leaf_code(uint32_t first, uint32_t count) {
  return (int32_t)~(first | ((count - 1u) << 27));
}

struct DevEntity {
  int32_t type;      // PRT_ENTITY_*
  int32_t material;
  int32_t ninv;      // normal_vector_inverse
  int32_t light;     // material isLight()
  int32_t has_albedo;// POLYGON with per-face albedo in tri_alb
  int32_t pad_;
  double p0[3];      // sphere position / uvplane org
  double radius;
  double u_axis[3], v_axis[3], normal[3];  // uvplane (normalised, normal = u x v)
  double u_len, v_len, area;
};

// Fourth word of a triangle record's third float4: the entity's material
// index and flags (the first two hold the entity and face indices).
#define TRI_SHADE_MAT_MASK 0x00FFFFFFu
#define TRI_SHADE_NINV 0x01000000u    // normal_vector_inverse < 0
#define TRI_SHADE_LIGHT 0x02000000u   // material isLight()
#define TRI_SHADE_ALBEDO 0x04000000u  // per-face albedo in tri_alb

struct DevMaterial {
  double color[3];
  double emission[3];
  double roughness;
  double ior;
  int32_t type;
  int32_t is_light;
};

// Everything a kernel needs about the scene and the render settings.
struct Params {
  // camera (CameraScreenPram, render.h:100-146)
  double cam_pos[3], screen_x[3], screen_y[3], center[3];
  double iw, ih;
  int32_t W, H, SS, samples;
  int32_t spp, depth, depth_limit, nee;
  int32_t max_vertices, n_lights, n_analytic, n_nodes;
  uint32_t key0, key1;
  double smp, sensor_response;
  float cutoff;
  int32_t n_materials;
  double background[3], env_light[3];
  // scene arrays
  const DevNode* nodes;
  const float* tris;         // read as float4 on the device
  const float* tri_nrm;
  const float* tri_alb;
  const float* tri_alb_rr;   // albedo for Russian-roulette probabilities (null: tri_alb)
  const DevEntity* ents;
  const DevMaterial* mats;
  const int32_t* analytic;   // entity indices of spheres / uvplanes
  const int32_t* lights;     // light entity indices, power-descending
  uint32_t* stack_ovf;       // per-thread traversal stack overflow
  int32_t ovf_threads;
  int32_t n_tris;
  // Metropolis renders (null otherwise): the batch's "pixels" are Markov
  // chains; chain c's primary samples of the current proposal are
  // mlt_u[c * mlt_stride ...], MLT_CAMERA_WORDS camera words, then 4 words
  // per path vertex for the first mlt_depths vertices; the proposal's image
  // element goes to mlt_pix[c]
  const uint32_t* mlt_u;
  uint32_t* mlt_pix;
  uint32_t* mlt_flags;  // per chain: 1 if the camera ray returned a light directly (ERPT), or null
  int32_t mlt_stride;
  int32_t mlt_depths;
  int32_t ovf_words;    // words of one thread's stack overflow area
  int32_t n_top;        // nodes 0..n_top-1 are the top PRT_TOP_LEVELS BVH levels (kept in LDS)
};
// camera words of a Metropolis proposal: pixel x, y, sub-pixel x, y, tent
// filter x, y, two spare (render.h:262-300)
#define MLT_CAMERA_WORDS 8
