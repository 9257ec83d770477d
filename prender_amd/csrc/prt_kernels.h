// prt_kernels.h — queue layouts and kernel launchers of libprt.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"

// counters (uint64) accumulated by the kernels
enum {
  CTR_NODES = 0,   // BVH nodes fetched
  CTR_TRIS = 1,    // triangle tests
  CTR_EXT = 2,     // extension rays queued
  CTR_SHADOW = 3,  // shadow rays queued
  CTR_PICKUP = 4,  // specular pickups taken from a child hit
  CTR_TRUNC = 5,   // children dropped by the vertex cap
  CTR_DROP = 6,    // appends beyond a queue's capacity (must stay 0)
  CTR_MIXED = 7,   // batched renders: extra passes of waves holding rays of several scenes
  CTR_SNODES = 8,  // BVH nodes fetched by shadow rays
  CTR_STRIS = 9,   // triangle tests of shadow rays
  CTR_REPLAY_MISMATCH = 10,  // adjoint: samples whose replayed radiance differs from the primal
  CTR_TRACED = 11,   // closest-hit queries finished by k_extend (must equal the rays queued)
  CTR_STRACED = 12,  // shadow rays finished by k_shadow (must equal CTR_SHADOW)
  CTR_QUEUED = 13,   // rays of the ray queues listed for k_extend (primary + extension)
  CTR_MLT_ACCEPT = 14,  // Metropolis proposals accepted
  CTR_COUNT = 15
};

// A path segment a lane of the persistent renderer keeps for later: a glass
// vertex at depth <= 2 emits two children; the lane follows the first and
// keeps the second here.  A lane's path is a binary tree split at depths 0,
// 1 and 2 walked depth first, so at most one kept segment per split depth.
struct PendRec {
  double o[3], d[3];
  float t[3], p[3], n[3];
  uint32_t meta;
  uint32_t pad_[2];
};
static_assert(sizeof(PendRec) == 96, "pending segment record must be 96 bytes");
#define PRT_PEND_DEPTH 3

struct prt_ray_dev {
  double org[3];
  double dir[3];
};
struct prt_hit_dev {
  double t, u, v;
  int32_t entity, prim;
};

// Structure-of-arrays ray queue (one entry = one path segment to trace).
struct RayQ {
  double *ox, *oy, *oz, *dx, *dy, *dz;
  float *tr, *tg, *tb;   // path throughput
  float *pr, *pg, *pb;   // pickup weight (mirror / glass parent with NEE)
  uint32_t* slot;        // batch-local camera sample
  uint32_t* meta;        // depth | branch << 8 | META_NEE
  float *nx, *ny, *nz;   // orienting normal of a diffuse parent (NEE input)
  uint32_t cap;          // capacity (entries)
};

// meta bit: the ray leaves a diffuse vertex whose next-event estimation is
// still to be done (by k_nee, from the ray's origin, throughput and nx/ny/nz)
#define META_NEE 0x80000000u
#define META_BRANCH(m) (((m) >> 8) & 0xFFFFu)

// NEE shadow rays: nudged origin, direction, the light's own hit distance,
// and the contribution added when nothing nearer blocks it.
struct ShadowQ {
  double *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
  float *cr, *cg, *cb;
  uint32_t* slot;
  int32_t* light;        // light entity sampled
  uint32_t cap;
};

// Path-replay adjoint (inactive when on == 0).  The replay re-traces the
// primal paths with the same random numbers; at a vertex of `entity` the
// derivative of the sample radiance with respect to the face albedo is the
// radiance the path still gathers from there on, minus the vertex's own
// emission, divided by the albedo.
struct Adjoint {
  int32_t on;
  int32_t entity;             // POLYGON entity whose per-face albedo is differentiated
  int32_t light;              // light entity whose emission is differentiated
  int32_t pad_;
  const float* dimage;        // dLoss/dimage, W*H*3 in image row order
  const uint64_t* ltot;       // primal sample radiance (fixed point)
  double* grad_alb;           // n_faces*3 of `entity`
  double* grad_e;             // 3
  double scale;               // d image / d sample radiance = sensor_response * smp
};

// A batch: pixels [pixel_base, pixel_base + n_pixels) of the y*W + x order,
// all spp samples of each.
struct Batch {
  int64_t pixel_base;
  int64_t n_pixels;
};

// Slab queues of the per-scene renderer.  k_shade and k_nee run
// PRT_SLAB_WAVES waves, each over a static range [c0, c1) of input chunks;
// wave w writes its outputs to its slab, entries from fan * 64 * c0 on (fan:
// most outputs per input entry), counting them in LDS, and stores the slab
// (start, count) once at its end, so no counter is shared between waves
// (same-address atomics serialise across the device).  k_slab_index
// turns the counts into descriptors of 64-entry chunks, none spanning two
// slabs, which the consuming kernels walk: k_shade and k_nee in static
// per-wave ranges, k_extend and k_shadow through PRT_FETCH_COUNTERS claim
// counters (each 128 bytes apart, one sub-range of the chunks each).
#ifndef PRT_SLAB_WAVES
#define PRT_SLAB_WAVES 4096  // a multiple of 4, at most PRT_SLAB_INDEX_MAX
#endif
#define PRT_SLAB_INDEX_MAX 8192
// slabs (or slab writers) listed by one block of the chunk-list kernels; one
// wave scans them, so at most 64
#define PRT_INDEX_SLABS 64
static_assert(PRT_INDEX_SLABS <= 64, "one wave scans a block's slabs");
#define PRT_FETCH_COUNTERS 64
#define PRT_FETCH_STRIDE 32
struct ChunkList {
  const uint2* desc;   // [c] = (queue index of the chunk's first entry, its entries 1..64)
  const uint32_t* n;   // number of chunks (device)
};

hipError_t launch_trace(const Params& P, const prt_ray_dev* rays, int64_t n, prt_hit_dev* hits, uint64_t* ctr,
                        int max_blocks, hipStream_t s);
// Words of the library's file-scope device array that holds the counters of
// the allocation-free trace entry point (one copy per GPU).
#define PRT_NOALLOC_SCRATCH_WORDS 32
#define PRT_NOALLOC_MAX_DEVICES 64
static_assert(CTR_COUNT <= PRT_NOALLOC_SCRATCH_WORDS, "trace counters must fit the allocation-free scratch");
// Device address of that array on `device`, which must be current: looked
// up with hipGetSymbolAddress on the first call for a device and kept.
hipError_t noalloc_scratch(int device, uint64_t** out);
hipError_t launch_gen_primary(const Params& P, const Batch& B, const RayQ& q, hipStream_t s);
hipError_t launch_slab_index(const uint2* slabs, uint32_t n_slabs, uint2* desc, uint32_t max_chunks, uint32_t* n_chunks,
                             uint32_t* n_entries, uint64_t* ctr, int ctr_entries, hipStream_t s);
hipError_t launch_extend(const Params& P, const RayQ& q, const ChunkList& cl, int64_t n_host, uint32_t* fetch,
                         double* hit_t, int32_t* hit_id, uint64_t* ctr, int max_blocks, hipStream_t s);
hipError_t launch_shade(const Params& P, const Batch& B, int primary, const RayQ& qin, const ChunkList& cl,
                        const double* hit_t, const int32_t* hit_id, const RayQ& qout, uint32_t fan, uint2* slabs,
                        uint64_t* accum, const Adjoint& A, uint64_t* ctr, hipStream_t s);
hipError_t launch_nee(const Params& P, const Batch& B, const RayQ& q, const ChunkList& cl, const ShadowQ& sq,
                      uint2* slabs, uint64_t* ctr, hipStream_t s);
hipError_t launch_shadow(const Params& P, const Batch& B, const ShadowQ& q, const ChunkList& cl, uint32_t* fetch,
                         uint64_t* accum, const Adjoint& A, uint64_t* ctr, int max_blocks, hipStream_t s);
hipError_t launch_resolve(const Params& P, const Batch& B, const uint64_t* accum, void* rgb, void* var, int f64,
                          hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
hipError_t launch_fill_f32(float* p, float v, int64_t n, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s);
hipError_t launch_scatter_albedo(float* tri_alb, const uint32_t* slot_of_tri, int64_t first, int64_t n_faces,
                                 const float* albedo, hipStream_t s);
hipError_t launch_grad_to_f32(const double* g, float* out, int64_t n, hipStream_t s);
hipError_t launch_count_mismatch(const uint64_t* a, const uint64_t* b, int64_t n, uint64_t* ctr, hipStream_t s);
