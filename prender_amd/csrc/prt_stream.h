// prt_stream.h — device state of the multi-scene streaming renderer.
//
// Many scenes render through one shared work queue of pixel tiles.  A tile
// is a run of consecutive pixels of one scene with all their camera samples
// (at most T = 1 << tile_shift samples).  A fixed pool of tile slots holds
// the tiles in flight: a slot owns T fixed-point sample accumulators and a
// count of the tile's live rays (queued path segments and shadow rays).
//
// Ray queues are binned by scene: each iteration, refill lays the next queue
// out as one region per scene, sized by an exact bound on the children that
// scene's vertices can spawn (two per vertex with glass, else one), and the
// kernels walk a padded prefix table so that every 64-wide chunk of work —
// one wave — holds rays of one scene only and reads that scene's parameters
// with scalar loads.
//
// Every iteration runs, on one stream:
//   refill  (1 block)  assign free slots to the next tiles of the work queue;
//                      lay out the next queue's scene regions; build the
//                      current queue's work table
//   gen                unique sub-pixel camera rays of the new tiles, after the
//                      scene regions of the current queue
//   extend             closest hit of the current queue
//   shade              one vertex per queued ray and per new camera sample;
//                      children go to their scene's region of the next queue
//   scan    (1 block)  prefix table of the next queue's regions (NEE input)
//   nee / shadow       light samples of the diffuse vertices, any-hit tests
//   resolve            tiles with no live ray left: fold samples into their
//                      scene's image, clear the slot, free it
// so the GPU always holds rays from as many tiles as there are slots, of
// whatever scenes those tiles belong to.  Two iterations (one per queue
// parity) are captured as one hipGraph and replayed until refill reports that
// the work queue is drained and no tile is in flight.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"

// One work item: pixels [pixel_base, pixel_base + n_pixels) (y*W + x order) of
// scene `scene`, all spp samples of each.
struct TileDesc {
  int32_t scene;
  int32_t n_pixels;
  int64_t pixel_base;
};

// What a slot holds (written by refill; tile < 0: free).  spp is copied from
// the scene so that a ray's pixel and sample index are known before the
// scene's parameters are read.
struct SlotInfo {
  int32_t scene;
  int32_t tile;
  int64_t pixel_base;
  int32_t spp;
  int32_t pad_;
};

// A tile assigned by this iteration's refill: its slot, and where its unique
// camera rays sit in the primary region of the current queue (ray_base is a
// multiple of 64).
struct NewTile {
  uint32_t slot;
  uint32_t ray_base;
  uint32_t n_rays;     // n_pixels * SS^2
  uint32_t n_samples;  // n_pixels * spp
  uint32_t spp;
  uint32_t ss2;        // SS^2 camera rays per pixel
  int32_t scene;
  uint32_t pad_;
};

// One entry of a work table: `cnt` rays of one scene at queue positions
// [off, off + cnt), occupying roundup64(cnt) entries of the padded index space.
struct WorkEntry {
  uint32_t off;
  uint32_t cnt;
  int32_t scene;
  uint32_t pad_;
};

// Control block (device memory; the host reads it after each graph replay).
struct StreamCtl {
  uint32_t next_tile;      // first tile of the work queue not yet assigned
  uint32_t free_count;     // entries of free_list
  uint32_t n_new;          // tiles assigned by this iteration's refill
  uint32_t n_cont;         // rays carried over in the current queue's scene regions
  uint32_t n_prim;         // unused (gen adds the new tiles' camera rays to total_prim)
  uint32_t status;         // 0 running, 1 finished, 2 inconsistent state, 3 queue capacity
  uint32_t iter;           // iterations run
  uint32_t cnt_sq;         // shadow queue count
  uint32_t fetch[4];       // work-fetch counters of extend, shadow, shade and NEE
  uint32_t region_end[2];  // end of the scene regions of queue p (its primary region starts there)
  uint32_t work_width;     // padded index space of the current queue's work table
  uint32_t cont_width;     // ... of its scene regions alone (the first n_scenes entries)
  uint32_t nee_width;      // padded index space of the next queue's scene regions
  uint32_t n_work;         // entries of the work table (n_scenes + n_new)
  uint64_t total_prim;     // camera rays traversed
  uint64_t total_cont;     // bounce rays traversed
};

struct Stream {
  const Params* scenes;     // [n_scenes] per-scene parameters
  const TileDesc* tiles;    // [n_tiles] the work queue, in order
  float* const* out_rgb;    // [n_scenes] W*H*3 device images
  float* const* out_var;    // [n_scenes] (entries may be null)
  const int32_t* scene_fan; // [n_scenes] most children one vertex can spawn (2 with glass, else 1)
  SlotInfo* slot_info;      // [n_slots]
  int32_t* live;            // [n_slots] live rays of the slot's tile
  uint32_t* free_list;      // [n_slots]
  NewTile* new_tiles;       // [n_slots]
  uint32_t* seg_off;        // [2][n_scenes] region start of scene s in queue p
  uint32_t* seg_cap;        // [2][n_scenes] region capacity
  uint32_t* seg_cnt;        // [2][n_scenes] rays appended to the region
  uint32_t* new_samp;       // [n_scenes] camera samples of this iteration's new tiles
  WorkEntry* work;          // [n_scenes + n_slots] work table of the current queue
  uint32_t* work_pref;      // [n_scenes + n_slots + 1] its padded prefix
  uint32_t* nee_pref;       // [n_scenes + 1] padded prefix of the next queue's scene regions
  StreamCtl* ctl;
  uint64_t* accum;          // [n_slots << tile_shift][3] fixed point
  uint64_t* ctr;            // [CTR_COUNT]
  int32_t n_scenes, n_tiles, n_slots, tile_shift;
  uint32_t q_cap, sq_cap;
};

hipError_t launch_stream_init(const Stream& st, hipStream_t s);
hipError_t launch_stream_refill(const Stream& st, int cur, hipStream_t s);
hipError_t launch_stream_gen(const Stream& st, const RayQ& q, int cur, int blocks, hipStream_t s);
hipError_t launch_stream_extend(const Stream& st, const RayQ& q, double* hit_t, int32_t* hit_id, int stack_lds,
                                int blocks, hipStream_t s);
hipError_t launch_stream_shade(const Stream& st, const RayQ& qin, const double* hit_t, const int32_t* hit_id,
                               const RayQ& qout, int nxt, int blocks, hipStream_t s);
hipError_t launch_stream_scan(const Stream& st, int nxt, hipStream_t s);
hipError_t launch_stream_nee(const Stream& st, const RayQ& q, int nxt, const ShadowQ& sq, int blocks, hipStream_t s);
hipError_t launch_stream_shadow(const Stream& st, const ShadowQ& sq, int stack_lds, int blocks, hipStream_t s);
hipError_t launch_stream_resolve(const Stream& st, int blocks, hipStream_t s);
