// prt_stream.h — device state of the multi-scene streaming renderer.
//
// Many scenes render through one shared work queue of pixel tiles.  A tile
// is a run of consecutive pixels of one scene with all their camera samples
// (at most T = 1 << tile_shift samples).  A fixed pool of tile slots holds
// the tiles in flight: a slot owns T fixed-point sample accumulators and a
// count of the tile's live rays (queued path segments and shadow rays).
//
// Every iteration runs, on one stream:
//   refill  (1 block)  assign free slots to the next tiles of the work queue
//   gen                write the new tiles' unique sub-pixel camera rays after
//                      the rays carried over from the previous iteration
//   extend             closest hit of the whole queue
//   shade              one vertex per queued ray and per new camera sample;
//                      children go to the other queue
//   nee / shadow       light samples of the diffuse vertices, any-hit tests
//   resolve            tiles whose live count reached zero: fold samples into
//                      their scene's image, clear the slot, free it
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

// What a slot holds (written by refill).
struct SlotInfo {
  int32_t scene;
  int32_t tile;
  int64_t pixel_base;
};

// A tile assigned by this iteration's refill: its slot, and where its unique
// camera rays sit after the carried-over rays of the queue.
struct NewTile {
  uint32_t slot;
  uint32_t ray_base;
  uint32_t n_rays;     // n_pixels * SS^2
  uint32_t n_samples;  // n_pixels * spp
};

// Control block (device memory; the host reads it after each graph replay).
struct StreamCtl {
  uint32_t next_tile;   // first tile of the work queue not yet assigned
  uint32_t free_count;  // entries of free_list
  uint32_t done_count;  // tiles finished this iteration (resolve's input)
  uint32_t n_new;       // tiles assigned by this iteration's refill
  uint32_t n_cont;      // rays carried over at the head of the current queue
  uint32_t n_prim;      // unique camera rays appended after them
  uint32_t status;      // 0 running, 1 finished, 2 inconsistent state, 3 queue capacity
  uint32_t iter;        // iterations run
  uint32_t cnt[2];      // ray queue counts
  uint32_t cnt_sq;      // shadow queue count
  uint32_t fetch[2];    // work-fetch counters of extend and shadow
  uint32_t pad_;
  uint64_t total_prim;  // camera rays traversed
  uint64_t total_cont;  // bounce rays traversed
};

struct Stream {
  const Params* scenes;   // [n_scenes] per-scene parameters
  const TileDesc* tiles;  // [n_tiles] the work queue, in order
  float* const* out_rgb;  // [n_scenes] W*H*3 device images
  float* const* out_var;  // [n_scenes] (entries may be null)
  SlotInfo* slot_info;    // [n_slots]
  int32_t* live;          // [n_slots] live rays of the slot's tile
  uint32_t* free_list;    // [n_slots]
  uint32_t* done_list;    // [n_slots]
  NewTile* new_tiles;     // [n_slots]
  StreamCtl* ctl;
  uint64_t* accum;        // [n_slots << tile_shift][3] fixed point
  uint64_t* ctr;          // [CTR_COUNT]
  int32_t n_scenes, n_tiles, n_slots, tile_shift;
  uint32_t q_cap, sq_cap;
};

hipError_t launch_stream_init(const Stream& st, hipStream_t s);
hipError_t launch_stream_refill(const Stream& st, int cur, hipStream_t s);
hipError_t launch_stream_gen(const Stream& st, const RayQ& q, int blocks, hipStream_t s);
hipError_t launch_stream_extend(const Stream& st, const RayQ& q, int cur, double* hit_t, int32_t* hit_id,
                                int stack_lds, int blocks, hipStream_t s);
hipError_t launch_stream_shade(const Stream& st, const RayQ& qin, const double* hit_t, const int32_t* hit_id,
                               const RayQ& qout, int nxt, int blocks, hipStream_t s);
hipError_t launch_stream_nee(const Stream& st, const RayQ& q, int nxt, const ShadowQ& sq, int blocks, hipStream_t s);
hipError_t launch_stream_shadow(const Stream& st, const ShadowQ& sq, int stack_lds, int blocks, hipStream_t s);
hipError_t launch_stream_resolve(const Stream& st, int blocks, hipStream_t s);
