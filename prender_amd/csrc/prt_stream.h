// prt_stream.h — device state of the multi-scene streaming renderer.
//
// Many scenes render through one shared work queue of pixel tiles.  A tile
// is a run of consecutive pixels of one scene with all their camera samples
// (at most T = 1 << tile_shift samples).  A fixed pool of tile slots holds
// the tiles in flight: a slot owns T fixed-point sample accumulators and a
// count of the tile's live rays (queued path segments; shadow rays are not
// counted, see below).
//
// Queues are slab queues, as in the per-scene renderer: the shade and NEE
// kernels run PRT_SLAB_WAVES waves, and wave w writes its outputs into its
// own slab of the output queue, counting them in LDS (no counter is shared
// between waves).  A wave walks its input in runs of one scene and records
// each run (scene, slab range), so the index kernel can list the queue as
// 64-entry chunks of one scene each: every wave of a consuming kernel holds
// rays of one scene and reads that scene's parameters with scalar loads.
//
// Every iteration runs:
//   refill  (1 block)  assign free slots to the next tiles of the work queue;
//                      append the chunks of their camera rays to the current
//                      queue's chunk list; clear the claim counters of the
//                      extend launch and of this parity's shadow launch
//   gen                unique sub-pixel camera rays of the new tiles, in the
//                      primary region after the slabs
//   extend             closest hit of every chunk of the current queue
//   shade              one vertex per queued ray and per new camera sample;
//                      children go to the wave's slab of the next queue
//   index              chunk list of the next queue (NEE input, and the next
//                      iteration's extend / shade input)
//   nee / index / shadow  light samples of the diffuse vertices into shadow
//                      slabs, their chunk list, any-hit tests
//   resolve            tiles with no live ray left: fold samples into their
//                      scene's image, clear the slot, free it
// so the GPU always holds rays from as many tiles as there are slots, of
// whatever scenes those tiles belong to.  The nee / index / shadow kernels
// of iteration i may run on a second stream, after iteration i's shade and
// alongside iteration i+1's refill / gen / extend; resolve of iteration i
// then waits for the shadow kernel of iteration i-1 only.  That is enough:
// a tile whose live count is zero after shade i has no ray in the next queue,
// so NEE i gave it no shadow ray, and its last shadow rays came from NEE i-1
// or earlier.  Shadow rays therefore stay out of the live count, and the
// shadow claim counters have one region per queue parity (refill of
// iteration i clears parity i's region, whose previous user, shadow i-2,
// has finished).  Two iterations (one per queue parity) are launched at a
// time (directly, or replayed as one captured hipGraph) until refill reports
// that the work queue is drained and no tile is in flight.
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
// camera rays sit in the primary region (ray_base is a multiple of 64).
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

// A run of one scene's entries that a slab writer appended: queue entries
// [off, off + cnt).
struct RunRec {
  uint32_t off;
  uint32_t cnt;
  int32_t scene;
  uint32_t pad_;
};

// A chunk of a queue: entries [base, base + cnt), cnt <= 64, of one scene.
struct ChunkDesc {
  uint32_t base;
  uint32_t cnt;
  int32_t scene;
  uint32_t pad_;
};

// Control block (device memory; the host reads it once per group of
// iteration pairs).
struct StreamCtl {
  uint32_t next_tile;        // first tile of the work queue not yet assigned
  uint32_t free_count;       // entries of free_list
  uint32_t n_new;            // tiles assigned by this iteration's refill
  uint32_t n_cont;           // rays carried over in the current queue
  uint32_t status;           // 0 running, 1 finished, 2 inconsistent state, 3 queue capacity
  uint32_t iter;             // iterations run
  uint32_t cont_chunks[2];   // chunks of the slabs of queue p (index kernel)
  uint32_t cont_entries[2];  // rays in them
  uint32_t ext_chunks;       // chunks of the current queue: the slabs', then the new tiles' camera rays
  uint32_t sh_chunks;        // chunks of the shadow queue
  uint32_t sh_entries;
  uint32_t pad_;
  uint64_t total_prim;       // camera rays traversed
  uint64_t total_cont;       // bounce rays traversed
};

struct Stream {
  const Params* scenes;     // [n_scenes] per-scene parameters
  const TileDesc* tiles;    // [n_tiles] the work queue, in order
  float* const* out_rgb;    // [n_scenes] W*H*3 device images
  float* const* out_var;    // [n_scenes] (entries may be null)
  SlotInfo* slot_info;      // [n_slots]
  int32_t* live;            // [n_slots] live rays of the slot's tile
  uint32_t* free_list;      // [n_slots]
  NewTile* new_tiles;       // [n_slots]
  RunRec* ray_runs;         // [PRT_SLAB_WAVES * max_runs] runs of the shade waves
  uint32_t* ray_nruns;      // [PRT_SLAB_WAVES]
  RunRec* sh_runs;          // [PRT_SLAB_WAVES * max_runs] runs of the NEE waves
  uint32_t* sh_nruns;       // [PRT_SLAB_WAVES]
  ChunkDesc* desc[2];       // [max_desc] chunk list of queue p
  ChunkDesc* sh_desc;       // [max_sh_desc] chunk list of the shadow queue
  uint32_t* fetch;          // [3][PRT_FETCH_COUNTERS * PRT_FETCH_STRIDE] claim counters of extend, and of
                            // the shadow launches of queue parity 0 and 1
  StreamCtl* ctl;
  uint64_t* accum;          // [n_slots << tile_shift][3] fixed point
  uint64_t* ctr;            // [CTR_COUNT]
  int32_t n_scenes, n_tiles, n_slots, tile_shift;
  uint32_t slab_ray;        // entries per shade slab
  uint32_t slab_sh;         // entries per NEE slab
  uint32_t max_runs;        // run records per slab writer
  uint32_t prim_base;       // start of the primary region (= PRT_SLAB_WAVES * slab_ray)
  uint32_t prim_cap;        // its entries
  uint32_t max_desc, max_sh_desc;
  int32_t nee_max_lights;   // most lights of any scene (picks how k_stream_nee reads light records)
};

// The persistent form of the batched renderer (prt_persist.hip): one launch
// renders every camera sample of the batch.  Camera samples are numbered
// scene by scene (scene s owns samples sample_base[s] ..., pixel-major, spp
// per pixel) and cut into 64-sample chunks (scene s owns chunks
// chunk_base[s] .. chunk_base[s+1]-1; a scene's last chunk may be short).
struct Persist {
  const Params* scenes;         // [n_scenes]
  const uint32_t* chunk_base;   // [n_scenes + 1]
  const uint64_t* sample_base;  // [n_scenes + 1]; sample_base[n_scenes] = all samples
  uint32_t* next;               // claim counter: first chunk not yet handed out
  uint64_t* accum;              // [all samples][3] fixed point, written once per sample
  PendRec* pend;                // [grid threads][PRT_PEND_DEPTH] pending glass segments
  uint64_t* ctr;                // [CTR_COUNT]
  float* const* out_rgb;        // [n_scenes] W*H*3 device images
  float* const* out_var;        // [n_scenes] (entries may be null)
  uint32_t n_chunks;
  uint32_t claim;               // chunks per claim
  int32_t n_scenes;
  int32_t pad_;
};

// stack_lds: 8, 16 or (at 2 waves, where two 66-KB blocks fit a CU) 32 LDS
// stack entries; waves: waves per SIMD its registers are budgeted for (2..4).
hipError_t launch_persist(const Persist& p, int stack_lds, int waves, int blocks, hipStream_t s);
// Resident 256-thread blocks per CU of that kernel instance.
hipError_t persist_blocks_per_cu(int stack_lds, int waves, int* blocks);
hipError_t launch_persist_resolve(const Persist& p, int64_t n_pixels_total, const int64_t* d_pixel_base,
                                  hipStream_t s);

hipError_t launch_stream_init(const Stream& st, hipStream_t s);
hipError_t launch_stream_refill(const Stream& st, int cur, hipStream_t s);
hipError_t launch_stream_gen(const Stream& st, const RayQ& q, int cur, int blocks, hipStream_t s);
hipError_t launch_stream_extend(const Stream& st, const RayQ& q, int cur, double* hit_t, int32_t* hit_id,
                                int stack_lds, int blocks, hipStream_t s);
hipError_t launch_stream_shade(const Stream& st, const RayQ& qin, const double* hit_t, const int32_t* hit_id,
                               const RayQ& qout, int nxt, hipStream_t s);
hipError_t launch_stream_index(const Stream& st, int which, hipStream_t s);  // which: 0 / 1 ray queue, 2 shadow
hipError_t launch_stream_nee(const Stream& st, const RayQ& q, int nxt, const ShadowQ& sq, hipStream_t s);
hipError_t launch_stream_shadow(const Stream& st, const ShadowQ& sq, int par, int stack_lds, int blocks,
                                hipStream_t s);
hipError_t launch_stream_resolve(const Stream& st, int blocks, hipStream_t s);
