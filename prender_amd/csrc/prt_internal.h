// prt_internal.h — host-side internals of libprt shared by the C ABI
// translation units (prt_api.cpp: scenes, queries, per-scene renders, the
// adjoint; prt_batch.cpp: batched multi-scene renders and their tuner).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/prt.h"
#include "prt_device.h"
#include "prt_kernels.h"

namespace prt {

// message of the last failure on this thread (prt_last_error)
extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// A hipEvent owned by one call: destroyed on every return path.
struct ScopedEvent {
  hipEvent_t e = nullptr;
  ScopedEvent() = default;
  ScopedEvent(const ScopedEvent&) = delete;
  ScopedEvent& operator=(const ScopedEvent&) = delete;
  ~ScopedEvent() {
    if (e) (void)hipEventDestroy(e);
  }
};

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::prt::fail(PRT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// (Re)allocates b to at least `bytes` (contents are not kept).
int dev_alloc(DevBuf& b, size_t bytes);

template <class T>
int dev_upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  int rc = dev_alloc(b, std::max<size_t>(v.size() * sizeof(T), 256));
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return PRT_OK;
}

// Ray / shadow queues as 17 / 12 separately allocated SoA columns.
enum { RAYQ_COLS = 17, SHADOWQ_COLS = 12 };
int alloc_ray_queue(DevBuf* cols, int64_t cap);
size_t ray_entry_bytes();     // device bytes per ray-queue entry
size_t shadow_entry_bytes();  // device bytes per shadow-queue entry
int alloc_shadow_queue(DevBuf* cols, int64_t cap);
RayQ ray_queue(DevBuf* b, int64_t cap);
ShadowQ shadow_queue(DevBuf* b, int64_t cap);

// Per-scene render workspace sized for a batch of samples.  The ray and
// shadow queues are slab queues (prt_kernels.h), each with its chunk
// descriptors.
struct Workspace {
  int64_t batch_samples = 0;  // camera samples per batch
  int64_t queue_cap = 0;      // entries allocated per ray queue
  int64_t shadow_cap = 0;     // entries allocated in the shadow queue
  DevBuf q[2][RAYQ_COLS];     // two ray queues (SoA)
  DevBuf hit_t, hit_id;
  DevBuf sq[SHADOWQ_COLS];
  DevBuf sq2[SHADOWQ_COLS];   // second shadow queue (NEE scenes, when a depth's shadow rays are traced
                              // while the next depth's NEE fills the other queue)
  DevBuf desc[4];             // chunk descriptors (uint2) of q[0], q[1], the shadow queue, the second one
  DevBuf slabs;               // uint2 [3][PRT_SLAB_WAVES] slab (start, count) of q[0], q[1], the NEE output
  DevBuf accum, accum2;       // fixed-point sample radiance (accum2: adjoint primal)
  DevBuf counts;              // uint32 [16]: chunks of q0, q1, sq, sq2 at 0..3, entries at 4..7
  DevBuf fetch;               // uint32 claim counters, PRT_FETCH_COUNTERS * PRT_FETCH_STRIDE per launch
  DevBuf ctr;                 // uint64 [CTR_COUNT]
  uint32_t* h_counts = nullptr;
  void release();
};
// launches whose claim counters one fill clears
enum { FETCH_GROUP = 64, FETCH_LAUNCHES = 1024 };

// HIP events kept for reuse by timed renders (creating an event per mark
// costs more host time than the late, short path depths leave).
struct EventPool {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  int get(hipEvent_t* e) {
    if (used == ev.size()) {
      hipEvent_t x;
      HIPCHK(hipEventCreate(&x));
      ev.push_back(x);
    }
    *e = ev[used++];
    return PRT_OK;
  }
  void release() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    used = 0;
  }
  ~EventPool() { release(); }
};

// Device-time spans between successive marks, summed per category.
struct Timer {
  bool on = false;
  hipStream_t s = nullptr;
  EventPool own;
  EventPool* pool = &own;
  std::vector<std::pair<hipEvent_t, int>> marks;  // event + category of the span that ends here
  hipEvent_t start{};
  Timer() = default;
  Timer(const Timer&) = delete;
  Timer& operator=(const Timer&) = delete;
  // events come from `events` (kept by the caller across calls) or, if null,
  // from a pool of this timer's own
  int begin(bool enable, hipStream_t st, EventPool* events = nullptr) {
    on = enable;
    s = st;
    if (events) pool = events;
    pool->used = 0;
    if (!on) return PRT_OK;
    int rc;
    if ((rc = pool->get(&start))) return rc;
    HIPCHK(hipEventRecord(start, s));
    return PRT_OK;
  }
  int mark(int cat) {
    if (!on) return PRT_OK;
    hipEvent_t e;
    int rc;
    if ((rc = pool->get(&e))) return rc;
    HIPCHK(hipEventRecord(e, s));
    marks.push_back({e, cat});
    return PRT_OK;
  }
  // a span on another stream, overlapping the marks of the main stream
  std::vector<std::pair<hipEvent_t, int>> side;  // begin (category -1) and end events
  int begin_side(hipStream_t st) { return side_event(st, -1); }
  int end_side(hipStream_t st, int cat) { return side_event(st, cat); }
  int side_event(hipStream_t st, int cat) {
    if (!on) return PRT_OK;
    hipEvent_t e;
    int rc;
    if ((rc = pool->get(&e))) return rc;
    HIPCHK(hipEventRecord(e, st));
    side.push_back({e, cat});
    return PRT_OK;
  }
  // After both streams have completed: adds each category's device time to
  // cat_ms, sets *union_ms to the time during which a span of a category in
  // union_mask (bit per category) ran on either stream, and returns the main
  // stream's elapsed time (side spans overlap it).
  double collect(double* cat_ms, double* union_ms = nullptr, unsigned union_mask = 0) {
    if (union_ms) *union_ms = 0.0;
    if (!on) return 0.0;
    auto at = [&](hipEvent_t e) {
      float ms = 0;
      return hipEventElapsedTime(&ms, start, e) == hipSuccess ? (double)ms : -1.0;
    };
    std::vector<std::pair<double, double>> iv;  // spans of the union categories
    double wall = 0.0, t_prev = 0.0;
    for (auto& m : marks) {
      const double t = at(m.first);
      if (t < 0) continue;
      cat_ms[m.second] += t - t_prev;
      if (union_mask >> m.second & 1u) iv.push_back({t_prev, t});
      wall = t;
      t_prev = t;
    }
    for (size_t i = 0; i + 1 < side.size(); i += 2) {
      const double t0 = at(side[i].first), t1 = at(side[i + 1].first);
      if (t0 < 0 || t1 < 0) continue;
      cat_ms[side[i + 1].second] += t1 - t0;
      if (union_mask >> side[i + 1].second & 1u) iv.push_back({t0, t1});
    }
    if (union_ms) {
      std::sort(iv.begin(), iv.end());
      double end = -1.0;
      for (auto& v : iv) {
        const double b = std::max(v.first, end);
        if (v.second > b) *union_ms += v.second - b;
        end = std::max(end, v.second);
      }
    }
    marks.clear();
    side.clear();
    pool->used = 0;  // the events are complete and free for the next call
    return wall;
  }
};
enum { T_EXTEND = 0, T_SHADE = 1, T_SHADOW = 2, T_OTHER = 3, T_EXTEND0 = 4, T_NEE = 5, T_NCAT = 6 };

// Workspace of the multi-scene streaming renderer (prt_batch.cpp).
struct StreamWorkspace;
void release_stream_workspace(StreamWorkspace* w);

// Render memory of one device, shared by every scene handle on it: the ray
// and shadow queues, accumulators and counters of the per-scene renderer and
// the batched renderer's workspace, each sized for the largest call so far.
// Render calls are synchronous (they return after the device is done), so
// the pool serves one call at a time: an entry point that uses it holds its
// lock for the whole call (recursive: entry points call each other).  It is
// released when the last scene handle on the device is destroyed.
struct DevicePool {
  std::recursive_mutex m;
  Workspace ws;
  StreamWorkspace* sws = nullptr;
  int scenes = 0;  // live scene handles on the device (guarded by the registry's lock)
};
// The pool of `device`, counting one more user (acquire) or one fewer
// (release, which frees its memory when the count reaches zero).
DevicePool* acquire_device_pool(int device);
void release_device_pool(int device);

// Chain and seed-path state of Metropolis renders (prt_mlt.cpp).
struct MltWorkspace {
  DevBuf cur, prop, pix_new, pix_old, f_old, img;
  DevBuf seed_f, seed_lum, seed_pix, seed_cdf, tile_sums;
  // energy redistribution: initial paths of a batch and per-chain run state
  DevBuf path_u, path_f, path_lum, path_pix, path_direct, path_count, path_first, pix_now, run, sums;
  void release() {
    for (DevBuf* b : {&cur, &prop, &pix_new, &pix_old, &f_old, &img, &seed_f, &seed_lum, &seed_pix, &seed_cdf,
                      &tile_sums, &path_u, &path_f, &path_lum, &path_pix, &path_direct, &path_count, &path_first,
                      &pix_now, &run, &sums})
      b->release();
  }
};

// Counters of a call, summed over its batches.
struct CallTotals {
  uint64_t primary_rays = 0;
  int64_t launches = 0;
};

// Adds the traversal / shading counters c[CTR_*] to *st.
void fill_stats(prt_stats* st, const uint64_t* c);
// Sizes the scene's workspace for batches of `batch_samples` camera samples.
int ensure_workspace(prt_scene* s, int64_t batch_samples, bool two_accum);
// One batch of camera samples (or Metropolis proposals, when P.mlt_u is set)
// through the wavefront; sample radiance accumulates into accum (fixed point).
int run_batch(prt_scene* s, const Params& P, const Batch& B, uint64_t* accum, const Adjoint& A, Timer& tm,
              CallTotals& tot);

}  // namespace prt

struct prt_scene {
  int device = 0;
  int cu_count = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;              // shadow traversal, overlapping the next depth
  hipEvent_t ev_nee[2] = {nullptr, nullptr};  // last NEE that read ray queue 0 / 1
  hipEvent_t ev_shadow = nullptr, ev_shade = nullptr;
  hipEvent_t ev_sh[2] = {nullptr, nullptr};    // last shadow launch that read shadow queue 0 / 1
  Params P{};  // scene + sampling part; per-call fields are patched in
  prt::DevBuf nodes, tris, tri_nrm, tri_alb, ents, mats, analytic, lights, ovf;
  std::vector<DevEntity> h_ents;
  std::vector<DevMaterial> h_mats;
  std::vector<int32_t> h_lights;
  std::vector<uint32_t> slot_of_tri;  // global input triangle -> leaf slot
  std::vector<int64_t> ent_first_tri; // per entity: first global input triangle
  std::vector<int64_t> ent_n_tris;
  int64_t n_tris = 0;
  int bvh_depth = 0;
  int n_top = 0;  // BVH nodes of the top levels (numbered first)
  double bvh_ms = 0;
  bool has_glass = false;
  int max_trav_blocks = 0;
  prt::DevicePool* pool = nullptr;      // the device's shared render memory
  prt::DevBuf out_rgb, out_var;         // staging for host-output renders
  prt::DevBuf d_slot_of_tri;            // input triangle -> leaf slot (per-face albedo updates)
  prt::DevBuf grad_buf;                 // adjoint accumulators (double)
  prt::DevBuf tri_alb_rr;               // held albedo for Russian-roulette probabilities
  prt::MltWorkspace mlt;                // Metropolis renders
  prt::EventPool timer_events;          // per-kernel timing of this scene's renders
};
