// prt_batch.cpp — batched rendering of many scenes through one shared work
// queue of pixel tiles (prt_render_batch_device) and the measurement-driven
// choice of its tile / in-flight / stack / grid sizes (prt_tune_batch).
//
// The device side is prt_stream.hip; every iteration is refill, gen, extend,
// shade, nee, shadow, resolve on the lead scene's stream (nee and shadow on
// its second stream when PRT_BATCH_OVERLAP is set), launched two iterations
// at a time directly or (use_graph = 1) as one captured hipGraph.  Once per
// group of four such pairs the host reads the control block to learn whether
// the work queue is drained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "prt_internal.h"
#include "prt_stream.h"

// 1: NEE and shadow rays of an iteration run on a second stream, alongside
// the next iteration's camera-ray generation and closest-hit rays
#ifndef PRT_BATCH_OVERLAP
#define PRT_BATCH_OVERLAP 1
#endif

namespace prt {

struct StreamWorkspace {
  int64_t n_slots = 0, tile_shift = 0, q_cap = 0, sq_cap = 0;
  DevBuf q[2][RAYQ_COLS], hit_t, hit_id, sq[SHADOWQ_COLS];
  DevBuf accum, ctr, ctl, slot_info, live, free_list, new_tiles;
  DevBuf scenes, tiles, outs, ovf, runs, nruns, desc[2], sh_desc, fetch;
  // persistent engine: per-sample accumulators, pending glass segments per
  // thread, the chunk / sample / pixel bases of the scenes and the claim counter
  DevBuf p_accum, p_pend, p_bases;
  int64_t ovf_threads = 0;
  int ovf_words = 0;
  StreamCtl* h_ctl = nullptr;  // pinned, two copies (double-buffered status reads)
  void release() {
    for (auto& a : q)
      for (auto& b : a) b.release();
    for (auto& b : sq) b.release();
    DevBuf* all[] = {&hit_t,  &hit_id, &accum, &ctr,     &ctl,     &slot_info, &live,    &free_list, &new_tiles,
                     &scenes, &tiles,  &outs,  &ovf,     &runs,    &nruns,     &desc[0], &desc[1],   &sh_desc,
                     &fetch,  &p_accum, &p_pend, &p_bases};
    for (DevBuf* b : all) b->release();
    if (h_ctl) (void)hipHostFree(h_ctl);
    h_ctl = nullptr;
    n_slots = tile_shift = q_cap = sq_cap = ovf_threads = 0;
    ovf_words = 0;
  }
};

// Slab geometry of a batch (prt_stream.h): PRT_SLAB_WAVES slab writers,
// slab_ray / slab_sh entries per shade / NEE slab, max_runs run records per
// writer, the primary region after the slabs.
struct StreamGeometry {
  int64_t slab_ray = 0, slab_sh = 0, max_runs = 0, prim_base = 0, prim_cap = 0, q_cap = 0, sq_cap = 0;
  int64_t max_desc = 0, max_sh_desc = 0;
};

void release_stream_workspace(StreamWorkspace* w) {
  if (!w) return;
  w->release();
  delete w;
}

}  // namespace prt

using namespace prt;

namespace {

// Defaults used where prt_batch_opts leaves a field at 0.
constexpr int kDefaultTileSamples = 4096;
// Automatic samples in flight (inflight_tiles = 0): the most, up to 2^24 and
// up to the batch's camera samples, whose tile-queue workspace fits 3/8 of
// the device memory.  More keep the traversal kernels' grids full for longer
// and need fewer iterations: cfg3 549, 630 and 682 M samples/s at 2^22, 2^23
// and 2^24 (about 97 GB of workspace with glass), cfg4 351, 363, 369.
constexpr int64_t kMaxAutoInflight = (int64_t)1 << 24;
// 12 LDS stack entries by default: that instance of the traversal kernels
// runs 5 waves per SIMD (cfg3 478 -> 504 M samples/s against 16)
#ifndef PRT_BATCH_STACK_LDS
#define PRT_BATCH_STACK_LDS 12
#endif
constexpr int kDefaultStackLds = PRT_BATCH_STACK_LDS;
constexpr int kDefaultTravBlocksPerCu = 8;
constexpr int kMaxTravBlocksPerCu = 16;
constexpr int kReplaysPerGroup = 4;  // pairs of iterations (or graph replays) between status reads
constexpr int64_t kMaxIterations = 2000000;

struct BatchConfig {
  int tile_shift = 12;
  int n_slots = 512;
  int stack_lds = kDefaultStackLds;
  int trav_blocks_per_cu = kDefaultTravBlocksPerCu;
  bool use_graph = false;
  int tile_stride = 1;
  int max_vertices = 128;
  bool timing = false;
  int persistent = 0;  // 0 tile queue, 1 one persistent launch, 2 split by NEE
  int persist_blocks_per_cu = 2;
  uint64_t seed = 0;
};

int ceil_log2(int64_t v) {
  int s = 0;
  while (((int64_t)1 << s) < v) ++s;
  return s;
}

int64_t auto_inflight(prt_scene* const* scenes, int n, const BatchConfig& c);

int resolve_config(prt_scene* const* scenes, int n, const prt_batch_opts* o, BatchConfig& c) {
  prt_batch_opts opts{};
  if (o) opts = *o;
  int max_spp = 1;
  for (int i = 0; i < n; ++i) max_spp = std::max(max_spp, scenes[i]->P.spp);
  const int64_t ts = opts.tile_samples > 0 ? opts.tile_samples : kDefaultTileSamples;
  // at least 64 samples per tile: a wave's chunk of camera samples lies in one tile
  c.tile_shift = std::max(std::max(ceil_log2(ts), ceil_log2(max_spp)), 6);
  if (c.tile_shift > 24) return fail(PRT_E_INVALID, "tile_samples above 2^24");
  const int64_t T = (int64_t)1 << c.tile_shift;
  c.n_slots = opts.inflight_tiles > 0 ? opts.inflight_tiles : (int)std::max<int64_t>(1, auto_inflight(scenes, n, c) / T);
  if ((int64_t)c.n_slots * T > ((int64_t)1 << 28)) return fail(PRT_E_INVALID, "more than 2^28 samples in flight");
  c.stack_lds = opts.stack_lds > 0 ? opts.stack_lds : kDefaultStackLds;
  if (c.stack_lds != 8 && c.stack_lds != 12 && c.stack_lds != 16 && c.stack_lds != 32)
    return fail(PRT_E_INVALID, "stack_lds must be 8, 12, 16 or 32");
  c.trav_blocks_per_cu = opts.trav_blocks_per_cu > 0 ? opts.trav_blocks_per_cu : kDefaultTravBlocksPerCu;
  if (c.trav_blocks_per_cu > kMaxTravBlocksPerCu) return fail(PRT_E_INVALID, "trav_blocks_per_cu above 16");
  c.use_graph = opts.use_graph > 0;  // 0 / -1 = plain launches
  c.timing = opts.timing != 0 && !c.use_graph;
  c.tile_stride = std::max(1, opts.tile_stride);
  c.max_vertices = opts.max_depth > 0 ? std::min(opts.max_depth, 255) : 128;
  c.seed = opts.seed;
  if (opts.persistent < 0 || opts.persistent > 2) return fail(PRT_E_INVALID, "persistent must be 0, 1 or 2");
  c.persistent = opts.persistent;
  // the persistent kernel's register budget (waves per SIMD): persist_waves,
  // else in mode 1 trav_blocks_per_cu (mode 2 sizes the tile-queue part's
  // grid with it), else the compiler's own budget (2)
  if (opts.persist_waves != 0 && (opts.persist_waves < 2 || opts.persist_waves > 4))
    return fail(PRT_E_INVALID, "persist_waves must be 0, 2, 3 or 4");
  c.persist_blocks_per_cu = opts.persist_waves > 0                               ? opts.persist_waves
                            : (opts.persistent == 1 && opts.trav_blocks_per_cu > 0) ? opts.trav_blocks_per_cu
                                                                                   : 2;
  if (c.persistent && opts.tile_stride > 1) return fail(PRT_E_INVALID, "the persistent engine renders every sample");
  return PRT_OK;
}

// Slab sizing.  A camera sample has at most 8 path segments in flight with
// glass (splits at depths 0..2), else 1, so a queue holds at most rays =
// 8x / 1x the samples in flight; shade's input is those plus one vertex per
// new camera sample (at most the samples in flight), in 64-entry chunks plus
// at most one partly filled chunk per run (a run per scene change in a
// writer's range: a few per writer with the scene-major work queue).  A
// writer gets at most cw chunks and writes at most fan (2 with glass, else
// 1) children or 3 shadow rays per entry.  Anything beyond is dropped and
// reported, never written out of bounds.
StreamGeometry stream_geometry(const BatchConfig& c, bool glass) {
  StreamGeometry g;
  const int64_t W = PRT_SLAB_WAVES;
  const int64_t T = (int64_t)1 << c.tile_shift;
  const int64_t inflight = (int64_t)c.n_slots * T;
  const int64_t rays = (glass ? 8 : 1) * inflight;
  const int64_t cw = ((rays + inflight) / 64 + W - 1) / W + 4;
  g.slab_ray = (glass ? 2 : 1) * 64 * cw;
  g.slab_sh = 3 * 64 * cw;
  g.max_runs = cw;
  g.prim_base = W * g.slab_ray;
  g.prim_cap = inflight;  // a tile's camera rays (n_pixels * SS^2 <= T) padded to 64 (T is a multiple of 64)
  g.q_cap = g.prim_base + g.prim_cap;
  g.sq_cap = W * g.slab_sh;
  // a slab's runs each end in at most one partly filled chunk
  g.max_desc = W * (g.slab_ray / 64 + g.max_runs) + g.prim_cap / 64;
  g.max_sh_desc = W * (g.slab_sh / 64 + g.max_runs);
  return g;
}

// Device bytes of the tile queue's workspace that grow with the samples in
// flight: two ray queues and the hits, the shadow queue, the accumulators
// and the chunk lists.
int64_t stream_workspace_bytes(const BatchConfig& c, bool glass) {
  const StreamGeometry g = stream_geometry(c, glass);
  const int64_t inflight = (int64_t)c.n_slots << c.tile_shift;
  return g.q_cap * (int64_t)(2 * ray_entry_bytes() + 12) + g.sq_cap * (int64_t)shadow_entry_bytes() +
         inflight * 24 + (2 * g.max_desc + g.max_sh_desc) * (int64_t)sizeof(ChunkDesc);
}

int64_t auto_inflight(prt_scene* const* scenes, int n, const BatchConfig& c) {
  uint64_t total = 0;
  bool glass = false;
  for (int i = 0; i < n; ++i) {
    total += (uint64_t)scenes[i]->P.W * scenes[i]->P.H * scenes[i]->P.spp;
    glass |= scenes[i]->has_glass;
  }
  size_t dev_total = 0;
  if (hipDeviceTotalMem(&dev_total, scenes[0]->device) != hipSuccess) dev_total = 0;
  const int64_t budget = dev_total ? (int64_t)(dev_total / 8 * 3) : ((int64_t)1 << 33);
  const int64_t T = (int64_t)1 << c.tile_shift;
  for (int64_t v = kMaxAutoInflight; v > T; v >>= 1) {
    if ((uint64_t)v > total) continue;
    BatchConfig t = c;
    t.n_slots = (int)(v / T);
    if (stream_workspace_bytes(t, glass) <= budget) return v;
  }
  return T;
}

int ensure_stream_workspace(prt_scene* lead, const BatchConfig& c, const StreamGeometry& g, int n_scenes,
                            int64_t n_tiles, int ovf_words) {
  if (!lead->pool->sws) lead->pool->sws = new (std::nothrow) StreamWorkspace();
  if (!lead->pool->sws) return fail(PRT_E_NOMEM, "out of host memory");
  StreamWorkspace& w = *lead->pool->sws;
  const int64_t T = (int64_t)1 << c.tile_shift;
  const int64_t inflight = (int64_t)c.n_slots * T;
  if (g.q_cap >= ((int64_t)1 << 32) || g.sq_cap >= ((int64_t)1 << 32))
    return fail(PRT_E_INVALID, "queue capacity above 2^32");
  int rc;
  if (w.q_cap < g.q_cap) {
    w.q_cap = 0;  // until every column holds g.q_cap entries (a failed allocation released its buffer)
    for (int qi = 0; qi < 2; ++qi)
      if ((rc = alloc_ray_queue(w.q[qi], g.q_cap))) return rc;
    if ((rc = dev_alloc(w.hit_t, (size_t)g.q_cap * 8))) return rc;
    if ((rc = dev_alloc(w.hit_id, (size_t)g.q_cap * 4))) return rc;
    w.q_cap = g.q_cap;
  }
  if (w.sq_cap < g.sq_cap) {
    w.sq_cap = 0;
    if ((rc = alloc_shadow_queue(w.sq, g.sq_cap))) return rc;
    w.sq_cap = g.sq_cap;
  }
  const size_t W = PRT_SLAB_WAVES;
  if ((rc = dev_alloc(w.accum, (size_t)inflight * 3 * sizeof(uint64_t)))) return rc;
  if ((rc = dev_alloc(w.ctr, CTR_COUNT * sizeof(uint64_t)))) return rc;
  if ((rc = dev_alloc(w.ctl, sizeof(StreamCtl)))) return rc;
  if ((rc = dev_alloc(w.slot_info, (size_t)c.n_slots * sizeof(SlotInfo)))) return rc;
  if ((rc = dev_alloc(w.live, (size_t)c.n_slots * sizeof(int32_t)))) return rc;
  if ((rc = dev_alloc(w.free_list, (size_t)c.n_slots * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(w.new_tiles, (size_t)c.n_slots * sizeof(NewTile)))) return rc;
  if ((rc = dev_alloc(w.scenes, (size_t)n_scenes * sizeof(Params)))) return rc;
  if ((rc = dev_alloc(w.tiles, (size_t)n_tiles * sizeof(TileDesc)))) return rc;
  if ((rc = dev_alloc(w.outs, (size_t)n_scenes * 2 * sizeof(float*)))) return rc;
  if ((rc = dev_alloc(w.runs, 2 * W * (size_t)g.max_runs * sizeof(RunRec)))) return rc;  // shade's, NEE's
  if ((rc = dev_alloc(w.nruns, 2 * W * sizeof(uint32_t)))) return rc;
  for (int p = 0; p < 2; ++p)
    if ((rc = dev_alloc(w.desc[p], (size_t)g.max_desc * sizeof(ChunkDesc)))) return rc;
  if ((rc = dev_alloc(w.sh_desc, (size_t)g.max_sh_desc * sizeof(ChunkDesc)))) return rc;
  if ((rc = dev_alloc(w.fetch, (size_t)3 * PRT_FETCH_COUNTERS * PRT_FETCH_STRIDE * sizeof(uint32_t)))) return rc;
  const int64_t ovf_threads = (int64_t)std::max(1, lead->cu_count) * kMaxTravBlocksPerCu * 256;
  if ((rc = dev_alloc(w.ovf, (size_t)PRT_OVF_REGIONS * ovf_threads * ovf_words * sizeof(uint32_t))))
    return rc;
  w.ovf_threads = ovf_threads;
  w.ovf_words = ovf_words;
  if (!w.h_ctl) HIPCHK(hipHostMalloc((void**)&w.h_ctl, 2 * sizeof(StreamCtl), hipHostMallocDefault));
  w.n_slots = c.n_slots;
  w.tile_shift = c.tile_shift;
  return PRT_OK;
}

struct BatchResult {
  double ms = 0;
  uint64_t samples = 0;
};

// One batched render with a resolved configuration.
int run_stream(prt_scene* const* scenes, int n, const BatchConfig& c, float* const* d_rgb, float* const* d_var,
               prt_stats* stats, BatchResult* res) {
  prt_scene* lead = scenes[0];
  HIPCHK(hipSetDevice(lead->device));
  const int64_t T = (int64_t)1 << c.tile_shift;
  // work queue: scene-major runs of T / spp pixels; every tile_stride-th tile
  std::vector<TileDesc> tiles;
  bool glass = false, nee = false;
  uint64_t samples = 0;
  int64_t tile_counter = 0;
  for (int i = 0; i < n; ++i) {
    const Params& P = scenes[i]->P;
    const bool scene_nee = P.nee && P.n_lights > 0;
    glass |= scenes[i]->has_glass;
    nee |= scene_nee;
    const int64_t npix = (int64_t)P.W * P.H;
    const int64_t per = std::max<int64_t>(1, T / P.spp);
    for (int64_t p = 0; p < npix; p += per, ++tile_counter) {
      if (tile_counter % c.tile_stride != 0) continue;
      TileDesc td;
      td.scene = i;
      td.n_pixels = (int32_t)std::min(per, npix - p);
      td.pixel_base = p;
      tiles.push_back(td);
      samples += (uint64_t)td.n_pixels * P.spp;
    }
  }
  if (tiles.size() >= ((size_t)1 << 31)) return fail(PRT_E_INVALID, "too many tiles");
  int rc;
  const StreamGeometry geo = stream_geometry(c, glass);
  int ovf_words = 2;  // the deepest BVH of the batch sizes the shared overflow area
  for (int i = 0; i < n; ++i) ovf_words = std::max(ovf_words, (int)scenes[i]->P.ovf_words);
  if ((rc = ensure_stream_workspace(lead, c, geo, n, (int64_t)tiles.size(), ovf_words))) return rc;
  StreamWorkspace& w = *lead->pool->sws;
  hipStream_t s = lead->stream;

  std::vector<Params> params(n);
  std::vector<float*> outs(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    Params P = scenes[i]->P;
    P.key0 = (uint32_t)c.seed;
    P.key1 = (uint32_t)(c.seed >> 32);
    P.max_vertices = c.max_vertices;
    P.stack_ovf = (uint32_t*)w.ovf.p;
    P.ovf_threads = (int32_t)w.ovf_threads;
    P.ovf_words = w.ovf_words;
    params[i] = P;
    outs[i] = d_rgb[i];
    outs[n + i] = d_var ? d_var[i] : nullptr;
  }
  // the lead scene's stream must see the other scenes' uploads and updates
  for (int i = 1; i < n; ++i)
    if (scenes[i]->stream != s) HIPCHK(hipStreamSynchronize(scenes[i]->stream));
  HIPCHK(hipMemcpyAsync(w.scenes.p, params.data(), params.size() * sizeof(Params), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(w.tiles.p, tiles.data(), tiles.size() * sizeof(TileDesc), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(w.outs.p, outs.data(), outs.size() * sizeof(float*), hipMemcpyHostToDevice, s));

  Stream S{};
  S.scenes = (const Params*)w.scenes.p;
  S.tiles = (const TileDesc*)w.tiles.p;
  S.out_rgb = (float* const*)w.outs.p;
  S.out_var = (float* const*)w.outs.p + n;
  S.slot_info = (SlotInfo*)w.slot_info.p;
  S.live = (int32_t*)w.live.p;
  S.free_list = (uint32_t*)w.free_list.p;
  S.new_tiles = (NewTile*)w.new_tiles.p;
  S.ray_runs = (RunRec*)w.runs.p;
  S.sh_runs = S.ray_runs + (size_t)PRT_SLAB_WAVES * geo.max_runs;
  S.ray_nruns = (uint32_t*)w.nruns.p;
  S.sh_nruns = S.ray_nruns + PRT_SLAB_WAVES;
  S.desc[0] = (ChunkDesc*)w.desc[0].p;
  S.desc[1] = (ChunkDesc*)w.desc[1].p;
  S.sh_desc = (ChunkDesc*)w.sh_desc.p;
  S.fetch = (uint32_t*)w.fetch.p;
  S.ctl = (StreamCtl*)w.ctl.p;
  S.accum = (uint64_t*)w.accum.p;
  S.ctr = (uint64_t*)w.ctr.p;
  S.n_scenes = n;
  S.nee_max_lights = 0;
  for (int i = 0; i < n; ++i) S.nee_max_lights = std::max(S.nee_max_lights, (int32_t)scenes[i]->P.n_lights);
  S.n_tiles = (int32_t)tiles.size();
  S.n_slots = c.n_slots;
  S.tile_shift = c.tile_shift;
  S.slab_ray = (uint32_t)geo.slab_ray;
  S.slab_sh = (uint32_t)geo.slab_sh;
  S.max_runs = (uint32_t)geo.max_runs;
  S.prim_base = (uint32_t)geo.prim_base;
  S.prim_cap = (uint32_t)geo.prim_cap;
  S.max_desc = (uint32_t)geo.max_desc;
  S.max_sh_desc = (uint32_t)geo.max_sh_desc;
  RayQ q[2] = {ray_queue(w.q[0], w.q_cap), ray_queue(w.q[1], w.q_cap)};
  ShadowQ sq = shadow_queue(w.sq, w.sq_cap);
  double* hit_t = (double*)w.hit_t.p;
  int32_t* hit_id = (int32_t*)w.hit_id.p;
  const int cu = std::max(1, lead->cu_count);
  const int trav_blocks = cu * c.trav_blocks_per_cu;
  const int work_blocks = cu * 8;

  // NEE and shadow rays on the lead scene's second stream (prt_stream.h)
  const bool overlap = PRT_BATCH_OVERLAP && nee;
  hipStream_t s2 = overlap ? lead->stream2 : s;
  hipEvent_t e0 = nullptr, e1 = nullptr, ev[2] = {nullptr, nullptr}, ev_shade = nullptr,
             ev_sh[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // every return after this point goes through cleanup (the events that were
  // created, the graph)
  auto cleanup = [&]() {
    if (overlap) (void)hipStreamSynchronize(s2);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : {e0, e1, ev[0], ev[1], ev_shade, ev_sh[0], ev_sh[1]})
      if (e) (void)hipEventDestroy(e);
  };
  {
    hipError_t ce = hipEventCreate(&e0);
    if (ce == hipSuccess) ce = hipEventCreate(&e1);
    if (ce == hipSuccess) ce = hipEventCreate(&ev[0]);
    if (ce == hipSuccess) ce = hipEventCreate(&ev[1]);
    if (ce == hipSuccess) ce = hipEventCreateWithFlags(&ev_shade, hipEventDisableTiming);
    if (ce == hipSuccess) ce = hipEventCreateWithFlags(&ev_sh[0], hipEventDisableTiming);
    if (ce == hipSuccess) ce = hipEventCreateWithFlags(&ev_sh[1], hipEventDisableTiming);
    if (ce != hipSuccess) {
      cleanup();
      return fail(PRT_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(ce));
    }
  }
  Timer tm;
  int sh_pending = -1;  // parity of the last shadow launch not yet waited for on s
  // one iteration; with timing on, an event after each kernel (plain launches)
  auto iteration = [&](int cur) -> hipError_t {
    const int nxt = cur ^ 1;
    hipError_t e;
    auto mark = [&](int cat) { return tm.mark(cat) == PRT_OK ? hipSuccess : hipErrorUnknown; };
    auto side = [&](int cat) {
      return (cat < 0 ? tm.begin_side(s2) : tm.end_side(s2, cat)) == PRT_OK ? hipSuccess : hipErrorUnknown;
    };
    if ((e = launch_stream_refill(S, cur, s)) != hipSuccess) return e;
    if ((e = launch_stream_gen(S, q[cur], cur, work_blocks, s)) != hipSuccess) return e;
    if ((e = mark(T_OTHER)) != hipSuccess) return e;
    if ((e = launch_stream_extend(S, q[cur], cur, hit_t, hit_id, c.stack_lds, trav_blocks, s)) != hipSuccess)
      return e;
    if ((e = mark(T_EXTEND)) != hipSuccess) return e;
    if ((e = launch_stream_shade(S, q[cur], hit_t, hit_id, q[nxt], nxt, s)) != hipSuccess) return e;
    if ((e = launch_stream_index(S, nxt, s)) != hipSuccess) return e;
    if ((e = mark(T_SHADE)) != hipSuccess) return e;
    if (nee && overlap) {
      if ((e = hipEventRecord(ev_shade, s)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(s2, ev_shade, 0)) != hipSuccess) return e;
      if ((e = side(-1)) != hipSuccess) return e;
      if ((e = launch_stream_nee(S, q[nxt], nxt, sq, s2)) != hipSuccess) return e;
      if ((e = launch_stream_index(S, 2, s2)) != hipSuccess) return e;
      if ((e = side(T_NEE)) != hipSuccess) return e;
      if ((e = side(-1)) != hipSuccess) return e;
      if ((e = launch_stream_shadow(S, sq, cur, c.stack_lds, trav_blocks, s2)) != hipSuccess) return e;
      if ((e = side(T_SHADOW)) != hipSuccess) return e;
      if ((e = hipEventRecord(ev_sh[cur], s2)) != hipSuccess) return e;
      // resolve waits for the previous iteration's shadow launch only
      if (sh_pending == nxt && (e = hipStreamWaitEvent(s, ev_sh[nxt], 0)) != hipSuccess) return e;
      sh_pending = cur;
    } else if (nee) {
      if ((e = launch_stream_nee(S, q[nxt], nxt, sq, s)) != hipSuccess) return e;
      if ((e = launch_stream_index(S, 2, s)) != hipSuccess) return e;
      if ((e = mark(T_NEE)) != hipSuccess) return e;
      if ((e = launch_stream_shadow(S, sq, cur, c.stack_lds, trav_blocks, s)) != hipSuccess) return e;
      if ((e = mark(T_SHADOW)) != hipSuccess) return e;
    }
    if ((e = launch_stream_resolve(S, cu * 2, s)) != hipSuccess) return e;
    return mark(T_OTHER);
  };
  // s waits for the last shadow launch (end of a captured graph, end of the render)
  auto join = [&]() -> hipError_t {
    if (sh_pending < 0) return hipSuccess;
    const hipError_t e = hipStreamWaitEvent(s, ev_sh[sh_pending], 0);
    sh_pending = -1;
    return e;
  };

  hipError_t he = hipEventRecord(e0, s);
  if (he == hipSuccess) he = launch_fill_u64(S.accum, 0ull, (int64_t)c.n_slots * T * 3, s);
  if (he == hipSuccess) he = launch_fill_u64(S.ctr, 0ull, CTR_COUNT, s);
  if (he == hipSuccess) he = launch_stream_init(S, s);
  if (he != hipSuccess) {
    cleanup();
    return fail(PRT_E_HIP, std::string("stream init: ") + hipGetErrorString(he));
  }
  if ((rc = tm.begin(c.timing, s, &lead->timer_events))) {
    cleanup();
    return rc;
  }
  if (c.use_graph) {
    he = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (he == hipSuccess) {
      hipError_t e_it = iteration(0);
      if (e_it == hipSuccess) e_it = iteration(1);
      if (e_it == hipSuccess) e_it = join();  // the capture ends on s
      he = hipStreamEndCapture(s, &graph);
      if (he == hipSuccess) he = e_it;
    }
    if (he == hipSuccess) he = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (he != hipSuccess) {
      cleanup();
      return fail(PRT_E_HIP, std::string("hipGraph capture of the batch iteration: ") + hipGetErrorString(he));
    }
  }
  // replay in groups; read the status of group g-1 while group g runs
  int64_t iters = 0;
  int g = 0;
  int status = 0;
  while (true) {
    for (int r = 0; r < kReplaysPerGroup && he == hipSuccess; ++r) {
      if (c.use_graph) {
        he = hipGraphLaunch(exec, s);
      } else {
        he = iteration(0);
        if (he == hipSuccess) he = iteration(1);
      }
      iters += 2;
    }
    if (he == hipSuccess) he = hipMemcpyAsync(&w.h_ctl[g & 1], S.ctl, sizeof(StreamCtl), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipEventRecord(ev[g & 1], s);
    if (he == hipSuccess && g >= 1) {
      he = hipEventSynchronize(ev[(g - 1) & 1]);
      status = (int)w.h_ctl[(g - 1) & 1].status;
    }
    if (he != hipSuccess) {
      (void)hipStreamSynchronize(s);
      cleanup();
      return fail(PRT_E_HIP, std::string("batch iteration: ") + hipGetErrorString(he));
    }
    ++g;
    if (status != 0) break;
    if (iters > kMaxIterations) {
      (void)hipStreamSynchronize(s);
      cleanup();
      return fail(PRT_E_INVALID, "batched render did not finish within the iteration limit");
    }
  }
  he = join();
  if (he == hipSuccess) he = hipEventRecord(e1, s);
  uint64_t ctrh[CTR_COUNT];
  StreamCtl final_ctl;
  if (he == hipSuccess) he = hipMemcpyAsync(ctrh, S.ctr, sizeof(ctrh), hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipMemcpyAsync(&final_ctl, S.ctl, sizeof(StreamCtl), hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
  double cat[T_NCAT] = {0, 0, 0, 0, 0, 0};
  double trav = 0.0;
  tm.collect(cat, &trav, 1u << T_EXTEND | 1u << T_SHADOW);
  cleanup();
  if (he != hipSuccess) return fail(PRT_E_HIP, std::string("batch completion: ") + hipGetErrorString(he));
  if (status == 2) return fail(PRT_E_INVALID, "batched render stalled: tiles in flight without rays");
  if (status == 3) return fail(PRT_E_INVALID, "batched render: ray queue capacity exceeded");
  if (ctrh[CTR_DROP] != 0)
    return fail(PRT_E_INVALID, "ray queue overflow (" + std::to_string(ctrh[CTR_DROP]) + " rays)");
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->node_visits = ctrh[CTR_NODES] + ctrh[CTR_SNODES];
    stats->tri_tests = ctrh[CTR_TRIS] + ctrh[CTR_STRIS];
    stats->shadow_node_visits = ctrh[CTR_SNODES];
    stats->shadow_tri_tests = ctrh[CTR_STRIS];
    stats->pickup_hits = ctrh[CTR_PICKUP];
    stats->truncated_paths = ctrh[CTR_TRUNC];
    stats->mixed_scene_passes = ctrh[CTR_MIXED];
    stats->camera_samples = samples;
    stats->primary_rays = final_ctl.total_prim;
    stats->extension_rays = final_ctl.total_cont;
    stats->shadow_rays = ctrh[CTR_SHADOW];
    stats->ms_total = ms;
    stats->ms_extend = cat[T_EXTEND];
    stats->ms_shade = cat[T_SHADE];
    stats->ms_nee = cat[T_NEE];
    stats->ms_shadow = cat[T_SHADOW];
    stats->ms_traversal = trav;
    stats->ms_other = cat[T_OTHER];
    stats->launches_extend = (int64_t)final_ctl.iter;
    stats->bytes_extend = 128 * ctrh[CTR_NODES] + 48 * ctrh[CTR_TRIS] + 48 * (final_ctl.total_prim + final_ctl.total_cont);
    stats->bytes_shadow = 128 * ctrh[CTR_SNODES] + 48 * ctrh[CTR_STRIS] + 48 * ctrh[CTR_SHADOW];
    stats->bytes_algorithmic = stats->bytes_extend + stats->bytes_shadow;
  }
  if (res) {
    res->ms = ms;
    res->samples = samples;
  }
  return PRT_OK;
}

// Chunks of 64 samples a wave claims at a time in the persistent engine
// (256 samples: one claim atomic per 256 camera samples)
#ifndef PRT_PERSIST_CLAIM
#define PRT_PERSIST_CLAIM 4
#endif
// 32: the 2-wave persistent kernel keeps 32 stack entries in LDS (two
// 66-KB blocks per CU; fewer spills: cfg4's persistent half 215 -> 212 ms);
// 0: the batch's stack_lds decides between 8 and 16
#ifndef PRT_PERSIST_STACK
#define PRT_PERSIST_STACK 32
#endif
constexpr uint64_t kMaxPersistSamples = (uint64_t)1 << 28;

// One batched render as a single persistent launch (prt_persist.hip), then
// one fold of every pixel.  The scenes' parameters, the images' pointers and
// the traversal overflow area live in the lead scene's stream workspace, as
// for the tile-queue engine.
int run_persistent(prt_scene* const* scenes, int n, const BatchConfig& c, float* const* d_rgb, float* const* d_var,
                   prt_stats* stats, BatchResult* res) {
  prt_scene* lead = scenes[0];
  HIPCHK(hipSetDevice(lead->device));
  // chunk / sample / pixel bases of the scenes (host layout of Persist)
  std::vector<uint32_t> chunk_base(n + 1);
  std::vector<uint64_t> sample_base(n + 1);
  std::vector<int64_t> pixel_base(n + 1);
  uint64_t samples = 0, chunks = 0;
  int64_t pixels = 0;
  int ovf_words = 2;
  for (int i = 0; i < n; ++i) {
    const Params& P = scenes[i]->P;
    const uint64_t ns = (uint64_t)P.W * P.H * P.spp;
    chunk_base[i] = (uint32_t)chunks;
    sample_base[i] = samples;
    pixel_base[i] = pixels;
    chunks += (ns + 63) / 64;
    samples += ns;
    pixels += (int64_t)P.W * P.H;
    ovf_words = std::max(ovf_words, (int)P.ovf_words);
    if (chunks >= ((uint64_t)1 << 31)) return fail(PRT_E_INVALID, "too many camera samples in the batch");
  }
  chunk_base[n] = (uint32_t)chunks;
  sample_base[n] = samples;
  pixel_base[n] = pixels;
  if (samples > kMaxPersistSamples)
    return fail(PRT_E_INVALID, "the persistent engine takes at most 2^28 camera samples per batch (" +
                                   std::to_string(samples) + " asked)");
  // one 256-thread block per CU puts one wave on each SIMD: blocks per CU
  // picks the register budget (2 = the compiler's own, about 230 VGPRs;
  // 3 and 4 spill part of the state), capped by what fits
  const int waves = std::min(4, std::max(2, c.persist_blocks_per_cu));
  // LDS stack of the persistent kernel: PRT_PERSIST_STACK when set, else from
  // the batch's stack_lds (8, or 16 for 12..32)
  const int stack = PRT_PERSIST_STACK == 32 && waves <= 2 ? 32 : (c.stack_lds <= 8 ? 8 : 16);
  int occ = 0;
  HIPCHK(persist_blocks_per_cu(stack, waves, &occ));
  const int cu = std::max(1, lead->cu_count);
  const int per_cu = std::max(1, std::min(occ, waves));
  const int blocks = cu * per_cu;
  const int64_t threads = (int64_t)blocks * 256;

  if (!lead->pool->sws) lead->pool->sws = new (std::nothrow) StreamWorkspace();
  if (!lead->pool->sws) return fail(PRT_E_NOMEM, "out of host memory");
  StreamWorkspace& w = *lead->pool->sws;
  int rc;
  const int64_t ovf_threads = (int64_t)cu * kMaxTravBlocksPerCu * 256;
  if ((int64_t)w.ovf.bytes < (int64_t)PRT_OVF_REGIONS * ovf_threads * ovf_words * (int64_t)sizeof(uint32_t) ||
      w.ovf_words < ovf_words) {
    if ((rc = dev_alloc(w.ovf, (size_t)PRT_OVF_REGIONS * ovf_threads * ovf_words * sizeof(uint32_t)))) return rc;
    w.ovf_threads = ovf_threads;
    w.ovf_words = ovf_words;
  }
  if (threads > w.ovf_threads) return fail(PRT_E_INTERNAL, "persistent grid larger than the overflow area");
  if ((rc = dev_alloc(w.ctr, CTR_COUNT * sizeof(uint64_t)))) return rc;
  if ((rc = dev_alloc(w.scenes, (size_t)n * sizeof(Params)))) return rc;
  if ((rc = dev_alloc(w.outs, (size_t)n * 2 * sizeof(float*)))) return rc;
  if ((rc = dev_alloc(w.p_accum, (size_t)samples * 3 * sizeof(uint64_t)))) return rc;
  if ((rc = dev_alloc(w.p_pend, (size_t)threads * PRT_PEND_DEPTH * sizeof(PendRec)))) return rc;
  const size_t bases_bytes = (size_t)(n + 1) * (sizeof(uint32_t) + sizeof(uint64_t) + sizeof(int64_t)) + 64;
  if ((rc = dev_alloc(w.p_bases, bases_bytes))) return rc;
  hipStream_t s = lead->stream;

  std::vector<Params> params(n);
  std::vector<float*> outs(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    Params P = scenes[i]->P;
    P.key0 = (uint32_t)c.seed;
    P.key1 = (uint32_t)(c.seed >> 32);
    P.max_vertices = c.max_vertices;
    P.stack_ovf = (uint32_t*)w.ovf.p;
    P.ovf_threads = (int32_t)w.ovf_threads;
    P.ovf_words = w.ovf_words;
    params[i] = P;
    outs[i] = d_rgb[i];
    outs[n + i] = d_var ? d_var[i] : nullptr;
  }
  // bases: sample_base (u64, 8-aligned), pixel_base (i64), chunk_base (u32), then the claim counter
  char* b = (char*)w.p_bases.p;
  uint64_t* d_sample = (uint64_t*)b;
  int64_t* d_pixel = (int64_t*)(b + (n + 1) * sizeof(uint64_t));
  uint32_t* d_chunk = (uint32_t*)(b + (n + 1) * (sizeof(uint64_t) + sizeof(int64_t)));
  uint32_t* d_next = d_chunk + (n + 1);
  for (int i = 1; i < n; ++i)
    if (scenes[i]->stream != s) HIPCHK(hipStreamSynchronize(scenes[i]->stream));
  HIPCHK(hipMemcpyAsync(w.scenes.p, params.data(), params.size() * sizeof(Params), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(w.outs.p, outs.data(), outs.size() * sizeof(float*), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_sample, sample_base.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_pixel, pixel_base.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_chunk, chunk_base.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));

  Persist P{};
  P.scenes = (const Params*)w.scenes.p;
  P.chunk_base = d_chunk;
  P.sample_base = d_sample;
  P.next = d_next;
  P.accum = (uint64_t*)w.p_accum.p;
  P.pend = (PendRec*)w.p_pend.p;
  P.ctr = (uint64_t*)w.ctr.p;
  P.out_rgb = (float* const*)w.outs.p;
  P.out_var = (float* const*)w.outs.p + n;
  P.n_chunks = (uint32_t)chunks;
  P.claim = PRT_PERSIST_CLAIM;
  P.n_scenes = n;

  ScopedEvent e0, e1, e2;
  HIPCHK(hipEventCreate(&e0.e));
  HIPCHK(hipEventCreate(&e1.e));
  HIPCHK(hipEventCreate(&e2.e));
  // counters and the claim counter start at zero (fill kernels, not memsets)
  HIPCHK(launch_fill_u64(P.ctr, 0ull, CTR_COUNT, s));
  HIPCHK(launch_fill_u32(d_next, 0u, 1, s));
  HIPCHK(hipEventRecord(e0.e, s));
  HIPCHK(launch_persist(P, stack, waves, blocks, s));
  HIPCHK(hipEventRecord(e1.e, s));
  HIPCHK(launch_persist_resolve(P, pixels, d_pixel, s));
  HIPCHK(hipEventRecord(e2.e, s));
  uint64_t ctrh[CTR_COUNT];
  HIPCHK(hipMemcpyAsync(ctrh, P.ctr, sizeof(ctrh), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms_render = 0, ms_resolve = 0;
  HIPCHK(hipEventElapsedTime(&ms_render, e0.e, e1.e));
  HIPCHK(hipEventElapsedTime(&ms_resolve, e1.e, e2.e));
  if (ctrh[CTR_DROP] != 0)
    return fail(PRT_E_INTERNAL, "persistent render: " + std::to_string(ctrh[CTR_DROP]) + " glass segments dropped");
  // every sample's camera ray and every emitted child was traced exactly once
  if (ctrh[CTR_QUEUED] != samples || ctrh[CTR_TRACED] != ctrh[CTR_QUEUED] + ctrh[CTR_EXT])
    return fail(PRT_E_INTERNAL, "persistent render traced " + std::to_string(ctrh[CTR_TRACED]) + " segments for " +
                                    std::to_string(ctrh[CTR_QUEUED]) + " samples (of " + std::to_string(samples) +
                                    ") and " + std::to_string(ctrh[CTR_EXT]) + " children");
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->node_visits = ctrh[CTR_NODES] + ctrh[CTR_SNODES];
    stats->tri_tests = ctrh[CTR_TRIS] + ctrh[CTR_STRIS];
    stats->shadow_node_visits = ctrh[CTR_SNODES];
    stats->shadow_tri_tests = ctrh[CTR_STRIS];
    stats->pickup_hits = ctrh[CTR_PICKUP];
    stats->truncated_paths = ctrh[CTR_TRUNC];
    stats->camera_samples = samples;
    stats->primary_rays = ctrh[CTR_QUEUED];  // one camera ray per sample in this engine
    stats->extension_rays = ctrh[CTR_EXT];
    stats->shadow_rays = ctrh[CTR_SHADOW];
    stats->ms_total = ms_render + ms_resolve;
    stats->ms_extend = ms_render;  // the one launch does every stage
    stats->ms_traversal = ms_render;
    stats->ms_other = ms_resolve;
    stats->launches_extend = 1;
    stats->bytes_extend = 128 * ctrh[CTR_NODES] + 48 * ctrh[CTR_TRIS] + 48 * ctrh[CTR_TRACED];
    stats->bytes_shadow = 128 * ctrh[CTR_SNODES] + 48 * ctrh[CTR_STRIS] + 48 * ctrh[CTR_SHADOW];
    stats->bytes_algorithmic = stats->bytes_extend + stats->bytes_shadow;
  }
  if (res) {
    res->ms = ms_render + ms_resolve;
    res->samples = samples;
  }
  return PRT_OK;
}

// Split mode: the scenes without next-event estimation in one persistent
// launch (waves follow whole paths: no queue round trips), the others through
// the replayed tile queue (whose traversal kernels keep their lanes busy
// with shadow rays from many vertices); one after the other, each group on
// its first scene's stream.  Statistics are summed.
int run_split(prt_scene* const* scenes, int n, const BatchConfig& c, float* const* d_rgb, float* const* d_var,
              prt_stats* stats, BatchResult* res) {
  std::vector<prt_scene*> grp[2];
  std::vector<float*> rgb[2], var[2];
  for (int i = 0; i < n; ++i) {
    const int g = (scenes[i]->P.nee && scenes[i]->P.n_lights > 0) ? 1 : 0;
    grp[g].push_back(scenes[i]);
    rgb[g].push_back(d_rgb[i]);
    var[g].push_back(d_var ? d_var[i] : nullptr);
  }
  prt_stats part[2];
  BatchResult r[2];
  int rc;
  BatchConfig cp = c;
  cp.persistent = 1;
  BatchConfig cs = c;
  cs.persistent = 0;
  cs.trav_blocks_per_cu = c.trav_blocks_per_cu;
  if (!grp[0].empty() &&
      (rc = run_persistent(grp[0].data(), (int)grp[0].size(), cp, rgb[0].data(), var[0].data(), &part[0], &r[0])))
    return rc;
  if (!grp[1].empty() &&
      (rc = run_stream(grp[1].data(), (int)grp[1].size(), cs, rgb[1].data(), var[1].data(), &part[1], &r[1])))
    return rc;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    for (int g = 0; g < 2; ++g) {
      if (grp[g].empty()) continue;
      const prt_stats& p = part[g];
      stats->camera_samples += p.camera_samples;
      stats->primary_rays += p.primary_rays;
      stats->extension_rays += p.extension_rays;
      stats->shadow_rays += p.shadow_rays;
      stats->pickup_hits += p.pickup_hits;
      stats->node_visits += p.node_visits;
      stats->tri_tests += p.tri_tests;
      stats->truncated_paths += p.truncated_paths;
      stats->bytes_algorithmic += p.bytes_algorithmic;
      stats->ms_total += p.ms_total;
      stats->ms_extend += p.ms_extend;
      stats->ms_shade += p.ms_shade;
      stats->ms_shadow += p.ms_shadow;
      stats->ms_nee += p.ms_nee;
      stats->ms_other += p.ms_other;
      stats->ms_traversal += p.ms_traversal;
      stats->launches_extend += p.launches_extend;
      stats->shadow_node_visits += p.shadow_node_visits;
      stats->shadow_tri_tests += p.shadow_tri_tests;
      stats->bytes_extend += p.bytes_extend;
      stats->bytes_shadow += p.bytes_shadow;
      stats->mixed_scene_passes += p.mixed_scene_passes;
    }
  }
  if (res) {
    res->ms = (grp[0].empty() ? 0.0 : r[0].ms) + (grp[1].empty() ? 0.0 : r[1].ms);
    res->samples = (grp[0].empty() ? 0 : r[0].samples) + (grp[1].empty() ? 0 : r[1].samples);
  }
  return PRT_OK;
}

int check_batch_args(prt_scene* const* scenes, int n, float* const* d_rgb) {
  if (!scenes || n <= 0 || !d_rgb) return fail(PRT_E_INVALID, "bad arguments");
  for (int i = 0; i < n; ++i) {
    if (!scenes[i] || !d_rgb[i]) return fail(PRT_E_INVALID, "null scene or output in the batch");
    if (scenes[i]->device != scenes[0]->device) return fail(PRT_E_INVALID, "batched scenes must share one device");
  }
  return PRT_OK;
}

}  // namespace

extern "C" {

int prt_render_batch_device(prt_scene* const* scenes, int n_scenes, const prt_batch_opts* opts, float* const* d_rgb,
                            float* const* d_var, prt_stats* stats) {
  int rc;
  if ((rc = check_batch_args(scenes, n_scenes, d_rgb))) return rc;
  std::lock_guard<std::recursive_mutex> pool_lock(scenes[0]->pool->m);
  BatchConfig c;
  if ((rc = resolve_config(scenes, n_scenes, opts, c))) return rc;
  if (c.persistent == 1) return run_persistent(scenes, n_scenes, c, d_rgb, d_var, stats, nullptr);
  // the automatic in-flight size is halved (down to 2^20 samples) when its
  // workspace cannot be allocated, e.g. beside the caller's own allocations
  const bool auto_size = !opts || opts->inflight_tiles <= 0;
  while (true) {
    rc = c.persistent == 2 ? run_split(scenes, n_scenes, c, d_rgb, d_var, stats, nullptr)
                           : run_stream(scenes, n_scenes, c, d_rgb, d_var, stats, nullptr);
    if (rc != PRT_E_NOMEM || !auto_size || c.n_slots <= 1 ||
        ((int64_t)c.n_slots << c.tile_shift) <= ((int64_t)1 << 20))
      return rc;
    c.n_slots /= 2;
  }
}

// Coordinate descent over tile size, LDS stack depth, traversal grid and
// launch form (the samples in flight stay the caller's or the automatic
// size), each trial rendering every k-th tile so that a trial costs about
// 1/k of the whole batch; ranked by device time per camera sample.  Then
// the persistent and split forms on the whole batch.
int prt_tune_batch(prt_scene* const* scenes, int n_scenes, const prt_batch_opts* base, float* const* d_rgb,
                   float* const* d_var, prt_batch_opts* best, double* best_ns_per_sample) {
  int rc;
  if ((rc = check_batch_args(scenes, n_scenes, d_rgb))) return rc;
  if (!best) return fail(PRT_E_INVALID, "null result");
  std::lock_guard<std::recursive_mutex> pool_lock(scenes[0]->pool->m);
  prt_batch_opts cur{};
  if (base) cur = *base;
  uint64_t total = 0;
  for (int i = 0; i < n_scenes; ++i) total += (uint64_t)scenes[i]->P.W * scenes[i]->P.H * scenes[i]->P.spp;
  if (cur.tile_stride <= 0) {
    // about 16M camera samples per trial (at least the whole batch / 16)
    cur.tile_stride = (int)std::max<uint64_t>(1, std::min<uint64_t>(16, total / ((uint64_t)1 << 24)));
  }
  if (cur.tile_samples <= 0) cur.tile_samples = kDefaultTileSamples;
  if (cur.stack_lds <= 0) cur.stack_lds = kDefaultStackLds;
  if (cur.trav_blocks_per_cu <= 0) cur.trav_blocks_per_cu = kDefaultTravBlocksPerCu;
  // samples in flight: the caller's, else the automatic size (the most that
  // fits the workspace budget, which measured fastest at every size tried)
  const int64_t inflight = cur.inflight_tiles > 0 ? (int64_t)cur.inflight_tiles * cur.tile_samples : 0;
  auto trial = [&](prt_batch_opts o, int64_t infl, double& ns) -> int {
    o.inflight_tiles = infl > 0 ? (int32_t)std::max<int64_t>(1, infl / o.tile_samples) : 0;
    BatchConfig c;
    int r = resolve_config(scenes, n_scenes, &o, c);
    if (r) return r;
    BatchResult res;
    r = run_stream(scenes, n_scenes, c, d_rgb, d_var, nullptr, &res);
    if (r) return r;
    ns = res.samples ? res.ms * 1e6 / (double)res.samples : 1e30;
    return PRT_OK;
  };
  double best_ns = 0;
  if ((rc = trial(cur, inflight, best_ns))) return rc;  // also allocates the workspace
  if ((rc = trial(cur, inflight, best_ns))) return rc;
  const int tiles[] = {1024, 4096, 16384};
  const int stacks[] = {8, 12, 16, 32};
  const int blocks[] = {4, 8, 16};
  for (int v : tiles) {
    if (v == cur.tile_samples) continue;
    prt_batch_opts o = cur;
    o.tile_samples = v;
    double ns;
    if ((rc = trial(o, inflight, ns))) return rc;
    if (ns < best_ns) { best_ns = ns; cur = o; }
  }
  for (int v : stacks) {
    if (v == cur.stack_lds) continue;
    prt_batch_opts o = cur;
    o.stack_lds = v;
    double ns;
    if ((rc = trial(o, inflight, ns))) return rc;
    if (ns < best_ns) { best_ns = ns; cur = o; }
  }
  for (int v : blocks) {
    if (v == cur.trav_blocks_per_cu) continue;
    prt_batch_opts o = cur;
    o.trav_blocks_per_cu = v;
    double ns;
    if ((rc = trial(o, inflight, ns))) return rc;
    if (ns < best_ns) { best_ns = ns; cur = o; }
  }
  {  // the other launch form: graph replay or plain launches
    prt_batch_opts o = cur;
    o.use_graph = cur.use_graph > 0 ? 0 : 1;
    double ns;
    if ((rc = trial(o, inflight, ns))) return rc;
    if (ns < best_ns) { best_ns = ns; cur = o; }
  }
  cur.inflight_tiles = inflight > 0 ? (int32_t)std::max<int64_t>(1, inflight / cur.tile_samples) : 0;
  cur.tile_stride = base ? base->tile_stride : 0;
  // the persistent engine (every sample, one launch) at each register
  // budget, when the batch fits it and no subsample was asked for
  if (total <= kMaxPersistSamples && (!base || base->tile_stride <= 1)) {
    const prt_batch_opts queued = cur;
    for (int wv = 2; wv <= 4; ++wv) {
      prt_batch_opts o = queued;
      o.persistent = 1;
      o.trav_blocks_per_cu = wv;
      BatchConfig c;
      if ((rc = resolve_config(scenes, n_scenes, &o, c))) return rc;
      BatchResult res;
      if ((rc = run_persistent(scenes, n_scenes, c, d_rgb, d_var, nullptr, &res))) return rc;
      const double ns = res.samples ? res.ms * 1e6 / (double)res.samples : 1e30;
      if (ns < best_ns) { best_ns = ns; cur = o; }
    }
    // split: scenes without NEE persistent (at each register budget), the
    // others through the tuned tile queue
    for (int wv = 2; wv <= 4; ++wv) {
      prt_batch_opts o = queued;
      o.persistent = 2;
      o.persist_waves = wv;
      BatchConfig c;
      if ((rc = resolve_config(scenes, n_scenes, &o, c))) return rc;
      BatchResult res;
      if ((rc = run_split(scenes, n_scenes, c, d_rgb, d_var, nullptr, &res))) return rc;
      const double ns = res.samples ? res.ms * 1e6 / (double)res.samples : 1e30;
      if (ns < best_ns) { best_ns = ns; cur = o; }
    }
  }
  *best = cur;
  if (best_ns_per_sample) *best_ns_per_sample = best_ns;
  return PRT_OK;
}

}  // extern "C"
