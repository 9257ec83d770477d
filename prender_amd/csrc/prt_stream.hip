// prt_stream.hip — gfx950 kernels of the multi-scene streaming renderer
// (the iteration structure is described in prt_stream.h).
//
// Extend, shade and NEE walk padded work tables whose entries are 64-aligned
// runs of one scene's rays, so a wave's scene is uniform and its parameters
// are read with scalar loads.  The shadow queue is shared by all scenes (each
// wave appends its shadow rays contiguously); its kernel runs each distinct
// scene of a wave in turn.  The per-vertex arithmetic is the per-scene
// renderer's (prt_shade.h), and a camera sample's random numbers depend only
// on (pixel, sample, depth, branch), so a tile renders bit-identically to the
// same pixels of a per-scene render.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"
#include "prt_math.h"
#include "prt_shade.h"
#include "prt_stream.h"
#include "prt_trace.h"

__device__ __forceinline__ uint32_t round64(uint32_t n) { return (n + 63u) & ~63u; }

// Parameters of scene s (s wave-uniform) through a constant-address-space
// pointer, so that their loads are scalar loads as kernel arguments' are.
__device__ __forceinline__ const Params& scene_params(const Params* scenes, int s) {
  const uint64_t a = (uint64_t)(scenes + s);
  const uint64_t au = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
                      (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
  const __attribute__((address_space(4))) Params* Pc = (const __attribute__((address_space(4))) Params*)au;
  return *(const Params*)Pc;
}

// Entry e of a padded prefix table with pref[e] <= x < pref[e + 1]
// (pref[0] = 0, pref[n] > x); x wave-uniform.
__device__ __forceinline__ uint32_t find_entry(const uint32_t* __restrict__ pref, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pref[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Calls f(scenes[s]) for the lanes whose scene is s, one distinct s at a time
// (s wave-uniform); lanes with sid < 0 take no part.  Returns the number of
// passes (distinct scenes).  Every lane of the wave must reach this call.
template <class F>
__device__ __forceinline__ uint32_t per_scene(const Params* __restrict__ scenes, int sid, F&& f) {
  bool pend = sid >= 0;
  uint64_t m = __ballot(pend);
  uint32_t passes = 0;
  while (m) {
    ++passes;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int s = __builtin_amdgcn_readfirstlane(__shfl(sid, leader, 64));
    // the parameters' address is formed from the uniform s before the lane
    // test, so the compiler cannot substitute the per-lane sid (equal inside
    // the branch) and the loads stay scalar
    const Params& P = scene_params(scenes, s);
    if (pend && sid == s) {
      f(P);
      pend = false;
    }
    m = __ballot(pend);
  }
  return passes;
}

// Extra waterfall passes of a wave (waves holding rays of several scenes),
// summed per wave and added to the counter at the end of a kernel.
__device__ __forceinline__ void flush_mixed(uint64_t* ctr, uint32_t extra) {
  if (lane_id() == 0 && extra) atomicAdd((unsigned long long*)&ctr[CTR_MIXED], (unsigned long long)extra);
}

// Adds delta to the live-ray count of tile slot ts (one atomic per distinct
// slot of the wave, without a return value, so no wave waits on it).  A ray's
// own retirement and the children it spawned are one update, so a count is
// zero at the end of an iteration exactly when no ray of its tile is left;
// resolve looks for those slots.  Every lane of the wave must reach this call.
__device__ __forceinline__ void live_update(const Stream& st, int ts, int delta) {
  bool pend = ts >= 0 && delta != 0;
  uint64_t m = __ballot(pend);
  while (m) {
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int t0 = __builtin_amdgcn_readfirstlane(__shfl(ts, leader, 64));
    const bool mine = pend && ts == t0;
    const int sum = (int)(uint32_t)wave_sum_u32((uint32_t)(mine ? delta : 0));
    if ((int)lane_id() == leader && sum != 0) (void)atomicAdd(&st.live[t0], sum);
    pend = pend && !mine;
    m = __ballot(pend);
  }
}

// Exclusive prefix of val(i), i < n, over one 256-thread block: each thread
// owns a contiguous chunk [b, e) and gets the prefix at b; every thread gets
// the total.  Every thread of the block must call it.
struct ChunkScan {
  uint32_t b, e, base, total;
};
template <class F>
__device__ __forceinline__ ChunkScan block_scan(uint32_t n, F val, uint32_t* s_part /* [257] */) {
  ChunkScan r;
  const uint32_t tid = threadIdx.x;
  const uint32_t chunk = (n + 255u) / 256u;
  r.b = min(tid * chunk, n);
  r.e = min(r.b + chunk, n);
  uint32_t sum = 0;
  for (uint32_t i = r.b; i < r.e; ++i) sum += val(i);
  s_part[tid] = sum;
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (int t = 0; t < 256; ++t) {
      const uint32_t v = s_part[t];
      s_part[t] = acc;
      acc += v;
    }
    s_part[256] = acc;
  }
  __syncthreads();
  r.base = s_part[tid];
  r.total = s_part[256];
  __syncthreads();
  return r;
}

__global__ void k_stream_init(Stream st) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = gid; i < st.n_slots; i += stride) {
    st.free_list[i] = (uint32_t)i;
    st.live[i] = 0;
    st.slot_info[i].tile = -1;  // free
  }
  for (int i = gid; i < 2 * st.n_scenes; i += stride) {
    st.seg_off[i] = 0;
    st.seg_cap[i] = 0;
    st.seg_cnt[i] = 0;
  }
  if (gid == 0) {
    StreamCtl* c = st.ctl;
    c->next_tile = 0;
    c->free_count = (uint32_t)st.n_slots;
    c->n_new = c->n_cont = c->n_prim = 0;
    c->status = 0;
    c->iter = 0;
    c->cnt_sq = 0;
    c->fetch[0] = c->fetch[1] = c->fetch[2] = c->fetch[3] = 0;
    c->region_end[0] = c->region_end[1] = 0;
    c->work_width = c->cont_width = c->nee_width = c->n_work = 0;
    c->total_prim = c->total_cont = 0;
  }
}

// One block: hand the free slots to the next tiles of the work queue, lay out
// the next queue's scene regions and build the current queue's work table;
// reset the per-iteration counters.
__global__ __launch_bounds__(256) void k_stream_refill(Stream st, int cur) {
  __shared__ uint32_t s_part[257];
  __shared__ uint32_t s_hdr[4];
  StreamCtl* c = st.ctl;
  const int nxt = cur ^ 1;
  const uint32_t tid = threadIdx.x;
  const uint32_t NS = (uint32_t)st.n_scenes;
  if (tid == 0) {
    const uint32_t free_n = min(c->free_count, (uint32_t)st.n_slots), next = c->next_tile;
    const uint32_t left = (uint32_t)st.n_tiles - next;
    s_hdr[0] = free_n < left ? free_n : left;
    s_hdr[1] = free_n;
    s_hdr[2] = next;
    s_hdr[3] = c->region_end[cur];  // the primary region of the current queue starts here
  }
  for (uint32_t s = tid; s < NS; s += 256) st.new_samp[s] = 0;
  __syncthreads();
  const uint32_t n_want = s_hdr[0], free_n = s_hdr[1], next = s_hdr[2], prim_base = s_hdr[3];

  // 1. camera rays of the candidate tiles (64-aligned per tile)
  auto tile_rays = [&](uint32_t k) {
    const TileDesc td = st.tiles[next + k];
    const Params& P = st.scenes[td.scene];
    return round64((uint32_t)td.n_pixels * (uint32_t)(P.SS * P.SS));
  };
  const ChunkScan tr = block_scan(n_want, tile_rays, s_part);
  const bool fits_prim = (uint64_t)prim_base + tr.total <= st.q_cap;
  const uint32_t n_new = fits_prim ? n_want : 0u;
  if (fits_prim) {
    uint32_t base = tr.base;
    for (uint32_t k = tr.b; k < tr.e; ++k) {
      const uint32_t tile = next + k;
      const TileDesc td = st.tiles[tile];
      const Params& P = st.scenes[td.scene];
      const uint32_t slot = st.free_list[free_n - 1u - k];
      NewTile nt;
      nt.slot = slot;
      nt.ray_base = base;
      nt.n_rays = (uint32_t)td.n_pixels * (uint32_t)(P.SS * P.SS);
      nt.n_samples = (uint32_t)td.n_pixels * (uint32_t)P.spp;
      nt.spp = (uint32_t)P.spp;
      nt.ss2 = (uint32_t)(P.SS * P.SS);
      nt.scene = td.scene;
      nt.pad_ = 0;
      st.new_tiles[k] = nt;
      SlotInfo si;
      si.scene = td.scene;
      si.tile = (int32_t)tile;
      si.pixel_base = td.pixel_base;
      si.spp = P.spp;
      si.pad_ = 0;
      st.slot_info[slot] = si;
      st.live[slot] = (int32_t)nt.n_samples;
      atomicAdd(&st.new_samp[td.scene], nt.n_samples);
      base += round64(nt.n_rays);
    }
  }
  __syncthreads();

  // 2. scene regions of the next queue: bound = fan * (carried-over rays +
  //    new camera samples), the vertices this iteration shades
  const uint32_t* cnt_cur = st.seg_cnt + cur * NS;
  const uint32_t* cap_cur = st.seg_cap + cur * NS;
  auto bound = [&](uint32_t s) {
    const uint32_t cnt = min(cnt_cur[s], cap_cur[s]);
    return round64((uint32_t)st.scene_fan[s] * (cnt + st.new_samp[s]));
  };
  const ChunkScan rb = block_scan(NS, bound, s_part);
  // regions past the queue's end are never handed out: with no room, every
  // region gets capacity 0 (shade then drops and counts the children, and the
  // status below reports the overflow)
  const bool fits_next = rb.total <= st.q_cap;
  {
    uint32_t base = rb.base;
    for (uint32_t s = rb.b; s < rb.e; ++s) {
      const uint32_t w = bound(s);
      st.seg_off[nxt * NS + s] = fits_next ? base : 0u;
      st.seg_cap[nxt * NS + s] = fits_next ? w : 0u;
      st.seg_cnt[nxt * NS + s] = 0;
      base += w;
    }
  }

  // 3. work table of the current queue: the scene regions, then the new tiles
  const uint32_t n_work = NS + n_new;
  auto width = [&](uint32_t e) {
    if (e < NS) return round64(min(cnt_cur[e], cap_cur[e]));
    return round64(st.new_tiles[e - NS].n_rays);
  };
  const ChunkScan wt = block_scan(n_work, width, s_part);
  {
    uint32_t base = wt.base;
    for (uint32_t e = wt.b; e < wt.e; ++e) {
      WorkEntry we;
      if (e < NS) {
        we.off = st.seg_off[cur * NS + e];
        we.cnt = min(cnt_cur[e], cap_cur[e]);
        we.scene = (int32_t)e;
      } else {
        const NewTile nt = st.new_tiles[e - NS];
        we.off = prim_base + nt.ray_base;
        we.cnt = nt.n_rays;
        we.scene = nt.scene;
      }
      we.pad_ = 0;
      st.work[e] = we;
      st.work_pref[e] = base;
      base += round64(we.cnt);
    }
    if (tid == 255) st.work_pref[n_work] = wt.total;
  }
  // carried-over rays (sum of the scene regions' counts)
  auto cont = [&](uint32_t s) { return min(cnt_cur[s], cap_cur[s]); };
  const ChunkScan cs = block_scan(NS, cont, s_part);
  if (tid == 0) {
    const uint32_t n_cont = cs.total;
    const uint32_t cont_width = wt.total - (fits_prim ? tr.total : 0u);  // = work_pref[n_scenes]
    c->n_new = n_new;
    c->n_cont = n_cont;
    c->n_prim = 0;  // counted by gen into total_prim
    c->free_count = free_n - n_new;
    c->next_tile = next + n_new;
    c->region_end[nxt] = fits_next ? rb.total : 0u;
    c->work_width = wt.total;
    c->cont_width = cont_width;
    c->n_work = n_work;
    c->cnt_sq = 0;
    c->fetch[0] = c->fetch[1] = c->fetch[2] = c->fetch[3] = 0;
    c->total_cont += n_cont;
    c->iter += 1;
    if (c->status < 2) {
      if (!fits_prim || !fits_next) c->status = 3;
      else if (n_cont == 0 && n_new == 0)
        c->status = (next == (uint32_t)st.n_tiles && free_n == (uint32_t)st.n_slots) ? 1u : 2u;
      else c->status = 0;
    }
  }
}

// Unique sub-pixel camera rays of the new tiles (one wave per tile), in the
// primary region of queue `cur`.
__global__ __launch_bounds__(256) void k_stream_gen(Stream st, RayQ q, int cur) {
  const uint32_t n_new = st.ctl->n_new, prim_base = st.ctl->region_end[cur];
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  uint32_t rays = 0;
  for (uint32_t k = wave; k < n_new; k += n_waves) {
    const NewTile nt = st.new_tiles[k];
    const SlotInfo si = st.slot_info[nt.slot];
    const Params& P = scene_params(st.scenes, nt.scene);
    const uint32_t ss2 = nt.ss2;
    for (uint32_t r = lane_id(); r < nt.n_rays; r += 64) {
      const int64_t pix = si.pixel_base + r / ss2;
      const int sub = (int)(r % ss2);
      d3 o, d;
      camera_ray(P, (int)(pix % P.W), (int)(pix / P.W), sub % P.SS, sub / P.SS, o, d);
      const uint32_t i = prim_base + nt.ray_base + r;
      q.ox[i] = o.x; q.oy[i] = o.y; q.oz[i] = o.z;
      q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
      q.slot[i] = nt.slot << st.tile_shift;
      q.meta[i] = 0u;
    }
    if (lane_id() == 0) rays += nt.n_rays;
  }
  if (lane_id() == 0 && rays) atomicAdd((unsigned long long*)&st.ctl->total_prim, (unsigned long long)rays);
}

// Waves take FETCH_CHUNKS 64-wide chunks of a padded index space per atomic
// and walk them in runs of one table entry (one scene), so the scene's
// parameters are loop-invariant within a run and their loads are hoisted.
#define FETCH_CHUNKS 4u

__device__ __forceinline__ uint32_t fetch_chunks(uint32_t* counter) {
  uint32_t c = 0;
  if (lane_id() == 0) c = atomicAdd(counter, FETCH_CHUNKS);
  return (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)c, 0, 64));
}

// Closest hit for the current queue through its work table.
template <int NL>
__global__ __launch_bounds__(256) void k_stream_extend(Stream st, RayQ q, double* __restrict__ hit_t,
                                                       int32_t* __restrict__ hit_id) {
  PRT_DECLARE_STACK_N(S, st.scenes[0], NL);
  const uint32_t n_chunks = st.ctl->work_width >> 6, n_work = st.ctl->n_work;
  uint32_t* fetch = &st.ctl->fetch[0];
  TravStats ts{0, 0, 0};
  while (true) {
    uint32_t c = fetch_chunks(fetch);
    if (c >= n_chunks) break;
    const uint32_t c_end = min(c + FETCH_CHUNKS, n_chunks);
    while (c < c_end) {
      const uint32_t e = find_entry(st.work_pref, n_work, c << 6);
      const WorkEntry we = st.work[e];
      const uint32_t e_base = st.work_pref[e];
      const uint32_t run_end = min(st.work_pref[e + 1] >> 6, c_end);
      const Params& P = scene_params(st.scenes, we.scene);
      for (; c < run_end; ++c) {
        const uint32_t local = (c << 6) - e_base + lane_id();
        if (local < we.cnt) {
          const uint32_t i = we.off + local;
          double t, u, v;
          int32_t id;
          closest_hit(P, mk(q.ox[i], q.oy[i], q.oz[i]), mk(q.dx[i], q.dy[i], q.dz[i]), t, id, u, v, S, ts);
          hit_t[i] = t;
          hit_id[i] = id;
        }
      }
    }
  }
  flush_stats(st.ctr, CTR_NODES, ts);
}

// One vertex per carried-over ray (scene regions of queue cur, through the
// work table) and per camera sample of the new tiles (T padded entries per
// tile); children go to their scene's region of queue nxt.
__global__ __launch_bounds__(256) void k_stream_shade(Stream st, RayQ qin, const double* __restrict__ hit_t,
                                                      const int32_t* __restrict__ hit_id, RayQ qout, int nxt) {
  const uint32_t cont_chunks = st.ctl->cont_width >> 6, n_new = st.ctl->n_new;
  const uint32_t prim_base = st.ctl->region_end[nxt ^ 1];
  const uint32_t NS = (uint32_t)st.n_scenes;
  const uint32_t shift = (uint32_t)st.tile_shift, tmask = (1u << shift) - 1u;
  const uint32_t n_chunks = cont_chunks + (n_new << (shift - 6));  // T >= 64: T/64 chunks per new tile
  ShadeCtx C;
  C.qout = qout;
  C.truncated = C.n_ext = C.pickups = C.dropped = 0;
  {
    uint32_t c, c_end;
    wave_range(n_chunks, c, c_end);
    while (c < c_end) {
      // a run: one scene region of the carried-over rays, or one new tile
      const bool prim = c >= cont_chunks;
      int scene;
      uint32_t run_end, e_base = 0;
      WorkEntry we{0, 0, 0, 0};
      NewTile nt{};
      if (!prim) {
        const uint32_t e = find_entry(st.work_pref, NS, c << 6);
        we = st.work[e];
        e_base = st.work_pref[e];
        run_end = min(st.work_pref[e + 1] >> 6, c_end);
        scene = we.scene;
      } else {
        const uint32_t k = (c - cont_chunks) >> (shift - 6);
        nt = st.new_tiles[k];
        run_end = min(cont_chunks + ((k + 1) << (shift - 6)), c_end);
        scene = nt.scene;
      }
      const Params& P = scene_params(st.scenes, scene);
      C.count_out = &st.seg_cnt[nxt * NS + scene];
      C.off = st.seg_off[nxt * NS + scene];
      C.cap = st.seg_cap[nxt * NS + scene];
      for (; c < run_end; ++c) {
        bool valid;
        uint32_t slot = 0, ray = 0, k = 0, branch = 1;
        int depth = 0;
        int64_t pixel = 0;
        d3 T = mk(1.0, 1.0, 1.0), Pw = mk(0.0, 0.0, 0.0);
        if (!prim) {
          const uint32_t local = (c << 6) - e_base + lane_id();
          valid = local < we.cnt;
          if (valid) {
            ray = we.off + local;
            slot = qin.slot[ray];
            T = mk(qin.tr[ray], qin.tg[ray], qin.tb[ray]);
            Pw = mk(qin.pr[ray], qin.pg[ray], qin.pb[ray]);
            const uint32_t meta = qin.meta[ray];
            depth = (int)(meta & 0xFFu);
            branch = META_BRANCH(meta);
            const SlotInfo si = st.slot_info[slot >> shift];
            const uint32_t r = slot & tmask;
            k = r % (uint32_t)si.spp;
            pixel = si.pixel_base + r / (uint32_t)si.spp;
          }
        } else {
          const uint32_t r = (((c - cont_chunks) << 6) & tmask) + lane_id();
          valid = r < nt.n_samples;
          if (valid) {
            slot = (nt.slot << shift) | r;
            const uint32_t pl = r / nt.spp;
            k = r % nt.spp;
            ray = prim_base + nt.ray_base + pl * nt.ss2 + k % nt.ss2;
            pixel = st.slot_info[nt.slot].pixel_base + pl;
          }
        }
        int delta = 0;
        if (valid) {
          C.slot = slot;
          const uint32_t e0 = C.n_ext;
          VertexAdj va;
          const d3 add = shade_vertex(P, C, mk(qin.ox[ray], qin.oy[ray], qin.oz[ray]),
                                      mk(qin.dx[ray], qin.dy[ray], qin.dz[ray]), hit_t[ray], hit_id[ray], T, Pw,
                                      (uint32_t)pixel, k, depth, branch, va);
          delta = (int)(C.n_ext - e0) - 1;
          accum_add(&st.accum[3 * (size_t)slot + 0], add.x);
          accum_add(&st.accum[3 * (size_t)slot + 1], add.y);
          accum_add(&st.accum[3 * (size_t)slot + 2], add.z);
        }
        live_update(st, valid ? (int)(slot >> shift) : -1, delta);
      }
    }
  }
  const uint64_t s_pick = wave_sum_u32(C.pickups), s_tr = wave_sum_u32(C.truncated);
  const uint64_t s_ext = wave_sum_u32(C.n_ext), s_drop = wave_sum_u32(C.dropped);
  if (lane_id() == 0) {
    if (s_pick) atomicAdd((unsigned long long*)&st.ctr[CTR_PICKUP], (unsigned long long)s_pick);
    if (s_tr) atomicAdd((unsigned long long*)&st.ctr[CTR_TRUNC], (unsigned long long)s_tr);
    if (s_ext) atomicAdd((unsigned long long*)&st.ctr[CTR_EXT], (unsigned long long)s_ext);
    if (s_drop) atomicAdd((unsigned long long*)&st.ctr[CTR_DROP], (unsigned long long)s_drop);
  }
}

// One block: padded prefix of the scene regions of queue nxt (NEE's input).
__global__ __launch_bounds__(256) void k_stream_scan(Stream st, int nxt) {
  __shared__ uint32_t s_part[257];
  const uint32_t NS = (uint32_t)st.n_scenes;
  const uint32_t* cnt = st.seg_cnt + nxt * NS;
  const uint32_t* cap = st.seg_cap + nxt * NS;
  auto width = [&](uint32_t s) { return round64(min(cnt[s], cap[s])); };
  const ChunkScan sc = block_scan(NS, width, s_part);
  uint32_t base = sc.base;
  for (uint32_t s = sc.b; s < sc.e; ++s) {
    st.nee_pref[s] = base;
    base += width(s);
  }
  if (threadIdx.x == 255) st.nee_pref[NS] = sc.total;
  if (threadIdx.x == 0) st.ctl->nee_width = sc.total;
}

// Light samples of the diffuse vertices behind the rays of queue nxt.
__global__ __launch_bounds__(256) void k_stream_nee(Stream st, RayQ q, int nxt, ShadowQ sq) {
  const uint32_t n_chunks = st.ctl->nee_width >> 6;
  const uint32_t NS = (uint32_t)st.n_scenes;
  const uint32_t shift = (uint32_t)st.tile_shift, tmask = (1u << shift) - 1u;
  uint32_t emitted = 0, dropped = 0;
  {
    uint32_t c, c_end;
    wave_range(n_chunks, c, c_end);
    while (c < c_end) {
      const uint32_t s = find_entry(st.nee_pref, NS, c << 6);
      const uint32_t s_base = st.nee_pref[s];
      const uint32_t run_end = min(st.nee_pref[s + 1] >> 6, c_end);
      const uint32_t cnt = min(st.seg_cnt[nxt * NS + s], st.seg_cap[nxt * NS + s]);
      const uint32_t off = st.seg_off[nxt * NS + s];
      const Params& P = scene_params(st.scenes, (int)s);
      for (; c < run_end; ++c) {
        const uint32_t local = (c << 6) - s_base + lane_id();
        bool have = false;
        uint32_t slot = 0, w[9];
        d3 x, nrm, ct;
        if (local < cnt) {
          const uint32_t i = off + local;
          const uint32_t meta = q.meta[i];
          if (meta & META_NEE) {
            have = true;
            slot = q.slot[i];
            const SlotInfo si = st.slot_info[slot >> shift];
            const uint32_t r = slot & tmask;
            const uint32_t k = r % (uint32_t)si.spp;
            const int64_t pixel = si.pixel_base + r / (uint32_t)si.spp;
            nee_words(P, (uint32_t)pixel, k, (int)(meta & 0xFFu) - 1, META_BRANCH(meta), w);
            x = mk(q.ox[i], q.oy[i], q.oz[i]);
            nrm = mk(q.nx[i], q.ny[i], q.nz[i]);
            ct = mk(q.tr[i], q.tg[i], q.tb[i]);
          }
        }
        ShadowRay s0, s1, s2;
        int32_t l0 = -1, l1 = -1, l2 = -1;
        const bool v0 = have && nee_one(P, w, 0, x, nrm, ct, s0, l0);
        const bool v1 = have && nee_one(P, w, 1, x, nrm, ct, s1, l1);
        const bool v2 = have && nee_one(P, w, 2, x, nrm, ct, s2, l2);
        uint32_t n_sh = (uint32_t)v0 + (uint32_t)v1 + (uint32_t)v2;
        uint32_t j = wave_append_n(&st.ctl->cnt_sq, n_sh);
        if (j + n_sh > sq.cap) {
          dropped += n_sh;
          n_sh = 0;
        } else {
          if (v0) put_shadow(sq, j++, s0, slot, l0);
          if (v1) put_shadow(sq, j++, s1, slot, l1);
          if (v2) put_shadow(sq, j++, s2, slot, l2);
        }
        emitted += n_sh;
        live_update(st, n_sh > 0 ? (int)(slot >> shift) : -1, (int)n_sh);
      }
    }
  }
  const uint64_t s_sh = wave_sum_u32(emitted), s_drop = wave_sum_u32(dropped);
  if (lane_id() == 0) {
    if (s_sh) atomicAdd((unsigned long long*)&st.ctr[CTR_SHADOW], (unsigned long long)s_sh);
    if (s_drop) atomicAdd((unsigned long long*)&st.ctr[CTR_DROP], (unsigned long long)s_drop);
  }
}

// Any-hit tests of the shared shadow queue; unoccluded rays add their
// contribution.
template <int NL>
__global__ __launch_bounds__(256) void k_stream_shadow(Stream st, ShadowQ sq) {
  PRT_DECLARE_STACK_N(S, st.scenes[0], NL);
  const uint32_t n = min(st.ctl->cnt_sq, sq.cap);
  uint32_t* fetch = &st.ctl->fetch[1];
  const uint32_t shift = (uint32_t)st.tile_shift;
  TravStats ts{0, 0, 0};
  uint32_t extra = 0;
  while (true) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = (uint32_t)__shfl((int)base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane_id();
    uint32_t slot = 0;
    int sid = -1;
    d3 o = mk(0.0, 0.0, 0.0), d = o;
    double tmax = 0.0;
    if (i < n) {
      slot = sq.slot[i];
      o = mk(sq.ox[i], sq.oy[i], sq.oz[i]);
      d = mk(sq.dx[i], sq.dy[i], sq.dz[i]);
      tmax = sq.tmax[i];
      sid = st.slot_info[slot >> shift].scene;
    }
    bool lit = false;
    const uint32_t passes =
        per_scene(st.scenes, sid, [&](const Params& P) { lit = !occluded(P, o, d, tmax, S, ts); });
    extra += passes > 1 ? passes - 1 : 0;
    if (lit) {
      accum_add(&st.accum[3 * (size_t)slot + 0], sq.cr[i]);
      accum_add(&st.accum[3 * (size_t)slot + 1], sq.cg[i]);
      accum_add(&st.accum[3 * (size_t)slot + 2], sq.cb[i]);
    }
    live_update(st, sid >= 0 ? (int)(slot >> shift) : -1, -1);
  }
  flush_stats(st.ctr, CTR_SNODES, ts);
  flush_mixed(st.ctr, extra);
}

// Finished tiles (busy slots whose live count is zero after this iteration's
// kernels): fold every pixel's samples into its scene's image, clear the
// slot's accumulators and return the slot to the free list.
__global__ __launch_bounds__(256) void k_stream_resolve(Stream st) {
  const uint32_t shift = (uint32_t)st.tile_shift;
  for (uint32_t ts = blockIdx.x; ts < (uint32_t)st.n_slots; ts += gridDim.x) {
    const SlotInfo si = st.slot_info[ts];
    if (si.tile < 0 || st.live[ts] != 0) continue;  // free, or rays still in flight
    const Params& P = scene_params(st.scenes, si.scene);
    const TileDesc td = st.tiles[si.tile];
    float* rgb = st.out_rgb[si.scene];
    float* var = st.out_var[si.scene];
    for (int p = threadIdx.x; p < td.n_pixels; p += blockDim.x) {
      const int64_t pix = td.pixel_base + p;
      const int x = (int)(pix % P.W), y = (int)(pix / P.W);
      const size_t ii = (size_t)(P.H - y - 1) * P.W + x;  // render.h:1234
      uint64_t* a = st.accum + 3 * (((size_t)ts << shift) + (size_t)p * P.spp);
      double c[3], v[3];
      fold_pixel(P, a, c, v);
      for (int k = 0; k < 3; ++k) {
        rgb[3 * ii + k] = (float)c[k];
        if (var) var[3 * ii + k] = (float)v[k];
      }
      for (int k = 0; k < 3 * P.spp; ++k) a[k] = 0ull;
    }
    if (threadIdx.x == 0) {
      st.slot_info[ts].tile = -1;
      const uint32_t f = atomicAdd(&st.ctl->free_count, 1u);
      if (f < (uint32_t)st.n_slots) st.free_list[f] = ts;
    }
  }
}

// ------------------------------------------------------------ launchers
hipError_t launch_stream_init(const Stream& st, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_init, dim3(64), dim3(256), 0, s, st);
  return hipGetLastError();
}
hipError_t launch_stream_refill(const Stream& st, int cur, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_refill, dim3(1), dim3(256), 0, s, st, cur);
  return hipGetLastError();
}
hipError_t launch_stream_gen(const Stream& st, const RayQ& q, int cur, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_gen, dim3(blocks), dim3(256), 0, s, st, q, cur);
  return hipGetLastError();
}
hipError_t launch_stream_extend(const Stream& st, const RayQ& q, double* hit_t, int32_t* hit_id, int stack_lds,
                                int blocks, hipStream_t s) {
  if (stack_lds <= 8)
    hipLaunchKernelGGL(k_stream_extend<8>, dim3(blocks), dim3(256), 0, s, st, q, hit_t, hit_id);
  else if (stack_lds <= 16)
    hipLaunchKernelGGL(k_stream_extend<16>, dim3(blocks), dim3(256), 0, s, st, q, hit_t, hit_id);
  else
    hipLaunchKernelGGL(k_stream_extend<32>, dim3(blocks), dim3(256), 0, s, st, q, hit_t, hit_id);
  return hipGetLastError();
}
hipError_t launch_stream_shade(const Stream& st, const RayQ& qin, const double* hit_t, const int32_t* hit_id,
                               const RayQ& qout, int nxt, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_shade, dim3(blocks), dim3(256), 0, s, st, qin, hit_t, hit_id, qout, nxt);
  return hipGetLastError();
}
hipError_t launch_stream_scan(const Stream& st, int nxt, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_scan, dim3(1), dim3(256), 0, s, st, nxt);
  return hipGetLastError();
}
hipError_t launch_stream_nee(const Stream& st, const RayQ& q, int nxt, const ShadowQ& sq, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_nee, dim3(blocks), dim3(256), 0, s, st, q, nxt, sq);
  return hipGetLastError();
}
hipError_t launch_stream_shadow(const Stream& st, const ShadowQ& sq, int stack_lds, int blocks, hipStream_t s) {
  if (stack_lds <= 8)
    hipLaunchKernelGGL(k_stream_shadow<8>, dim3(blocks), dim3(256), 0, s, st, sq);
  else if (stack_lds <= 16)
    hipLaunchKernelGGL(k_stream_shadow<16>, dim3(blocks), dim3(256), 0, s, st, sq);
  else
    hipLaunchKernelGGL(k_stream_shadow<32>, dim3(blocks), dim3(256), 0, s, st, sq);
  return hipGetLastError();
}
hipError_t launch_stream_resolve(const Stream& st, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_resolve, dim3(blocks), dim3(256), 0, s, st);
  return hipGetLastError();
}
