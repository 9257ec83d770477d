// prt_stream.hip — gfx950 kernels of the multi-scene streaming renderer
// (the iteration structure is described in prt_stream.h).
//
// Every kernel that reads a queue walks its chunk list, whose chunks each
// hold rays of one scene, so a wave's scene is uniform and its parameters are
// read with scalar loads.  Shade and NEE take static chunk ranges and write
// into their own slabs (no shared counter); extend and shadow claim chunks
// through sub-range counters.  The per-vertex arithmetic is the per-scene
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

// chunks per claim of the traversal kernels
#define FETCH_CHUNKS 2u


// Chunk c of a chunk list (c wave-uniform), read with scalar loads.
__device__ __forceinline__ ChunkDesc chunk_at(const ChunkDesc* list, uint32_t c) {
  const uint64_t a = (uint64_t)(list + c);
  const uint64_t au = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
                      (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
  return *(const ChunkDesc*)(const __attribute__((address_space(4))) ChunkDesc*)au;
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

// Run bookkeeping of a slab writer: the wave's entries in its slab are
// counted in LDS; a run of one scene starts where that count stood when the
// scene changed.  Runs beyond max_runs (never, with the host's sizing) are
// dropped and counted.  All state is wave-uniform.
struct RunLog {
  RunRec* recs;
  uint32_t max_runs, n, dropped;
  uint32_t slab_off;  // queue index of the slab's first entry
  uint32_t start;     // slab count at the start of the open run
  int scene;          // scene of the open run (-1: none)
  __device__ __forceinline__ void close(uint32_t count) {
    if (scene < 0) return;
    const uint32_t c = count - start;
    if (c > 0) {
      if (n < max_runs) {
        if (lane_id() == 0) recs[n] = RunRec{slab_off + start, c, scene, 0u};
        ++n;
      } else {
        dropped += c;
      }
    }
  }
  // called with the scene of the next chunk, before that chunk's appends
  __device__ __forceinline__ void enter(int s, uint32_t count) {
    if (s == scene) return;
    close(count);
    scene = s;
    start = count;
  }
};

__global__ void k_stream_init(Stream st) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = gid; i < st.n_slots; i += stride) {
    st.free_list[i] = (uint32_t)i;
    st.live[i] = 0;
    st.slot_info[i].tile = -1;  // free
  }
  if (gid == 0) {
    StreamCtl* c = st.ctl;
    c->next_tile = 0;
    c->free_count = (uint32_t)st.n_slots;
    c->n_new = c->n_cont = 0;
    c->status = 0;
    c->iter = 0;
    c->cont_chunks[0] = c->cont_chunks[1] = 0;
    c->cont_entries[0] = c->cont_entries[1] = 0;
    c->ext_chunks = c->sh_chunks = c->sh_entries = 0;
    c->total_prim = c->total_cont = 0;
  }
}

// One block: hand the free slots to the next tiles of the work queue, append
// the chunks of their camera rays to queue cur's chunk list, clear the claim
// counters and update the status.
__global__ __launch_bounds__(256) void k_stream_refill(Stream st, int cur) {
  __shared__ uint32_t s_part[257];
  __shared__ uint32_t s_hdr[4];
  StreamCtl* c = st.ctl;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    const uint32_t free_n = min(c->free_count, (uint32_t)st.n_slots), next = c->next_tile;
    const uint32_t left = (uint32_t)st.n_tiles - next;
    s_hdr[0] = free_n < left ? free_n : left;
    s_hdr[1] = free_n;
    s_hdr[2] = next;
    s_hdr[3] = c->cont_chunks[cur];
  }
  for (uint32_t k = tid; k < 2u * PRT_FETCH_COUNTERS; k += 256) {
    // the extend launch's region, then this parity's shadow region
    const uint32_t region = k < PRT_FETCH_COUNTERS ? 0u : 1u + (uint32_t)cur;
    st.fetch[((size_t)region * PRT_FETCH_COUNTERS + k % PRT_FETCH_COUNTERS) * PRT_FETCH_STRIDE] = 0u;
  }
  __syncthreads();
  const uint32_t n_want = s_hdr[0], free_n = s_hdr[1], next = s_hdr[2], cont_chunks = s_hdr[3];

  // 1. camera rays of the candidate tiles (64-aligned per tile)
  auto tile_rays = [&](uint32_t k) {
    const TileDesc td = st.tiles[next + k];
    const Params& P = st.scenes[td.scene];
    return round64((uint32_t)td.n_pixels * (uint32_t)(P.SS * P.SS));
  };
  const ChunkScan tr = block_scan(n_want, tile_rays, s_part);
  const bool fits = tr.total <= st.prim_cap && cont_chunks + (tr.total >> 6) <= st.max_desc;
  const uint32_t n_new = fits ? n_want : 0u;
  if (fits) {
    uint32_t base = tr.base;
    ChunkDesc* desc = st.desc[cur];
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
      // 2. the tile's camera-ray chunks, after the slabs' chunks
      for (uint32_t j = 0; j < nt.n_rays; j += 64) {
        ChunkDesc d;
        d.base = st.prim_base + base + j;
        d.cnt = min(64u, nt.n_rays - j);
        d.scene = td.scene;
        d.pad_ = 0;
        desc[cont_chunks + ((base + j) >> 6)] = d;
      }
      base += round64(nt.n_rays);
    }
  }
  if (tid == 0) {
    const uint32_t n_cont = c->cont_entries[cur];
    c->n_new = n_new;
    c->n_cont = n_cont;
    c->ext_chunks = cont_chunks + (fits ? tr.total >> 6 : 0u);
    c->free_count = free_n - n_new;
    c->next_tile = next + n_new;
    c->total_cont += n_cont;
    c->iter += 1;
    if (c->status < 2) {
      if (!fits) c->status = 3;
      else if (n_cont == 0 && n_new == 0)
        c->status = (next == (uint32_t)st.n_tiles && free_n == (uint32_t)st.n_slots) ? 1u : 2u;
      else c->status = 0;
    }
  }
}

// Unique sub-pixel camera rays of the new tiles (one wave per tile), in the
// primary region.
__global__ __launch_bounds__(256) void k_stream_gen(Stream st, RayQ q, int cur) {
  const uint32_t n_new = st.ctl->n_new;
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
      const uint32_t i = st.prim_base + nt.ray_base + r;
      q.ox[i] = o.x; q.oy[i] = o.y; q.oz[i] = o.z;
      q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
      q.slot[i] = nt.slot << st.tile_shift;
      q.meta[i] = 0u;
    }
    if (lane_id() == 0) rays += nt.n_rays;
  }
  if (lane_id() == 0 && rays) atomicAdd((unsigned long long*)&st.ctl->total_prim, (unsigned long long)rays);
}

// Waves per SIMD the traversal kernels' registers are budgeted for: the
// 12-entry LDS stack lets five blocks share a CU, so that instance is built
// for 5 waves; the others keep the compiler's budget.
#define PRT_STREAM_WAVES(NL) ((NL) == 12 ? 5 : 1)

// Closest hit for every chunk of queue cur (slab chunks, then camera rays).
template <int NL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PRT_STREAM_WAVES(NL)))) void k_stream_extend(Stream st, RayQ q, int cur, double* __restrict__ hit_t,
                                                       int32_t* __restrict__ hit_id) {
  PRT_DECLARE_STACK_N(S, st.scenes[0], NL, 0);
  // the top-level node copy serves a batch of one scene only
  PRT_TOP_CACHE(S, st.scenes[0].nodes, st.n_scenes == 1 ? st.scenes[0].n_top : 0);
  const ChunkDesc* desc = st.desc[cur];
  const ClaimRange cr = claim_range(st.fetch, st.ctl->ext_chunks);
  TravStats ts{0, 0, 0};
  uint32_t c, c1;
  while (claim(cr, FETCH_CHUNKS, c, c1)) {
    for (; c < c1; ++c) {
      const ChunkDesc dc = chunk_at(desc, c);
      const Params& P = scene_params(st.scenes, dc.scene);
      if (lane_id() < dc.cnt) {
        const uint32_t i = dc.base + lane_id();
        double t, u, v;
        int32_t id;
        closest_hit(P, mk(q.ox[i], q.oy[i], q.oz[i]), mk(q.dx[i], q.dy[i], q.dz[i]), t, id, u, v, S, ts);
        hit_t[i] = t;
        hit_id[i] = id;
      }
    }
  }
  flush_stats(st.ctr, CTR_NODES, ts);
}

// One vertex per ray of queue cur's slab chunks and per camera sample of the
// new tiles (T / 64 chunks per tile); children go to this wave's slab of
// queue nxt, logged as runs of one scene.
__global__ __launch_bounds__(256) void k_stream_shade(Stream st, RayQ qin, const double* __restrict__ hit_t,
                                                      const int32_t* __restrict__ hit_id, RayQ qout, int nxt) {
  __shared__ uint32_t s_cnt[4];  // entries in each wave's slab
  const int cur = nxt ^ 1;
  const uint32_t wib = threadIdx.x >> 6;
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (lane_id() == 0) s_cnt[wib] = 0;
  const uint32_t cont_chunks = st.ctl->cont_chunks[cur], n_new = st.ctl->n_new;
  const uint32_t shift = (uint32_t)st.tile_shift, tmask = (1u << shift) - 1u;
  const uint32_t n_chunks = cont_chunks + (n_new << (shift - 6));  // T >= 64: T/64 chunks per new tile
  const ChunkDesc* desc = st.desc[cur];
  ShadeCtx C;
  C.qout = qout;
  C.count_out = &s_cnt[wib];
  C.off = wave * st.slab_ray;
  C.cap = st.slab_ray;
  C.truncated = C.n_ext = C.pickups = C.dropped = 0;
  C.direct = 0;
  RunLog log{st.ray_runs + (size_t)wave * st.max_runs, st.max_runs, 0u, 0u, C.off, 0u, -1};
  uint32_t c, c_end;
  wave_range(n_chunks, c, c_end);
  for (; c < c_end; ++c) {
    const bool prim = c >= cont_chunks;
    ChunkDesc dc{0, 0, 0, 0};
    NewTile nt{};
    int scene;
    if (!prim) {
      dc = chunk_at(desc, c);
      scene = dc.scene;
    } else {
      nt = st.new_tiles[(c - cont_chunks) >> (shift - 6)];
      scene = nt.scene;
    }
    log.enter(scene, min(s_cnt[wib], st.slab_ray));
    const Params& P = scene_params(st.scenes, scene);
    bool valid;
    uint32_t slot = 0, ray = 0, k = 0, branch = 1;
    int depth = 0;
    int64_t pixel = 0;
    d3 T = mk(1.0, 1.0, 1.0), Pw = mk(0.0, 0.0, 0.0);
    if (!prim) {
      valid = lane_id() < dc.cnt;
      if (valid) {
        ray = dc.base + lane_id();
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
        ray = st.prim_base + nt.ray_base + pl * nt.ss2 + k % nt.ss2;
        pixel = st.slot_info[nt.slot].pixel_base + pl;
      }
    }
    int delta = 0;
    if (valid) {
      C.slot = slot;
      const uint32_t e0 = C.n_ext;
      VertexAdj va;
      const d3 add = shade_vertex(P, P.mats, C, mk(qin.ox[ray], qin.oy[ray], qin.oz[ray]),
                                  mk(qin.dx[ray], qin.dy[ray], qin.dz[ray]), hit_t[ray], hit_id[ray], T, Pw,
                                  (uint32_t)pixel, k, depth, branch, va);
      delta = (int)(C.n_ext - e0) - 1;
      accum_add(&st.accum[3 * (size_t)slot + 0], add.x);
      accum_add(&st.accum[3 * (size_t)slot + 1], add.y);
      accum_add(&st.accum[3 * (size_t)slot + 2], add.z);
    }
    live_update(st, valid ? (int)(slot >> shift) : -1, delta);
  }
  log.close(min(s_cnt[wib], st.slab_ray));
  const uint64_t s_pick = wave_sum_u32(C.pickups), s_tr = wave_sum_u32(C.truncated);
  const uint64_t s_ext = wave_sum_u32(C.n_ext), s_drop = wave_sum_u32(C.dropped);
  if (lane_id() == 0) {
    st.ray_nruns[wave] = log.n < st.max_runs ? log.n : st.max_runs;
    if (s_pick) atomicAdd((unsigned long long*)&st.ctr[CTR_PICKUP], (unsigned long long)s_pick);
    if (s_tr) atomicAdd((unsigned long long*)&st.ctr[CTR_TRUNC], (unsigned long long)s_tr);
    if (s_ext) atomicAdd((unsigned long long*)&st.ctr[CTR_EXT], (unsigned long long)s_ext);
    if (s_drop + log.dropped)
      atomicAdd((unsigned long long*)&st.ctr[CTR_DROP], (unsigned long long)(s_drop + log.dropped));
  }
}

// The chunk list of a queue from its slab writers' runs (which 0 / 1: ray
// queue p; 2: the shadow queue), in writer order, and its chunk and entry
// totals.  Block g lists the chunks of the PRT_INDEX_SLABS writers from
// g * PRT_INDEX_SLABS on: it sums the chunks of every writer before them over
// their run records, scans its own writers' chunk counts in one wave, and
// writes their descriptors, a wave per writer and a lane per chunk.  Block 0
// sums every writer instead and stores the totals.  Chunks beyond the list's
// capacity (never, with the host's sizing) are dropped and counted.
__global__ __launch_bounds__(256) void k_stream_index(Stream st, int which) {
  __shared__ uint32_t s_part[2][4];
  __shared__ uint32_t s_first[PRT_INDEX_SLABS];
  const bool sh = which == 2;
  const RunRec* runs = sh ? st.sh_runs : st.ray_runs;
  const uint32_t* nruns = sh ? st.sh_nruns : st.ray_nruns;
  ChunkDesc* desc = sh ? st.sh_desc : st.desc[which];
  const uint32_t max_desc = sh ? st.max_sh_desc : st.max_desc;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
  const uint32_t W = PRT_SLAB_WAVES;
  const uint32_t w0 = blockIdx.x * (uint32_t)PRT_INDEX_SLABS;  // < W by the grid
  const uint32_t nw = min(W - w0, (uint32_t)PRT_INDEX_SLABS);
  const bool totals = blockIdx.x == 0;
  const uint32_t lim = totals ? W : w0;
  uint32_t ch = 0, ent = 0;
  for (uint32_t w = tid; w < lim; w += 256u) {
    const uint32_t n = min(nruns[w], st.max_runs);
    for (uint32_t r = 0; r < n; ++r) {
      const uint32_t cnt = runs[(size_t)w * st.max_runs + r].cnt;
      ch += (cnt + 63u) >> 6;
      ent += cnt;
    }
  }
  ch = (uint32_t)wave_sum_u32(ch);
  ent = (uint32_t)wave_sum_u32(ent);
  if (lane == 0) {
    s_part[0][wv] = ch;
    s_part[1][wv] = ent;
  }
  if (wv == 0) {
    // exclusive scan of this block's writers' chunk counts
    uint32_t n = 0;
    if (lane < nw) {
      const uint32_t w = w0 + lane, nr = min(nruns[w], st.max_runs);
      for (uint32_t r = 0; r < nr; ++r) n += (runs[(size_t)w * st.max_runs + r].cnt + 63u) >> 6;
    }
    uint32_t incl = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
      if (lane >= (uint32_t)o) incl += y;
    }
    s_first[lane] = incl - n;
  }
  __syncthreads();
  const uint32_t sum_ch = s_part[0][0] + s_part[0][1] + s_part[0][2] + s_part[0][3];
  uint32_t lost = 0;
  for (uint32_t i = wv; i < nw; i += 4u) {
    const uint32_t w = w0 + i, nr = min(nruns[w], st.max_runs);
    uint32_t base = (totals ? 0u : sum_ch) + s_first[i];
    for (uint32_t r = 0; r < nr; ++r) {
      const RunRec rr = runs[(size_t)w * st.max_runs + r];
      const uint32_t nch = (rr.cnt + 63u) >> 6;
      for (uint32_t j = lane; j < nch; j += 64u) {
        const uint32_t c = base + j, m = min(64u, rr.cnt - (j << 6));
        if (c < max_desc) {
          ChunkDesc d;
          d.base = rr.off + (j << 6);
          d.cnt = m;
          d.scene = rr.scene;
          d.pad_ = 0;
          desc[c] = d;
        } else {
          lost += m;
        }
      }
      base += nch;
    }
  }
  const uint64_t s_lost = wave_sum_u32(lost);
  if (lane == 0 && s_lost) atomicAdd((unsigned long long*)&st.ctr[CTR_DROP], (unsigned long long)s_lost);
  if (totals && tid == 0) {
    StreamCtl* c = st.ctl;
    const uint32_t n_ch = min(sum_ch, max_desc);
    const uint32_t tot_ent = s_part[1][0] + s_part[1][1] + s_part[1][2] + s_part[1][3];
    if (sh) {
      c->sh_chunks = n_ch;
      c->sh_entries = tot_ent;
    } else {
      c->cont_chunks[which] = n_ch;
      c->cont_entries[which] = tot_ent;
    }
  }
}

// Light samples of the diffuse vertices behind the rays of queue nxt's slab
// chunks; shadow rays go to this wave's shadow slab, logged as runs.
// MODE (NeeMode): NEE_ONE, no scene of the batch has more than one light and
// a wave reads the light of its chunk's scene into wave-uniform registers;
// NEE_LDS, no scene has more than NEE_LDS_LIGHTS lights and a wave copies its
// chunk's scene's lights into its LDS slots; NEE_GLOBAL, light records are
// read per sample.  Either copy is refreshed when the chunk's scene changes.
template <int MODE>
__global__ __launch_bounds__(256) void k_stream_nee(Stream st, RayQ q, int nxt, ShadowQ sq) {
  __shared__ uint32_t s_cnt[4];  // entries in each wave's shadow slab
  __shared__ LightSlot s_lights[MODE == NEE_LDS ? 4 * NEE_LDS_LIGHTS : 1];
  const uint32_t wib = threadIdx.x >> 6;
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (lane_id() == 0) s_cnt[wib] = 0;
  const uint32_t shift = (uint32_t)st.tile_shift, tmask = (1u << shift) - 1u;
  const ChunkDesc* desc = st.desc[nxt];
  const uint32_t off = wave * st.slab_sh;
  RunLog log{st.sh_runs + (size_t)wave * st.max_runs, st.max_runs, 0u, 0u, off, 0u, -1};
  uint32_t emitted = 0, dropped = 0;
  uint32_t c, c_end;
  wave_range(st.ctl->cont_chunks[nxt], c, c_end);
  UniformLight U{};
  LightSlot* slots = s_lights + (MODE == NEE_LDS ? wib * NEE_LDS_LIGHTS : 0);
  int32_t u_scene = -1;
  // an entry's fields are all read before its NEE flag is tested (one
  // memory round trip less per chunk)
  for (; c < c_end; ++c) {
    const ChunkDesc dc = chunk_at(desc, c);
    log.enter(dc.scene, min(s_cnt[wib], st.slab_sh));
    const Params& P = scene_params(st.scenes, dc.scene);
    if (MODE != NEE_GLOBAL && (int32_t)dc.scene != u_scene) {
      u_scene = (int32_t)dc.scene;
      if (MODE == NEE_ONE && P.n_lights == 1) U = uniform_light(P, P.lights[0]);
      if (MODE == NEE_LDS) load_light_slots(P, slots);
    }
    bool have = false;
    uint32_t slot = 0, w[9];
    d3 x, nrm, ct;
    if (lane_id() < dc.cnt) {
      const uint32_t i = dc.base + lane_id();
      const uint32_t meta = q.meta[i];
      slot = q.slot[i];
      x = mk(q.ox[i], q.oy[i], q.oz[i]);
      nrm = mk(q.nx[i], q.ny[i], q.nz[i]);
      ct = mk(q.tr[i], q.tg[i], q.tb[i]);
      if (meta & META_NEE) {
        have = true;
        const SlotInfo si = st.slot_info[slot >> shift];
        const uint32_t r = slot & tmask;
        const uint32_t k = r % (uint32_t)si.spp;
        const int64_t pixel = si.pixel_base + r / (uint32_t)si.spp;
        nee_words(P, (uint32_t)pixel, k, (int)(meta & 0xFFu) - 1, META_BRANCH(meta), w);
      }
    }
    // one light sample at a time, appended through the LDS count: only one
    // shadow ray's registers are live
    uint32_t n_sh = 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      ShadowRay sr;
      int32_t l = -1;
      const bool v = have && (MODE == NEE_ONE   ? nee_one_single(U, w, m, x, nrm, ct, sr, l)
                              : MODE == NEE_LDS ? nee_one_slots(P, slots, w, m, x, nrm, ct, sr, l)
                                                : nee_one(P, w, m, x, nrm, ct, sr, l));
      const uint32_t j = wave_append_bit(&s_cnt[wib], v);
      if (v) {
        if (j < st.slab_sh) {
          put_shadow(sq, off + j, sr, slot, l);
          ++n_sh;
        } else {
          ++dropped;
        }
      }
    }
    emitted += n_sh;
  }
  log.close(min(s_cnt[wib], st.slab_sh));
  const uint64_t s_sh = wave_sum_u32(emitted), s_drop = wave_sum_u32(dropped);
  if (lane_id() == 0) {
    st.sh_nruns[wave] = log.n < st.max_runs ? log.n : st.max_runs;
    if (s_sh) atomicAdd((unsigned long long*)&st.ctr[CTR_SHADOW], (unsigned long long)s_sh);
    if (s_drop + log.dropped)
      atomicAdd((unsigned long long*)&st.ctr[CTR_DROP], (unsigned long long)(s_drop + log.dropped));
  }
}

// Any-hit tests of the shadow queue's chunks; unoccluded rays add their
// contribution.
template <int NL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PRT_STREAM_WAVES(NL)))) void k_stream_shadow(
    Stream st, ShadowQ sq, int par) {
  PRT_DECLARE_STACK_N(S, st.scenes[0], NL, 1);
  PRT_TOP_CACHE(S, st.scenes[0].nodes, st.n_scenes == 1 ? st.scenes[0].n_top : 0);
  const ClaimRange cr =
      claim_range(st.fetch + (size_t)(1 + par) * PRT_FETCH_COUNTERS * PRT_FETCH_STRIDE, st.ctl->sh_chunks);
  TravStats ts{0, 0, 0};
  uint32_t c, c1;
  while (claim(cr, FETCH_CHUNKS, c, c1)) {
    for (; c < c1; ++c) {
      const ChunkDesc dc = chunk_at(st.sh_desc, c);
      const Params& P = scene_params(st.scenes, dc.scene);
      if (lane_id() < dc.cnt) {
        const uint32_t i = dc.base + lane_id();
        const uint32_t slot = sq.slot[i];
        if (!occluded(P, mk(sq.ox[i], sq.oy[i], sq.oz[i]), mk(sq.dx[i], sq.dy[i], sq.dz[i]), sq.tmax[i], S, ts)) {
          accum_add(&st.accum[3 * (size_t)slot + 0], sq.cr[i]);
          accum_add(&st.accum[3 * (size_t)slot + 1], sq.cg[i]);
          accum_add(&st.accum[3 * (size_t)slot + 2], sq.cb[i]);
        }
      }
    }
  }
  flush_stats(st.ctr, CTR_SNODES, ts);
}

// Finished tiles (busy slots whose live count is zero after this iteration's
// shade; launched after the shadow kernels of the earlier iterations): fold
// every pixel's samples into its scene's image, clear the slot's
// accumulators and return the slot to the free list.  One thread reads the
// slot's record for the whole block and the slot is released only after
// every wave of the block has finished its pixels: a wave that read the
// record after another wave had already released the slot would skip its
// pixels, leaving them unwritten and their accumulators uncleared for the
// slot's next tile.
__global__ __launch_bounds__(256) void k_stream_resolve(Stream st) {
  __shared__ SlotInfo s_si;
  __shared__ int32_t s_live;
  const uint32_t shift = (uint32_t)st.tile_shift;
  for (uint32_t ts = blockIdx.x; ts < (uint32_t)st.n_slots; ts += gridDim.x) {
    if (threadIdx.x == 0) {
      s_si = st.slot_info[ts];
      s_live = st.live[ts];
    }
    __syncthreads();
    const SlotInfo si = s_si;
    const bool done = si.tile >= 0 && s_live == 0;  // busy, and no ray of its tile left
    __syncthreads();  // s_si / s_live are rewritten for the next slot only after every thread read them
    if (!done) continue;
    const Params& P = scene_params(st.scenes, si.scene);
    const TileDesc td = st.tiles[si.tile];
    float* rgb = st.out_rgb[si.scene];
    float* var = st.out_var[si.scene];
    // three threads per pixel, one per channel
    uint64_t* tile_acc = st.accum + 3 * ((size_t)ts << shift);
    for (int j = threadIdx.x; j < 3 * td.n_pixels; j += blockDim.x) {
      const int p = j / 3, ch = j - 3 * p;
      const int64_t pix = td.pixel_base + p;
      const int x = (int)(pix % P.W), y = (int)(pix / P.W);
      const size_t ii = (size_t)(P.H - y - 1) * P.W + x;  // render.h:1234
      double c, v;
      fold_channel(P, tile_acc + 3 * (size_t)p * P.spp, ch, c, v);
      rgb[3 * ii + ch] = (float)c;
      if (var) var[3 * ii + ch] = (float)v;
    }
    __syncthreads();  // every pixel of the tile is folded
    for (int j = threadIdx.x; j < 3 * td.n_pixels * P.spp; j += blockDim.x) tile_acc[j] = 0ull;
    __syncthreads();  // and cleared
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
hipError_t launch_stream_extend(const Stream& st, const RayQ& q, int cur, double* hit_t, int32_t* hit_id,
                                int stack_lds, int blocks, hipStream_t s) {
  if (stack_lds <= 8)
    hipLaunchKernelGGL(k_stream_extend<8>, dim3(blocks), dim3(256), 0, s, st, q, cur, hit_t, hit_id);
  else if (stack_lds <= 12)
    hipLaunchKernelGGL(k_stream_extend<12>, dim3(blocks), dim3(256), 0, s, st, q, cur, hit_t, hit_id);
  else if (stack_lds <= 16)
    hipLaunchKernelGGL(k_stream_extend<16>, dim3(blocks), dim3(256), 0, s, st, q, cur, hit_t, hit_id);
  else
    hipLaunchKernelGGL(k_stream_extend<32>, dim3(blocks), dim3(256), 0, s, st, q, cur, hit_t, hit_id);
  return hipGetLastError();
}
hipError_t launch_stream_shade(const Stream& st, const RayQ& qin, const double* hit_t, const int32_t* hit_id,
                               const RayQ& qout, int nxt, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_shade, dim3(PRT_SLAB_WAVES / 4), dim3(256), 0, s, st, qin, hit_t, hit_id, qout, nxt);
  return hipGetLastError();
}
hipError_t launch_stream_index(const Stream& st, int which, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_index, dim3((PRT_SLAB_WAVES + PRT_INDEX_SLABS - 1) / PRT_INDEX_SLABS), dim3(256), 0, s, st,
                     which);
  return hipGetLastError();
}
hipError_t launch_stream_nee(const Stream& st, const RayQ& q, int nxt, const ShadowQ& sq, hipStream_t s) {
  if (st.nee_max_lights <= 1)
    hipLaunchKernelGGL(k_stream_nee<NEE_ONE>, dim3(PRT_SLAB_WAVES / 4), dim3(256), 0, s, st, q, nxt, sq);
  else if (st.nee_max_lights <= NEE_LDS_LIGHTS)
    hipLaunchKernelGGL(k_stream_nee<NEE_LDS>, dim3(PRT_SLAB_WAVES / 4), dim3(256), 0, s, st, q, nxt, sq);
  else
    hipLaunchKernelGGL(k_stream_nee<NEE_GLOBAL>, dim3(PRT_SLAB_WAVES / 4), dim3(256), 0, s, st, q, nxt, sq);
  return hipGetLastError();
}
hipError_t launch_stream_shadow(const Stream& st, const ShadowQ& sq, int par, int stack_lds, int blocks,
                                hipStream_t s) {
  if (par != 0 && par != 1) return hipErrorInvalidValue;
  if (stack_lds <= 8)
    hipLaunchKernelGGL(k_stream_shadow<8>, dim3(blocks), dim3(256), 0, s, st, sq, par);
  else if (stack_lds <= 12)
    hipLaunchKernelGGL(k_stream_shadow<12>, dim3(blocks), dim3(256), 0, s, st, sq, par);
  else if (stack_lds <= 16)
    hipLaunchKernelGGL(k_stream_shadow<16>, dim3(blocks), dim3(256), 0, s, st, sq, par);
  else
    hipLaunchKernelGGL(k_stream_shadow<32>, dim3(blocks), dim3(256), 0, s, st, sq, par);
  return hipGetLastError();
}
hipError_t launch_stream_resolve(const Stream& st, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_resolve, dim3(blocks), dim3(256), 0, s, st);
  return hipGetLastError();
}
