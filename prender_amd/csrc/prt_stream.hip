// prt_stream.hip — gfx950 kernels of the multi-scene streaming renderer
// (the iteration structure is described in prt_stream.h).
//
// Rays of different scenes share the queues.  A wave's rays usually come from
// one tile (refill writes tiles contiguously and the appends keep a wave's
// children together), so the kernels run each distinct scene of a wave in
// turn ("waterfall") with that scene's parameters at a wave-uniform address.
// The per-vertex arithmetic is the per-scene renderer's (prt_shade.h), and a
// camera sample's random numbers depend only on (pixel, sample, depth,
// branch), so a tile renders bit-identically to the same pixels of a
// per-scene render.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"
#include "prt_math.h"
#include "prt_shade.h"
#include "prt_stream.h"
#include "prt_trace.h"

// Calls f(scenes[s]) for the lanes whose scene is s, one distinct s at a time
// (s wave-uniform); lanes with sid < 0 take no part.  Every lane of the wave
// must reach this call.
template <class F>
__device__ __forceinline__ void per_scene(const Params* __restrict__ scenes, int sid, F&& f) {
  bool pend = sid >= 0;
  uint64_t m = __ballot(pend);
  while (m) {
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int s = __builtin_amdgcn_readfirstlane(__shfl(sid, leader, 64));
    // the address is formed from the uniform s before the lane test (so the
    // compiler cannot substitute the per-lane sid, equal inside the branch)
    // and lives in the constant address space: the parameters are read
    // with scalar loads, as kernel arguments are
    const uint64_t a = (uint64_t)(scenes + s);
    const uint64_t au = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    const __attribute__((address_space(4))) Params* Pc = (const __attribute__((address_space(4))) Params*)au;
    if (pend && sid == s) {
      f(*(const Params*)Pc);
      pend = false;
    }
    m = __ballot(pend);
  }
}

// Adds delta to the live-ray count of tile slot ts (one atomic per distinct
// slot of the wave).  The update that brings a count to zero queues the slot
// for resolve: a ray's own retirement and the children it spawned are one
// update, so a count cannot reach zero while any ray of its tile exists.
// Every lane of the wave must reach this call.
__device__ __forceinline__ void live_update(const Stream& st, int ts, int delta) {
  bool pend = ts >= 0 && delta != 0;
  uint64_t m = __ballot(pend);
  while (m) {
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int t0 = __builtin_amdgcn_readfirstlane(__shfl(ts, leader, 64));
    const bool mine = pend && ts == t0;
    const int sum = (int)(uint32_t)wave_sum_u32((uint32_t)(mine ? delta : 0));
    if ((int)lane_id() == leader && sum != 0) {
      const int old = atomicAdd(&st.live[t0], sum);
      if (old + sum == 0) {
        const uint32_t k = atomicAdd(&st.ctl->done_count, 1u);
        if (k < (uint32_t)st.n_slots) st.done_list[k] = (uint32_t)t0;  // a slot finishes once per iteration
      }
    }
    pend = pend && !mine;
    m = __ballot(pend);
  }
}

__global__ void k_stream_init(Stream st) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = gid; i < st.n_slots; i += gridDim.x * blockDim.x) {
    st.free_list[i] = (uint32_t)i;
    st.live[i] = 0;
  }
  if (gid == 0) {
    StreamCtl* c = st.ctl;
    c->next_tile = 0;
    c->free_count = (uint32_t)st.n_slots;
    c->done_count = 0;
    c->n_new = c->n_cont = c->n_prim = 0;
    c->status = 0;
    c->iter = 0;
    c->cnt[0] = c->cnt[1] = 0;
    c->cnt_sq = 0;
    c->fetch[0] = c->fetch[1] = 0;
    c->pad_ = 0;
    c->total_prim = c->total_cont = 0;
  }
}

// One block: hand the free slots to the next tiles of the work queue and lay
// out their camera rays after the carried-over rays of queue `cur`; reset the
// per-iteration counters.
__global__ __launch_bounds__(256) void k_stream_refill(Stream st, int cur) {
  __shared__ uint32_t s_part[256];
  __shared__ uint32_t s_hdr[5];
  StreamCtl* c = st.ctl;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    const uint32_t free_n = min(c->free_count, (uint32_t)st.n_slots), next = c->next_tile;
    uint32_t n_cont = c->cnt[cur];
    if (n_cont > st.q_cap) n_cont = st.q_cap;  // only after an overflow, which the drop counter reports
    const uint32_t left = (uint32_t)st.n_tiles - next;
    s_hdr[0] = free_n < left ? free_n : left;
    s_hdr[1] = n_cont;
    s_hdr[2] = free_n;
    s_hdr[3] = next;
  }
  __syncthreads();
  const uint32_t n_new = s_hdr[0], n_cont = s_hdr[1], free_n = s_hdr[2], next = s_hdr[3];
  const uint32_t chunk = (n_new + 255u) / 256u;
  const uint32_t b = min(tid * chunk, n_new), e = min(b + chunk, n_new);
  uint32_t rays = 0;
  for (uint32_t k = b; k < e; ++k) {
    const TileDesc td = st.tiles[next + k];
    const Params& P = st.scenes[td.scene];
    rays += (uint32_t)td.n_pixels * (uint32_t)(P.SS * P.SS);
  }
  s_part[tid] = rays;
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < 256; ++i) {
      const uint32_t v = s_part[i];
      s_part[i] = acc;
      acc += v;
    }
    s_hdr[4] = acc;
  }
  __syncthreads();
  const uint32_t n_prim = s_hdr[4];
  const bool fits = (uint64_t)n_cont + n_prim <= st.q_cap;
  if (fits) {
    uint32_t base = s_part[tid];
    for (uint32_t k = b; k < e; ++k) {
      const uint32_t tile = next + k;
      const TileDesc td = st.tiles[tile];
      const Params& P = st.scenes[td.scene];
      const uint32_t slot = st.free_list[free_n - 1u - k];
      NewTile nt;
      nt.slot = slot;
      nt.ray_base = base;
      nt.n_rays = (uint32_t)td.n_pixels * (uint32_t)(P.SS * P.SS);
      nt.n_samples = (uint32_t)td.n_pixels * (uint32_t)P.spp;
      st.new_tiles[k] = nt;
      SlotInfo si;
      si.scene = td.scene;
      si.tile = (int32_t)tile;
      si.pixel_base = td.pixel_base;
      st.slot_info[slot] = si;
      st.live[slot] = (int32_t)nt.n_samples;
      base += nt.n_rays;
    }
  }
  if (tid == 0) {
    const uint32_t taken = fits ? n_new : 0u;
    const uint32_t prim = fits ? n_prim : 0u;
    c->n_new = taken;
    c->n_cont = n_cont;
    c->n_prim = prim;
    c->free_count = free_n - taken;
    c->next_tile = next + taken;
    c->cnt[cur] = n_cont + prim;
    c->cnt[cur ^ 1] = 0;
    c->cnt_sq = 0;
    c->fetch[0] = c->fetch[1] = 0;
    c->done_count = 0;
    c->total_prim += prim;
    c->total_cont += n_cont;
    c->iter += 1;
    if (c->status < 2) {
      if (!fits) c->status = 3;
      else if (n_cont == 0 && taken == 0)
        c->status = (next == (uint32_t)st.n_tiles && free_n == (uint32_t)st.n_slots) ? 1u : 2u;
      else c->status = 0;
    }
  }
}

// Unique sub-pixel camera rays of the new tiles (one wave per tile).
__global__ __launch_bounds__(256) void k_stream_gen(Stream st, RayQ q) {
  const uint32_t n_new = st.ctl->n_new, n_cont = st.ctl->n_cont;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t k = wave; k < n_new; k += n_waves) {
    const NewTile nt = st.new_tiles[k];
    const SlotInfo si = st.slot_info[nt.slot];
    const Params& P = st.scenes[si.scene];
    const uint32_t ss2 = (uint32_t)(P.SS * P.SS);
    for (uint32_t r = lane_id(); r < nt.n_rays; r += 64) {
      const int64_t pix = si.pixel_base + r / ss2;
      const int sub = (int)(r % ss2);
      d3 o, d;
      camera_ray(P, (int)(pix % P.W), (int)(pix / P.W), sub % P.SS, sub / P.SS, o, d);
      const uint32_t i = n_cont + nt.ray_base + r;
      q.ox[i] = o.x; q.oy[i] = o.y; q.oz[i] = o.z;
      q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
      q.slot[i] = nt.slot << st.tile_shift;
      q.meta[i] = 0u;
    }
  }
}

// Closest hit for the whole current queue.
template <int NL>
__global__ __launch_bounds__(256) void k_stream_extend(Stream st, RayQ q, int cur, double* __restrict__ hit_t,
                                                       int32_t* __restrict__ hit_id) {
  PRT_DECLARE_STACK_N(S, st.scenes[0], NL);
  const uint32_t n = min(st.ctl->cnt[cur], q.cap);
  uint32_t* fetch = &st.ctl->fetch[0];
  TravStats ts{0, 0};
  while (true) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = (uint32_t)__shfl((int)base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane_id();
    const int sid = i < n ? st.slot_info[q.slot[i] >> st.tile_shift].scene : -1;
    per_scene(st.scenes, sid, [&](const Params& P) {
      double t, u, v;
      int32_t id;
      closest_hit(P, mk(q.ox[i], q.oy[i], q.oz[i]), mk(q.dx[i], q.dy[i], q.dz[i]), t, id, u, v, S, ts);
      hit_t[i] = t;
      hit_id[i] = id;
    });
  }
  flush_stats(st.ctr, CTR_NODES, ts);
}

// One vertex per carried-over ray and per camera sample of the new tiles.
__global__ __launch_bounds__(256) void k_stream_shade(Stream st, RayQ qin, const double* __restrict__ hit_t,
                                                      const int32_t* __restrict__ hit_id, RayQ qout, int nxt) {
  const uint32_t n_cont = st.ctl->n_cont, n_new = st.ctl->n_new;
  const uint32_t shift = (uint32_t)st.tile_shift, tmask = (1u << shift) - 1u;
  const uint32_t total = n_cont + (n_new << shift);
  const uint32_t n_round = (total + 63u) & ~63u;
  const uint32_t stride = gridDim.x * blockDim.x;
  ShadeCtx C;
  C.qout = qout;
  C.count_out = &st.ctl->cnt[nxt];
  C.truncated = C.n_ext = C.pickups = C.dropped = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
    int sid = -1;
    uint32_t slot = 0;
    bool prim = false;
    NewTile nt{0, 0, 0, 0};
    if (i < n_cont) {
      slot = qin.slot[i];
      sid = st.slot_info[slot >> shift].scene;
    } else if (i < total) {
      const uint32_t j = i - n_cont;
      nt = st.new_tiles[j >> shift];
      const uint32_t r = j & tmask;
      if (r < nt.n_samples) {
        prim = true;
        slot = (nt.slot << shift) | r;
        sid = st.slot_info[nt.slot].scene;
      }
    }
    d3 add = mk(0.0, 0.0, 0.0);
    int delta = 0;
    per_scene(st.scenes, sid, [&](const Params& P) {
      const uint32_t r = slot & tmask;
      const uint32_t pl = r / (uint32_t)P.spp, k = r % (uint32_t)P.spp;
      const int64_t pixel = st.slot_info[slot >> shift].pixel_base + pl;
      d3 o, d, T, Pw;
      int depth;
      uint32_t branch, ray;
      if (prim) {
        const uint32_t ss2 = (uint32_t)(P.SS * P.SS);
        ray = n_cont + nt.ray_base + pl * ss2 + k % ss2;
        T = mk(1.0, 1.0, 1.0);
        Pw = mk(0.0, 0.0, 0.0);
        depth = 0;
        branch = 1;
      } else {
        ray = i;
        T = mk(qin.tr[i], qin.tg[i], qin.tb[i]);
        Pw = mk(qin.pr[i], qin.pg[i], qin.pb[i]);
        const uint32_t meta = qin.meta[i];
        depth = (int)(meta & 0xFFu);
        branch = META_BRANCH(meta);
      }
      o = mk(qin.ox[ray], qin.oy[ray], qin.oz[ray]);
      d = mk(qin.dx[ray], qin.dy[ray], qin.dz[ray]);
      C.slot = slot;
      const uint32_t e0 = C.n_ext;
      VertexAdj va;
      add = shade_vertex(P, C, o, d, hit_t[ray], hit_id[ray], T, Pw, (uint32_t)pixel, k, depth, branch, va);
      delta = (int)(C.n_ext - e0) - 1;
    });
    if (sid >= 0) {
      accum_add(&st.accum[3 * (size_t)slot + 0], add.x);
      accum_add(&st.accum[3 * (size_t)slot + 1], add.y);
      accum_add(&st.accum[3 * (size_t)slot + 2], add.z);
    }
    live_update(st, sid >= 0 ? (int)(slot >> shift) : -1, delta);
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

// Light samples of the diffuse vertices behind the rays of queue `nxt`.
__global__ __launch_bounds__(256) void k_stream_nee(Stream st, RayQ q, int nxt, ShadowQ sq) {
  const uint32_t n = min(st.ctl->cnt[nxt], q.cap);
  const uint32_t n_round = (n + 63u) & ~63u;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t shift = (uint32_t)st.tile_shift, tmask = (1u << shift) - 1u;
  uint32_t emitted = 0, dropped = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
    ShadowRay s0, s1, s2;
    bool v0 = false, v1 = false, v2 = false;
    int32_t l0 = -1, l1 = -1, l2 = -1;
    uint32_t slot = 0, meta = 0;
    int sid = -1;
    if (i < n) {
      meta = q.meta[i];
      if (meta & META_NEE) {
        slot = q.slot[i];
        sid = st.slot_info[slot >> shift].scene;
      }
    }
    per_scene(st.scenes, sid, [&](const Params& P) {
      const uint32_t r = slot & tmask;
      const uint32_t pl = r / (uint32_t)P.spp, k = r % (uint32_t)P.spp;
      const int64_t pixel = st.slot_info[slot >> shift].pixel_base + pl;
      nee_vertex(P, (uint32_t)pixel, k, (int)(meta & 0xFFu) - 1, META_BRANCH(meta), mk(q.ox[i], q.oy[i], q.oz[i]),
                 mk(q.nx[i], q.ny[i], q.nz[i]), mk(q.tr[i], q.tg[i], q.tb[i]), s0, s1, s2, v0, v1, v2, l0, l1, l2);
    });
    uint32_t cnt = (uint32_t)v0 + (uint32_t)v1 + (uint32_t)v2;
    uint32_t j = wave_append_n(&st.ctl->cnt_sq, cnt);
    if (j + cnt > sq.cap) {
      dropped += cnt;
      cnt = 0;
    } else {
      if (v0) put_shadow(sq, j++, s0, slot, l0);
      if (v1) put_shadow(sq, j++, s1, slot, l1);
      if (v2) put_shadow(sq, j++, s2, slot, l2);
      emitted += cnt;
    }
    live_update(st, cnt > 0 ? (int)(slot >> shift) : -1, (int)cnt);
  }
  const uint64_t s_sh = wave_sum_u32(emitted), s_drop = wave_sum_u32(dropped);
  if (lane_id() == 0) {
    if (s_sh) atomicAdd((unsigned long long*)&st.ctr[CTR_SHADOW], (unsigned long long)s_sh);
    if (s_drop) atomicAdd((unsigned long long*)&st.ctr[CTR_DROP], (unsigned long long)s_drop);
  }
}

// Any-hit tests of the shadow queue; unoccluded rays add their contribution.
template <int NL>
__global__ __launch_bounds__(256) void k_stream_shadow(Stream st, ShadowQ sq) {
  PRT_DECLARE_STACK_N(S, st.scenes[0], NL);
  const uint32_t n = min(st.ctl->cnt_sq, sq.cap);
  uint32_t* fetch = &st.ctl->fetch[1];
  const uint32_t shift = (uint32_t)st.tile_shift;
  TravStats ts{0, 0};
  while (true) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = (uint32_t)__shfl((int)base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane_id();
    uint32_t slot = 0;
    int sid = -1;
    if (i < n) {
      slot = sq.slot[i];
      sid = st.slot_info[slot >> shift].scene;
    }
    bool lit = false;
    per_scene(st.scenes, sid, [&](const Params& P) {
      lit = !occluded(P, mk(sq.ox[i], sq.oy[i], sq.oz[i]), mk(sq.dx[i], sq.dy[i], sq.dz[i]), sq.tmax[i], S, ts);
    });
    if (lit) {
      accum_add(&st.accum[3 * (size_t)slot + 0], sq.cr[i]);
      accum_add(&st.accum[3 * (size_t)slot + 1], sq.cg[i]);
      accum_add(&st.accum[3 * (size_t)slot + 2], sq.cb[i]);
    }
    live_update(st, sid >= 0 ? (int)(slot >> shift) : -1, -1);
  }
  flush_stats(st.ctr, CTR_SNODES, ts);
}

// Finished tiles: fold every pixel's samples into its scene's image, clear
// the slot's accumulators and return the slot to the free list.
__global__ __launch_bounds__(256) void k_stream_resolve(Stream st) {
  const uint32_t nd = min(st.ctl->done_count, (uint32_t)st.n_slots);
  const uint32_t shift = (uint32_t)st.tile_shift;
  for (uint32_t dn = blockIdx.x; dn < nd; dn += gridDim.x) {
    const uint32_t ts = st.done_list[dn];
    const SlotInfo si = st.slot_info[ts];
    const Params& P = st.scenes[si.scene];
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
hipError_t launch_stream_gen(const Stream& st, const RayQ& q, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_gen, dim3(blocks), dim3(256), 0, s, st, q);
  return hipGetLastError();
}
hipError_t launch_stream_extend(const Stream& st, const RayQ& q, int cur, double* hit_t, int32_t* hit_id,
                                int stack_lds, int blocks, hipStream_t s) {
  if (stack_lds <= 8)
    hipLaunchKernelGGL(k_stream_extend<8>, dim3(blocks), dim3(256), 0, s, st, q, cur, hit_t, hit_id);
  else if (stack_lds <= 16)
    hipLaunchKernelGGL(k_stream_extend<16>, dim3(blocks), dim3(256), 0, s, st, q, cur, hit_t, hit_id);
  else
    hipLaunchKernelGGL(k_stream_extend<32>, dim3(blocks), dim3(256), 0, s, st, q, cur, hit_t, hit_id);
  return hipGetLastError();
}
hipError_t launch_stream_shade(const Stream& st, const RayQ& qin, const double* hit_t, const int32_t* hit_id,
                               const RayQ& qout, int nxt, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_shade, dim3(blocks), dim3(256), 0, s, st, qin, hit_t, hit_id, qout, nxt);
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
