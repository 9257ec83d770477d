// prt_kernels.hip — gfx950 kernels of the per-scene wavefront path tracer.
//
// Per batch of camera samples: k_gen_primary writes the unique sub-pixel
// rays, then per path depth k_extend (closest hit), k_shade (one radiance()
// vertex per ray: RR, BSDF choice, emission, child rays), k_nee (next-event
// estimation of the diffuse vertices, appending shadow rays) and k_shadow
// (any hit, adding unoccluded light to the sample).  k_resolve folds the
// samples of each pixel into the image.  The vertex and NEE arithmetic
// (radiance.h:25-556, brdf.h:408-953, directLight_radiance.h:19-556) lives in
// prt_shade.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "prt_device.h"
#include "prt_kernels.h"
#include "prt_math.h"
#include "prt_shade.h"
#include "prt_trace.h"

// Register budgets (waves per SIMD) per kernel class; 0 leaves the choice to
// the compiler.  The per-scene traversal kernels run 5 waves per SIMD (the
// compiler's own budget, ~118 VGPRs, allows 4; a 12-entry LDS stack lets 5
// blocks share a CU): cfg2 367 -> 376 M samples/s, mostly from the shadow
// rays sharing the chip with NEE better.
#ifndef PRT_WAVES_TRAVERSE
#define PRT_WAVES_TRAVERSE 5
#endif
#ifndef PRT_WAVES_SHADE
#define PRT_WAVES_SHADE 0
#endif
#ifndef PRT_WAVES_NEE
#define PRT_WAVES_NEE 0
#endif
// the per-scene shadow-ray kernel: its register budget and LDS stack depth
#ifndef PRT_WAVES_SHADOW
#define PRT_WAVES_SHADOW PRT_WAVES_TRAVERSE
#endif
#ifndef PRT_STACK_LDS_SHADOW
#define PRT_STACK_LDS_SHADOW PRT_STACK_LDS
#endif
static_assert(PRT_STACK_LDS_SHADOW >= PRT_STACK_LDS_MIN, "the overflow area is sized for the smallest LDS stack");
#define PRT_WAVES_ATTR(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1)))

// ------------------------------------------------------------ trace API kernel
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_TRAVERSE) void k_trace(
    Params P, const prt_ray_dev* __restrict__ rays, int64_t n, prt_hit_dev* __restrict__ hits, uint64_t* ctr) {
  PRT_DECLARE_STACK(S, P, 0);
  PRT_TOP_CACHE(S, P.nodes, P.n_top);
  TravStats ts{0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const prt_ray_dev r = rays[i];
    double t, u, v;
    int32_t id;
    closest_hit(P, mk(r.org[0], r.org[1], r.org[2]), mk(r.dir[0], r.dir[1], r.dir[2]), t, id, u, v, S, ts);
    prt_hit_dev h;
    if (id >= 0) {
      const float4* tr = reinterpret_cast<const float4*>(P.tris) + (size_t)id * 3;
      h.entity = __float_as_int(tr[0].w);
      h.prim = __float_as_int(tr[1].w);
      h.t = t; h.u = u; h.v = v;
    } else if (id <= -2) {
      h.entity = -2 - id;
      h.prim = -1;
      h.t = t; h.u = -PS_INF; h.v = -PS_INF;
    } else {
      h.entity = -1; h.prim = -1; h.t = PS_INF; h.u = -PS_INF; h.v = -PS_INF;
    }
    hits[i] = h;
  }
  flush_stats(ctr, CTR_NODES, ts);
}

// ------------------------------------------------------------ generate
__global__ __launch_bounds__(256) void k_gen_primary(Params P, Batch B, RayQ q) {
  const int ss2 = P.SS * P.SS;
  const int64_t n = B.n_pixels * ss2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pix = B.pixel_base + i / ss2;
    const int sub = (int)(i % ss2);
    d3 o, d;
    camera_ray(P, (int)(pix % P.W), (int)(pix / P.W), sub % P.SS, sub / P.SS, o, d);
    q.ox[i] = o.x; q.oy[i] = o.y; q.oz[i] = o.z;
    q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
  }
}

// Image element (render.h:1234: row-flipped) of batch-local camera sample slot.
__device__ __forceinline__ size_t sample_image_index(const Params& P, const Batch& B, uint32_t slot) {
  const int64_t pix = B.pixel_base + slot / P.spp;
  const int x = (int)(pix % P.W), y = (int)(pix / P.W);
  return (size_t)(P.H - y - 1) * P.W + x;
}

// ------------------------------------------------------------ extend / shadow
// Closest hit for every ray of a chunked queue.
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_TRAVERSE) void k_extend(
    Params P, RayQ q, ChunkList cl, uint32_t* fetch, double* __restrict__ hit_t, int32_t* __restrict__ hit_id,
    uint64_t* ctr) {
  PRT_DECLARE_STACK(S, P, 0);
  PRT_TOP_CACHE(S, P.nodes, P.n_top);
  const uint32_t n_chunks = *cl.n;
  TravStats ts{0, 0, 0};
  traverse_pool<false>(
      P, n_chunks, cl.desc, fetch, S, ts,
      [&](uint32_t j, d3& o, d3& d, double& tmax) {
        o = mk(q.ox[j], q.oy[j], q.oz[j]);
        d = mk(q.dx[j], q.dy[j], q.dz[j]);
        tmax = PS_INF;
      },
      [&](uint32_t j, double t, int32_t id, bool) {
        hit_t[j] = t;
        hit_id[j] = id;
      });
  flush_stats(ctr, CTR_NODES, ts, CTR_TRACED);
}

// Any hit for NEE shadow rays; an unoccluded ray adds its contribution to the
// camera sample's accumulator.
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_SHADOW) void k_shadow(
    Params P, Batch B, ShadowQ q, ChunkList cl, uint32_t* fetch, uint64_t* __restrict__ accum, Adjoint A,
    uint64_t* ctr) {
  PRT_DECLARE_STACK_N(S, P, PRT_STACK_LDS_SHADOW, 1);
  PRT_TOP_CACHE(S, P.nodes, P.n_top);
  const uint32_t n_chunks = *cl.n;
  TravStats ts{0, 0, 0};
  double ge[3] = {0.0, 0.0, 0.0};
  auto lit = [&](uint32_t i) {
    const uint32_t s = q.slot[i];
    const float c[3] = {q.cr[i], q.cg[i], q.cb[i]};
    accum_add(&accum[3 * (size_t)s + 0], c[0]);
    accum_add(&accum[3 * (size_t)s + 1], c[1]);
    accum_add(&accum[3 * (size_t)s + 2], c[2]);
    if (A.on && q.light[i] == A.light) {
      // d(contribution)/d(emission) = contribution / emission
      const double* E = P.mats[P.ents[A.light].material].emission;
      const size_t px = sample_image_index(P, B, s);
      for (int k = 0; k < 3; ++k)
        if (E[k] != 0.0) ge[k] += (double)A.dimage[3 * px + k] * A.scale * (double)c[k] / E[k];
    }
  };
  traverse_pool<true>(
      P, n_chunks, cl.desc, fetch, S, ts,
      [&](uint32_t j, d3& o, d3& d, double& tmax) {
        o = mk(q.ox[j], q.oy[j], q.oz[j]);
        d = mk(q.dx[j], q.dy[j], q.dz[j]);
        tmax = q.tmax[j];
      },
      [&](uint32_t j, double, int32_t, bool blocked) {
        if (!blocked) lit(j);
      });
  flush_stats(ctr, CTR_SNODES, ts, CTR_STRACED);
  if (A.on) {
    for (int k = 0; k < 3; ++k) {
      const double w = wave_sum_f64(ge[k]);
      if (lane_id() == 0 && w != 0.0) atomicAdd(&A.grad_e[k], w);
    }
  }
}

// ------------------------------------------------------------ shade
// Shade one queue (or, with primary = 1, every camera sample of the batch
// against the hit of its sub-pixel's primary ray).  ADJ: with the adjoint
// bookkeeping of a path replay (a separate instance, so forward renders
// carry none of its registers).
// MLDS: the scene's materials (at most PRT_LDS_MATERIALS) are copied into
// the block's LDS and read from there.
#define PRT_LDS_MATERIALS 64
template <bool ADJ, bool MLDS>
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_SHADE) void k_shade(
    Params P, Batch B, int primary, RayQ qin, ChunkList cl, const double* __restrict__ hit_t,
    const int32_t* __restrict__ hit_id, RayQ qout, uint32_t fan, uint2* slabs, uint64_t* __restrict__ accum,
    Adjoint A, uint64_t* ctr) {
  __shared__ uint32_t s_cnt[4];  // children appended by each wave of the block
  __shared__ DevMaterial s_mats[MLDS ? PRT_LDS_MATERIALS : 1];
  const uint32_t wib = threadIdx.x >> 6;
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (lane_id() == 0) s_cnt[wib] = 0;
  if (MLDS) {
    for (int k = (int)threadIdx.x; k < P.n_materials; k += (int)blockDim.x) s_mats[k] = P.mats[k];
    __syncthreads();
  }
  const DevMaterial* mats = MLDS ? s_mats : P.mats;
  const uint32_t n_samples = (uint32_t)(B.n_pixels * P.spp);
  const uint32_t n_chunks = primary ? (n_samples + 63u) >> 6 : *cl.n;
  const int ss2 = P.SS * P.SS;
  uint32_t c, c_end;
  wave_range(n_chunks, c, c_end);
  // this wave's slab: fan children per entry of its chunks
  ShadeCtx C;
  C.qout = qout;
  C.count_out = &s_cnt[wib];
  C.off = fan * 64u * c;
  C.cap = fan * 64u * (c_end - c);
  C.truncated = C.n_ext = C.pickups = C.dropped = 0;
  double ge[3] = {0.0, 0.0, 0.0};
  // chunk descriptors are read two chunks ahead and (forward renders: the
  // adjoint instance has no registers to spare) the hits (t, id) of a ray one
  // chunk ahead, so that a chunk's triangle reads, which depend on the hit,
  // go out with its ray reads instead of after them
  constexpr bool PF = !ADJ;
  uint2 dn1 = make_uint2(0u, 0u), dn2 = make_uint2(0u, 0u);
  double hn_t = 0.0;
  int32_t hn_id = -1;
  if (!primary && c < c_end) {
    dn1 = cl.desc[c];
    if (PF && lane_id() < dn1.y) {
      hn_t = hit_t[dn1.x + lane_id()];
      hn_id = hit_id[dn1.x + lane_id()];
    }
    if (PF && c + 1 < c_end) dn2 = cl.desc[c + 1];
  }
  for (; c < c_end; ++c) {
    uint32_t i;
    double hc_t = 0.0;
    int32_t hc_id = -1;
    if (primary) {
      i = (c << 6) + lane_id();
      if (i >= n_samples) continue;
    } else {
      const uint2 dc = dn1;
      if (PF) {
        hc_t = hn_t;
        hc_id = hn_id;
        dn1 = dn2;
        if (c + 1 < c_end && lane_id() < dn1.y) {
          hn_t = hit_t[dn1.x + lane_id()];
          hn_id = hit_id[dn1.x + lane_id()];
        }
        if (c + 2 < c_end) dn2 = cl.desc[c + 2];
      } else if (c + 1 < c_end) {
        dn1 = cl.desc[c + 1];
      }
      if (lane_id() >= dc.y) continue;
      i = dc.x + lane_id();
    }
    d3 o, d, T, Pw;
    double t;
    int32_t id;
    int depth;
    uint32_t branch, slot;
    if (primary) {
      slot = i;
      const uint32_t pl = i / P.spp, k = i % P.spp;
      const int sub = (int)(k % ss2);
      const int64_t pix = B.pixel_base + pl;
      const uint32_t hi = pl * ss2 + sub;
      if (P.mlt_u != nullptr) {  // a Metropolis proposal's own camera ray
        o = mk(qin.ox[hi], qin.oy[hi], qin.oz[hi]);
        d = mk(qin.dx[hi], qin.dy[hi], qin.dz[hi]);
      } else {
        camera_ray(P, (int)(pix % P.W), (int)(pix / P.W), sub % P.SS, sub / P.SS, o, d);
      }
      t = hit_t[hi];
      id = hit_id[hi];
      T = mk(1.0, 1.0, 1.0);
      Pw = mk(0.0, 0.0, 0.0);
      depth = 0;
      branch = 1;
    } else {
      slot = qin.slot[i];
      o = mk(qin.ox[i], qin.oy[i], qin.oz[i]);
      d = mk(qin.dx[i], qin.dy[i], qin.dz[i]);
      T = mk(qin.tr[i], qin.tg[i], qin.tb[i]);
      Pw = mk(qin.pr[i], qin.pg[i], qin.pb[i]);
      const uint32_t meta = qin.meta[i];
      depth = (int)(meta & 0xFFu);
      branch = META_BRANCH(meta);
      t = PF ? hc_t : hit_t[i];
      id = PF ? hc_id : hit_id[i];
    }
    C.slot = slot;
    C.direct = 0u;
    const uint32_t pl = slot / P.spp, k = slot % P.spp;
    VertexAdj va;
    const d3 add = shade_vertex(P, mats, C, o, d, t, id, T, Pw, (uint32_t)(B.pixel_base + pl), k, depth, branch, va);
    if (primary && P.mlt_flags != nullptr) P.mlt_flags[slot] = C.direct;
    if (ADJ && A.on) {
      const size_t px = sample_image_index(P, B, slot);
      const d3 g = mk(A.dimage[3 * px] * A.scale, A.dimage[3 * px + 1] * A.scale, A.dimage[3 * px + 2] * A.scale);
      if (va.albedo && va.entity == A.entity && va.face >= 0) {
        // radiance still to come from this vertex on (no glass splits: one
        // path per sample), without the vertex's own emission
        const uint64_t* lt = A.ltot + 3 * (size_t)slot;
        const uint64_t* lb = accum + 3 * (size_t)slot;
        const double rest[3] = {accum_get(lt) - accum_get(lb) - va.pre.x - va.own.x,
                                accum_get(lt + 1) - accum_get(lb + 1) - va.pre.y - va.own.y,
                                accum_get(lt + 2) - accum_get(lb + 2) - va.pre.z - va.own.z};
        const double rho[3] = {va.rho.x, va.rho.y, va.rho.z};
        const double gg[3] = {g.x, g.y, g.z};
        for (int c = 0; c < 3; ++c)
          if (rho[c] > 0.0 && rest[c] != 0.0) atomicAdd(&A.grad_alb[3 * (size_t)va.face + c], gg[c] * rest[c] / rho[c]);
      }
      if (va.entity == A.light) {
        ge[0] += g.x * va.add_e.x;
        ge[1] += g.y * va.add_e.y;
        ge[2] += g.z * va.add_e.z;
      }
    }
    accum_add(&accum[3 * (size_t)slot + 0], add.x);
    accum_add(&accum[3 * (size_t)slot + 1], add.y);
    accum_add(&accum[3 * (size_t)slot + 2], add.z);
  }
  if (ADJ && A.on) {
    for (int c = 0; c < 3; ++c) {
      const double w = wave_sum_f64(ge[c]);
      if (lane_id() == 0 && w != 0.0) atomicAdd(&A.grad_e[c], w);
    }
  }
  const uint64_t s_pick = wave_sum_u32(C.pickups), s_tr = wave_sum_u32(C.truncated);
  const uint64_t s_ext = wave_sum_u32(C.n_ext), s_drop = wave_sum_u32(C.dropped);
  if (lane_id() == 0) {
    slabs[wave] = make_uint2(C.off, min(s_cnt[wib], C.cap));
    if (s_pick) atomicAdd((unsigned long long*)&ctr[CTR_PICKUP], (unsigned long long)s_pick);
    if (s_tr) atomicAdd((unsigned long long*)&ctr[CTR_TRUNC], (unsigned long long)s_tr);
    if (s_ext) atomicAdd((unsigned long long*)&ctr[CTR_EXT], (unsigned long long)s_ext);
    if (s_drop) atomicAdd((unsigned long long*)&ctr[CTR_DROP], (unsigned long long)s_drop);
  }
}

// ------------------------------------------------------------ NEE
// NEE for the diffuse vertices that produced the rays of queue q (their
// origin, throughput T*colour/rr and orienting normal), with the parent
// vertex's random numbers.  Shadow rays go to the wave's slab of the shadow
// queue.
// MODE (NeeMode): NEE_ONE, the scene has exactly one light, whose entity
// record is read once into wave-uniform registers; NEE_LDS, each wave copies
// the scene's lights (at most NEE_LDS_LIGHTS) into its LDS slots; NEE_GLOBAL,
// light records are read from global memory per sample.  Separate instances
// keep each path's registers out of the others.
template <int MODE>
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_NEE) void k_nee(
    Params P, Batch B, RayQ q, ChunkList cl, ShadowQ sq, uint2* slabs, uint64_t* ctr) {
  __shared__ uint32_t s_cnt[4];  // shadow rays appended by each wave of the block
  __shared__ LightSlot s_lights[MODE == NEE_LDS ? 4 * NEE_LDS_LIGHTS : 1];
  const uint32_t wib = threadIdx.x >> 6;
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (lane_id() == 0) s_cnt[wib] = 0;
  uint32_t emitted = 0, dropped = 0;
  uint32_t c, c_end;
  wave_range(*cl.n, c, c_end);
  // this wave's slab: three shadow rays per entry of its chunks
  const uint32_t off = 3u * 64u * c, slab_cap = 3u * 64u * (c_end - c);
  UniformLight U{};
  LightSlot* slots = s_lights + (MODE == NEE_LDS ? wib * NEE_LDS_LIGHTS : 0);
  if (MODE == NEE_ONE && c < c_end) U = uniform_light(P, P.lights[0]);
  if (MODE == NEE_LDS) load_light_slots(P, slots);
  // the next chunk's descriptor is read one chunk ahead, and an entry's
  // fields are all read before its NEE flag is tested: one memory round
  // trip per chunk instead of three dependent ones
  uint2 dn = c < c_end ? cl.desc[c] : make_uint2(0u, 0u);
  for (; c < c_end; ++c) {
    const uint2 dc = dn;
    if (c + 1 < c_end) dn = cl.desc[c + 1];
    bool have = false;
    uint32_t slot = 0, w[9];
    d3 x, nrm, ct;
    if (lane_id() < dc.y) {
      const uint32_t i = dc.x + lane_id();
      const uint32_t meta = q.meta[i];
      slot = q.slot[i];
      x = mk(q.ox[i], q.oy[i], q.oz[i]);
      nrm = mk(q.nx[i], q.ny[i], q.nz[i]);
      ct = mk(q.tr[i], q.tg[i], q.tb[i]);
      if (meta & META_NEE) {
        have = true;
        const uint32_t pl = slot / P.spp, k = slot % P.spp;
        nee_words(P, (uint32_t)(B.pixel_base + pl), k, (int)(meta & 0xFFu) - 1, META_BRANCH(meta), w);
      }
    }
    // one light sample at a time, appended through the LDS count: only one
    // shadow ray's registers are live
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      ShadowRay sr;
      int32_t l = -1;
      const bool v = have && (MODE == NEE_ONE   ? nee_one_single(U, w, m, x, nrm, ct, sr, l)
                              : MODE == NEE_LDS ? nee_one_slots(P, slots, w, m, x, nrm, ct, sr, l)
                                                : nee_one(P, w, m, x, nrm, ct, sr, l));
      const uint32_t j = wave_append_bit(&s_cnt[wib], v);
      if (v) {
        if (j < slab_cap) {
          put_shadow(sq, off + j, sr, slot, l);
          ++emitted;
        } else {
          ++dropped;
        }
      }
    }
  }
  const uint64_t s_sh = wave_sum_u32(emitted), s_drop = wave_sum_u32(dropped);
  if (lane_id() == 0) {
    slabs[wave] = make_uint2(off, min(s_cnt[wib], slab_cap));
    if (s_sh) atomicAdd((unsigned long long*)&ctr[CTR_SHADOW], (unsigned long long)s_sh);
    if (s_drop) atomicAdd((unsigned long long*)&ctr[CTR_DROP], (unsigned long long)s_drop);
  }
}

// ------------------------------------------------------------ slab index
// Chunk descriptors of a slab queue: slab s holds slabs[s].y entries from
// queue index slabs[s].x on, ceil(that / 64) chunks, numbered slab by slab.
// Block g lists the chunks of the PRT_INDEX_SLABS slabs from
// g * PRT_INDEX_SLABS on: it sums the chunks of all slabs before them (every
// block reads those slab counts, 8 bytes each, mostly from L2), scans its own
// slabs' chunk counts in one wave, and writes their descriptors, a wave per
// slab and a lane per chunk.  Block 0, whose range starts at slab 0, sums
// every slab instead and stores the number of chunks and of entries (and adds
// the entries to ctr[ctr_entries] unless that is < 0).  Chunks beyond
// max_chunks (never, with the host's sizing) are dropped and counted.
__global__ __launch_bounds__(256) void k_slab_index(const uint2* __restrict__ slabs, uint32_t n_slabs,
                                                    uint2* __restrict__ desc, uint32_t max_chunks,
                                                    uint32_t* n_chunks, uint32_t* n_entries, uint64_t* ctr,
                                                    int ctr_entries) {
  __shared__ uint32_t s_part[2][4];
  __shared__ uint32_t s_first[PRT_INDEX_SLABS];
  const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
  const uint32_t s0 = blockIdx.x * (uint32_t)PRT_INDEX_SLABS;  // < n_slabs by the grid
  const uint32_t ns = min(n_slabs - s0, (uint32_t)PRT_INDEX_SLABS);
  const bool totals = blockIdx.x == 0;
  const uint32_t lim = totals ? n_slabs : s0;
  uint32_t ch = 0, ent = 0;
  for (uint32_t s = tid; s < lim; s += 256u) {
    const uint32_t c = slabs[s].y;
    ch += (c + 63u) >> 6;
    ent += c;
  }
  ch = (uint32_t)wave_sum_u32(ch);
  ent = (uint32_t)wave_sum_u32(ent);
  if (lane == 0) {
    s_part[0][wv] = ch;
    s_part[1][wv] = ent;
  }
  if (wv == 0) {
    // exclusive scan of this block's slabs' chunk counts
    const uint32_t c = lane < ns ? slabs[s0 + lane].y : 0u;
    const uint32_t n = (c + 63u) >> 6;
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
  const uint32_t before = totals ? 0u : sum_ch;
  for (uint32_t i = wv; i < ns; i += 4u) {
    const uint2 sl = slabs[s0 + i];
    const uint32_t nch = (sl.y + 63u) >> 6, base = before + s_first[i];
    for (uint32_t j = lane; j < nch; j += 64u) {
      const uint32_t c = base + j;
      if (c < max_chunks) desc[c] = make_uint2(sl.x + (j << 6), min(64u, sl.y - (j << 6)));
    }
  }
  if (totals && tid == 0) {
    const uint32_t tot_ent = s_part[1][0] + s_part[1][1] + s_part[1][2] + s_part[1][3];
    *n_chunks = min(sum_ch, max_chunks);
    *n_entries = tot_ent;
    if (ctr_entries >= 0 && tot_ent) atomicAdd((unsigned long long*)&ctr[ctr_entries], (unsigned long long)tot_ent);
    if (sum_ch > max_chunks)
      atomicAdd((unsigned long long*)&ctr[CTR_DROP], (unsigned long long)(sum_ch - max_chunks) * 64ull);
  }
}

// ------------------------------------------------------------ resolve
// Per pixel of the batch: fold its samples into the image (and the variance
// of the pixel mean).  With f64 set, out_rgb / out_var point to doubles.
__global__ __launch_bounds__(256) void k_resolve(Params P, Batch B, const uint64_t* __restrict__ accum,
                                                 void* __restrict__ out_rgb, void* __restrict__ out_var, int f64) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n_pixels) return;
  const int64_t pix = B.pixel_base + i;
  const int x = (int)(pix % P.W), y = (int)(pix / P.W);
  const size_t image_index = (size_t)(P.H - y - 1) * P.W + x;
  double rgb[3], var[3];
  fold_pixel(P, accum + 3 * (size_t)i * P.spp, rgb, var);
  for (int c = 0; c < 3; ++c) {
    if (f64) {
      static_cast<double*>(out_rgb)[image_index * 3 + c] = rgb[c];
      if (out_var) static_cast<double*>(out_var)[image_index * 3 + c] = var[c];
    } else {
      static_cast<float*>(out_rgb)[image_index * 3 + c] = (float)rgb[c];
      if (out_var) static_cast<float*>(out_var)[image_index * 3 + c] = (float)var[c];
    }
  }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
// Per-face albedo of one POLYGON entity -> the leaf-ordered triangle records.
__global__ void k_scatter_albedo(float* tri_alb, const uint32_t* slot_of_tri, int64_t first, int64_t n_faces,
                                 const float* albedo) {
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n_faces; f += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t s = slot_of_tri[first + f];
    tri_alb[3 * (size_t)s + 0] = albedo[3 * f + 0];
    tri_alb[3 * (size_t)s + 1] = albedo[3 * f + 1];
    tri_alb[3 * (size_t)s + 2] = albedo[3 * f + 2];
  }
}
__global__ void k_grad_to_f32(const double* g, float* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)g[i];
}
// The replay must rebuild the primal sample radiance bit for bit.
__global__ void k_count_mismatch(const uint64_t* a, const uint64_t* b, int64_t n, uint64_t* ctr) {
  uint32_t bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bad += a[i] != b[i];
  const uint64_t w = wave_sum_u32(bad);
  if (lane_id() == 0 && w) atomicAdd((unsigned long long*)&ctr[CTR_REPLAY_MISMATCH], (unsigned long long)w);
}

// ------------------------------------------------------------ launchers
static int grid_for(int64_t n, int per_block = 256, int max_blocks = 1 << 16) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

hipError_t launch_trace(const Params& P, const prt_ray_dev* rays, int64_t n, prt_hit_dev* hits, uint64_t* ctr,
                        int max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_trace, dim3(grid_for(n, 256, max_blocks)), dim3(256), 0, s, P, rays, n, hits, ctr);
  return hipGetLastError();
}

// Counters of prt_trace_closest_device_noalloc: a fixed device array instead
// of a per-scene allocation, so that the call allocates nothing.  Each GPU
// holds its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch), for
// kernel arguments and copies alike.
static __device__ __align__(256) uint64_t scratch[PRT_NOALLOC_SCRATCH_WORDS];

hipError_t noalloc_scratch(int device, uint64_t** out) {
  static std::mutex mu;
  static uint64_t* addr[PRT_NOALLOC_MAX_DEVICES] = {};
  if (device < 0 || device >= PRT_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (addr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    addr[device] = static_cast<uint64_t*>(p);
  }
  *out = addr[device];
  return hipSuccess;
}

hipError_t launch_gen_primary(const Params& P, const Batch& B, const RayQ& q, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_primary, dim3(grid_for(B.n_pixels * P.SS * P.SS)), dim3(256), 0, s, P, B, q);
  return hipGetLastError();
}
hipError_t launch_extend(const Params& P, const RayQ& q, const ChunkList& cl, int64_t n_host, uint32_t* fetch,
                         double* hit_t, int32_t* hit_id, uint64_t* ctr, int max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_extend, dim3(grid_for(n_host, 256, max_blocks)), dim3(256), 0, s, P, q, cl, fetch, hit_t,
                     hit_id, ctr);
  return hipGetLastError();
}
hipError_t launch_shadow(const Params& P, const Batch& B, const ShadowQ& q, const ChunkList& cl, uint32_t* fetch,
                         uint64_t* accum, const Adjoint& A, uint64_t* ctr, int max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow, dim3(max_blocks), dim3(256), 0, s, P, B, q, cl, fetch, accum, A, ctr);
  return hipGetLastError();
}
hipError_t launch_shade(const Params& P, const Batch& B, int primary, const RayQ& qin, const ChunkList& cl,
                        const double* hit_t, const int32_t* hit_id, const RayQ& qout, uint32_t fan, uint2* slabs,
                        uint64_t* accum, const Adjoint& A, uint64_t* ctr, hipStream_t s) {
  const dim3 grid(PRT_SLAB_WAVES / 4);
  const bool mlds = P.n_materials <= PRT_LDS_MATERIALS;
  if (A.on)
    hipLaunchKernelGGL((k_shade<true, false>), grid, dim3(256), 0, s, P, B, primary, qin, cl, hit_t, hit_id, qout, fan,
                       slabs, accum, A, ctr);
  else if (mlds)
    hipLaunchKernelGGL((k_shade<false, true>), grid, dim3(256), 0, s, P, B, primary, qin, cl, hit_t, hit_id, qout, fan,
                       slabs, accum, A, ctr);
  else
    hipLaunchKernelGGL((k_shade<false, false>), grid, dim3(256), 0, s, P, B, primary, qin, cl, hit_t, hit_id, qout, fan,
                       slabs, accum, A, ctr);
  return hipGetLastError();
}
hipError_t launch_nee(const Params& P, const Batch& B, const RayQ& q, const ChunkList& cl, const ShadowQ& sq,
                      uint2* slabs, uint64_t* ctr, hipStream_t s) {
  if (P.n_lights == 1)
    hipLaunchKernelGGL(k_nee<NEE_ONE>, dim3(PRT_SLAB_WAVES / 4), dim3(256), 0, s, P, B, q, cl, sq, slabs, ctr);
  else if (P.n_lights <= NEE_LDS_LIGHTS)
    hipLaunchKernelGGL(k_nee<NEE_LDS>, dim3(PRT_SLAB_WAVES / 4), dim3(256), 0, s, P, B, q, cl, sq, slabs, ctr);
  else
    hipLaunchKernelGGL(k_nee<NEE_GLOBAL>, dim3(PRT_SLAB_WAVES / 4), dim3(256), 0, s, P, B, q, cl, sq, slabs, ctr);
  return hipGetLastError();
}
hipError_t launch_slab_index(const uint2* slabs, uint32_t n_slabs, uint2* desc, uint32_t max_chunks, uint32_t* n_chunks,
                             uint32_t* n_entries, uint64_t* ctr, int ctr_entries, hipStream_t s) {
  if (n_slabs == 0 || n_slabs > PRT_SLAB_INDEX_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_slab_index, dim3((n_slabs + PRT_INDEX_SLABS - 1) / PRT_INDEX_SLABS), dim3(256), 0, s, slabs,
                     n_slabs, desc, max_chunks, n_chunks, n_entries, ctr, ctr_entries);
  return hipGetLastError();
}
hipError_t launch_resolve(const Params& P, const Batch& B, const uint64_t* accum, void* rgb, void* var, int f64,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_resolve, dim3(grid_for(B.n_pixels)), dim3(256), 0, s, P, B, accum, rgb, var, f64);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_f32(float* p, float v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t launch_scatter_albedo(float* tri_alb, const uint32_t* slot_of_tri, int64_t first, int64_t n_faces,
                                 const float* albedo, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter_albedo, dim3(grid_for(n_faces, 256, 4096)), dim3(256), 0, s, tri_alb, slot_of_tri,
                     first, n_faces, albedo);
  return hipGetLastError();
}
hipError_t launch_grad_to_f32(const double* g, float* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_to_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, g, out, n);
  return hipGetLastError();
}
hipError_t launch_count_mismatch(const uint64_t* a, const uint64_t* b, int64_t n, uint64_t* ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, a, b, n, ctr);
  return hipGetLastError();
}
