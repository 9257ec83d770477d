// prt_persist.hip — the persistent form of the batched renderer: one kernel
// launch renders every camera sample of every scene of a batch.
//
// Waves claim ranges of 64-sample chunks from one shared counter over the
// batch's chunk list (scene-major, prt_stream.h Persist) and keep each of
// their 64 lanes on one camera sample at a time.  Per loop trip a lane
// advances its sample's path by one segment: next-event estimation of the
// segment's origin vertex when the segment carries the NEE flag (three light
// samples, each with its any-hit shadow ray), the segment's closest hit, and
// one radiance() vertex (prt_shade.h), whose first child becomes the lane's
// next segment and whose second child (glass at depth <= 2) waits on the
// lane's pending stack.  When a path has no segment left the lane writes the
// sample's fixed-point sum to the sample's accumulator and takes the next
// sample.  Nothing goes through a ray queue, there is no launch per depth,
// and waves never wait for each other: a wave exits when the counter is past
// the last chunk and its lanes have finished.
//
// A wave works on one scene at a time, so the scene's parameters are
// wave-uniform scalar loads: a lane starts a sample of another scene only
// after every lane of the wave has finished the previous scene's paths.
//
// The arithmetic is the queued renderers' (same vertex and NEE functions,
// segment throughput / pickup weight / NEE normal rounded to float as a queue
// stores them, contributions added in 32.32 fixed point), a sample's random
// numbers depend only on (pixel, sample, depth, branch), and integer sums do
// not depend on order, so every image equals the per-scene render bit for
// bit.  k_persist_resolve then folds each pixel's samples (fold_pixel).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"
#include "prt_math.h"
#include "prt_shade.h"
#include "prt_stream.h"
#include "prt_trace.h"

// 32.32 fixed-point value that accum_add would add for v.
__device__ __forceinline__ uint64_t fix_of(double v) {
  return v > 0.0 ? (uint64_t)(v * PRT_FIX_ONE + 0.5) : 0ull;
}

// Scene owning chunk c (c wave-uniform): the last s with chunk_base[s] <= c.
__device__ __forceinline__ int32_t scene_of_chunk(const Persist& st, uint32_t c) {
  int32_t lo = 0, hi = st.n_scenes;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (st.chunk_base[mid] <= c) lo = mid;
    else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// WAVES: waves per SIMD the registers are budgeted for (2: the compiler's
// own choice, about 230 VGPRs; 3 and 4 spill some state to scratch)
template <int NL, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void k_persist(Persist st) {
  PRT_DECLARE_STACK_N(S, st.scenes[0], NL, 0);
  // the top-level node copy serves a batch of one scene only
  PRT_TOP_CACHE(S, st.scenes[0].nodes, st.n_scenes == 1 ? st.scenes[0].n_top : 0);
  const uint64_t below = (1ull << lane_id()) - 1ull;
  PendRec* pend = st.pend + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * PRT_PEND_DEPTH;
  TravStats tc{0, 0, 0}, tsh{0, 0, 0};
  uint32_t n_prim = 0, n_traced = 0, n_shadow = 0, truncated = 0, n_ext = 0, pickups = 0, dropped = 0;

  // wave-uniform work cursor: claimed chunks [cb, ce), next sample co of
  // chunk cb, which belongs to scene cs (-1: not yet looked up) and holds cn
  // samples; the wave's current scene
  uint32_t cb = 0, ce = 0, co = 0, cn = 0;
  int32_t cs = -1, scene = -1;
  bool claims_left = true;
  UniformLight U{};
  bool one_light = false;

  // the lane's sample and current segment
  bool act = false;
  d3 o = mk(0.0, 0.0, 0.0), d = o;
  float tr[3] = {1.0f, 1.0f, 1.0f}, pw[3] = {0.0f, 0.0f, 0.0f}, nr[3] = {0.0f, 0.0f, 0.0f};
  uint32_t meta = 0, k = 0;
  int64_t pixel = 0;
  uint64_t gs = 0;
  uint64_t acc[3] = {0ull, 0ull, 0ull};
  int npend = 0;

  while (true) {
    // ---- idle lanes take the next samples of the wave's scene
    uint64_t need = __ballot(!act);
    while (need) {
      if (cb >= ce) {
        if (!claims_left) break;
        uint32_t c0 = 0;
        if (lane_id() == 0) c0 = atomicAdd(st.next, st.claim);
        c0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl((int)c0, 0, 64));
        if (c0 >= st.n_chunks) {
          claims_left = false;
          break;
        }
        cb = c0;
        ce = min(c0 + st.claim, st.n_chunks);
        co = 0;
        cs = -1;
      }
      if (cs < 0) {
        cs = scene_of_chunk(st, cb);
        const uint64_t first = (uint64_t)(cb - st.chunk_base[cs]) * 64u;
        const uint64_t n_s = st.sample_base[cs + 1] - st.sample_base[cs];
        cn = (uint32_t)__builtin_amdgcn_readfirstlane((int)(n_s - first < 64u ? n_s - first : 64u));
      }
      if (cs != scene) {
        if (__any(act)) break;  // the wave finishes its scene's paths first
        scene = cs;
        const Params& P = scene_params(st.scenes, scene);
        one_light = P.nee && P.n_lights == 1;
        if (one_light) U = uniform_light(P, P.lights[0]);
      }
      const uint32_t take = min((uint32_t)__popcll(need), cn - co);
      const uint32_t rank = (uint32_t)__popcll(need & below);
      if (!act && rank < take) {
        const Params& P = scene_params(st.scenes, scene);
        const uint64_t r = (uint64_t)(cb - st.chunk_base[scene]) * 64u + co + rank;  // sample of the scene
        pixel = (int64_t)(r / (uint32_t)P.spp);
        k = (uint32_t)(r % (uint32_t)P.spp);
        const int sub = (int)(k % (uint32_t)(P.SS * P.SS));
        camera_ray(P, (int)(pixel % P.W), (int)(pixel / P.W), sub % P.SS, sub / P.SS, o, d);
        tr[0] = tr[1] = tr[2] = 1.0f;
        pw[0] = pw[1] = pw[2] = 0.0f;
        nr[0] = nr[1] = nr[2] = 0.0f;
        meta = rng_c2(0, 1u);
        gs = st.sample_base[scene] + r;
        acc[0] = acc[1] = acc[2] = 0ull;
        npend = 0;
        act = true;
        ++n_prim;
      }
      co += take;
      if (co >= cn) {
        ++cb;
        co = 0;
        cs = -1;
      }
      need = __ballot(!act);
    }
    if (!__any(act)) {
      if (!claims_left && cb >= ce) break;
      continue;
    }
    const Params& P = scene_params(st.scenes, scene);
    // ---- next-event estimation of the segment's origin vertex
    // (k_stream_nee / k_stream_shadow on this segment's queue entry)
    if (act && (meta & META_NEE)) {
      uint32_t w[9];
      nee_words(P, (uint32_t)pixel, k, (int)(meta & 0xFFu) - 1, META_BRANCH(meta), w);
      const d3 x = o, nrm = mk(nr[0], nr[1], nr[2]), ct = mk(tr[0], tr[1], tr[2]);
      for (int m = 0; m < 3; ++m) {
        ShadowRay sr;
        int32_t l = -1;
        const bool ok = one_light ? nee_one_single(U, w, m, x, nrm, ct, sr, l) : nee_one(P, w, m, x, nrm, ct, sr, l);
        if (ok) {
          ++n_shadow;
          if (!occluded(P, sr.o2, sr.dir, sr.tmax, S, tsh)) {
            acc[0] += fix_of(sr.c[0]);
            acc[1] += fix_of(sr.c[1]);
            acc[2] += fix_of(sr.c[2]);
          }
        }
      }
    }
    // ---- the segment's closest hit and its vertex
    if (act) {
      double t, u, v;
      int32_t id;
      closest_hit(P, o, d, t, id, u, v, S, tc);
      ++n_traced;
      LaneCtx C;
      C.has0 = false;
      C.pend = pend;
      C.npend = npend;
      C.truncated = C.n_ext = C.pickups = C.dropped = C.direct = 0;
      VertexAdj va;
      const d3 add = shade_vertex(P, P.mats, C, o, d, t, id, mk(tr[0], tr[1], tr[2]), mk(pw[0], pw[1], pw[2]),
                                  (uint32_t)pixel, k, (int)(meta & 0xFFu), META_BRANCH(meta), va);
      acc[0] += fix_of(add.x);
      acc[1] += fix_of(add.y);
      acc[2] += fix_of(add.z);
      npend = C.npend;
      truncated += C.truncated;
      n_ext += C.n_ext;
      pickups += C.pickups;
      dropped += C.dropped;
      if (C.has0) {
        o = C.o0;
        d = C.d0;
        for (int j = 0; j < 3; ++j) {
          tr[j] = C.t0[j];
          pw[j] = C.p0[j];
          nr[j] = C.n0[j];
        }
        meta = C.meta0;
      } else if (npend > 0) {
        const PendRec& r = pend[--npend];
        o = mk(r.o[0], r.o[1], r.o[2]);
        d = mk(r.d[0], r.d[1], r.d[2]);
        for (int j = 0; j < 3; ++j) {
          tr[j] = r.t[j];
          pw[j] = r.p[j];
          nr[j] = r.n[j];
        }
        meta = r.meta;
      } else {
        uint64_t* a = st.accum + 3 * gs;
        a[0] = acc[0];
        a[1] = acc[1];
        a[2] = acc[2];
        act = false;
      }
    }
  }
  flush_stats(st.ctr, CTR_NODES, tc);
  flush_stats(st.ctr, CTR_SNODES, tsh);
  const uint64_t s_prim = wave_sum_u32(n_prim), s_tr = wave_sum_u32(n_traced), s_sh = wave_sum_u32(n_shadow);
  const uint64_t s_ext = wave_sum_u32(n_ext), s_pick = wave_sum_u32(pickups), s_trunc = wave_sum_u32(truncated);
  const uint64_t s_drop = wave_sum_u32(dropped);
  if (lane_id() == 0) {
    if (s_prim) atomicAdd((unsigned long long*)&st.ctr[CTR_QUEUED], (unsigned long long)s_prim);
    if (s_tr) atomicAdd((unsigned long long*)&st.ctr[CTR_TRACED], (unsigned long long)s_tr);
    if (s_sh) atomicAdd((unsigned long long*)&st.ctr[CTR_SHADOW], (unsigned long long)s_sh);
    if (s_sh) atomicAdd((unsigned long long*)&st.ctr[CTR_STRACED], (unsigned long long)s_sh);
    if (s_ext) atomicAdd((unsigned long long*)&st.ctr[CTR_EXT], (unsigned long long)s_ext);
    if (s_pick) atomicAdd((unsigned long long*)&st.ctr[CTR_PICKUP], (unsigned long long)s_pick);
    if (s_trunc) atomicAdd((unsigned long long*)&st.ctr[CTR_TRUNC], (unsigned long long)s_trunc);
    if (s_drop) atomicAdd((unsigned long long*)&st.ctr[CTR_DROP], (unsigned long long)s_drop);
  }
}

// Every pixel of every scene of the batch: fold its samples' accumulators
// into the scene's image (fold_pixel: 1/spp scale, luminance clamp,
// sensor_response, variance of the mean).  pixel_base[s]: first pixel of
// scene s in the batch's pixel numbering ([n_scenes + 1]).
__global__ __launch_bounds__(256) void k_persist_resolve(Persist st, int64_t n_pix, const int64_t* pixel_base) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_pix; g += (int64_t)gridDim.x * blockDim.x) {
    int32_t lo = 0, hi = st.n_scenes;
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (pixel_base[mid] <= g) lo = mid;
      else hi = mid;
    }
    const Params& P = st.scenes[lo];
    const int64_t pix = g - pixel_base[lo];
    const int x = (int)(pix % P.W), y = (int)(pix / P.W);
    const size_t ii = (size_t)(P.H - y - 1) * P.W + x;  // render.h:1234
    double c[3], v[3];
    fold_pixel(P, st.accum + 3 * (st.sample_base[lo] + (uint64_t)pix * (uint64_t)P.spp), c, v);
    float* rgb = st.out_rgb[lo];
    float* var = st.out_var[lo];
    for (int j = 0; j < 3; ++j) {
      rgb[3 * ii + j] = (float)c[j];
      if (var) var[3 * ii + j] = (float)v[j];
    }
  }
}

// ------------------------------------------------------------ launchers
// The kernel instance for an LDS stack depth (8 or 16 entries) and a register
// budget of `waves` waves per SIMD (2, 3 or 4).
template <class F>
static hipError_t with_persist_kernel(int stack_lds, int waves, F&& f) {
  if (stack_lds <= 8) {
    if (waves <= 2) return f(k_persist<8, 2>);
    if (waves == 3) return f(k_persist<8, 3>);
    return f(k_persist<8, 4>);
  }
  if (stack_lds <= 16) {
    if (waves <= 2) return f(k_persist<16, 2>);
    if (waves == 3) return f(k_persist<16, 3>);
    return f(k_persist<16, 4>);
  }
  // 32 entries (64 KB of LDS per block) only at the 2-wave budget, where two
  // blocks per CU fit
  return f(k_persist<32, 2>);
}

hipError_t launch_persist(const Persist& p, int stack_lds, int waves, int blocks, hipStream_t s) {
  return with_persist_kernel(stack_lds, waves, [&](void (*k)(Persist)) {
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, s, p);
    return hipGetLastError();
  });
}

hipError_t persist_blocks_per_cu(int stack_lds, int waves, int* blocks) {
  return with_persist_kernel(stack_lds, waves, [&](void (*k)(Persist)) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, 256, 0);
  });
}

hipError_t launch_persist_resolve(const Persist& p, int64_t n_pixels_total, const int64_t* d_pixel_base,
                                  hipStream_t s) {
  int64_t g = (n_pixels_total + 255) / 256;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_persist_resolve, dim3((unsigned)g), dim3(256), 0, s, p, n_pixels_total, d_pixel_base);
  return hipGetLastError();
}
