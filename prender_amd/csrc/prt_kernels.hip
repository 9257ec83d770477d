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

#include "prt_device.h"
#include "prt_kernels.h"
#include "prt_math.h"
#include "prt_shade.h"
#include "prt_trace.h"

// Register budgets (waves per SIMD) per kernel class; 0 leaves the choice to
// the compiler.
#ifndef PRT_WAVES_TRAVERSE
#define PRT_WAVES_TRAVERSE 0
#endif
#ifndef PRT_WAVES_SHADE
#define PRT_WAVES_SHADE 0
#endif
#ifndef PRT_WAVES_NEE
#define PRT_WAVES_NEE 0
#endif
#define PRT_WAVES_ATTR(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1)))

// ------------------------------------------------------------ trace API kernel
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_TRAVERSE) void k_trace(
    Params P, const prt_ray_dev* __restrict__ rays, int64_t n, prt_hit_dev* __restrict__ hits, uint64_t* ctr) {
  PRT_DECLARE_STACK(S, P);
  TravStats ts{0, 0};
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
// Closest hit for every ray of a queue; each wave fetches 64 rays at a time.
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_TRAVERSE) void k_extend(
    Params P, RayQ q, const uint32_t* __restrict__ count_ptr, uint32_t* fetch, double* __restrict__ hit_t,
    int32_t* __restrict__ hit_id, uint64_t* ctr) {
  PRT_DECLARE_STACK(S, P);
  const uint32_t n = *count_ptr;
  TravStats ts{0, 0};
  while (true) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = (uint32_t)__shfl((int)base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane_id();
    if (i < n) {
      double t, u, v;
      int32_t id;
      closest_hit(P, mk(q.ox[i], q.oy[i], q.oz[i]), mk(q.dx[i], q.dy[i], q.dz[i]), t, id, u, v, S, ts);
      hit_t[i] = t;
      hit_id[i] = id;
    }
  }
  flush_stats(ctr, CTR_NODES, ts);
}

// Any hit for NEE shadow rays; an unoccluded ray adds its contribution to the
// camera sample's accumulator.
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_TRAVERSE) void k_shadow(
    Params P, Batch B, ShadowQ q, const uint32_t* __restrict__ count_ptr, uint32_t* fetch,
    uint64_t* __restrict__ accum, Adjoint A, uint64_t* ctr) {
  PRT_DECLARE_STACK(S, P);
  const uint32_t n = *count_ptr;
  TravStats ts{0, 0};
  double ge[3] = {0.0, 0.0, 0.0};
  while (true) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = (uint32_t)__shfl((int)base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane_id();
    if (i < n) {
      if (!occluded(P, mk(q.ox[i], q.oy[i], q.oz[i]), mk(q.dx[i], q.dy[i], q.dz[i]), q.tmax[i], S, ts)) {
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
      }
    }
  }
  flush_stats(ctr, CTR_SNODES, ts);
  if (A.on) {
    for (int k = 0; k < 3; ++k) {
      const double w = wave_sum_f64(ge[k]);
      if (lane_id() == 0 && w != 0.0) atomicAdd(&A.grad_e[k], w);
    }
  }
}

// ------------------------------------------------------------ shade
// Shade one queue (or, with primary = 1, every camera sample of the batch
// against the hit of its sub-pixel's primary ray).
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_SHADE) void k_shade(
    Params P, Batch B, int primary, RayQ qin, const uint32_t* __restrict__ count_in, const double* __restrict__ hit_t,
    const int32_t* __restrict__ hit_id, RayQ qout, uint32_t* count_out, uint64_t* __restrict__ accum, Adjoint A,
    uint64_t* ctr) {
  const uint32_t n = primary ? (uint32_t)(B.n_pixels * P.spp) : *count_in;
  const uint32_t stride = gridDim.x * blockDim.x;
  const int ss2 = P.SS * P.SS;
  ShadeCtx C;
  C.qout = qout;
  C.count_out = count_out;
  C.off = 0;
  C.cap = qout.cap;
  C.truncated = C.n_ext = C.pickups = C.dropped = 0;
  double ge[3] = {0.0, 0.0, 0.0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
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
      camera_ray(P, (int)(pix % P.W), (int)(pix / P.W), sub % P.SS, sub / P.SS, o, d);
      const uint32_t hi = pl * ss2 + sub;
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
      t = hit_t[i];
      id = hit_id[i];
    }
    C.slot = slot;
    const uint32_t pl = slot / P.spp, k = slot % P.spp;
    VertexAdj va;
    const d3 add = shade_vertex(P, C, o, d, t, id, T, Pw, (uint32_t)(B.pixel_base + pl), k, depth, branch, va);
    if (A.on) {
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
  if (A.on) {
    for (int c = 0; c < 3; ++c) {
      const double w = wave_sum_f64(ge[c]);
      if (lane_id() == 0 && w != 0.0) atomicAdd(&A.grad_e[c], w);
    }
  }
  const uint64_t s_pick = wave_sum_u32(C.pickups), s_tr = wave_sum_u32(C.truncated);
  const uint64_t s_ext = wave_sum_u32(C.n_ext), s_drop = wave_sum_u32(C.dropped);
  if (lane_id() == 0) {
    if (s_pick) atomicAdd((unsigned long long*)&ctr[CTR_PICKUP], (unsigned long long)s_pick);
    if (s_tr) atomicAdd((unsigned long long*)&ctr[CTR_TRUNC], (unsigned long long)s_tr);
    if (s_ext) atomicAdd((unsigned long long*)&ctr[CTR_EXT], (unsigned long long)s_ext);
    if (s_drop) atomicAdd((unsigned long long*)&ctr[CTR_DROP], (unsigned long long)s_drop);
  }
}

// ------------------------------------------------------------ NEE
// NEE for the diffuse vertices that produced the rays of queue q (their
// origin, throughput T*colour/rr and orienting normal), with the parent
// vertex's random numbers.  Each wave appends its shadow rays with one prefix
// sum and one atomic.
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_NEE) void k_nee(
    Params P, Batch B, RayQ q, const uint32_t* __restrict__ count_ptr, ShadowQ sq, uint32_t* count_sq, uint64_t* ctr) {
  const uint32_t n = *count_ptr;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t n_round = (n + 63u) & ~63u;
  uint32_t emitted = 0, dropped = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
    ShadowRay s0, s1, s2;
    bool v0 = false, v1 = false, v2 = false;
    int32_t l0 = -1, l1 = -1, l2 = -1;
    uint32_t slot = 0;
    if (i < n) {
      const uint32_t meta = q.meta[i];
      if (meta & META_NEE) {
        slot = q.slot[i];
        const uint32_t pl = slot / P.spp, k = slot % P.spp;
        nee_vertex(P, (uint32_t)(B.pixel_base + pl), k, (int)(meta & 0xFFu) - 1, META_BRANCH(meta),
                   mk(q.ox[i], q.oy[i], q.oz[i]), mk(q.nx[i], q.ny[i], q.nz[i]), mk(q.tr[i], q.tg[i], q.tb[i]), s0,
                   s1, s2, v0, v1, v2, l0, l1, l2);
      }
    }
    const uint32_t cnt = (uint32_t)v0 + (uint32_t)v1 + (uint32_t)v2;
    uint32_t j = wave_append_n(count_sq, cnt);
    if (j + cnt > sq.cap) {
      dropped += cnt;
      continue;
    }
    if (v0) put_shadow(sq, j++, s0, slot, l0);
    if (v1) put_shadow(sq, j++, s1, slot, l1);
    if (v2) put_shadow(sq, j++, s2, slot, l2);
    emitted += cnt;
  }
  const uint64_t s_sh = wave_sum_u32(emitted), s_drop = wave_sum_u32(dropped);
  if (lane_id() == 0) {
    if (s_sh) atomicAdd((unsigned long long*)&ctr[CTR_SHADOW], (unsigned long long)s_sh);
    if (s_drop) atomicAdd((unsigned long long*)&ctr[CTR_DROP], (unsigned long long)s_drop);
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
hipError_t launch_gen_primary(const Params& P, const Batch& B, const RayQ& q, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_primary, dim3(grid_for(B.n_pixels * P.SS * P.SS)), dim3(256), 0, s, P, B, q);
  return hipGetLastError();
}
hipError_t launch_extend(const Params& P, const RayQ& q, const uint32_t* count, int64_t n_host, uint32_t* fetch,
                         double* hit_t, int32_t* hit_id, uint64_t* ctr, int max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_extend, dim3(grid_for(n_host, 256, max_blocks)), dim3(256), 0, s, P, q, count, fetch, hit_t,
                     hit_id, ctr);
  return hipGetLastError();
}
hipError_t launch_shadow(const Params& P, const Batch& B, const ShadowQ& q, const uint32_t* count, int64_t n_max,
                         uint32_t* fetch, uint64_t* accum, const Adjoint& A, uint64_t* ctr, int max_blocks,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_shadow, dim3(grid_for(n_max, 256, max_blocks)), dim3(256), 0, s, P, B, q, count, fetch, accum,
                     A, ctr);
  return hipGetLastError();
}
hipError_t launch_shade(const Params& P, const Batch& B, int primary, const RayQ& qin, const uint32_t* count_in,
                        int64_t n_host, const double* hit_t, const int32_t* hit_id, const RayQ& qout,
                        uint32_t* count_out, uint64_t* accum, const Adjoint& A, uint64_t* ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_shade, dim3(grid_for(n_host)), dim3(256), 0, s, P, B, primary, qin, count_in, hit_t, hit_id,
                     qout, count_out, accum, A, ctr);
  return hipGetLastError();
}
hipError_t launch_nee(const Params& P, const Batch& B, const RayQ& q, const uint32_t* count, int64_t n_max,
                      const ShadowQ& sq, uint32_t* count_sq, uint64_t* ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_nee, dim3(grid_for(n_max, 256, 8192)), dim3(256), 0, s, P, B, q, count, sq, count_sq, ctr);
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
