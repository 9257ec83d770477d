// prt_kernels.hip — gfx950 kernels of the wavefront path tracer.
//
// Per batch of camera samples: k_gen_primary writes the unique sub-pixel
// rays, then per path depth k_extend (closest hit), k_shade (one radiance()
// vertex per ray: RR, BSDF choice, emission, child rays), k_nee (next-event
// estimation of the diffuse vertices, appending shadow rays) and k_shadow
// (any hit, adding unoccluded light to the sample).  k_resolve folds the
// samples of each pixel into the image.
//
// The light transport follows radiance() (radiance.h:25-556) and the BRDFs
// (brdf.h:408-953) turned from recursion into forward throughput: a vertex
// returns Env + E + direct + W * L(next), so each path segment carries T, adds
// T*(Env + E) and the NEE term, and its child carries T*W.  The "pickup" ray
// of mirror/glass vertices (direct_light_perfect_refrection,
// directLight_radiance.h:516-556) is the same ray as the child, so the child's
// hit supplies it: the child records a pickup weight and adds weight * emission
// when it lands on a light.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"
#include "prt_math.h"
#include "prt_trace.h"

// Register budgets (waves per SIMD) per kernel class; 0 leaves the choice to
// the compiler.  Chosen by measurement (the tuning sweep builds variants
// with -DPRT_WAVES_*).
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
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_TRAVERSE) void k_trace(Params P, const prt_ray_dev* __restrict__ rays, int64_t n,
                                               prt_hit_dev* __restrict__ hits, uint64_t* ctr) {
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
// Camera ray of sub-pixel (sx, sy) of pixel (x, y) (render.h:1248-1269:
// r = (s + 0.5)/SS, no jitter).
__device__ __forceinline__ void camera_ray(const Params& P, int x, int y, int sx, int sy, d3& o, d3& d) {
  const double rate = 1.0 / P.SS;
  const double r1 = (sx + 0.5 + 0.0) * rate;
  const double r2 = (sy + 0.5 + 0.0) * rate;
  const d3 sp = ld3(P.center) + ld3(P.screen_x) * ((r1 + x) * P.iw - 0.5) + ld3(P.screen_y) * ((r2 + y) * P.ih - 0.5);
  o = ld3(P.cam_pos);
  d = normalize(sp - o);
}

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
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_TRAVERSE) void k_extend(Params P, RayQ q, const uint32_t* __restrict__ count_ptr,
                                                uint32_t* fetch, double* __restrict__ hit_t,
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
struct ShadeCtx {
  RayQ qout;
  uint32_t* count_out;
  uint32_t slot;
  uint32_t truncated, n_ext, pickups, dropped;
};

// What the adjoint needs to know about one shaded vertex.
struct VertexAdj {
  int32_t entity;  // hit entity (-1: miss)
  int32_t face;    // face of a POLYGON hit
  bool albedo;     // the continuation and NEE of this vertex scale with its colour
  d3 rho;          // that colour
  d3 own;          // radiance added at this vertex that does not scale with it (T*(Env+E))
  d3 pre;          // pickup added for the parent vertex (belongs to the radiance before this vertex)
  d3 add_e;        // radiance added here that is linear in the hit entity's emission, / emission
};

__device__ __forceinline__ void emit_child(ShadeCtx& C, d3 o, d3 d, d3 ct, d3 pw, uint32_t meta, d3 nrm) {
  const uint32_t j = agg_append(C.count_out);
  if (j >= C.qout.cap) {
    C.dropped++;
    return;
  }
  C.qout.ox[j] = o.x; C.qout.oy[j] = o.y; C.qout.oz[j] = o.z;
  C.qout.dx[j] = d.x; C.qout.dy[j] = d.y; C.qout.dz[j] = d.z;
  C.qout.tr[j] = (float)ct.x; C.qout.tg[j] = (float)ct.y; C.qout.tb[j] = (float)ct.z;
  C.qout.pr[j] = (float)pw.x; C.qout.pg[j] = (float)pw.y; C.qout.pb[j] = (float)pw.z;
  C.qout.slot[j] = C.slot;
  C.qout.meta[j] = meta;
  if (meta & META_NEE) {
    C.qout.nx[j] = (float)nrm.x; C.qout.ny[j] = (float)nrm.y; C.qout.nz[j] = (float)nrm.z;
  }
  C.n_ext++;
}

// One path vertex: the radiance() body for the hit (or miss) of ray (o, d).
// Returns the radiance added to the camera sample; child rays go to the
// output queue, and a diffuse vertex marks its child for k_nee.
__device__ __forceinline__ d3 shade_vertex(const Params& P, ShadeCtx& C, d3 o, d3 d, double t, int32_t id, d3 T,
                                           d3 Pw, uint32_t pixel, uint32_t sample, int depth, uint32_t branch,
                                           VertexAdj& va) {
  d3 add = mk(0.0, 0.0, 0.0);
  const d3 zero = mk(0.0, 0.0, 0.0);
  va.entity = -1;
  va.face = -1;
  va.albedo = false;
  va.rho = va.own = va.pre = va.add_e = zero;
  if (depth >= P.max_vertices) {  // safety cap (the reference has no hard cap)
    C.truncated++;
    return add;
  }
  if (id == -1)  // miss: BackgroundColor (radiance.h:42)
    return vmul(T, ld3(P.background));
  const d3 o2 = o + d * 0.000001;
  d3 x, n;
  int ei;
  if (id >= 0) {
    const float4* tr = reinterpret_cast<const float4*>(P.tris) + (size_t)id * 3;
    ei = __float_as_int(tr[0].w);
    x = o2 + d * t;  // polygon.h:170
    n = mk(P.tri_nrm[3 * (size_t)id], P.tri_nrm[3 * (size_t)id + 1], P.tri_nrm[3 * (size_t)id + 2]);
  } else {
    ei = -2 - id;
    const DevEntity& e = P.ents[ei];
    if (e.type == ENT_SPHERE) {
      x = o2 + t * d;  // sphere.h:89-95
      n = normalize(x - ld3(e.p0));
    } else {
      x = d * t + o2;
      n = ld3(e.normal);
    }
  }
  const DevEntity& e = P.ents[ei];
  va.entity = ei;
  if (id >= 0) va.face = __float_as_int(reinterpret_cast<const float4*>(P.tris)[(size_t)id * 3 + 1].w);
  if (e.ninv < 0) n = n * -1.0;
  const DevMaterial& m = P.mats[e.material];
  d3 color;
  if (e.has_albedo && id >= 0)
    color = mk(P.tri_alb[3 * (size_t)id], P.tri_alb[3 * (size_t)id + 1], P.tri_alb[3 * (size_t)id + 2]);
  else
    color = ld3(m.color);
  const d3 E = ld3(m.emission);
  // pickup of the parent's mirror / glass vertex (directLight_radiance.h:516-556)
  if (e.light && (Pw.x != 0.0 || Pw.y != 0.0 || Pw.z != 0.0)) {
    add = add + vmul(Pw, E);
    va.pre = vmul(Pw, E);
    va.add_e = Pw;
    C.pickups++;
  }
  const d3 n_orient = dot(n, d) < 0.0 ? n : (-1.0 * n);
  // RGB_MAX = std::max(x, std::max(y, z)) (def.h:19); RR (radiance.h:77-126)
  d3 color_rr = color;
  if (P.tri_alb_rr && e.has_albedo && id >= 0)
    color_rr = mk(P.tri_alb_rr[3 * (size_t)id], P.tri_alb_rr[3 * (size_t)id + 1], P.tri_alb_rr[3 * (size_t)id + 2]);
  const d3 rw = E + color_rr;
  const double myz = (rw.y < rw.z) ? rw.z : rw.y;
  double rr = (rw.x < myz) ? myz : rw.x;
  if (depth > P.depth_limit) rr *= ldexp(1.0, -(depth - P.depth_limit));  // pow(0.5, k), exact
  uint32_t b0[4];
  const uint32_t c2 = rng_c2(depth, branch);
  philox(pixel, sample, c2, 0u, P.key0, P.key1, b0);
  if (depth > P.depth) {
    if (u01(b0[0]) >= rr) {  // terminate with the emission
      va.own = vmul(T, E);
      va.add_e = va.add_e + T;
      return add + vmul(T, E);
    }
  } else {
    rr = 1.0;
  }
  int kind;  // 0 diffuse, 1 specular, 2 refraction (radiance.h:514-551)
  if (m.type == REFL_DIFFUSE) kind = (u01(b0[1]) < m.roughness) ? 0 : 1;
  else if (m.type == REFL_SPECULAR) kind = (u01(b0[1]) < 1.0 - m.roughness) ? 1 : 0;
  else kind = 2;
  const uint32_t meta_same = rng_c2(depth + 1, branch);
  const double pk = P.nee ? 1.0 : 0.0;
  if (kind == 0) {
    // BRDF_Diffuse, brdf.h:408-700
    if (e.light && e.type == ENT_UVPLANE && dot(n, d) >= 0.0) return add;
    if (e.light) {
      if (!P.nee || depth == 0) {
        add = add + vmul(T, E / rr / (1.0 - 0.0));
        va.own = vmul(T, E / rr);
        va.add_e = va.add_e + T / rr;
      }
      return add;
    }
    add = add + vmul(T, ld3(P.env_light) + E);
    va.own = vmul(T, ld3(P.env_light) + E);
    va.add_e = va.add_e + T;
    va.albedo = true;
    va.rho = color;
    // OrthonormalBasis (vector3d.h:244-253), cosine sampling (brdf.h:498-505)
    const d3 w = n_orient;
    const d3 uu = (fabs(w.x) > PS_EPS) ? normalize(cross(mk(0.0, 1.0, 0.0), w)) : normalize(cross(mk(1.0, 0.0, 0.0), w));
    const d3 vv = normalize(cross(w, uu));
    const double r1 = PS_TWOPI * u01(b0[2]);
    const double r2 = u01(b0[3]), r2s = sqrt(r2);
    double s1, c1;
    sincos(r1, &s1, &c1);
    const d3 dir = normalize((uu * c1 * r2s + vv * s1 * r2s + w * sqrt(1.0 - r2)));
    const d3 ct = vmul(T, color / rr / (1.0 - 0.0));
    const bool nee = P.nee && P.n_lights > 0;
    emit_child(C, x, dir, ct, zero, meta_same | (nee ? META_NEE : 0u), n_orient);
    return add;
  }
  add = add + vmul(T, ld3(P.env_light) + E);
  va.own = vmul(T, ld3(P.env_light) + E);
  va.add_e = va.add_e + T;
  va.albedo = true;
  va.rho = color;
  if (kind == 1) {
    // BRDF_Specular, brdf.h:702-761
    const d3 dir = normalize(d - n * 2.0 * dot(n, d));
    const d3 ct = vmul(T, color / rr / (1.0 - 0.0));
    emit_child(C, x, dir, ct, ct * pk, meta_same, zero);
    return add;
  }
  // BRDF_Refraction, brdf.h:763-953
  const d3 refl = normalize(d - n * 2.0 * dot(n, d));
  const bool into = dot(n, n_orient) > 0.0;
  const double nc = 1.0, nt = m.ior;
  const double nnt = into ? nc / nt : nt / nc;
  const double ddn = dot(d, n_orient);
  const double cos2t = 1.0 - nnt * nnt * (1.0 - ddn * ddn);
  if (cos2t < 0.0) {  // total internal reflection
    const d3 ct = vmul(T, color / rr / (1.0 - 0.0));
    emit_child(C, x, refl, ct, ct * pk, meta_same, zero);
    return add;
  }
  const d3 tdir = normalize(d * nnt - n * (into ? 1.0 : -1.0) * (ddn * nnt + sqrt(fabs(cos2t))));
  const double a = nt - nc, b = nt + nc;
  const double R0 = (a * a) / (b * b);
  const double c = 1.0 - (into ? -ddn : dot(tdir, -1.0 * n_orient));
  const double Re = R0 + (1.0 - R0) * pow(c, 5.0);
  const double q = into ? nc / nt : nt / nc;
  const double nnt2 = q * q;  // pow(q, 2.0), exact
  const double Tr = (1.0 - Re) * nnt2;
  const double probability = 0.25 + 0.5 * Re;
  if (depth > 2) {
    uint32_t b3[4];
    philox(pixel, sample, c2, 3u, P.key0, P.key1, b3);
    const bool take_refl = u01(b3[1]) < probability;  // dimension 13
    const d3 wgt = take_refl ? color / (probability * rr) / (1.0 - 0.0)
                             : color / ((1.0 - probability) * rr) / (1.0 - 0.0);
    const d3 tw = vmul(T, wgt);
    emit_child(C, x, take_refl ? refl : tdir, tw * (take_refl ? Re : Tr), tw * pk, meta_same, zero);
    return add;
  }
  // depth <= 2: trace both (reflection = branch*2, refraction = branch*2+1)
  const d3 tw = vmul(T, color / (rr) / (1.0 - 0.0));
  emit_child(C, x, refl, tw * Re, tw * pk, rng_c2(depth + 1, branch * 2u), zero);
  emit_child(C, x, tdir, tw * Tr, tw * pk, rng_c2(depth + 1, branch * 2u + 1u), zero);
  return add;
}

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
struct ShadowRay {
  d3 o2, dir;
  double tmax;
  float c[3];
};

// direct_radiance_sample + eval_direct_radiance (diffuse branch),
// directLight_radiance.h:19-387, for one light sample.  Produces a shadow ray
// only when the light's own first hit along it lies within 1e-3 of the
// sampled distance (the reference's acceptance test, line 141); k_shadow
// then checks that nothing nearer blocks it.
__device__ __forceinline__ bool nee_sample(const Params& P, int li, double ra, double rb, d3 x, d3 n_orient, d3 ct,
                                           ShadowRay& s) {
  const DevEntity& le = P.ents[li];
  d3 light_pos, light_normal;
  if (le.type == ENT_SPHERE) {
    const double r1 = PS_TWOPI * ra;
    const double r2 = 1.0 - 2.0 * rb;
    const double s22 = sqrt(1.0 - r2 * r2);
    double sn, cs;
    sincos(r1, &sn, &cs);
    light_pos = ld3(le.p0) + ((le.radius + PS_EPS) * mk(s22 * cs, s22 * sn, r2));
    light_normal = normalize(light_pos - ld3(le.p0));
  } else {
    const double tt = le.v_len * ra;
    const double ss = le.u_len * rb;
    light_pos = ld3(le.p0) + ld3(le.u_axis) * ss + ld3(le.v_axis) * tt + ld3(le.normal) * PS_EPS;
    light_normal = ld3(le.normal);
  }
  if (le.ninv < 0) light_normal = light_normal * -1.0;
  const double pdf = 1.0 / le.area;
  const d3 light_dir = normalize(light_pos - x);
  const double dist2 = sqr(light_pos - x);
  const double dot0 = dot(n_orient, light_dir);
  const double dot1 = dot(light_normal, -1.0 * light_dir);
  if (!(dot0 >= 0 && dot1 >= 0)) return false;
  const d3 o2 = x + light_dir * 0.000001;
  double tl;
  if (!analytic_t(le, o2, light_dir, tl)) return false;
  if (!(fabs(sqrt(dist2) - tl) < 1e-3)) return false;
  const double G = dot0 * dot1 / dist2;
  const d3 E = ld3(P.mats[le.material].emission);
  // (E * colour/pi) * G / pdf, averaged over the 3 samples, / rr, times T;
  // ct already holds T * colour / rr
  const d3 rad = E * (PS_INV_PI * G / pdf / 3.0);
  s.o2 = o2;
  s.dir = light_dir;
  s.tmax = tl;
  s.c[0] = (float)(ct.x * rad.x);
  s.c[1] = (float)(ct.y * rad.y);
  s.c[2] = (float)(ct.z * rad.z);
  return true;
}

__device__ __forceinline__ void put_shadow(ShadowQ& q, uint32_t j, const ShadowRay& s, uint32_t slot, int32_t light) {
  q.light[j] = light;
  q.ox[j] = s.o2.x; q.oy[j] = s.o2.y; q.oz[j] = s.o2.z;
  q.dx[j] = s.dir.x; q.dy[j] = s.dir.y; q.dz[j] = s.dir.z;
  q.tmax[j] = s.tmax;
  q.cr[j] = s.c[0]; q.cg[j] = s.c[1]; q.cb[j] = s.c[2];
  q.slot[j] = slot;
}

// NEE for the diffuse vertices that produced the rays of queue q (their
// origin, throughput T*colour/rr and orienting normal), with the parent
// vertex's random numbers (dimensions 4..12).  Each wave appends its shadow
// rays with one prefix sum and one atomic.
__global__ __launch_bounds__(256) PRT_WAVES_ATTR(PRT_WAVES_NEE) void k_nee(Params P, Batch B, RayQ q, const uint32_t* __restrict__ count_ptr,
                                             ShadowQ sq, uint32_t* count_sq, uint64_t* ctr) {
  const uint32_t n = *count_ptr;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t n_round = (n + 63u) & ~63u;
  uint32_t emitted = 0, dropped = 0;
  const int NL = P.n_lights;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
    ShadowRay s0, s1, s2;
    bool v0 = false, v1 = false, v2 = false;
    int32_t l0 = -1, l1 = -1, l2 = -1;
    uint32_t slot = 0;
    if (i < n) {
      const uint32_t meta = q.meta[i];
      if (meta & META_NEE) {
        slot = q.slot[i];
        const int depth = (int)(meta & 0xFFu) - 1;
        const uint32_t branch = META_BRANCH(meta);
        const uint32_t pl = slot / P.spp, k = slot % P.spp;
        const uint32_t pixel = (uint32_t)(B.pixel_base + pl);
        const uint32_t c2 = rng_c2(depth, branch);
        uint32_t b1[4], b2[4], b3[4];
        philox(pixel, k, c2, 1u, P.key0, P.key1, b1);
        philox(pixel, k, c2, 2u, P.key0, P.key1, b2);
        philox(pixel, k, c2, 3u, P.key0, P.key1, b3);
        const d3 x = mk(q.ox[i], q.oy[i], q.oz[i]);
        const d3 nrm = mk(q.nx[i], q.ny[i], q.nz[i]);
        const d3 ct = mk(q.tr[i], q.tg[i], q.tb[i]);
        // LightList::randomLight (light.h:146-151): index = (int)(u*(N-1)+0.5),
        // probability 1; dimensions 4+3i (pick), 5+3i and 6+3i (point)
        int i0 = (int)(u01(b1[0]) * (double)(NL - 1) + 0.5);
        int i1 = (int)(u01(b1[3]) * (double)(NL - 1) + 0.5);
        int i2 = (int)(u01(b2[2]) * (double)(NL - 1) + 0.5);
        i0 = i0 >= NL ? NL - 1 : i0;
        i1 = i1 >= NL ? NL - 1 : i1;
        i2 = i2 >= NL ? NL - 1 : i2;
        l0 = P.lights[i0];
        l1 = P.lights[i1];
        l2 = P.lights[i2];
        v0 = nee_sample(P, l0, u01(b1[1]), u01(b1[2]), x, nrm, ct, s0);
        v1 = nee_sample(P, l1, u01(b2[0]), u01(b2[1]), x, nrm, ct, s1);
        v2 = nee_sample(P, l2, u01(b2[3]), u01(b3[0]), x, nrm, ct, s2);
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
// Luminance of mlt.h:171-174 with the Vector3d dot pairing.
__device__ __forceinline__ double luminance(double r, double g, double b) {
  return (0.2126 * r + 0.0722 * b) + 0.7152 * g;
}

// Per pixel of the batch: v_k = L_k * smp, luminance clamp (render.h:1375-1384),
// image += v_k; then *= sensor_response (render.h:1405-1409) and the variance
// of the pixel mean.
// With f64 set, out_rgb / out_var point to doubles.
__global__ __launch_bounds__(256) void k_resolve(Params P, Batch B, const uint64_t* __restrict__ accum,
                                                 void* __restrict__ out_rgb, void* __restrict__ out_var, int f64) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n_pixels) return;
  const int64_t pix = B.pixel_base + i;
  const int x = (int)(pix % P.W), y = (int)(pix / P.W);
  const size_t image_index = (size_t)(P.H - y - 1) * P.W + x;
  double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  const uint64_t* a = accum + 3 * (size_t)i * P.spp;
  for (int k = 0; k < P.spp; ++k) {
    double v0 = accum_get(a + 3 * k) * P.smp, v1 = accum_get(a + 3 * k + 1) * P.smp, v2 = accum_get(a + 3 * k + 2) * P.smp;
    if (P.cutoff > 0.0f) {
      const float lv = (float)luminance(v0, v1, v2);
      if (lv > P.cutoff) {
        const float f = P.cutoff / lv;
        v0 *= f; v1 *= f; v2 *= f;
      }
    }
    s1[0] += v0; s1[1] += v1; s1[2] += v2;
    s2[0] += v0 * v0; s2[1] += v1 * v1; s2[2] += v2 * v2;
  }
  const double sr = P.sensor_response;
  const double nn = (double)P.spp;
  for (int c = 0; c < 3; ++c) {
    const double rgb = s1[c] * sr;
    const double var = nn > 1 ? fmax(0.0, (nn * s2[c] - s1[c] * s1[c]) / (nn - 1)) * sr * sr : 0.0;
    if (f64) {
      static_cast<double*>(out_rgb)[image_index * 3 + c] = rgb;
      if (out_var) static_cast<double*>(out_var)[image_index * 3 + c] = var;
    } else {
      static_cast<float*>(out_rgb)[image_index * 3 + c] = (float)rgb;
      if (out_var) static_cast<float*>(out_var)[image_index * 3 + c] = (float)var;
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
