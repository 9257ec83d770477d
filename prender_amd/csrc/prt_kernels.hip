// prt_kernels.hip — gfx950 kernels of the wavefront path tracer.
//
// Geometry is evaluated in double precision with the reference's operation
// order and no FMA contraction (this file is built with -ffp-contract=off):
// Moller-Trumbore (polygon.h:78-129), Sphere::intersect (sphere.h:68-96),
// UVPlane::intersect (plane.h:86-110, 220-240), the 1e-6 origin nudge of
// intersect_scene_ (scene.cpp:92).  Closest hits therefore carry the same t
// bits as the reference.  Box tests are float, as in the reference's QBVH
// (BoundingBox.h:137-178), with a conservative margin built into the boxes.
//
// The light transport follows radiance() (radiance.h:25-556) and the BRDFs
// (brdf.h:408-953) turned from recursion into forward throughput: a vertex
// returns Env + E + direct + W * L(next), so each path carries T and adds
// T*(Env + E) and the NEE term, and its child carries T*W.  The "pickup" ray
// of mirror/glass vertices (direct_light_perfect_refrection,
// directLight_radiance.h:516-556) follows the same ray as the child, so the
// child's hit supplies it: a child records a pickup weight and adds
// weight * emission when it lands on a light.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"

#define PS_PI 3.14159265358979323846264338327950288
#define PS_INV_PI 0.31830988618379067153776752674502872
#define PS_TWOPI 6.28318530717958647692528676655900576
#define PS_EPS 1.0e-6
#define PS_EPS14 1.0e-14
#define PS_INF 1.0e128

#define ENT_SPHERE 1
#define ENT_UVPLANE 3
#define ENT_POLYGON 5
#define REFL_DIFFUSE 0
#define REFL_SPECULAR 1
#define REFL_REFRACTION 2

// ------------------------------------------------------------ double vectors
struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 mk(double a, double b, double c) { d3 r; r.x = a; r.y = b; r.z = c; return r; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ d3 operator/(d3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ d3 vmul(d3 a, d3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
// Vector3d dot / sqr pairing (x*x' + z*z') + y*y' (vector3d.h:99-124, 195-218)
__device__ __forceinline__ double dot(d3 a, d3 b) { return (a.x * b.x + a.z * b.z) + a.y * b.y; }
__device__ __forceinline__ double sqr(d3 a) { return (a.x * a.x + a.z * a.z) + a.y * a.y; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
  return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
__device__ __forceinline__ d3 normalize(d3 a) { return a * (1.0 / sqrt(sqr(a))); }
__device__ __forceinline__ d3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }

// ------------------------------------------------------------ Philox4x32-10
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
__device__ __forceinline__ double u01(uint32_t w) { return (double)w * 2.3283064365386963e-10; }

// ------------------------------------------------------------ geometry
// Triangle::CheckIntersection2 + acceptance t > PS_EPS (polygon.h:78-166)
__device__ __forceinline__ bool mt_test(const float4 a, const float4 b, const float4 c, d3 o, d3 d, double& t,
                                        double& u, double& v) {
  const d3 v0 = mk(a.x, a.y, a.z);
  const d3 edge1 = mk(b.x, b.y, b.z) - v0;
  const d3 edge2 = mk(c.x, c.y, c.z) - v0;
  const d3 p = cross(d, edge2);
  const double det = dot(edge1, p);
  const d3 tv = o - v0;
  const d3 q = cross(tv, edge1);
  const double eps = PS_EPS * 0.1;
  double inv_det;
  if (det > eps) {
    u = dot(tv, p);
    if (u < 0.0 || u > det) return false;
    v = dot(d, q);
    if (v < 0.0 || u + v > det) return false;
  } else if (det < -eps) {
    u = dot(tv, p);
    if (u > 0.0 || u < det) return false;
    v = dot(d, q);
    if (v > 0.0 || u + v < det) return false;
  } else {
    return false;
  }
  inv_det = 1.0 / det;
  t = dot(edge2, q) * inv_det;
  if (!(t > PS_EPS)) return false;
  u *= inv_det;
  v *= inv_det;
  return true;
}

// Sphere::intersect, sphere.h:68-87
__device__ __forceinline__ bool sphere_t(const DevEntity& e, d3 o, d3 d, double& t) {
  const d3 p_o = ld3(e.p0) - o;
  const double b = dot(p_o, d);
  const double D4 = b * b - dot(p_o, p_o) + e.radius * e.radius;
  if (D4 < 0.0) return false;
  const double s = sqrt(D4);
  const double t1 = b - s, t2 = b + s;
  if (t1 < PS_EPS && t2 < PS_EPS) return false;
  t = (t1 > PS_EPS) ? t1 : t2;
  return true;
}

// UVPlane::intersect over Plane::intersect, plane.h:86-110, 220-240
__device__ __forceinline__ bool uvplane_t(const DevEntity& e, d3 o, d3 d, double& t) {
  d3 n2 = ld3(e.normal);
  if (e.ninv < 0) n2 = n2 * -1.0;
  const double dd = dot(d, n2);
  if (fabs(dd) < PS_EPS14) return false;
  t = dot((ld3(e.p0) - o), n2) / dd;
  const d3 pos = d * t + o;
  if (t < PS_EPS) return false;
  const d3 rel = pos - ld3(e.p0);
  const double hu = dot(rel, ld3(e.u_axis)) / e.u_len;
  const double hv = dot(rel, ld3(e.v_axis)) / e.v_len;
  if (hu < 0 || hu > e.u_len) return false;
  if (hv < 0 || hv > e.v_len) return false;
  return true;
}

__device__ __forceinline__ bool analytic_t(const DevEntity& e, d3 o, d3 d, double& t) {
  return e.type == ENT_SPHERE ? sphere_t(e, o, d, t) : uvplane_t(e, o, d, t);
}

// ------------------------------------------------------------ traversal
struct TravStats {
  uint32_t nodes, tris;
};

// Per-lane traversal stack: the first PRT_STACK_LDS entries in LDS (lane
// interleaved, bank-conflict free), deeper ones in a per-thread global area.
struct Stack {
  uint32_t* lds_node;  // [PRT_STACK_LDS][blockDim]
  float* lds_t;
  uint32_t* ovf;       // (PRT_STACK_TOTAL - PRT_STACK_LDS) * 2 words
  int stride;
  __device__ __forceinline__ void push(int sp, int32_t node, float t) {
    if (sp < PRT_STACK_LDS) {
      lds_node[sp * stride] = (uint32_t)node;
      lds_t[sp * stride] = t;
    } else {
      ovf[2 * (sp - PRT_STACK_LDS)] = (uint32_t)node;
      ovf[2 * (sp - PRT_STACK_LDS) + 1] = __float_as_uint(t);
    }
  }
  __device__ __forceinline__ void pop(int sp, int32_t& node, float& t) {
    if (sp < PRT_STACK_LDS) {
      node = (int32_t)lds_node[sp * stride];
      t = lds_t[sp * stride];
    } else {
      node = (int32_t)ovf[2 * (sp - PRT_STACK_LDS)];
      t = __uint_as_float(ovf[2 * (sp - PRT_STACK_LDS) + 1]);
    }
  }
};

__device__ __forceinline__ float fup(double t) {
  // a float >= t (the culling bound must never be below the double hit)
  float f = (float)t;
  return (double)f < t ? __uint_as_float(__float_as_uint(f) + 1u) : f;  // t > 0
}

// Closest (ANY = false) or any (ANY = true) hit over the BVH triangles with
// t in (1e-6, t_best).  o is the already-nudged origin.
template <bool ANY>
__device__ __forceinline__ bool traverse(const Params& P, d3 o, d3 d, double& t_best, int32_t& id_best, double& ub,
                                         double& vb, Stack& S, TravStats& ts) {
  if (P.n_nodes <= 0) return false;
  const float ox = (float)o.x, oy = (float)o.y, oz = (float)o.z;
  float ix = (float)(1.0 / d.x), iy = (float)(1.0 / d.y), iz = (float)(1.0 / d.z);
  if (__builtin_isinf(ix)) ix = __builtin_copysignf(1.0e30f, ix);
  if (__builtin_isinf(iy)) iy = __builtin_copysignf(1.0e30f, iy);
  if (__builtin_isinf(iz)) iz = __builtin_copysignf(1.0e30f, iz);
  const float oix = ox * ix, oiy = oy * iy, oiz = oz * iz;
  float tb = t_best < 1.0e30 ? fup(t_best) : __builtin_inff();
  bool found = false;
  int sp = 0;
  int32_t cur = 0;
  const float4* nodes4 = reinterpret_cast<const float4*>(P.nodes);
  const float4* tris4 = reinterpret_cast<const float4*>(P.tris);
  while (true) {
    if (cur >= 0) {
      const float4* nd = nodes4 + (size_t)cur * 8;
      const float4 lx = nd[0], ly = nd[1], lz = nd[2], hx = nd[3], hy = nd[4], hz = nd[5];
      const int4 ch = *reinterpret_cast<const int4*>(nd + 6);
      ts.nodes++;
      float tn[4];
      int32_t cc[4];
      const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w}, lza[4] = {lz.x, lz.y, lz.z, lz.w};
      const float hxa[4] = {hx.x, hx.y, hx.z, hx.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
      const int32_t cha[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float x0 = __builtin_fmaf(lxa[k], ix, -oix), x1 = __builtin_fmaf(hxa[k], ix, -oix);
        const float y0 = __builtin_fmaf(lya[k], iy, -oiy), y1 = __builtin_fmaf(hya[k], iy, -oiy);
        const float z0 = __builtin_fmaf(lza[k], iz, -oiz), z1 = __builtin_fmaf(hza[k], iz, -oiz);
        const float tnear = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
        const float tfar = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
        tn[k] = (tnear <= tfar) ? tnear : __builtin_inff();
        cc[k] = cha[k];
      }
      // sort the four (t, child) pairs ascending
#define PRT_CSWAP(a, b)                                                  \
  if (tn[b] < tn[a]) {                                                   \
    float tt = tn[a]; tn[a] = tn[b]; tn[b] = tt;                         \
    int32_t c_ = cc[a]; cc[a] = cc[b]; cc[b] = c_;                       \
  }
      PRT_CSWAP(0, 1) PRT_CSWAP(2, 3) PRT_CSWAP(0, 2) PRT_CSWAP(1, 3) PRT_CSWAP(1, 2)
#undef PRT_CSWAP
      if (tn[3] != __builtin_inff()) { S.push(sp++, cc[3], tn[3]); }
      if (tn[2] != __builtin_inff()) { S.push(sp++, cc[2], tn[2]); }
      if (tn[1] != __builtin_inff()) { S.push(sp++, cc[1], tn[1]); }
      if (tn[0] != __builtin_inff()) { cur = cc[0]; continue; }
    } else {
      const uint32_t code = ~(uint32_t)cur;
      const uint32_t first = code & 0x07FFFFFFu;
      const uint32_t count = (code >> 27) + 1u;
      for (uint32_t k = 0; k < count; ++k) {
        const float4* tr = tris4 + (size_t)(first + k) * 3;
        const float4 a = tr[0], b = tr[1], c = tr[2];
        ts.tris++;
        double t, u, v;
        if (mt_test(a, b, c, o, d, t, u, v) && t < t_best) {
          if (ANY) {
            id_best = (int32_t)(first + k);
            return true;
          }
          t_best = t;
          ub = u;
          vb = v;
          id_best = (int32_t)(first + k);
          tb = fup(t);
          found = true;
        }
      }
    }
    // pop the nearest pending entry that can still hold a closer hit
    bool got = false;
    while (sp > 0) {
      float t;
      S.pop(--sp, cur, t);
      if (t <= tb) { got = true; break; }
    }
    if (!got) break;
  }
  return found;
}

// Full closest hit over analytic entities then the BVH; o is the ray origin
// before the 1e-6 nudge.  id >= 0: triangle slot; id <= -2: analytic entity
// index -2-id; -1: miss.
__device__ __forceinline__ void closest_hit(const Params& P, d3 o_in, d3 d, double& t, int32_t& id, double& u,
                                            double& v, Stack& S, TravStats& ts) {
  const d3 o = o_in + d * 0.000001;
  t = PS_INF;
  id = -1;
  u = v = -PS_INF;
  for (int k = 0; k < P.n_analytic; ++k) {
    const int32_t ei = P.analytic[k];
    double tt;
    if (analytic_t(P.ents[ei], o, d, tt) && tt < t) {
      t = tt;
      id = -2 - ei;
    }
  }
  traverse<false>(P, o, d, t, id, u, v, S, ts);
}

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t wave_sum_u32(uint32_t x) {
  uint32_t s = x;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += (uint32_t)__shfl_xor((int)s, off, 64);
  return (uint64_t)s;
}

__device__ __forceinline__ void flush_stats(uint64_t* ctr, int idx_nodes, TravStats ts) {
  const uint64_t n = wave_sum_u32(ts.nodes), t = wave_sum_u32(ts.tris);
  if (lane_id() == 0) {
    if (n) atomicAdd((unsigned long long*)&ctr[idx_nodes], (unsigned long long)n);
    if (t) atomicAdd((unsigned long long*)&ctr[idx_nodes + 1], (unsigned long long)t);
  }
}

// ------------------------------------------------------------ trace API kernel
__global__ __launch_bounds__(256) void k_trace(Params P, const prt_ray_dev* __restrict__ rays, int64_t n,
                                               prt_hit_dev* __restrict__ hits, uint64_t* ctr) {
  __shared__ uint32_t s_node[PRT_STACK_LDS * 256];
  __shared__ float s_t[PRT_STACK_LDS * 256];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack S{s_node + threadIdx.x, s_t + threadIdx.x,
          P.stack_ovf + (size_t)(gtid % P.ovf_threads) * 2 * (PRT_STACK_TOTAL - PRT_STACK_LDS), (int)blockDim.x};
  TravStats ts{0, 0};
  for (int64_t i = gtid; i < n; i += (int64_t)gridDim.x * blockDim.x) {
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

// ------------------------------------------------------------ wavefront kernels
// Primary rays of the unique sub-pixel positions of a batch
// (render.h:1248-1269: r = (s + 0.5)/SS, no jitter).
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
  const int64_t n = (int64_t)B.n_pixels * ss2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pix = B.pixel_base + i / ss2;
    const int sub = (int)(i % ss2);
    const int x = (int)(pix % P.W), y = (int)(pix / P.W);
    d3 o, d;
    camera_ray(P, x, y, sub % P.SS, sub / P.SS, o, d);
    q.ox[i] = o.x; q.oy[i] = o.y; q.oz[i] = o.z;
    q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
  }
}

// Closest hit for every ray of a queue; waves fetch 64 rays at a time.
__global__ __launch_bounds__(256) void k_extend(Params P, RayQ q, const uint32_t* __restrict__ count_ptr,
                                                uint32_t* fetch, double* __restrict__ hit_t,
                                                int32_t* __restrict__ hit_id, uint64_t* ctr) {
  __shared__ uint32_t s_node[PRT_STACK_LDS * 256];
  __shared__ float s_t[PRT_STACK_LDS * 256];
  const uint32_t n = *count_ptr;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack S{s_node + threadIdx.x, s_t + threadIdx.x,
          P.stack_ovf + (size_t)(gtid % P.ovf_threads) * 2 * (PRT_STACK_TOTAL - PRT_STACK_LDS), (int)blockDim.x};
  TravStats ts{0, 0};
  while (true) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = __shfl(base, 0, 64);
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

// Any-hit for NEE shadow rays; unoccluded rays add their contribution to the
// sample accumulator.
__global__ __launch_bounds__(256) void k_shadow(Params P, ShadowQ q, const uint32_t* __restrict__ count_ptr,
                                                uint32_t* fetch, float* __restrict__ accum, uint64_t* ctr) {
  __shared__ uint32_t s_node[PRT_STACK_LDS * 256];
  __shared__ float s_t[PRT_STACK_LDS * 256];
  const uint32_t n = *count_ptr;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack S{s_node + threadIdx.x, s_t + threadIdx.x,
          P.stack_ovf + (size_t)(gtid % P.ovf_threads) * 2 * (PRT_STACK_TOTAL - PRT_STACK_LDS), (int)blockDim.x};
  TravStats ts{0, 0};
  while (true) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = __shfl(base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane_id();
    if (i < n) {
      const d3 o = mk(q.ox[i], q.oy[i], q.oz[i]);
      const d3 d = mk(q.dx[i], q.dy[i], q.dz[i]);
      double tmax = q.tmax[i];
      bool occluded = false;
      for (int k = 0; k < P.n_analytic && !occluded; ++k) {
        double tt;
        if (analytic_t(P.ents[P.analytic[k]], o, d, tt) && tt < tmax) occluded = true;
      }
      if (!occluded) {
        int32_t id = -1;
        double u, v;
        occluded = traverse<true>(P, o, d, tmax, id, u, v, S, ts);
      }
      if (!occluded) {
        const uint32_t s = q.slot[i];
        atomicAdd(&accum[3 * (size_t)s + 0], q.cr[i]);
        atomicAdd(&accum[3 * (size_t)s + 1], q.cg[i]);
        atomicAdd(&accum[3 * (size_t)s + 2], q.cb[i]);
      }
    }
  }
  flush_stats(ctr, CTR_NODES, ts);
}

// ------------------------------------------------------------ shading
// Append to a queue from divergent code: the active lanes of the wave share
// one atomic (ballot, rank by popcount of the lower lanes, leader adds the
// wave's total and broadcasts the base).
__device__ __forceinline__ uint32_t agg_append(uint32_t* counter) {
  const uint64_t mask = __ballot(1);
  const uint32_t lane = lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
  const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader, 64);
  return base + rank;
}

struct ShadeCtx {
  RayQ qout;
  uint32_t* count_out;
  ShadowQ sq;
  uint32_t* count_sq;
  uint32_t slot;
  int max_vertices;
  uint32_t truncated, n_ext, n_sh, pickups, dropped;
};

__device__ __forceinline__ void emit_child(ShadeCtx& C, d3 o, d3 d, float t0, float t1, float t2, float p0, float p1,
                                           float p2, uint32_t meta) {
  if ((int)(meta & 0xFFu) >= C.max_vertices) {
    C.truncated++;
    return;
  }
  const uint32_t j = agg_append(C.count_out);
  if (j >= C.qout.cap) {
    C.dropped++;
    return;
  }
  C.qout.ox[j] = o.x; C.qout.oy[j] = o.y; C.qout.oz[j] = o.z;
  C.qout.dx[j] = d.x; C.qout.dy[j] = d.y; C.qout.dz[j] = d.z;
  C.qout.tr[j] = t0; C.qout.tg[j] = t1; C.qout.tb[j] = t2;
  C.qout.pr[j] = p0; C.qout.pg[j] = p1; C.qout.pb[j] = p2;
  C.qout.slot[j] = C.slot;
  C.qout.meta[j] = meta;
  C.n_ext++;
}

// direct_radiance_sample + eval_direct_radiance (diffuse branch),
// directLight_radiance.h:19-387.  Emits a shadow ray only when the light's
// own first hit along the ray lies within 1e-3 of the sampled distance; the
// shadow kernel then checks that nothing is nearer.
__device__ __forceinline__ void nee_sample(const Params& P, ShadeCtx& C, int li, double ra, double rb, d3 x,
                                           d3 n_orient, d3 color, d3 T, double rr) {
  const DevEntity& le = P.ents[li];
  d3 light_pos, light_normal;
  double pdf;
  if (le.type == ENT_SPHERE) {
    const double r1 = PS_TWOPI * ra;
    const double r2 = 1.0 - 2.0 * rb;
    const double s22 = sqrt(1.0 - r2 * r2);
    light_pos = ld3(le.p0) + ((le.radius + PS_EPS) * mk(s22 * cos(r1), s22 * sin(r1), r2));
    light_normal = normalize(light_pos - ld3(le.p0));
    if (le.ninv < 0) light_normal = light_normal * -1.0;
    pdf = 1.0 / le.area;
  } else {
    const double t = le.v_len * ra;
    const double s = le.u_len * rb;
    light_pos = ld3(le.p0) + ld3(le.u_axis) * s + ld3(le.v_axis) * t + ld3(le.normal) * PS_EPS;
    light_normal = ld3(le.normal);
    if (le.ninv < 0) light_normal = light_normal * -1.0;
    pdf = 1.0 / le.area;
  }
  const d3 light_dir = normalize(light_pos - x);
  const double dist2 = sqr(light_pos - x);
  const double dot0 = dot(n_orient, light_dir);
  const double dot1 = dot(light_normal, -1.0 * light_dir);
  if (!(dot0 >= 0 && dot1 >= 0)) return;
  const d3 o2 = x + light_dir * 0.000001;
  double tl;
  if (!analytic_t(le, o2, light_dir, tl)) return;
  if (!(fabs(sqrt(dist2) - tl) < 1e-3)) return;
  const double G = dot0 * dot1 / dist2;
  const d3 E = ld3(P.mats[le.material].emission);
  const d3 rad = vmul(E, color * PS_INV_PI) * G / pdf;
  const uint32_t j = agg_append(C.count_sq);
  if (j >= C.sq.cap) {
    C.dropped++;
    return;
  }
  C.sq.ox[j] = o2.x; C.sq.oy[j] = o2.y; C.sq.oz[j] = o2.z;
  C.sq.dx[j] = light_dir.x; C.sq.dy[j] = light_dir.y; C.sq.dz[j] = light_dir.z;
  C.sq.tmax[j] = tl;
  // direct = (sum of the three samples) / 3 / rr, times the path throughput
  C.sq.cr[j] = (float)(T.x * (rad.x / 3 / rr));
  C.sq.cg[j] = (float)(T.y * (rad.y / 3 / rr));
  C.sq.cb[j] = (float)(T.z * (rad.z / 3 / rr));
  C.sq.slot[j] = C.slot;
  C.n_sh++;
}

// One path vertex: the radiance() body for the hit (or miss) of ray (o, d).
// Returns the radiance added to the camera sample; children and shadow rays
// go straight to their queues.
__device__ __forceinline__ d3 shade_vertex(const Params& P, ShadeCtx& C, d3 o, d3 d, double t, int32_t id, d3 T,
                                           d3 Pw, uint32_t pixel, uint32_t sample, int depth, uint32_t branch) {
  d3 add = mk(0.0, 0.0, 0.0);
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
    C.pickups++;
  }
  const d3 n_orient = dot(n, d) < 0.0 ? n : (-1.0 * n);
  // RGB_MAX = std::max(x, std::max(y, z)) (def.h:19)
  const d3 rw = E + color;
  const double myz = (rw.y < rw.z) ? rw.z : rw.y;
  double rr = (rw.x < myz) ? myz : rw.x;
  if (depth > P.depth_limit) rr *= pow(0.5, (double)(depth - P.depth_limit));
  uint32_t b0[4];
  const uint32_t c2 = ((uint32_t)depth & 0xFFu) | (branch << 8);
  philox(pixel, sample, c2, 0u, P.key0, P.key1, b0);
  if (depth > P.depth) {
    if (u01(b0[0]) >= rr) return add + vmul(T, E);  // radiance.h:106-116
  } else {
    rr = 1.0;
  }
  int kind;  // 0 diffuse, 1 specular, 2 refraction
  if (m.type == REFL_DIFFUSE) kind = (u01(b0[1]) < m.roughness) ? 0 : 1;
  else if (m.type == REFL_SPECULAR) kind = (u01(b0[1]) < 1.0 - m.roughness) ? 1 : 0;
  else kind = 2;
  const uint32_t meta_same = ((uint32_t)(depth + 1) & 0xFFu) | (branch << 8);
  const float pk = P.nee ? 1.0f : 0.0f;
  if (kind == 0) {
    // BRDF_Diffuse, brdf.h:408-700
    if (e.light && e.type == ENT_UVPLANE && dot(n, d) >= 0.0) return add;
    if (e.light) {
      if (!P.nee || depth == 0) add = add + vmul(T, E / rr / (1.0 - 0.0));
      return add;
    }
    add = add + vmul(T, ld3(P.env_light) + E);
    // OrthonormalBasis (vector3d.h:244-253), cosine sampling (brdf.h:498-505)
    const d3 w = n_orient;
    const d3 uu = (fabs(w.x) > PS_EPS) ? normalize(cross(mk(0.0, 1.0, 0.0), w)) : normalize(cross(mk(1.0, 0.0, 0.0), w));
    const d3 vv = normalize(cross(w, uu));
    const double r1 = PS_TWOPI * u01(b0[2]);
    const double r2 = u01(b0[3]), r2s = sqrt(r2);
    const d3 dir = normalize((uu * cos(r1) * r2s + vv * sin(r1) * r2s + w * sqrt(1.0 - r2)));
    const d3 ct = vmul(T, color / rr / (1.0 - 0.0));
    emit_child(C, x, dir, (float)ct.x, (float)ct.y, (float)ct.z, 0.0f, 0.0f, 0.0f, meta_same);
    if (P.nee && P.n_lights > 0) {
      uint32_t b1[4], b2[4], b3[4];
      philox(pixel, sample, c2, 1u, P.key0, P.key1, b1);
      philox(pixel, sample, c2, 2u, P.key0, P.key1, b2);
      philox(pixel, sample, c2, 3u, P.key0, P.key1, b3);
      // dims 4+3i (light pick), 5+3i and 6+3i (point on the light)
      const uint32_t pick[3] = {b1[0], b1[3], b2[2]};
      const uint32_t pa[3] = {b1[1], b2[0], b2[3]};
      const uint32_t pb[3] = {b1[2], b2[1], b3[0]};
      const int NL = P.n_lights;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        // LightList::randomLight, light.h:146-151 (probability 1)
        int index = (int)((u01(pick[i])) * (double)(NL - 1) + 0.5);
        if (index >= NL) index = NL - 1;
        if (index < 0) index = 0;
        nee_sample(P, C, P.lights[index], u01(pa[i]), u01(pb[i]), x, n_orient, color, T, rr);
      }
    }
    return add;
  }
  add = add + vmul(T, ld3(P.env_light) + E);
  if (kind == 1) {
    // BRDF_Specular, brdf.h:702-761
    const d3 dir = normalize(d - n * 2.0 * dot(n, d));
    const d3 ct = vmul(T, color / rr / (1.0 - 0.0));
    emit_child(C, x, dir, (float)ct.x, (float)ct.y, (float)ct.z, pk * (float)ct.x, pk * (float)ct.y,
               pk * (float)ct.z, meta_same);
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
    emit_child(C, x, refl, (float)ct.x, (float)ct.y, (float)ct.z, pk * (float)ct.x, pk * (float)ct.y,
               pk * (float)ct.z, meta_same);
    return add;
  }
  const d3 tdir = normalize(d * nnt - n * (into ? 1.0 : -1.0) * (ddn * nnt + sqrt(fabs(cos2t))));
  const double a = nt - nc, b = nt + nc;
  const double R0 = (a * a) / (b * b);
  const double c = 1.0 - (into ? -ddn : dot(tdir, -1.0 * n_orient));
  const double Re = R0 + (1.0 - R0) * pow(c, 5.0);
  const double nnt2 = pow(into ? nc / nt : nt / nc, 2.0);
  const double Tr = (1.0 - Re) * nnt2;
  const double probability = 0.25 + 0.5 * Re;
  if (depth > 2) {
    uint32_t b3[4];
    philox(pixel, sample, c2, 3u, P.key0, P.key1, b3);
    const bool take_refl = u01(b3[1]) < probability;  // dim 13
    const d3 wgt = take_refl ? color / (probability * rr) / (1.0 - 0.0)
                             : color / ((1.0 - probability) * rr) / (1.0 - 0.0);
    const d3 tw = vmul(T, wgt);
    const d3 ct = tw * (take_refl ? Re : Tr);
    emit_child(C, x, take_refl ? refl : tdir, (float)ct.x, (float)ct.y, (float)ct.z, pk * (float)tw.x,
               pk * (float)tw.y, pk * (float)tw.z, meta_same);
    return add;
  }
  // depth <= 2: trace both (reflection = branch*2, refraction = branch*2+1)
  const d3 tw = vmul(T, color / (rr) / (1.0 - 0.0));
  const d3 cr = tw * Re, ctt = tw * Tr;
  emit_child(C, x, refl, (float)cr.x, (float)cr.y, (float)cr.z, pk * (float)tw.x, pk * (float)tw.y, pk * (float)tw.z,
             ((uint32_t)(depth + 1) & 0xFFu) | ((branch * 2u) << 8));
  emit_child(C, x, tdir, (float)ctt.x, (float)ctt.y, (float)ctt.z, pk * (float)tw.x, pk * (float)tw.y,
             pk * (float)tw.z, ((uint32_t)(depth + 1) & 0xFFu) | ((branch * 2u + 1u) << 8));
  return add;
}

// Shade one queue (or, with primary = 1, every camera sample of the batch
// against the hit of its sub-pixel's primary ray).
__global__ __launch_bounds__(256) void k_shade(Params P, Batch B, int primary, RayQ qin,
                                               const uint32_t* __restrict__ count_in, const double* __restrict__ hit_t,
                                               const int32_t* __restrict__ hit_id, RayQ qout, uint32_t* count_out,
                                               ShadowQ sq, uint32_t* count_sq, float* __restrict__ accum,
                                               uint64_t* ctr) {
  const uint32_t n = primary ? (uint32_t)((int64_t)B.n_pixels * P.spp) : *count_in;
  const uint32_t stride = gridDim.x * blockDim.x;
  const int ss2 = P.SS * P.SS;
  ShadeCtx C;
  C.qout = qout;
  C.count_out = count_out;
  C.sq = sq;
  C.count_sq = count_sq;
  C.max_vertices = P.max_vertices;
  C.truncated = C.n_ext = C.n_sh = C.pickups = C.dropped = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    d3 o, d, T, Pw;
    double t;
    int32_t id;
    int depth;
    uint32_t branch;
    uint32_t slot;
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
      branch = meta >> 8;
      t = hit_t[i];
      id = hit_id[i];
    }
    C.slot = slot;
    const uint32_t pl = slot / P.spp, k = slot % P.spp;
    const d3 add = shade_vertex(P, C, o, d, t, id, T, Pw, (uint32_t)(B.pixel_base + pl), k, depth, branch);
    if (add.x != 0.0 || add.y != 0.0 || add.z != 0.0) {
      atomicAdd(&accum[3 * (size_t)slot + 0], (float)add.x);
      atomicAdd(&accum[3 * (size_t)slot + 1], (float)add.y);
      atomicAdd(&accum[3 * (size_t)slot + 2], (float)add.z);
    }
  }
  const uint64_t s_pick = wave_sum_u32(C.pickups), s_tr = wave_sum_u32(C.truncated);
  const uint64_t s_ext = wave_sum_u32(C.n_ext), s_sh = wave_sum_u32(C.n_sh);
  const uint64_t s_drop = wave_sum_u32(C.dropped);
  if (lane_id() == 0) {
    if (s_drop) atomicAdd((unsigned long long*)&ctr[CTR_DROP], (unsigned long long)s_drop);
    if (s_pick) atomicAdd((unsigned long long*)&ctr[CTR_PICKUP], (unsigned long long)s_pick);
    if (s_tr) atomicAdd((unsigned long long*)&ctr[CTR_TRUNC], (unsigned long long)s_tr);
    if (s_ext) atomicAdd((unsigned long long*)&ctr[CTR_EXT], (unsigned long long)s_ext);
    if (s_sh) atomicAdd((unsigned long long*)&ctr[CTR_SHADOW], (unsigned long long)s_sh);
  }
}

// Luminance of mlt.h:171-174 with the Vector3d dot pairing.
__device__ __forceinline__ double luminance(double r, double g, double b) {
  return (0.2126 * r + 0.0722 * b) + 0.7152 * g;
}

// Per pixel of the batch: v_k = L_k * smp, luminance clamp (render.h:1375-1384),
// image += v_k; then *= sensor_response (render.h:1405-1409) and the
// variance of the pixel mean.
__global__ __launch_bounds__(256) void k_resolve(Params P, Batch B, const float* __restrict__ accum,
                                                 float* __restrict__ out_rgb, float* __restrict__ out_var) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n_pixels) return;
  const int64_t pix = B.pixel_base + i;
  const int x = (int)(pix % P.W), y = (int)(pix / P.W);
  const size_t image_index = (size_t)(P.H - y - 1) * P.W + x;
  double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  const float* a = accum + 3 * (size_t)i * P.spp;
  for (int k = 0; k < P.spp; ++k) {
    double v0 = (double)a[3 * k] * P.smp, v1 = (double)a[3 * k + 1] * P.smp, v2 = (double)a[3 * k + 2] * P.smp;
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
  const double n = (double)P.spp;
  for (int c = 0; c < 3; ++c) {
    out_rgb[image_index * 3 + c] = (float)(s1[c] * sr);
    if (out_var) out_var[image_index * 3 + c] = (float)(n > 1 ? fmax(0.0, (n * s2[c] - s1[c] * s1[c]) / (n - 1)) * sr * sr : 0.0);
  }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
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
  hipLaunchKernelGGL(k_gen_primary, dim3(grid_for((int64_t)B.n_pixels * P.SS * P.SS)), dim3(256), 0, s, P, B, q);
  return hipGetLastError();
}
hipError_t launch_extend(const Params& P, const RayQ& q, const uint32_t* count, int64_t n_host, uint32_t* fetch,
                         double* hit_t, int32_t* hit_id, uint64_t* ctr, int max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_extend, dim3(grid_for(n_host, 256, max_blocks)), dim3(256), 0, s, P, q, count, fetch, hit_t,
                     hit_id, ctr);
  return hipGetLastError();
}
hipError_t launch_shadow(const Params& P, const ShadowQ& q, const uint32_t* count, int64_t n_max, uint32_t* fetch,
                         float* accum, uint64_t* ctr, int max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow, dim3(grid_for(n_max, 256, max_blocks)), dim3(256), 0, s, P, q, count, fetch, accum,
                     ctr);
  return hipGetLastError();
}
hipError_t launch_shade(const Params& P, const Batch& B, int primary, const RayQ& qin, const uint32_t* count_in,
                        int64_t n_host, const double* hit_t, const int32_t* hit_id, const RayQ& qout,
                        uint32_t* count_out, const ShadowQ& sq, uint32_t* count_sq, float* accum, uint64_t* ctr,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_shade, dim3(grid_for(n_host)), dim3(256), 0, s, P, B, primary, qin, count_in, hit_t, hit_id,
                     qout, count_out, sq, count_sq, accum, ctr);
  return hipGetLastError();
}
hipError_t launch_resolve(const Params& P, const Batch& B, const float* accum, float* rgb, float* var,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_resolve, dim3(grid_for(B.n_pixels)), dim3(256), 0, s, P, B, accum, rgb, var);
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
