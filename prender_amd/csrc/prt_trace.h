// prt_trace.h — BVH traversal on gfx950: closest hit and any hit over the
// 4-wide node records, with a per-lane stack in LDS.
//
// Semantics are those of intersect_scene_ (scene.cpp:83-214) over the
// reference's QBVH (QBVH.cpp:17-159): the 1e-6 origin nudge, hits accepted at
// t > 1e-6, the nearest hit kept.  The float box test differs from the
// reference's only in being conservative (boxes carry a margin and NaN slabs
// impose no bound), so it never prunes a box that holds the closest hit.
#pragma once
#include "prt_math.h"

struct TravStats {
  uint32_t nodes, tris;
};

// Per-lane traversal stack: the first NL entries in LDS (lane interleaved,
// bank-conflict free), deeper ones in a per-thread global area of
// (PRT_STACK_TOTAL - PRT_STACK_LDS_MIN) * 2 words.
template <int NL>
struct StackN {
  uint32_t* lds_node;  // [NL][blockDim]
  float* lds_t;
  uint32_t* ovf;       // (PRT_STACK_TOTAL - NL) * 2 words used
  int stride;
  __device__ __forceinline__ void push(int sp, int32_t node, float t) {
    if (sp < NL) {
      lds_node[sp * stride] = (uint32_t)node;
      lds_t[sp * stride] = t;
    } else {
      ovf[2 * (sp - NL)] = (uint32_t)node;
      ovf[2 * (sp - NL) + 1] = __float_as_uint(t);
    }
  }
  __device__ __forceinline__ void pop(int sp, int32_t& node, float& t) {
    if (sp < NL) {
      node = (int32_t)lds_node[sp * stride];
      t = lds_t[sp * stride];
    } else {
      node = (int32_t)ovf[2 * (sp - NL)];
      t = __uint_as_float(ovf[2 * (sp - NL) + 1]);
    }
  }
};
typedef StackN<PRT_STACK_LDS> Stack;

// Blocks of 256 threads; a thread's overflow area is picked by its global
// index modulo ovf_threads (the host sizes the area for the largest grid).
#define PRT_DECLARE_STACK_N(S, P, NL)                                                                        \
  __shared__ uint32_t s_node_[(NL) * 256];                                                                   \
  __shared__ float s_t_[(NL) * 256];                                                                         \
  StackN<NL> S{s_node_ + threadIdx.x, s_t_ + threadIdx.x,                                                    \
               (P).stack_ovf + (size_t)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) % (P).ovf_threads) * \
                                   2 * (PRT_STACK_TOTAL - PRT_STACK_LDS_MIN),                                \
               (int)blockDim.x}
#define PRT_DECLARE_STACK(S, P) PRT_DECLARE_STACK_N(S, P, PRT_STACK_LDS)

// a float >= t (t > 0): the culling bound never drops below the double hit
__device__ __forceinline__ float fup(double t) {
  const float f = (float)t;
  return (double)f < t ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

// Closest (ANY = false) or any (ANY = true) hit over the BVH triangles with
// t in (1e-6, t_best).  o is the already-nudged origin.
template <bool ANY, class Stk>
__device__ __forceinline__ bool traverse(const Params& P, d3 o, d3 d, double& t_best, int32_t& id_best, double& ub,
                                         double& vb, Stk& S, TravStats& ts) {
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
      const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w},
                  lza[4] = {lz.x, lz.y, lz.z, lz.w};
      const float hxa[4] = {hx.x, hx.y, hx.z, hx.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w},
                  hza[4] = {hz.x, hz.y, hz.z, hz.w};
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
#define PRT_CSWAP(a, b)                  \
  if (tn[b] < tn[a]) {                   \
    const float tt = tn[a];              \
    tn[a] = tn[b];                       \
    tn[b] = tt;                          \
    const int32_t c_ = cc[a];            \
    cc[a] = cc[b];                       \
    cc[b] = c_;                          \
  }
      PRT_CSWAP(0, 1) PRT_CSWAP(2, 3) PRT_CSWAP(0, 2) PRT_CSWAP(1, 3) PRT_CSWAP(1, 2)
#undef PRT_CSWAP
      if (tn[3] != __builtin_inff()) S.push(sp++, cc[3], tn[3]);
      if (tn[2] != __builtin_inff()) S.push(sp++, cc[2], tn[2]);
      if (tn[1] != __builtin_inff()) S.push(sp++, cc[1], tn[1]);
      if (tn[0] != __builtin_inff()) {
        cur = cc[0];
        continue;
      }
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
      if (t <= tb) {
        got = true;
        break;
      }
    }
    if (!got) break;
  }
  return found;
}

// Closest hit over analytic entities then the BVH; o_in is the ray origin
// before the 1e-6 nudge.  id >= 0: triangle slot; id <= -2: analytic entity
// -2-id; -1: miss.
template <class Stk>
__device__ __forceinline__ void closest_hit(const Params& P, d3 o_in, d3 d, double& t, int32_t& id, double& u,
                                            double& v, Stk& S, TravStats& ts) {
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

// Any blocker with t in (1e-6, tmax) from the already-nudged origin o.
template <class Stk>
__device__ __forceinline__ bool occluded(const Params& P, d3 o, d3 d, double tmax, Stk& S, TravStats& ts) {
  for (int k = 0; k < P.n_analytic; ++k) {
    double tt;
    if (analytic_t(P.ents[P.analytic[k]], o, d, tt) && tt < tmax) return true;
  }
  int32_t id = -1;
  double u, v;
  return traverse<true>(P, o, d, tmax, id, u, v, S, ts);
}

__device__ __forceinline__ void flush_stats(uint64_t* ctr, int idx_nodes, TravStats ts) {
  const uint64_t n = wave_sum_u32(ts.nodes), t = wave_sum_u32(ts.tris);
  if (lane_id() == 0) {
    if (n) atomicAdd((unsigned long long*)&ctr[idx_nodes], (unsigned long long)n);
    if (t) atomicAdd((unsigned long long*)&ctr[idx_nodes + 1], (unsigned long long)t);
  }
}
