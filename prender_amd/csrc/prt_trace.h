// prt_trace.h — BVH traversal on gfx950: closest hit and any hit over the
// 4-wide node records, with a per-lane stack in LDS.
//
// Semantics are those of intersect_scene_ (scene.cpp:83-214) over the
// reference's QBVH (QBVH.cpp:17-159): the 1e-6 origin nudge, hits accepted at
// t > 1e-6, the nearest hit kept.  The float box test differs from the
// reference's only in being conservative (boxes carry a margin and NaN slabs
// impose no bound), so it never prunes a box that holds the closest hit.
#pragma once
#include "prt_kernels.h"
#include "prt_math.h"

struct TravStats {
  uint32_t nodes, tris;
  uint32_t rays;  // queries finished (traverse_pool)
};

// Per-lane traversal stack of (node, t) entries: the first NL in LDS (lane
// interleaved, one 64-bit entry per level, bank-conflict free), deeper ones
// in a global overflow area of Params::ovf_words words per thread (enough for
// the scene's BVH depth), thread interleaved as well: entry k of every
// thread of a region lies together (k * ovf_threads + thread), so the lanes
// of a wave that spill at the same depth write 512 contiguous bytes instead
// of 64 lines.  The overflow accesses are volatile so that the compiler keeps
// them apart from the LDS ones: merged, both would become flat-address loads
// and stores, which wait on every outstanding memory operation.
// four floats as a native vector (copyable out of an address-space-qualified
// object, unlike float4)
typedef float prt_f4 __attribute__((ext_vector_type(4)));

template <int NL>
struct StackN {
  static constexpr int stride = 256;  // threads per block of every traversal kernel
  uint2* lds;                         // [NL][256]
  uint2* ovf;                         // this thread's first overflow entry
  size_t ostride;                     // entries between its consecutive overflow entries
  // (3 * BVH depth + 1 - NL overflow entries used at most)
  // the block's LDS copy of BVH nodes 0..n_top-1 (PRT_TOP_CACHE)
  const __attribute__((address_space(3))) prt_f4* top = nullptr;
  int n_top = 0;
  __device__ __forceinline__ void push(int sp, int32_t node, float t) {
    const uint2 e = make_uint2((uint32_t)node, __float_as_uint(t));
    if (sp < NL) {
      lds[sp * stride] = e;
    } else {
      volatile uint32_t* g = reinterpret_cast<volatile uint32_t*>(ovf + (size_t)(sp - NL) * ostride);
      g[0] = e.x;
      g[1] = e.y;
    }
  }
  __device__ __forceinline__ void pop(int sp, int32_t& node, float& t) {
    uint2 e;
    if (sp < NL) {
      e = lds[sp * stride];
    } else {
      const volatile uint32_t* g = reinterpret_cast<const volatile uint32_t*>(ovf + (size_t)(sp - NL) * ostride);
      e.x = g[0];
      e.y = g[1];
    }
    node = (int32_t)e.x;
    t = __uint_as_float(e.y);
  }
};
typedef StackN<PRT_STACK_LDS> Stack;

// Blocks of 256 threads; a thread's overflow area is picked by its global
// index modulo ovf_threads (the host sizes the area for the largest grid)
// within one of PRT_OVF_REGIONS regions: closest-hit kernels use region 0,
// shadow kernels region 1, because the two run at the same time on two
// streams.
#define PRT_DECLARE_STACK_N(S, P, NL, REGION)                                                             \
  __shared__ uint2 s_stack_[(NL) * 256];                                                                  \
  StackN<NL> S{s_stack_ + threadIdx.x,                                                                    \
               reinterpret_cast<uint2*>((P).stack_ovf +                                                   \
                                        (size_t)(REGION) * (size_t)(P).ovf_threads * (size_t)(P).ovf_words) + \
                   (size_t)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) % (P).ovf_threads),          \
               (size_t)(P).ovf_threads}
#define PRT_DECLARE_STACK(S, P, REGION) PRT_DECLARE_STACK_N(S, P, PRT_STACK_LDS, REGION)

// Copies BVH nodes 0..NTOP-1 (the top levels, numbered first by the builder;
// NTOP <= PRT_TOP_NODES) into the block's LDS for stack S.  Every thread of
// the block must reach it.
#define PRT_TOP_CACHE(S, NODES, NTOP)                                                        \
  __shared__ float4 s_top_[PRT_TOP_NODES * 8];                                               \
  {                                                                                          \
    const int nt_ = (NTOP);                                                                  \
    const float4* g_ = reinterpret_cast<const float4*>(NODES);                               \
    for (int k_ = (int)threadIdx.x; k_ < nt_ * 8; k_ += (int)blockDim.x) s_top_[k_] = g_[k_]; \
    __syncthreads();                                                                         \
    S.top = (const __attribute__((address_space(3))) prt_f4*)s_top_;                         \
    S.n_top = nt_;                                                                           \
  }

// a float >= t (t > 0): the culling bound never drops below the double hit
__device__ __forceinline__ float fup(double t) {
  const float f = (float)t;
  return (double)f < t ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

// Byte offsets within a 128-byte node record of the row holding each axis'
// entry planes, packed x | y << 8 | z << 16: the lo row (x 0, y 16, z 32)
// where the ray's inverse direction on that axis is >= 0, else the hi row
// (x 48, y 64, z 80).  The exit planes are in the other row, at the entry
// offset XOR 48, 80 or 112.  ix, iy, iz are never zero (|1 / d| >= 1 for a
// unit direction, and an infinite one is clamped to +-1e30).
__device__ __forceinline__ uint32_t entry_rows(float ix, float iy, float iz) {
  return (ix < 0.0f ? 48u : 0u) | ((iy < 0.0f ? 64u : 16u) << 8) | ((iz < 0.0f ? 80u : 32u) << 16);
}

// Slab test of the 4 children of node `cur` against the ray's float slabs:
// entry distances sorted ascending with their child references (+inf: miss).
// Each axis reads its entry and exit planes from the rows entry_rows picks:
// fma(c, i, -oi) is monotonic in the plane coordinate c and every child has
// lo <= hi (an empty slot lo = hi = +inf), so the entry plane's distance is
// exactly the smaller of the two, without a min / max per axis.
// The sort is branch-free: each distance (>= 0, sign cleared so -0 sorts as
// 0) becomes a 32-bit key whose bit pattern orders like the float, with the
// child slot in its two low bits, and a min/max network sorts the keys.  The
// distance read back from a key is rounded down by at most 3 ulp, so culling
// against it stays conservative; +inf stays +inf.
// Nodes of the top levels come from the block's LDS copy.
template <class Stk>
__device__ __forceinline__ void node_children(const Stk& S, const DevNode* nodes, int32_t cur, uint32_t rows,
                                              float ix, float iy, float iz, float oix, float oiy, float oiz, float tb,
                                              float tn[4], int32_t cc[4]) {
  const uint32_t nx = rows & 0xFFu, ny = (rows >> 8) & 0xFFu, nz = rows >> 16;
  const uint32_t fx = nx ^ 48u, fy = ny ^ 80u, fz = nz ^ 112u;
  float4 nd[7];  // entry x, y, z rows, exit x, y, z rows, children
  if (cur < S.n_top) {
    typedef const __attribute__((address_space(3))) char lds_c;
    typedef const __attribute__((address_space(3))) prt_f4 lds_f4;
    lds_c* t = reinterpret_cast<lds_c*>(S.top + (size_t)cur * 8);
    const uint32_t off[7] = {nx, ny, nz, fx, fy, fz, 96u};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const prt_f4 v = *reinterpret_cast<lds_f4*>(t + off[k]);
      nd[k] = make_float4(v.x, v.y, v.z, v.w);
    }
  } else {
    const char* g = reinterpret_cast<const char*>(nodes + cur);
    const uint32_t off[7] = {nx, ny, nz, fx, fy, fz, 96u};
#pragma unroll
    for (int k = 0; k < 7; ++k) nd[k] = *reinterpret_cast<const float4*>(g + off[k]);
  }
  const float4 ex = nd[0], ey = nd[1], ez = nd[2], xx = nd[3], xy = nd[4], xz = nd[5];
  const int4 ch = make_int4(__float_as_int(nd[6].x), __float_as_int(nd[6].y), __float_as_int(nd[6].z),
                            __float_as_int(nd[6].w));
  const float exa[4] = {ex.x, ex.y, ex.z, ex.w}, eya[4] = {ey.x, ey.y, ey.z, ey.w}, eza[4] = {ez.x, ez.y, ez.z, ez.w};
  const float xxa[4] = {xx.x, xx.y, xx.z, xx.w}, xya[4] = {xy.x, xy.y, xy.z, xy.w}, xza[4] = {xz.x, xz.y, xz.z, xz.w};
  uint32_t key[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float x0 = __builtin_fmaf(exa[k], ix, -oix), x1 = __builtin_fmaf(xxa[k], ix, -oix);
    const float y0 = __builtin_fmaf(eya[k], iy, -oiy), y1 = __builtin_fmaf(xya[k], iy, -oiy);
    const float z0 = __builtin_fmaf(eza[k], iz, -oiz), z1 = __builtin_fmaf(xza[k], iz, -oiz);
    const float tnear = fmaxf(fmaxf(x0, y0), fmaxf(z0, 0.0f));
    const float tfar = fminf(fminf(x1, y1), fminf(z1, tb));
    const uint32_t tk = (tnear <= tfar) ? (__float_as_uint(tnear) & 0x7FFFFFFCu) : 0x7F800000u;
    key[k] = tk | (uint32_t)k;
  }
#define PRT_KSORT(a, b)                                \
  {                                                    \
    const uint32_t lo_ = min(key[a], key[b]);          \
    key[b] = max(key[a], key[b]);                      \
    key[a] = lo_;                                      \
  }
  PRT_KSORT(0, 1) PRT_KSORT(2, 3) PRT_KSORT(0, 2) PRT_KSORT(1, 3) PRT_KSORT(1, 2)
#undef PRT_KSORT
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t slot = key[j] & 3u;
    tn[j] = key[j] >= 0x7F800000u ? __builtin_inff() : __uint_as_float(key[j] & ~3u);
    cc[j] = (slot & 2u) ? ((slot & 1u) ? ch.w : ch.z) : ((slot & 1u) ? ch.y : ch.x);
  }
}

// Pushes children 3, 2, 1 of a sorted node test (farthest first) that were
// hit.
template <class Stk>
__device__ __forceinline__ void push_children(Stk& S, int& sp, const float tn[4], const int32_t cc[4]) {
#pragma unroll
  for (int j = 3; j >= 1; --j)
    if (tn[j] != __builtin_inff()) S.push(sp++, cc[j], tn[j]);
}

// While-while traversal with speculative leaf postponement (Aila & Laine,
// "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009): a lane
// walks internal nodes until it meets a leaf, parks that leaf and keeps
// walking until every lane of the wave holds a parked leaf (or is done);
// then the whole wave tests triangles together.  Node visits and triangle
// tests thus run with most lanes active instead of interleaving a node
// step of some lanes with a leaf loop of others in every trip.
//
// intersect_scene_ (scene.cpp:83-214) tests the entities in index order and
// keeps a hit only when strictly nearer, so of hits at exactly the same t the
// one of the lowest entity index wins (a box standing on the ground mesh
// meets it in a shared plane).  Within one mesh the QBVH keeps the first
// triangle it finds; that order is the reference tree's, which this BVH does
// not reproduce.
__device__ __forceinline__ bool nearer_hit(double t, int32_t ent, double t_best, int32_t ent_best) {
  return t < t_best || (t == t_best && ent < ent_best);
}
// Entity of a hit id (>= 0 triangle slot, <= -2 analytic entity, -1 none:
// above every entity).
__device__ __forceinline__ int32_t hit_entity(const float4* tris4, int32_t id) {
  return id >= 0 ? __float_as_int(tris4[(size_t)id * 3].w) : (id <= -2 ? -2 - id : 0x7FFFFFFF);
}

// Every triangle a ray tests is one whose box the conservative culling
// leaves open (a box whose entry equals the best t stays open), so the
// closest hit is the one over all triangles, with the tie rule of
// nearer_hit.
#define PRT_TRAV_DONE 0x7FFFFFFF
template <bool ANY, class Stk>
__device__ __forceinline__ bool traverse_ww(const Params& P, d3 o, d3 d, double& t_best, int32_t& id_best, double& ub,
                                            double& vb, Stk& S, TravStats& ts) {
  if (P.n_nodes <= 0) return false;
  const float ox = (float)o.x, oy = (float)o.y, oz = (float)o.z;
  float ix = (float)(1.0 / d.x), iy = (float)(1.0 / d.y), iz = (float)(1.0 / d.z);
  if (__builtin_isinf(ix)) ix = __builtin_copysignf(1.0e30f, ix);
  if (__builtin_isinf(iy)) iy = __builtin_copysignf(1.0e30f, iy);
  if (__builtin_isinf(iz)) iz = __builtin_copysignf(1.0e30f, iz);
  const float oix = ox * ix, oiy = oy * iy, oiz = oz * iz;
  const uint32_t rows = entry_rows(ix, iy, iz);
  float tb = t_best < 1.0e30 ? fup(t_best) : __builtin_inff();
  bool found = false;
  int sp = 0;
  const float4* tris4 = reinterpret_cast<const float4*>(P.tris);
  int32_t e_best = ANY ? 0 : hit_entity(tris4, id_best);
  // cur: internal node (0 <= cur < PRT_TRAV_DONE), leaf code (< 0) or
  // PRT_TRAV_DONE; lf: parked leaf (< 0) or none (>= 0)
  int32_t cur = 0, lf = 0;
  // the nearest pending entry that can still hold a closer hit
  auto pop_next = [&]() -> int32_t {
    while (sp > 0) {
      int32_t node;
      float t;
      S.pop(--sp, node, t);
      if (t <= tb) return node;
    }
    return PRT_TRAV_DONE;
  };
  while (true) {
    while ((uint32_t)cur < (uint32_t)PRT_TRAV_DONE) {
      float tn[4];
      int32_t cc[4];
      node_children(S, P.nodes, cur, rows, ix, iy, iz, oix, oiy, oiz, tb, tn, cc);
      ts.nodes++;
      push_children(S, sp, tn, cc);
      cur = (tn[0] != __builtin_inff()) ? cc[0] : pop_next();
      if (cur < 0 && lf >= 0) {  // park the first leaf and keep walking
        lf = cur;
        cur = pop_next();
      }
      if (!__any(lf >= 0)) break;  // every lane still walking holds a leaf
    }
    while (lf < 0) {
      const uint32_t code = ~(uint32_t)lf;
      const uint32_t first = code & 0x07FFFFFFu;
      const uint32_t count = (code >> 27) + 1u;
      // triangle k + 1 is loaded before triangle k is tested
      const float4* tr0 = tris4 + (size_t)first * 3;
      float4 na = tr0[0], nb = tr0[1], nc = tr0[2];
      for (uint32_t k = 0; k < count; ++k) {
        const float4 a = na, b = nb, c = nc;
        if (k + 1 < count) {
          const float4* tr = tris4 + (size_t)(first + k + 1) * 3;
          na = tr[0];
          nb = tr[1];
          nc = tr[2];
        }
        ts.tris++;
        double t, u, v;
        if (mt_test(a, b, c, o, d, t, u, v) &&
            (ANY ? t < t_best : nearer_hit(t, __float_as_int(a.w), t_best, e_best))) {
          if (ANY) {
            id_best = (int32_t)(first + k);
            return true;
          }
          t_best = t;
          ub = u;
          vb = v;
          id_best = (int32_t)(first + k);
          e_best = __float_as_int(a.w);
          tb = fup(t);
          found = true;
        }
      }
      lf = cur;  // a second leaf met while walking is tested next
      if (cur < 0) cur = pop_next();
    }
    if (cur == PRT_TRAV_DONE) break;
  }
  return found;
}

// Closest (ANY = false) or any (ANY = true) hit over the BVH triangles with
// t in (1e-6, t_best).  o is the already-nudged origin.
template <bool ANY, class Stk>
__device__ __forceinline__ bool traverse(const Params& P, d3 o, d3 d, double& t_best, int32_t& id_best, double& ub,
                                         double& vb, Stk& S, TravStats& ts) {
  return traverse_ww<ANY>(P, o, d, t_best, id_best, ub, vb, S, ts);
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
    if (analytic_t(P.ents[ei], o, d, tt) && nearer_hit(tt, ei, t, id <= -2 ? -2 - id : 0x7FFFFFFF)) {
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

// ------------------------------------------------ chunk claims
// The chunks [0, n_chunks) of a queue are split into PRT_FETCH_COUNTERS
// sub-ranges (fewer when the grid has fewer waves: every sub-range needs a
// wave that drains it), each with its own claim counter on its own 128-byte
// line, so no counter is shared by the whole device.  A wave draws from the
// sub-range wave % n_sub.
struct ClaimRange {
  uint32_t* counter;
  uint32_t b, e;
};
#ifndef PRT_XCD_CLAIMS
#define PRT_XCD_CLAIMS 0
#endif
__device__ __forceinline__ ClaimRange claim_range(uint32_t* fetch, uint32_t n_chunks) {
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t n_sub = n_waves < (uint32_t)PRT_FETCH_COUNTERS ? n_waves : (uint32_t)PRT_FETCH_COUNTERS;
  uint32_t sub = wave % n_sub;
  if (PRT_XCD_CLAIMS && n_sub == (uint32_t)PRT_FETCH_COUNTERS && (gridDim.x & 7u) == 0) {
    // blocks go round-robin to the 8 XCDs: the waves of XCD x draw from
    // sub-ranges x * n_sub / 8 .. (x + 1) * n_sub / 8 - 1, one eighth of
    // the queue, so each XCD's L2 sees the rays of one part of the image
    const uint32_t per = n_sub / 8u, xcd = blockIdx.x & 7u;
    const uint32_t local = (blockIdx.x >> 3) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    sub = (uint32_t)__builtin_amdgcn_readfirstlane((int)(xcd * per + local % per));
  }
  ClaimRange r;
  r.counter = fetch + (size_t)sub * PRT_FETCH_STRIDE;
  r.b = (uint32_t)((uint64_t)sub * n_chunks / n_sub);
  r.e = (uint32_t)((uint64_t)(sub + 1) * n_chunks / n_sub);
  return r;
}
// Claims up to k chunks [c0, c1) of the wave's sub-range (one atomic);
// false once it is drained.  Every lane of the wave must call it.
__device__ __forceinline__ bool claim(const ClaimRange& r, uint32_t k, uint32_t& c0, uint32_t& c1) {
  uint32_t x = 0;
  if (lane_id() == 0) x = atomicAdd(r.counter, k);
  x = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)x, 0, 64));
  c0 = r.b + x;
  if (c0 >= r.e) return false;
  c1 = min(c0 + k, r.e);
  return true;
}

// ------------------------------------------------ queue traversal with refill
//
// Closest or any hit for the rays of a chunked queue (chunk c holds desc[c].y
// rays from queue index desc[c].x on), the while-while loop of traverse_ww
// with lanes refilled as their rays finish.  A wave claims PRT_FETCH_RAYS /
// 64 chunks per atomic from its claim sub-range (claim_range), keeps the
// claimed entries not yet started as a pool, and hands
// them to idle lanes at the top of each round once at least PRT_REFILL_MIN
// lanes are idle.  Without refill a wave lives as long as its slowest ray
// and the other lanes idle meanwhile.  Each ray's own traversal is
// traverse_ww's, so its hit is the same.
//
// Load(j, o, d, tmax): ray j's origin (before the 1e-6 nudge for a closest
// hit, already nudged for an any hit), direction and t bound.
// Done(j, t, id, blocked): the finished ray (closest hit t / id as
// closest_hit; any hit: blocked).
#ifndef PRT_REFILL_MIN
#define PRT_REFILL_MIN 16
#endif
#ifndef PRT_FETCH_RAYS
#define PRT_FETCH_RAYS 128
#endif
template <bool ANY, class Stk, class Load, class Done>
__device__ __forceinline__ void traverse_pool(const Params& P, uint32_t n_chunks, const uint2* __restrict__ desc,
                                              uint32_t* fetch, Stk& S, TravStats& ts, Load&& load, Done&& done) {
  const float4* tris4 = reinterpret_cast<const float4*>(P.tris);
  const uint32_t lane = lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  constexpr uint32_t kClaim = (uint32_t)PRT_FETCH_RAYS / 64u;  // chunks per claim
  const ClaimRange cr = claim_range(fetch, n_chunks);
  d3 o = mk(0.0, 0.0, 0.0), d = o;
  float ix = 0.0f, iy = 0.0f, iz = 0.0f, oix = 0.0f, oiy = 0.0f, oiz = 0.0f, tb = 0.0f;
  uint32_t rows = 0;
  double t_best = 0.0;
  int32_t id_best = -1, e_best = 0x7FFFFFFF, cur = PRT_TRAV_DONE, lf = 0;
  int sp = 0;
  uint32_t ri = 0;
  bool act = false, blocked = false;
  // wave-uniform: claimed entries pool_base .. pool_base + pool_n - 1 (entry
  // g = chunk * 64 + lane slot) not yet started; more: the sub-range may hold
  // unclaimed chunks
  uint32_t pool_base = 0, pool_n = 0;
  bool more = cr.b < cr.e;
  auto pop_next = [&]() -> int32_t {
    while (sp > 0) {
      int32_t node;
      float t;
      S.pop(--sp, node, t);
      if (t <= tb) return node;
    }
    return PRT_TRAV_DONE;
  };
  while (true) {
    // ---- refill idle lanes
    const uint64_t idle = __ballot(!act);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= (uint32_t)PRT_REFILL_MIN && (pool_n > 0 || more)) {
      const uint32_t rank = (uint32_t)__popcll(idle & below);
      const uint32_t take = n_idle < pool_n ? n_idle : pool_n;
      int64_t g = -1;
      if (!act && rank < take) g = pool_base + rank;
      pool_base += take;
      pool_n -= take;
      if (n_idle > take && more) {
        uint32_t c0, c1;
        if (!claim(cr, kClaim, c0, c1)) {
          more = false;
        } else {
          const uint32_t ent = (c1 - c0) * 64u;
          const uint32_t need = n_idle - take, used = need < ent ? need : ent;
          if (!act && rank >= take && rank - take < ent) g = c0 * 64u + (rank - take);
          pool_base = c0 * 64u + used;
          pool_n = ent - used;
          if (c1 >= cr.e) more = false;
        }
      }
      if (g >= 0) {
        const uint2 dc = desc[(uint32_t)g >> 6];
        if (((uint32_t)g & 63u) < dc.y) {
          // start ray ri: analytic entities first, as closest_hit / occluded
          ri = dc.x + ((uint32_t)g & 63u);
          act = true;
          blocked = false;
          d3 oo;
          double tmax;
          load(ri, oo, d, tmax);
          o = ANY ? oo : oo + d * 0.000001;
          t_best = tmax;
          id_best = -1;
          e_best = 0x7FFFFFFF;
          for (int k = 0; k < P.n_analytic; ++k) {
            const int32_t ei = P.analytic[k];
            double tt;
            if (analytic_t(P.ents[ei], o, d, tt) && (ANY ? tt < t_best : nearer_hit(tt, ei, t_best, e_best))) {
              if (ANY) {
                blocked = true;
                break;
              }
              t_best = tt;
              id_best = -2 - ei;
              e_best = ei;
            }
          }
          lf = 0;
          sp = 0;
          cur = PRT_TRAV_DONE;
          if (!blocked && P.n_nodes > 0) {
            ix = (float)(1.0 / d.x);
            iy = (float)(1.0 / d.y);
            iz = (float)(1.0 / d.z);
            if (__builtin_isinf(ix)) ix = __builtin_copysignf(1.0e30f, ix);
            if (__builtin_isinf(iy)) iy = __builtin_copysignf(1.0e30f, iy);
            if (__builtin_isinf(iz)) iz = __builtin_copysignf(1.0e30f, iz);
            oix = (float)o.x * ix;
            oiy = (float)o.y * iy;
            oiz = (float)o.z * iz;
            rows = entry_rows(ix, iy, iz);
            tb = t_best < 1.0e30 ? fup(t_best) : __builtin_inff();
            cur = 0;
          }
        }
      }
    }
    if (!__any(act)) {
      if (pool_n == 0 && !more) break;  // nothing running and nothing left to start
      continue;                         // the claimed entries were padding
    }
    // ---- walk internal nodes until every walking lane holds a leaf
    while ((uint32_t)cur < (uint32_t)PRT_TRAV_DONE) {
      float tn[4];
      int32_t cc[4];
      node_children(S, P.nodes, cur, rows, ix, iy, iz, oix, oiy, oiz, tb, tn, cc);
      ts.nodes++;
      push_children(S, sp, tn, cc);
      cur = (tn[0] != __builtin_inff()) ? cc[0] : pop_next();
      if (cur < 0 && lf >= 0) {
        lf = cur;
        cur = pop_next();
      }
      if (!__any(lf >= 0)) break;
    }
    // ---- test the parked leaves (triangle k + 1 is loaded before triangle k
    // is tested, so a leaf costs one memory round trip rather than one per
    // triangle)
    while (lf < 0) {
      const uint32_t code = ~(uint32_t)lf;
      const uint32_t first = code & 0x07FFFFFFu;
      const uint32_t count = (code >> 27) + 1u;
      const float4* tr0 = tris4 + (size_t)first * 3;
      float4 na = tr0[0], nb = tr0[1], nc = tr0[2];
      for (uint32_t k = 0; k < count; ++k) {
        const float4 a = na, b = nb, c = nc;
        if (k + 1 < count) {
          const float4* tr = tris4 + (size_t)(first + k + 1) * 3;
          na = tr[0];
          nb = tr[1];
          nc = tr[2];
        }
        ts.tris++;
        double t, u, v;
        if (mt_test(a, b, c, o, d, t, u, v) &&
            (ANY ? t < t_best : nearer_hit(t, __float_as_int(a.w), t_best, e_best))) {
          if (ANY) {
            blocked = true;
            break;
          }
          t_best = t;
          id_best = (int32_t)(first + k);
          e_best = __float_as_int(a.w);
          tb = fup(t);
        }
      }
      if (ANY && blocked) {
        lf = 0;
        cur = PRT_TRAV_DONE;
      } else {
        lf = cur;
        if (cur < 0) cur = pop_next();
      }
    }
    // ---- hand over finished rays
    if (act && cur == PRT_TRAV_DONE && lf >= 0) {
      done(ri, t_best, id_best, blocked);
      act = false;
      ts.rays++;
    }
  }
}

__device__ __forceinline__ void flush_stats(uint64_t* ctr, int idx_nodes, TravStats ts, int idx_rays = -1) {
  const uint64_t n = wave_sum_u32(ts.nodes), t = wave_sum_u32(ts.tris);
  const uint64_t r = idx_rays >= 0 ? wave_sum_u32(ts.rays) : 0;
  if (lane_id() == 0) {
    if (n) atomicAdd((unsigned long long*)&ctr[idx_nodes], (unsigned long long)n);
    if (t) atomicAdd((unsigned long long*)&ctr[idx_nodes + 1], (unsigned long long)t);
    if (r) atomicAdd((unsigned long long*)&ctr[idx_rays], (unsigned long long)r);
  }
}
