// prt_math.h — device math shared by the kernels: double vectors with the
// reference's operation order, Philox4x32-10, and the primitive tests.
//
// Everything here is evaluated without FMA contraction (the library is built
// with -ffp-contract=off), so Moller-Trumbore (polygon.h:78-129),
// Sphere::intersect (sphere.h:68-96) and UVPlane::intersect
// (plane.h:86-110, 220-240) return the reference's double results bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"

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

struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 mk(double a, double b, double c) {
  d3 r;
  r.x = a; r.y = b; r.z = c;
  return r;
}
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
// normalize = v * (1 / length) (vector3d.h:162-165)
__device__ __forceinline__ d3 normalize(d3 a) { return a * (1.0 / sqrt(sqr(a))); }
__device__ __forceinline__ d3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }

// Philox4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53 / 0xCD9E8D57,
// key increments 0x9E3779B9 / 0xBB67AE85.
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

// Per-vertex random-number counter: (pixel, sample, depth | branch << 8, block)
__device__ __forceinline__ uint32_t rng_c2(int depth, uint32_t branch) { return ((uint32_t)depth & 0xFFu) | (branch << 8); }

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
  const double inv_det = 1.0 / det;
  t = dot(edge2, q) * inv_det;
  if (!(t > PS_EPS)) return false;
  u *= inv_det;
  v *= inv_det;
  return true;
}

// sin(2 pi u) and cos(2 pi u) for u in [0, 1].  The quadrant q = rint(4u) is
// split off exactly (u is a multiple of 2^-32, so u - q/4 is exact), leaving
// y = 2 pi (u - q/4) in [-pi/4, pi/4], where the Taylor series cut after
// y^15 / 15! (sine) and y^16 / 16! (cosine) are accurate to below half an
// ulp; the quadrant then rotates (sin y, cos y).  Differs from sincos(2 pi u)
// by a few ulps at most, with no large-argument reduction path (which holds
// most of a general sincos's registers).
__device__ __forceinline__ void sincos_2pi(double u, double& s, double& c) {
  const double q = rint(4.0 * u);
  const double y = PS_TWOPI * (u - 0.25 * q);
  const double z = y * y;
  const double sy =
      y + y * z *
              (-0.16666666666666666 +
               z * (0.008333333333333333 +
                    z * (-0.0001984126984126984 +
                         z * (2.7557319223985893e-06 +
                              z * (-2.505210838544172e-08 + z * (1.6059043836821613e-10 + z * -7.647163731819816e-13))))));
  const double cy =
      1.0 + z * (-0.5 +
                 z * (0.041666666666666664 +
                      z * (-0.001388888888888889 +
                           z * (2.48015873015873e-05 +
                                z * (-2.755731922398589e-07 +
                                     z * (2.08767569878681e-09 +
                                          z * (-1.1470745597729725e-11 + z * 4.779477332387385e-14)))))));
  const int k = (int)q & 3;
  s = (k == 0) ? sy : (k == 1) ? cy : (k == 2) ? -sy : -cy;
  c = (k == 0) ? cy : (k == 1) ? -sy : (k == 2) ? -cy : sy;
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

// ------------------------------------------------------------ sample accumulator
// Per-sample radiance is summed in 32.32 fixed point with 64-bit integer
// atomics: the sum is exact and independent of the order in which the
// shade and shadow kernels add to a sample, so renders are bit-reproducible
// (which the finite-difference gradient checks rely on).
#define PRT_FIX_ONE 4294967296.0
#define PRT_FIX_INV 2.3283064365386963e-10

__device__ __forceinline__ void accum_add(uint64_t* a, double v) {
  if (v > 0.0) atomicAdd((unsigned long long*)a, (unsigned long long)(v * PRT_FIX_ONE + 0.5));
}
__device__ __forceinline__ double accum_get(const uint64_t* a) { return (double)(*a) * PRT_FIX_INV; }

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

__device__ __forceinline__ uint64_t wave_sum_u32(uint32_t x) {
  uint32_t s = x;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += (uint32_t)__shfl_xor((int)s, off, 64);
  return (uint64_t)s;
}

// Exclusive prefix sum of n over the (converged) wave; one atomic per wave
// reserves the wave's range in *counter.  Every lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_append_n(uint32_t* counter, uint32_t n) {
  const uint32_t lane = lane_id();
  uint32_t incl = n;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
  uint32_t base = 0;
  if (total > 0) {
    if (lane == 63) base = atomicAdd(counter, total);
    base = (uint32_t)__shfl((int)base, 63, 64);
  }
  return base + incl - n;
}

// Append one entry per lane with v set (the wave converged): one atomic per
// wave, rank among the set lanes.
__device__ __forceinline__ uint32_t wave_append_bit(uint32_t* counter, bool v) {
  const uint64_t m = __ballot(v);
  const uint32_t total = (uint32_t)__popcll(m);
  uint32_t base = 0;
  if (total > 0) {
    if (lane_id() == 0) base = atomicAdd(counter, total);
    base = (uint32_t)__shfl((int)base, 0, 64);
  }
  return base + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
}

// Static contiguous range [c0, c1) of n_chunks chunks for this wave of the
// grid, for kernels whose work per chunk is too short to pay for a fetch
// atomic.
__device__ __forceinline__ void wave_range(uint32_t n_chunks, uint32_t& c0, uint32_t& c1) {
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t per = (n_chunks + n_waves - 1) / n_waves;
  c0 = min(wave * per, n_chunks);
  c1 = min(c0 + per, n_chunks);
}

// Append from divergent code: the active lanes share one atomic (ballot,
// rank = popcount of the lower active lanes, the leader adds the total and
// broadcasts the base).
__device__ __forceinline__ uint32_t agg_append(uint32_t* counter) {
  const uint64_t mask = __ballot(1);
  const uint32_t lane = lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
  const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, (int)leader, 64);
  return base + rank;
}
