// prt_shade.h — one path vertex and its next-event estimation, shared by the
// per-scene wavefront kernels (prt_kernels.hip) and the multi-scene streaming
// kernels (prt_stream.hip).
//
// The light transport follows radiance() (radiance.h:25-556) and the BRDFs
// (brdf.h:408-953) turned from recursion into forward throughput: a vertex
// returns Env + E + direct + W * L(next), so each path segment carries T, adds
// T*(Env + E) and the NEE term, and its child carries T*W.  The "pickup" ray
// of mirror/glass vertices (direct_light_perfect_refrection,
// directLight_radiance.h:516-556) is the same ray as the child, so the child's
// hit supplies it: the child records a pickup weight and adds weight * emission
// when it lands on a light.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"
#include "prt_math.h"

// Parameters of scene s (s wave-uniform) through a constant-address-space
// pointer, so that their loads are scalar loads as kernel arguments' are.
__device__ __forceinline__ const Params& scene_params(const Params* scenes, int s) {
  const uint64_t a = (uint64_t)(scenes + s);
  const uint64_t au = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
                      (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
  const __attribute__((address_space(4))) Params* Pc = (const __attribute__((address_space(4))) Params*)au;
  return *(const Params*)Pc;
}

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

// Random words 0..3 of a path vertex: Russian roulette, BSDF choice, the two
// direction numbers (counter (pixel, sample, depth | branch << 8, 0)).  In a
// Metropolis render the first mlt_depths vertices read them from the chain's
// primary samples instead (the chain is the batch "pixel"); the reference's
// primary samples are exactly these next01() draws (radiance.h:108, 517,
// 534; brdf.h:511-512, 866), while light sampling draws Next01() from an
// independent stream (directLight_radiance.h:343-359).
__device__ __forceinline__ void vertex_words(const Params& P, uint32_t pixel, uint32_t sample, int depth,
                                             uint32_t branch, uint32_t b[4]) {
  if (P.mlt_u != nullptr && depth < P.mlt_depths) {
    const uint4 v = *reinterpret_cast<const uint4*>(P.mlt_u + (size_t)pixel * (uint32_t)P.mlt_stride +
                                                    MLT_CAMERA_WORDS + 4 * depth);
    b[0] = v.x; b[1] = v.y; b[2] = v.z; b[3] = v.w;
    return;
  }
  philox(pixel, sample, rng_c2(depth, branch), 0u, P.key0, P.key1, b);
}

// Children go to qout positions off + j, j < cap, with j from *count_out.
struct ShadeCtx {
  RayQ qout;
  uint32_t* count_out;
  uint32_t off, cap;
  uint32_t slot;
  uint32_t truncated, n_ext, pickups, dropped;
  uint32_t direct;  // set when a camera ray's diffuse hit on a light is returned with NEE on
                    // (BRDF_Diffuse's *hit_direct_light = 1, brdf.h:443-452)
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
  uint32_t j = agg_append(C.count_out);
  if (j >= C.cap) {
    C.dropped++;
    return;
  }
  j += C.off;
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

// The children of one vertex in the persistent renderer: the first becomes
// the lane's next segment (registers), a second goes to the lane's pending
// stack.  Throughput, pickup weight and NEE normal are rounded to float as a
// queue entry stores them (emit_child above), so a segment continues
// bit-identically to the queued renderers.
struct LaneCtx {
  d3 o0, d0;
  float t0[3], p0[3], n0[3];
  uint32_t meta0;
  bool has0;
  PendRec* pend;
  int npend;
  uint32_t truncated, n_ext, pickups, dropped;
  uint32_t direct;
};

__device__ __forceinline__ void emit_child(LaneCtx& C, d3 o, d3 d, d3 ct, d3 pw, uint32_t meta, d3 nrm) {
  const float t[3] = {(float)ct.x, (float)ct.y, (float)ct.z};
  const float p[3] = {(float)pw.x, (float)pw.y, (float)pw.z};
  const float n[3] = {(float)nrm.x, (float)nrm.y, (float)nrm.z};
  if (!C.has0) {
    C.has0 = true;
    C.o0 = o;
    C.d0 = d;
    for (int k = 0; k < 3; ++k) {
      C.t0[k] = t[k];
      C.p0[k] = p[k];
      C.n0[k] = n[k];
    }
    C.meta0 = meta;
  } else if (C.npend < PRT_PEND_DEPTH) {
    PendRec& r = C.pend[C.npend++];
    r.o[0] = o.x; r.o[1] = o.y; r.o[2] = o.z;
    r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z;
    for (int k = 0; k < 3; ++k) {
      r.t[k] = t[k];
      r.p[k] = p[k];
      r.n[k] = n[k];
    }
    r.meta = meta;
  } else {
    C.dropped++;
    return;
  }
  C.n_ext++;
}

// One path vertex: the radiance() body for the hit (or miss) of ray (o, d).
// Returns the radiance added to the camera sample; child rays go to the
// context's sink (an output queue, or a lane's own segments), and a diffuse
// vertex marks its child for next-event estimation.
// mats: the material table (P.mats, or a block's LDS copy of it).
template <class Ctx>
__device__ __forceinline__ d3 shade_vertex(const Params& P, const DevMaterial* mats, Ctx& C, d3 o, d3 d,
                                           double t, int32_t id, d3 T, d3 Pw, uint32_t pixel, uint32_t sample,
                                           int depth, uint32_t branch, VertexAdj& va) {
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
  // the hit entity's material and flags: a triangle carries them in its
  // record (TRI_SHADE_*), so its material is read without the entity record
  uint32_t mat;
  bool ninv_neg, is_light, has_alb, uvplane;
  if (id >= 0) {
    const float4* tr = reinterpret_cast<const float4*>(P.tris) + (size_t)id * 3;
    ei = __float_as_int(tr[0].w);
    va.face = __float_as_int(tr[1].w);
    const uint32_t sw = __float_as_uint(tr[2].w);
    mat = sw & TRI_SHADE_MAT_MASK;
    ninv_neg = (sw & TRI_SHADE_NINV) != 0;
    is_light = (sw & TRI_SHADE_LIGHT) != 0;
    has_alb = (sw & TRI_SHADE_ALBEDO) != 0;
    uvplane = false;
    x = o2 + d * t;  // polygon.h:170
    n = mk(P.tri_nrm[3 * (size_t)id], P.tri_nrm[3 * (size_t)id + 1], P.tri_nrm[3 * (size_t)id + 2]);
  } else {
    ei = -2 - id;
    const DevEntity& e = P.ents[ei];
    mat = (uint32_t)e.material;
    ninv_neg = e.ninv < 0;
    is_light = e.light != 0;
    has_alb = false;
    uvplane = e.type == ENT_UVPLANE;
    if (e.type == ENT_SPHERE) {
      x = o2 + t * d;  // sphere.h:89-95
      n = normalize(x - ld3(e.p0));
    } else {
      x = d * t + o2;
      n = ld3(e.normal);
    }
  }
  va.entity = ei;
  if (ninv_neg) n = n * -1.0;
  const DevMaterial& m = mats[mat];
  d3 color;
  if (has_alb)
    color = mk(P.tri_alb[3 * (size_t)id], P.tri_alb[3 * (size_t)id + 1], P.tri_alb[3 * (size_t)id + 2]);
  else
    color = ld3(m.color);
  const d3 E = ld3(m.emission);
  // pickup of the parent's mirror / glass vertex (directLight_radiance.h:516-556)
  if (is_light && (Pw.x != 0.0 || Pw.y != 0.0 || Pw.z != 0.0)) {
    add = add + vmul(Pw, E);
    va.pre = vmul(Pw, E);
    va.add_e = Pw;
    C.pickups++;
  }
  const d3 n_orient = dot(n, d) < 0.0 ? n : (-1.0 * n);
  // RGB_MAX = std::max(x, std::max(y, z)) (def.h:19); RR (radiance.h:77-126)
  d3 color_rr = color;
  if (P.tri_alb_rr && has_alb)
    color_rr = mk(P.tri_alb_rr[3 * (size_t)id], P.tri_alb_rr[3 * (size_t)id + 1], P.tri_alb_rr[3 * (size_t)id + 2]);
  const d3 rw = E + color_rr;
  const double myz = (rw.y < rw.z) ? rw.z : rw.y;
  double rr = (rw.x < myz) ? myz : rw.x;
  if (depth > P.depth_limit) rr *= ldexp(1.0, -(depth - P.depth_limit));  // pow(0.5, k), exact
  uint32_t b0[4];
  const uint32_t c2 = rng_c2(depth, branch);
  vertex_words(P, pixel, sample, depth, branch, b0);
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
    if (is_light && uvplane && dot(n, d) >= 0.0) return add;
    if (is_light) {
      if (P.nee && depth == 0) C.direct = 1u;
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
    // r1 = 2 pi u (brdf.h:498); sin r1, cos r1 without a general reduction
    const double r2 = u01(b0[3]), r2s = sqrt(r2);
    double s1, c1;
    sincos_2pi(u01(b0[2]), s1, c1);
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
  const double c_2 = c * c;
  const double Re = R0 + (1.0 - R0) * (c_2 * c_2 * c);  // pow(c, 5.0) to an ulp or two
  const double q = into ? nc / nt : nt / nc;
  const double nnt2 = q * q;  // pow(q, 2.0), exact
  const double Tr = (1.0 - Re) * nnt2;
  const double probability = 0.25 + 0.5 * Re;
  if (depth > 2) {
    // dimension 13; a Metropolis proposal uses the vertex's BSDF-choice word,
    // unused by glass, so that the choice is a primary sample
    uint32_t wg = b0[1];
    if (P.mlt_u == nullptr) {
      uint32_t b3[4];
      philox(pixel, sample, c2, 3u, P.key0, P.key1, b3);
      wg = b3[1];
    }
    const bool take_refl = u01(wg) < probability;
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

// ------------------------------------------------------------ NEE
struct ShadowRay {
  d3 o2, dir;
  double tmax;
  float c[3];
};

// direct_radiance_sample + eval_direct_radiance (diffuse branch),
// directLight_radiance.h:19-387, for one light sample.  Produces a shadow ray
// only when the light's own first hit along it lies within 1e-3 of the
// sampled distance (the reference's acceptance test, line 141); the shadow
// kernel then checks that nothing nearer blocks it.
// le: the light entity, E: its material's emission.
__device__ __forceinline__ bool nee_sample_e(const DevEntity& le, const d3 E, double ra, double rb, d3 x,
                                             d3 n_orient, d3 ct, ShadowRay& s) {
  d3 light_pos, light_normal;
  if (le.type == ENT_SPHERE) {
    // r1 = 2 pi ra
    const double r2 = 1.0 - 2.0 * rb;
    const double s22 = sqrt(1.0 - r2 * r2);
    double sn, cs;
    sincos_2pi(ra, sn, cs);
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

__device__ __forceinline__ bool nee_sample(const Params& P, int li, double ra, double rb, d3 x, d3 n_orient, d3 ct,
                                           ShadowRay& s) {
  const DevEntity& le = P.ents[li];
  return nee_sample_e(le, ld3(P.mats[le.material].emission), ra, rb, x, n_orient, ct, s);
}

__device__ __forceinline__ double uniform_f64(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// The entity and emission of light entity li (the same for every lane), read
// into wave-uniform registers: with a single light, NEE samples use these
// instead of re-reading the entity record per sample.
struct UniformLight {
  DevEntity e;
  d3 E;
  int32_t li;
};
__device__ __forceinline__ UniformLight uniform_light(const Params& P, int li) {
  UniformLight u;
  u.li = __builtin_amdgcn_readfirstlane(li);
  const DevEntity& g = P.ents[u.li];
  u.e.type = __builtin_amdgcn_readfirstlane(g.type);
  u.e.material = __builtin_amdgcn_readfirstlane(g.material);
  u.e.ninv = __builtin_amdgcn_readfirstlane(g.ninv);
  u.e.light = 0;
  u.e.has_albedo = 0;
  u.e.pad_ = 0;
  for (int k = 0; k < 3; ++k) {
    u.e.p0[k] = uniform_f64(g.p0[k]);
    u.e.u_axis[k] = uniform_f64(g.u_axis[k]);
    u.e.v_axis[k] = uniform_f64(g.v_axis[k]);
    u.e.normal[k] = uniform_f64(g.normal[k]);
  }
  u.e.radius = uniform_f64(g.radius);
  u.e.u_len = uniform_f64(g.u_len);
  u.e.v_len = uniform_f64(g.v_len);
  u.e.area = uniform_f64(g.area);
  const DevMaterial& m = P.mats[u.e.material];
  u.E = mk(uniform_f64(m.emission[0]), uniform_f64(m.emission[1]), uniform_f64(m.emission[2]));
  return u;
}

__device__ __forceinline__ void put_shadow(ShadowQ& q, uint32_t j, const ShadowRay& s, uint32_t slot, int32_t light) {
  q.light[j] = light;
  q.ox[j] = s.o2.x; q.oy[j] = s.o2.y; q.oz[j] = s.o2.z;
  q.dx[j] = s.dir.x; q.dy[j] = s.dir.y; q.dz[j] = s.dir.z;
  q.tmax[j] = s.tmax;
  q.cr[j] = s.c[0]; q.cg[j] = s.c[1]; q.cb[j] = s.c[2];
  q.slot[j] = slot;
}

// Random words of the three light samples of a diffuse vertex (dimensions
// 4..12, drawn with the parent vertex's counter): sample m picks its light
// with w[3m] and its point with w[3m+1], w[3m+2].
__device__ __forceinline__ void nee_words(const Params& P, uint32_t pixel, uint32_t sample, int depth,
                                          uint32_t branch, uint32_t w[9]) {
  const uint32_t c2 = rng_c2(depth, branch);
  uint32_t b1[4], b2[4], b3[4];
  philox(pixel, sample, c2, 1u, P.key0, P.key1, b1);
  philox(pixel, sample, c2, 2u, P.key0, P.key1, b2);
  philox(pixel, sample, c2, 3u, P.key0, P.key1, b3);
  w[0] = b1[0]; w[1] = b1[1]; w[2] = b1[2]; w[3] = b1[3];
  w[4] = b2[0]; w[5] = b2[1]; w[6] = b2[2]; w[7] = b2[3];
  w[8] = b3[0];
}

// Light sample m (0..2) of a diffuse vertex at x (orienting normal nrm,
// throughput ct = T*colour/rr): LightList::randomLight (light.h:146-151),
// index (int)(u*(N-1)+0.5) with probability 1, then nee_sample.  l receives
// the light's entity index.
__device__ __forceinline__ bool nee_one(const Params& P, const uint32_t w[9], int m, d3 x, d3 nrm, d3 ct,
                                        ShadowRay& s, int32_t& l) {
  const int NL = P.n_lights;
  // with one light the index is (int)(u * 0 + 0.5) = 0 for every u
  int i = (int)(u01(w[3 * m]) * (double)(NL - 1) + 0.5);
  i = i >= NL ? NL - 1 : i;
  l = P.lights[i];
  return nee_sample(P, l, u01(w[3 * m + 1]), u01(w[3 * m + 2]), x, nrm, ct, s);
}
// Light tables of the NEE kernels: up to NEE_LDS_LIGHTS lights of a scene
// (entity record, emission, entity index) copied into a wave's LDS slots, so
// that a light sample reads its light with one LDS access instead of three
// dependent global loads (light list, entity record, material).
enum NeeMode { NEE_GLOBAL = 0, NEE_ONE = 1, NEE_LDS = 2 };
#define NEE_LDS_LIGHTS 16
struct LightSlot {
  DevEntity e;
  d3 E;
  int32_t li, pad_;
};
// Lanes < n_lights of the calling wave (which must all be active) copy
// light `lane` of P; the slots are readable by the whole wave afterwards.
__device__ __forceinline__ void load_light_slots(const Params& P, LightSlot* slots) {
  const int k = (int)lane_id();
  if (k < P.n_lights) {
    const int32_t li = P.lights[k];
    const DevEntity e = P.ents[li];
    const DevMaterial& m = P.mats[e.material];
    slots[k].e = e;
    slots[k].E = ld3(m.emission);
    slots[k].li = li;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// nee_one with the scene's lights in LDS slots.
__device__ __forceinline__ bool nee_one_slots(const Params& P, const LightSlot* slots, const uint32_t w[9], int m,
                                              d3 x, d3 nrm, d3 ct, ShadowRay& s, int32_t& l) {
  const int NL = P.n_lights;
  int i = (int)(u01(w[3 * m]) * (double)(NL - 1) + 0.5);
  i = i >= NL ? NL - 1 : i;
  const LightSlot& L = slots[i];
  l = L.li;
  return nee_sample_e(L.e, L.E, u01(w[3 * m + 1]), u01(w[3 * m + 2]), x, nrm, ct, s);
}
// nee_one for a scene with exactly one light, whose entity U holds.
__device__ __forceinline__ bool nee_one_single(const UniformLight& U, const uint32_t w[9], int m, d3 x, d3 nrm,
                                               d3 ct, ShadowRay& s, int32_t& l) {
  l = U.li;
  return nee_sample_e(U.e, U.E, u01(w[3 * m + 1]), u01(w[3 * m + 2]), x, nrm, ct, s);
}

// Luminance of mlt.h:171-174 with the Vector3d dot pairing.
__device__ __forceinline__ double luminance(double r, double g, double b) {
  return (0.2126 * r + 0.0722 * b) + 0.7152 * g;
}

// Folds the spp samples of one pixel (a[3*k..3*k+2], fixed point): v_k = L_k *
// smp, luminance clamp (render.h:1375-1384), image += v_k, *= sensor_response
// (render.h:1405-1409), and the variance of the pixel mean.  The sums run in
// sample order; the accumulators are read 8 samples at a time, so that a
// pixel's thread waits for one memory round trip per 8 samples instead of one
// per sample.
__device__ __forceinline__ void fold_pixel(const Params& P, const uint64_t* a, double rgb[3], double var[3]) {
  double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  auto add = [&](uint64_t w0, uint64_t w1, uint64_t w2) {
    double v0 = ((double)w0 * PRT_FIX_INV) * P.smp, v1 = ((double)w1 * PRT_FIX_INV) * P.smp,
           v2 = ((double)w2 * PRT_FIX_INV) * P.smp;
    if (P.cutoff > 0.0f) {
      const float lv = (float)luminance(v0, v1, v2);
      if (lv > P.cutoff) {
        const float f = P.cutoff / lv;
        v0 *= f; v1 *= f; v2 *= f;
      }
    }
    s1[0] += v0; s1[1] += v1; s1[2] += v2;
    s2[0] += v0 * v0; s2[1] += v1 * v1; s2[2] += v2 * v2;
  };
  constexpr int kB = 8;
  int k = 0;
  for (; k + kB <= P.spp; k += kB) {
    uint64_t w[3 * kB];
#pragma unroll
    for (int j = 0; j < 3 * kB; ++j) w[j] = a[3 * k + j];
#pragma unroll
    for (int j = 0; j < kB; ++j) add(w[3 * j], w[3 * j + 1], w[3 * j + 2]);
  }
  for (; k < P.spp; ++k) add(a[3 * k], a[3 * k + 1], a[3 * k + 2]);
  const double sr = P.sensor_response;
  const double nn = (double)P.spp;
  for (int c = 0; c < 3; ++c) {
    rgb[c] = s1[c] * sr;
    var[c] = nn > 1 ? fmax(0.0, (nn * s2[c] - s1[c] * s1[c]) / (nn - 1)) * sr * sr : 0.0;
  }
}

// One channel of fold_pixel: the same values in the same order for channel
// ch (0..2), so that three threads can fold one pixel.  With a luminance
// cutoff every channel's clamp needs all three, otherwise only channel ch
// is read.
__device__ __forceinline__ void fold_channel(const Params& P, const uint64_t* a, int ch, double& rgb, double& var) {
  double s1 = 0, s2 = 0;
  constexpr int kB = 16;
  int k = 0;
  if (P.cutoff > 0.0f) {
    auto add = [&](uint64_t w0, uint64_t w1, uint64_t w2) {
      double v0 = ((double)w0 * PRT_FIX_INV) * P.smp, v1 = ((double)w1 * PRT_FIX_INV) * P.smp,
             v2 = ((double)w2 * PRT_FIX_INV) * P.smp;
      const float lv = (float)luminance(v0, v1, v2);
      if (lv > P.cutoff) {
        const float f = P.cutoff / lv;
        v0 *= f; v1 *= f; v2 *= f;
      }
      const double v = ch == 0 ? v0 : (ch == 1 ? v1 : v2);
      s1 += v;
      s2 += v * v;
    };
    for (; k + kB / 2 <= P.spp; k += kB / 2) {
      uint64_t w[3 * (kB / 2)];
#pragma unroll
      for (int j = 0; j < 3 * (kB / 2); ++j) w[j] = a[3 * k + j];
#pragma unroll
      for (int j = 0; j < kB / 2; ++j) add(w[3 * j], w[3 * j + 1], w[3 * j + 2]);
    }
    for (; k < P.spp; ++k) add(a[3 * k], a[3 * k + 1], a[3 * k + 2]);
  } else {
    auto add = [&](uint64_t w) {
      const double v = ((double)w * PRT_FIX_INV) * P.smp;
      s1 += v;
      s2 += v * v;
    };
    const uint64_t* ac = a + ch;
    for (; k + kB <= P.spp; k += kB) {
      uint64_t w[kB];
#pragma unroll
      for (int j = 0; j < kB; ++j) w[j] = ac[3 * (k + j)];
#pragma unroll
      for (int j = 0; j < kB; ++j) add(w[j]);
    }
    for (; k < P.spp; ++k) add(ac[3 * k]);
  }
  const double sr = P.sensor_response;
  const double nn = (double)P.spp;
  rgb = s1 * sr;
  var = nn > 1 ? fmax(0.0, (nn * s2 - s1 * s1) / (nn - 1)) * sr * sr : 0.0;
}
