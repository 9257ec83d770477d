// prt_oracle.cpp — FP64 CPU restatement of prender's path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path in
// prender_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only to compare against or to time as the CPU baseline.
// No product code path may call it.
//
// It restates, in double precision and with the reference's operation order,
// Render::Run (proj/render.h:1125-1455), CameraScreenPram (render.h:70-158),
// radiance (proj/radiance.h:25-556), BRDF_Diffuse / BRDF_Specular /
// BRDF_Refraction (proj/brdf.h:408-953), eval_direct_radiance /
// direct_radiance_sample / direct_light_perfect_refrection
// (proj/directLight_radiance.h:19-387, 516-556), LightList
// (proj/light.h:95-155), intersect_scene_ (proj/scene.cpp:83-214),
// Sphere::intersect (proj/sphere.h:68-103), Plane/UVPlane::intersect
// (proj/plane.h:86-110, 220-373), Triangle::CheckIntersection2
// (proj/polygon.h:78-129), the SAH BVH -> QBVH build and traversal
// (proj/BVH.cpp:119-348, proj/QBVH.cpp:17-323, proj/BoundingBox.h:137-178)
// and, as the HIP path runs it, primary-sample-space Metropolis light
// transport (proj/render.h:195-356, 360-780; proj/mlt.h:26-220).
//
// Deliberate differences from the reference (each one also applies to the
// HIP path, so the two stay comparable):
//   * RNG: the reference draws from a per-thread MT19937-64 (random.h:73-186)
//     whose Windows stream cannot be reproduced (LLP64 state truncation).
//     Here every random number is Philox4x32-10 keyed by
//     (seed; pixel, sample, depth|branch<<8, dim/4) — a counter-based stream
//     that the HIP kernels reproduce exactly, so both sides can run on
//     common random numbers.
//   * a safety cap on path vertices (default 128); the reference has none.
//   * lights of equal power keep entity order (stable sort); the reference
//     uses std::sort (light.h:99), whose tie order is library-dependent.
//   * POLYGON light entities are rejected when NEE is on: the reference
//     samples an uninitialised position for them (directLight_radiance.h:329).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/prt.h"

namespace orc {

// ---------------------------------------------------------------- constants
// proj/constant.h:5-40
static const double PS_PI = 3.14159265358979323846264338327950288;
static const double PS_INV_PI = 0.31830988618379067153776752674502872;
static const double PS_TWOPI = 6.28318530717958647692528676655900576;
static const double PS_FORPI = 12.5663706143591729538505735331180115;
static const double PS_RAD = PS_PI / 180.0;
static const double PS_INF = 1.0e128;
static const double PS_EPS = 1.0e-6;
static const double PS_EPS14 = 1.0e-14;

// ---------------------------------------------------------------- Vector3d
// proj/vector3d.h: dot/sqr use the SSE2 pairing (x*x' + z*z') + (y*y' + 0)
// (vector3d.h:99-124, 195-218); normalize = v * (1/length) (162-165).
struct V3 {
  double x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
};
static inline V3 operator+(const V3& a, const V3& b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(const V3& a, const V3& b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator*(const V3& a, double s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline V3 operator*(double s, const V3& a) { return V3(s * a.x, s * a.y, s * a.z); }
static inline V3 operator/(const V3& a, double s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline V3 mul(const V3& a, const V3& b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double dot(const V3& a, const V3& b) { return (a.x * b.x + a.z * b.z) + a.y * b.y; }
static inline double sqr(const V3& a) { return (a.x * a.x + a.z * a.z) + a.y * a.y; }
static inline double length(const V3& a) { return std::sqrt(sqr(a)); }
static inline V3 normalize(const V3& a) { return a * (1.0 / length(a)); }
static inline V3 cross(const V3& a, const V3& b) {
  return V3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
static inline double rgb_max(const V3& c) { return std::max(c.x, std::max(c.y, c.z)); }
static inline V3 from(const double* p) { return V3(p[0], p[1], p[2]); }

// OrthonormalBasis, vector3d.h:244-253
static inline void orthonormal_basis(const V3& n, V3& w, V3& u, V3& v) {
  w = n;
  if (std::fabs(w.x) > PS_EPS)
    u = normalize(cross(V3(0.0, 1.0, 0.0), w));
  else
    u = normalize(cross(V3(1.0, 0.0, 0.0), w));
  v = normalize(cross(w, u));
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
// SC'11).  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.
static inline void philox4x32_10(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Random-number dimensions used at one path vertex.
enum {
  DIM_RR = 0, DIM_MATERIAL = 1, DIM_DIFF_U1 = 2, DIM_DIFF_U2 = 3,
  DIM_NEE0 = 4,  // 4 + 3*i: light pick, 5 + 3*i / 6 + 3*i: point on light
  DIM_GLASS = 13
};

// Camera words of a Metropolis proposal before its per-vertex words.
static const int MLT_CAMERA_WORDS = 8;

struct PathKey {
  uint32_t key[2];
  uint32_t pixel, sample;
  uint32_t depth, branch;
  // Metropolis renders: the chain's primary samples (null otherwise); the
  // first mlt_depths vertices take dimensions 0..3 from them
  const uint32_t* mlt;
  int mlt_depths;
  // lazily filled Philox blocks for this vertex
  mutable uint32_t blk[4][4];
  mutable unsigned have;
};
static inline double u01_from(uint32_t w) { return (double)w * 2.3283064365386963e-10; }
static inline double rnd(const PathKey& pk, int dim) {
  if (pk.mlt) {
    // the glass reflect / refract choice (brdf.h:866) is a primary sample
    // too; it shares the vertex's BSDF-choice word, which glass leaves unused
    if (dim == DIM_GLASS) dim = DIM_MATERIAL;
    if (dim < 4 && (int)pk.depth < pk.mlt_depths) return u01_from(pk.mlt[MLT_CAMERA_WORDS + 4 * pk.depth + dim]);
  }
  const int b = dim >> 2;
  if (!(pk.have & (1u << b))) {
    uint32_t c[4] = {pk.pixel, pk.sample, (pk.depth & 0xFFu) | (pk.branch << 8), (uint32_t)b};
    philox4x32_10(c, pk.key, pk.blk[b]);
    pk.have |= 1u << b;
  }
  return u01_from(pk.blk[b][dim & 3]);
}
static inline PathKey vertex_key(const uint32_t key[2], uint32_t pixel, uint32_t sample, int depth,
                                 uint32_t branch) {
  PathKey pk;
  pk.key[0] = key[0]; pk.key[1] = key[1];
  pk.pixel = pixel; pk.sample = sample; pk.depth = (uint32_t)depth; pk.branch = branch;
  pk.mlt = nullptr; pk.mlt_depths = 0;
  pk.have = 0;
  return pk;
}

// ---------------------------------------------------------------- scene
struct Ray { V3 org, dir; };

struct Mat {
  V3 color, emission;
  double roughness, ior;
  int type;
  // Material::isLight, material.h:195-210
  bool is_light() const {
    return type == PRT_REFLECTION_DIFFUSE && length(color) < 0.001 && length(emission) > 0.001;
  }
};

struct Hit {
  double t, u, v;
  int entity, prim;
  V3 position, normal;
};

struct Stats {
  uint64_t rays_primary = 0, rays_ext = 0, rays_shadow = 0, rays_pickup = 0;
  uint64_t nodes = 0, tris = 0, truncated = 0;
  void add(const Stats& o) {
    rays_primary += o.rays_primary; rays_ext += o.rays_ext; rays_shadow += o.rays_shadow;
    rays_pickup += o.rays_pickup; nodes += o.nodes; tris += o.tris; truncated += o.truncated;
  }
};

// ---- QBVH restatement (BVH.cpp:119-348 + QBVH.cpp:161-323) ----
struct BBox { V3 mn, mx, c; };
// BoundingBox::CalcSurfaceArea, BoundingBox.h:29-33 (half area)
static inline double half_area(const V3& mn, const V3& mx) {
  double dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
  return dx * dy + dx * dz + dy * dz;
}
struct BNode {  // BVH_structure, BVH.h:19-28
  double box[2][3];
  int children[2];
  int axis;
  std::vector<int> objects;  // triangle ids (leaf only)
};
struct QNode {  // QBVH_structure, QBVH.h:53-61 (boxes kept as float[2][3][4])
  float bboxes[2][3][4];
  int64_t children[4];
  int axis_top, axis_left, axis_right;
};
static const int64_t Q_INVALID = -1;
static const int64_t Q_LEAF_BIT = (int64_t)1 << 62;

struct Mesh {
  std::vector<float> vert;      // nv*3
  std::vector<int32_t> idx;     // nf*3
  std::vector<float> fnorm;     // nf*3, f32 face normals (ObjLoader.cpp:577-700)
  std::vector<float> albedo;    // nf*3 or empty
  std::vector<BBox> tbox;       // padded triangle boxes (polygon.h:686-703)
  std::vector<QNode> qnodes;
  std::vector<int> leaf_objects; // -1 terminated lists (QBVH.cpp:299-306)
  bool root_is_leaf = false;
  int nf = 0;
  V3 vertex(int i) const { return V3(vert[3 * i], vert[3 * i + 1], vert[3 * i + 2]); }
};

struct Ent {
  int type = 0, mat = 0, ninv = 1;
  bool light = false;
  // sphere (sphere.h:9-30)
  V3 position; double radius = 0;
  // uvplane (plane.h:129-166)
  V3 org, u_axis, v_axis, normal; double u_length = 0, v_length = 0;
  double area = PS_INF;
  Mesh mesh;
};

struct Light { int entity; double power; };

struct Camera {
  V3 position, dir, up, screen_x, screen_y, center;
  double screen_width, screen_height, screen_dist, iw, ih;
  int width, height;
};

struct Scene {
  std::vector<Mat> mats;
  std::vector<Ent> ents;
  std::vector<Light> lights;
  Camera cam;
  int samples, supersamples, depth, depth_limit, nee;
  double luminance_cutoff, sensor_response;
  V3 background, env_light;
  int max_vertices = 128;
};

static thread_local std::string g_err;

// ----------------------------------------------------------- BVH (reference)
static void bbox_of(const Mesh& m, const std::vector<int>& objs, double mn[3], double mx[3]) {
  // CalcBoundingBoxOfObjects, BVH.cpp:136-177
  const BBox& b0 = m.tbox[objs[0]];
  mn[0] = b0.mn.x; mn[1] = b0.mn.y; mn[2] = b0.mn.z;
  mx[0] = b0.mx.x; mx[1] = b0.mx.y; mx[2] = b0.mx.z;
  for (size_t i = 1; i < objs.size(); ++i) {
    const BBox& b = m.tbox[objs[i]];
    if (b.mn.x < mn[0]) mn[0] = b.mn.x;
    if (b.mn.y < mn[1]) mn[1] = b.mn.y;
    if (b.mn.z < mn[2]) mn[2] = b.mn.z;
    if (b.mx.x > mx[0]) mx[0] = b.mx.x;
    if (b.mx.y > mx[1]) mx[1] = b.mx.y;
    if (b.mx.z > mx[2]) mx[2] = b.mx.z;
  }
}
static inline double cen(const BBox& b, int axis) { return axis == 0 ? b.c.x : (axis == 1 ? b.c.y : b.c.z); }

// BVH::Construct_internal with CONSTRUCTION_OBJECT_SAH, BVH.cpp:194-348.  The
// sweep reproduces the reference's bookkeeping: the cost at cutIndex uses the
// left box of elements [0..cutIndex] and right box of [cutIndex+1..N-1] but
// the element counts after the move (cutIndex+2, N-cutIndex-2); the split
// then keeps [0..bestIndex] on the left.
static void bvh_build(const Mesh& m, std::vector<BNode>& nodes, std::vector<int>& targets, int index) {
  bbox_of(m, targets, nodes[index].box[0], nodes[index].box[1]);
  const double area = half_area(V3(nodes[index].box[0][0], nodes[index].box[0][1], nodes[index].box[0][2]),
                                V3(nodes[index].box[1][0], nodes[index].box[1][1], nodes[index].box[1][2]));
  const double inv_area = 1.0 / area;
  const size_t N = targets.size();
  int best_axis = -1, best_index = -1;
  double best_cost = -1;
  std::vector<int> sorted;
  std::vector<double> right_area(N);
  for (int axis = 0; axis < 3; ++axis) {
    sorted = targets;
    std::sort(sorted.begin(), sorted.end(),
              [&](int a, int b) { return cen(m.tbox[a], axis) < cen(m.tbox[b], axis); });
    right_area[N - 1] = PS_INF;
    // right sweep: right_area[i] = area(box of i+1..N-1)
    V3 mn = m.tbox[sorted[N - 1]].mn, mx = m.tbox[sorted[N - 1]].mx;
    for (int i = (int)N - 2; i >= 0; --i) {
      right_area[i] = half_area(mn, mx);
      const BBox& b = m.tbox[sorted[i]];
      mn = V3(std::min(mn.x, b.mn.x), std::min(mn.y, b.mn.y), std::min(mn.z, b.mn.z));
      mx = V3(std::max(mx.x, b.mx.x), std::max(mx.y, b.mx.y), std::max(mx.z, b.mx.z));
    }
    const double leaf_cost = 1.0 * (double)N;
    mn = m.tbox[sorted[0]].mn; mx = m.tbox[sorted[0]].mx;
    for (int cut = 0; cut <= (int)N - 2; ++cut) {
      const double left_area = half_area(mn, mx);
      const BBox& b = m.tbox[sorted[cut + 1]];
      mn = V3(std::min(mn.x, b.mn.x), std::min(mn.y, b.mn.y), std::min(mn.z, b.mn.z));
      mx = V3(std::max(mx.x, b.mx.x), std::max(mx.y, b.mx.y), std::max(mx.z, b.mx.z));
      const double nl = (double)(cut + 2), nr = (double)((int)N - cut - 2);
      const double cost = 2 * 1.0 + (left_area * nl + right_area[cut] * nr) * inv_area * 1.0;
      if ((cost < best_cost || best_cost == -1) && cost < leaf_cost) {
        best_cost = cost; best_index = cut; best_axis = axis;
      }
    }
  }
  if (best_axis == -1) {  // MakeLeaf_internal, BVH.cpp:181-191
    nodes[index].children[0] = nodes[index].children[1] = -1;
    nodes[index].objects = targets;
    return;
  }
  std::vector<int> lefts = targets;
  std::sort(lefts.begin(), lefts.end(),
            [&](int a, int b) { return cen(m.tbox[a], best_axis) < cen(m.tbox[b], best_axis); });
  std::vector<int> rights(lefts.begin() + best_index + 1, lefts.end());
  lefts.resize(best_index + 1);
  nodes[index].axis = best_axis;
  const int c0 = (int)nodes.size();
  nodes[index].children[0] = c0;
  nodes[index].children[1] = c0 + 1;
  nodes.emplace_back();
  nodes.emplace_back();
  std::vector<int>().swap(targets);
  bvh_build(m, nodes, lefts, c0);
  bvh_build(m, nodes, rights, c0 + 1);
}

static bool bnode_leaf(const BNode& n) { return n.children[0] < 0; }

// QBVH::MakeLeaf_internal, QBVH.cpp:293-306
static int64_t qbvh_make_leaf(Mesh& m, const BNode& leaf) {
  const int64_t start = (int64_t)m.leaf_objects.size();
  for (int o : leaf.objects) m.leaf_objects.push_back(o);
  m.leaf_objects.push_back(-1);
  return Q_LEAF_BIT | start;
}

// QBVH::Construct_internal, QBVH.cpp:202-290
static void qbvh_collapse(Mesh& m, const std::vector<BNode>& b, size_t qi, int bi) {
  const BNode& node = b[bi];
  m.qnodes[qi].axis_top = node.axis;
  const BNode* ch[2] = {&b[node.children[0]], &b[node.children[1]]};
  m.qnodes[qi].axis_left = bnode_leaf(*ch[0]) ? 0 : ch[0]->axis;
  m.qnodes[qi].axis_right = bnode_leaf(*ch[1]) ? 0 : ch[1]->axis;
  auto put_box = [&](const BNode& src, int slot) {
    for (int k = 0; k < 3; ++k) {
      m.qnodes[qi].bboxes[0][k][slot] = (float)src.box[0][k];
      m.qnodes[qi].bboxes[1][k][slot] = (float)src.box[1][k];
    }
  };
  for (int i = 0; i < 2; ++i) {
    const BNode& child = *ch[i];
    if (bnode_leaf(child)) {
      const int slot = 2 * i;
      int64_t leaf = qbvh_make_leaf(m, child);
      m.qnodes[qi].children[slot] = leaf;
      m.qnodes[qi].children[slot + 1] = Q_INVALID;
      put_box(child, slot);
    } else {
      const BNode* gc[2] = {&b[child.children[0]], &b[child.children[1]]};
      for (int j = 0; j < 2; ++j) {
        const int slot = 2 * i + j;
        put_box(*gc[j], slot);
        if (bnode_leaf(*gc[j])) {
          int64_t leaf = qbvh_make_leaf(m, *gc[j]);
          m.qnodes[qi].children[slot] = leaf;
        } else {
          const size_t ni = m.qnodes.size();
          m.qnodes.emplace_back();
          std::memset(&m.qnodes[ni], 0, sizeof(QNode));
          m.qnodes[qi].children[slot] = (int64_t)ni;
          qbvh_collapse(m, b, ni, child.children[j]);
        }
      }
    }
  }
}

static void mesh_build(Mesh& m) {
  const int nf = m.nf;
  m.tbox.resize(nf);
  for (int f = 0; f < nf; ++f) {  // Triangle::CreateBoundingBox, polygon.h:686-703
    V3 mn(PS_INF, PS_INF, PS_INF), mx(-PS_INF, -PS_INF, -PS_INF);
    for (int k = 0; k < 3; ++k) {
      V3 v = m.vertex(m.idx[3 * f + k]);
      if (mn.x > v.x) mn.x = v.x;
      if (mn.y > v.y) mn.y = v.y;
      if (mn.z > v.z) mn.z = v.z;
      if (mx.x < v.x) mx.x = v.x;
      if (mx.y < v.y) mx.y = v.y;
      if (mx.z < v.z) mx.z = v.z;
    }
    V3 eps(0.01, 0.01, 0.01);
    m.tbox[f].mn = mn - eps;
    m.tbox[f].mx = mx + eps;
    m.tbox[f].c = (m.tbox[f].mn + m.tbox[f].mx) / 2.0;  // BoundingBox ctor, BoundingBox.h:14-18
  }
  // f32 face normals, ObjLoader.cpp:577-602 + 688-698
  m.fnorm.resize((size_t)nf * 3);
  for (int f = 0; f < nf; ++f) {
    const float* v0 = &m.vert[3 * m.idx[3 * f]];
    const float* v1 = &m.vert[3 * m.idx[3 * f + 1]];
    const float* v2 = &m.vert[3 * m.idx[3 * f + 2]];
    float dx1 = v1[0] - v0[0], dy1 = v1[1] - v0[1], dz1 = v1[2] - v0[2];
    float dx2 = v2[0] - v0[0], dy2 = v2[1] - v0[1], dz2 = v2[2] - v0[2];
    float nx = dy1 * dz2 - dz1 * dy2;
    float ny = dz1 * dx2 - dx1 * dz2;
    float nz = dx1 * dy2 - dy1 * dx2;
    float a = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (a != 0.0f) { nx /= a; ny /= a; nz /= a; }
    m.fnorm[3 * f] = nx; m.fnorm[3 * f + 1] = ny; m.fnorm[3 * f + 2] = nz;
  }
  // BVH + QBVH (QBVH::Construct, QBVH.cpp:161-186)
  std::vector<BNode> nodes;
  nodes.reserve((size_t)nf * 2 + 1);
  nodes.emplace_back();
  std::vector<int> all(nf);
  for (int f = 0; f < nf; ++f) all[f] = f;
  bvh_build(m, nodes, all, 0);
  m.qnodes.clear();
  m.leaf_objects.clear();
  if (bnode_leaf(nodes[0])) {
    // The QBVH build throws when the root is a leaf and the loader falls back
    // to the binary BVH (scene_env.h:1827-1836); with a single leaf that BVH
    // reports the nearest triangle of the leaf, i.e. a brute-force search.
    m.root_is_leaf = true;
    return;
  }
  m.qnodes.emplace_back();
  std::memset(&m.qnodes[0], 0, sizeof(QNode));
  qbvh_collapse(m, nodes, 0, 0);
}

// ---------------------------------------------------------------- intersect
// Triangle::CheckIntersection2, polygon.h:78-129 (det eps = PS_EPS*0.1)
static inline int mt_test(const Mesh& m, int f, const Ray& ray, double& u, double& v, double& t) {
  const V3 v0 = m.vertex(m.idx[3 * f]);
  const V3 edge1 = m.vertex(m.idx[3 * f + 1]) - v0;
  const V3 edge2 = m.vertex(m.idx[3 * f + 2]) - v0;
  const V3 p = cross(ray.dir, edge2);
  const double det = dot(edge1, p);
  const V3 tv = ray.org - v0;
  const V3 q = cross(tv, edge1);
  const double eps = PS_EPS * 0.1;
  double inv_det;
  if (det > eps) {
    inv_det = 1.0 / det;
    u = dot(tv, p);
    if (u < 0.0 || u > det) return 0;
    v = dot(ray.dir, q);
    if (v < 0.0 || u + v > det) return 0;
  } else if (det < -eps) {
    inv_det = 1.0 / det;
    u = dot(tv, p);
    if (u > 0.0 || u < det) return 0;
    v = dot(ray.dir, q);
    if (v > 0.0 || u + v < det) return 0;
  } else {
    return 0;
  }
  t = dot(edge2, q) * inv_det;
  u *= inv_det;
  v *= inv_det;
  return 1;
}

// Triangle::intersect acceptance (polygon.h:163-166): stat==1 && t > PS_EPS
static inline bool tri_intersect(const Mesh& m, int f, const Ray& ray, double& t, double& u, double& v) {
  return mt_test(m, f, ray, u, v, t) == 1 && t > PS_EPS;
}

// CheckIntersection4floatAABB, BoundingBox.h:137-178, with _mm_max_ps /
// _mm_min_ps semantics (second operand returned when either is NaN).
static inline float sse_max(float a, float b) { return a > b ? a : b; }
static inline float sse_min(float a, float b) { return a < b ? a : b; }
static inline void slab4(const QNode& n, const float org[3], const float inv[3], const int sign[3], bool res[4]) {
  for (int c = 0; c < 4; ++c) {
    float tmin = 0.0f, tmax = FLT_MAX;
    for (int k = 0; k < 3; ++k) {
      tmin = sse_max(tmin, (n.bboxes[sign[k]][k][c] - org[k]) * inv[k]);
      tmax = sse_min(tmax, (n.bboxes[1 - sign[k]][k][c] - org[k]) * inv[k]);
    }
    res[c] = tmax >= tmin;
  }
}

// QBVH::CheckIntersection + CheckIntersection_Leaf, QBVH.cpp:17-159
static bool qbvh_closest(const Mesh& m, const Ray& ray, double& best_t, double& bu, double& bv, int& bprim,
                         Stats& st) {
  if (m.root_is_leaf) {
    bool any = false;
    double nearest = -1;
    for (int f = 0; f < m.nf; ++f) {
      double t, u, v;
      st.tris++;
      if (tri_intersect(m, f, ray, t, u, v) && t > 0 && (nearest < 0 || t < nearest)) {
        nearest = t; bu = u; bv = v; bprim = f; any = true;
      }
    }
    best_t = nearest;
    return any;
  }
  const int sign[3] = {ray.dir.x >= 0 ? 0 : 1, ray.dir.y >= 0 ? 0 : 1, ray.dir.z >= 0 ? 0 : 1};
  const float org[3] = {(float)ray.org.x, (float)ray.org.y, (float)ray.org.z};
  const float inv[3] = {ray.dir.x == 0 ? (float)PS_INF : (float)(1.0f / ray.dir.x),
                        ray.dir.y == 0 ? (float)PS_INF : (float)(1.0f / ray.dir.y),
                        ray.dir.z == 0 ? (float)PS_INF : (float)(1.0f / ray.dir.z)};
  std::vector<int64_t> stack;
  stack.reserve(64);
  stack.push_back(0);
  double dist = -1.0;
  while (!stack.empty()) {
    const int64_t ni = stack.back();
    stack.pop_back();
    const QNode& node = m.qnodes[ni];
    st.nodes++;
    bool res[4];
    slab4(node, org, inv, sign, res);
    if (!(res[0] || res[1] || res[2] || res[3])) continue;
    int ordering[4];
    const bool is_left = sign[node.axis_top] == 0;
    const int lf = is_left ? 0 : 2;
    const int rf = (lf + 2) % 4;
    if (sign[node.axis_left] == 0) { ordering[lf] = 0; ordering[lf + 1] = 1; }
    else { ordering[lf] = 1; ordering[lf + 1] = 0; }
    if (sign[node.axis_right] == 0) { ordering[rf] = 2; ordering[rf + 1] = 3; }
    else { ordering[rf] = 3; ordering[rf + 1] = 2; }
    for (int i = 0; i < 4; ++i) {
      const int c = ordering[i];
      const int64_t ch = node.children[c];
      if (!res[c] || ch == Q_INVALID) continue;
      if (ch & Q_LEAF_BIT) {
        const int64_t start = ch ^ Q_LEAF_BIT;
        double nearest = -1, lt = 0, lu = 0, lv = 0;
        int lp = -1;
        for (int64_t k = start; m.leaf_objects[k] >= 0; ++k) {
          double t, u, v;
          st.tris++;
          const int f = m.leaf_objects[k];
          if (tri_intersect(m, f, ray, t, u, v) && t > 0 && (nearest < 0 || t < nearest)) {
            nearest = t; lt = t; lu = u; lv = v; lp = f;
          }
        }
        if (nearest >= 0 && lt > PS_EPS) {
          if (dist < 0 || dist > lt) {
            dist = lt; best_t = lt; bu = lu; bv = lv; bprim = lp;
          }
        }
      } else {
        stack.push_back(ch);
      }
    }
  }
  return dist > 0;
}

// Brute force over all faces with the leaf predicate (strictly nearer, first
// face index wins ties): the reference's own invariant, scene.cpp:122-141.
static bool brute_closest(const Mesh& m, const Ray& ray, double& best_t, double& bu, double& bv, int& bprim,
                          Stats& st) {
  double nearest = -1;
  for (int f = 0; f < m.nf; ++f) {
    double t, u, v;
    st.tris++;
    if (tri_intersect(m, f, ray, t, u, v) && (nearest < 0 || t < nearest)) {
      nearest = t; bu = u; bv = v; bprim = f;
    }
  }
  best_t = nearest;
  return nearest > 0;
}

// Sphere::intersect, sphere.h:68-96
static inline bool sphere_intersect(const Ent& e, const Ray& ray, double& t) {
  const V3 p_o = e.position - ray.org;
  const double b = dot(p_o, ray.dir);
  const double D4 = b * b - dot(p_o, p_o) + e.radius * e.radius;
  if (D4 < 0.0) return false;
  const double s = std::sqrt(D4);
  const double t1 = b - s, t2 = b + s;
  if (t1 < PS_EPS && t2 < PS_EPS) return false;
  t = (t1 > PS_EPS) ? t1 : t2;
  return true;
}

// UVPlane::intersect over Plane::intersect, plane.h:86-110 and 220-240
static inline bool uvplane_intersect(const Ent& e, const Ray& ray, double& t, V3& nrm) {
  V3 normal2 = e.normal;
  if (e.ninv < 0) normal2 = e.normal * -1.0;
  const double d = dot(ray.dir, normal2);
  if (std::fabs(d) < PS_EPS14) return false;
  t = dot((e.org - ray.org), normal2) / d;
  const V3 pos = ray.dir * t + ray.org;
  if (t < PS_EPS) return false;
  const double hu = dot(pos - e.org, e.u_axis) / e.u_length;
  const double hv = dot(pos - e.org, e.v_axis) / e.v_length;
  if (hu < 0 || hu > e.u_length) return false;
  if (hv < 0 || hv > e.v_length) return false;
  nrm = normal2;
  return true;
}

enum { MODE_QBVH = 0, MODE_BRUTE = 1 };

// intersect_scene_, scene.cpp:83-214 (linear entity loop, 1e-6 nudge,
// strictly nearer replaces)
static bool intersect_scene(const Scene& S, const Ray& ray_in, Hit& hit, Stats& st, int mode = MODE_QBVH) {
  const Ray ray{ray_in.org + ray_in.dir * 0.000001, ray_in.dir};
  hit.t = PS_INF;
  hit.entity = -1;
  hit.prim = -1;
  hit.u = hit.v = -PS_INF;
  const int n = (int)S.ents.size();
  for (int i = 0; i < n; ++i) {
    const Ent& e = S.ents[i];
    double t = 0.0;
    if (e.type == PRT_ENTITY_POLYGON) {
      double u = 0, v = 0;
      int prim = -1;
      bool ok = (mode == MODE_BRUTE) ? brute_closest(e.mesh, ray, t, u, v, prim, st)
                                     : qbvh_closest(e.mesh, ray, t, u, v, prim, st);
      if (ok && t < hit.t) {
        hit.t = t; hit.u = u; hit.v = v; hit.entity = i; hit.prim = prim;
        hit.position = ray.org + ray.dir * t;  // polygon.h:170
        const float* fn = &e.mesh.fnorm[3 * prim];
        hit.normal = V3(fn[0], fn[1], fn[2]);
        if (e.ninv < 0) hit.normal = hit.normal * -1.0;
      }
    } else if (e.type == PRT_ENTITY_SPHERE) {
      if (sphere_intersect(e, ray, t) && t < hit.t) {
        hit.t = t; hit.entity = i; hit.prim = -1; hit.u = hit.v = -PS_INF;
        hit.position = ray.org + t * ray.dir;  // sphere.h:89-95
        hit.normal = normalize(hit.position - e.position);
        if (e.ninv < 0) hit.normal = hit.normal * -1.0;
      }
    } else if (e.type == PRT_ENTITY_UVPLANE) {
      V3 nrm;
      if (uvplane_intersect(e, ray, t, nrm) && t < hit.t) {
        hit.t = t; hit.entity = i; hit.prim = -1; hit.u = hit.v = -PS_INF;
        hit.position = ray.dir * t + ray.org;
        hit.normal = nrm;
      }
    }
  }
  return hit.entity >= 0;
}

// ---------------------------------------------------------------- transport
// Adjoint accumulators of one thread (the CPU restatement of the GPU's
// path-replay adjoint): d(loss)/d(face albedo) of one POLYGON entity and
// d(loss)/d(emission) of one light, with Russian-roulette probabilities held
// fixed.  g is d(loss)/d(image) at the sample's pixel times the image's
// derivative with respect to the sample radiance.
struct Adj {
  int entity = -1, light = -1;
  V3 g;
  std::vector<double> grad_alb;  // n_faces * 3
  double grad_e[3] = {0, 0, 0};
};

struct Ctx {
  const Scene* S;
  uint32_t key[2];
  uint32_t pixel, sample;
  Stats st;
  Adj* ad = nullptr;
  const uint32_t* mlt = nullptr;  // Metropolis proposal's primary samples
  int mlt_depths = 0;
  bool direct_hit = false;        // the camera ray's diffuse hit on a light was returned (NEE on)
};

static V3 radiance(Ctx& C, const Ray& ray, int depth, uint32_t branch, const V3& T);

// d(loss) += g . T . dR/dp for a term R linear in the parameter p: adds
// g * T * R / p per channel.
static inline void adj_emission(Ctx& C, const V3& T, const V3& term_over_e) {
  Adj& a = *C.ad;
  a.grad_e[0] += a.g.x * T.x * term_over_e.x;
  a.grad_e[1] += a.g.y * T.y * term_over_e.y;
  a.grad_e[2] += a.g.z * T.z * term_over_e.z;
}
static inline V3 safe_div(const V3& a, const V3& b) {
  return V3(b.x != 0 ? a.x / b.x : 0.0, b.y != 0 ? a.y / b.y : 0.0, b.z != 0 ? a.z / b.z : 0.0);
}
// A vertex whose continuation and NEE scale with its colour rho: R - Env - E
// is linear in rho.
static inline void adj_albedo(Ctx& C, const Hit& h, const V3& T, const V3& R, const Mat& m) {
  Adj& a = *C.ad;
  if (h.entity != a.entity || h.prim < 0) return;
  const V3 rest = R - C.S->env_light - m.emission;
  const V3 d = safe_div(rest, m.color);
  double* g = &a.grad_alb[3 * (size_t)h.prim];
  g[0] += a.g.x * T.x * d.x;
  g[1] += a.g.y * T.y * d.y;
  g[2] += a.g.z * T.z * d.z;
}

static inline Mat hit_material(const Scene& S, const Hit& h) {
  const Ent& e = S.ents[h.entity];
  Mat m = S.mats[e.mat];
  if (e.type == PRT_ENTITY_POLYGON && !e.mesh.albedo.empty()) {
    const float* a = &e.mesh.albedo[3 * (size_t)h.prim];
    m.color = V3(a[0], a[1], a[2]);
  }
  return m;
}

// direct_light_perfect_refrection, directLight_radiance.h:516-556 (infinity,
// parallel and spot lights are out of scope)
static V3 pickup(Ctx& C, const V3& pos, const V3& dir, int* hit_entity = nullptr) {
  Hit h;
  C.st.rays_pickup++;
  if (hit_entity) *hit_entity = -1;
  if (intersect_scene(*C.S, Ray{pos, dir}, h, C.st) && C.S->ents[h.entity].light) {
    if (hit_entity) *hit_entity = h.entity;
    return hit_material(*C.S, h).emission;
  }
  return V3();
}

// direct_radiance_sample + eval_direct_radiance (diffuse branch),
// directLight_radiance.h:19-387
static V3 direct_sample(Ctx& C, const PathKey& pk, int i, const Light& L, const V3& x, const V3& n_orient,
                        const V3& color) {
  const Scene& S = *C.S;
  const Ent& le = S.ents[L.entity];
  V3 light_pos, light_normal;
  double pdf;
  if (le.type == PRT_ENTITY_SPHERE) {
    const double r1 = PS_TWOPI * rnd(pk, DIM_NEE0 + 3 * i + 1);
    const double r2 = 1.0 - 2.0 * rnd(pk, DIM_NEE0 + 3 * i + 2);
    const double s22 = std::sqrt(1.0 - r2 * r2);
    light_pos = le.position + ((le.radius + PS_EPS) * V3(s22 * std::cos(r1), s22 * std::sin(r1), r2));
    light_normal = normalize(light_pos - le.position);
    if (le.ninv < 0) light_normal = light_normal * -1.0;
    pdf = 1.0 / le.area;
  } else {  // UVPLANE
    const double t = le.v_length * rnd(pk, DIM_NEE0 + 3 * i + 1);
    const double s = le.u_length * rnd(pk, DIM_NEE0 + 3 * i + 2);
    light_pos = le.org + le.u_axis * s + le.v_axis * t + le.normal * PS_EPS;
    light_normal = le.normal;
    if (le.ninv < 0) light_normal = light_normal * -1.0;
    pdf = 1.0 / le.area;
  }
  const V3 emission = S.mats[le.mat].emission;
  const V3 light_dir = normalize(light_pos - x);
  const double dist2 = sqr(light_pos - x);
  const double dot0 = dot(n_orient, light_dir);
  const double dot1 = dot(light_normal, -1.0 * light_dir);
  if (dot0 >= 0 && dot1 >= 0) {
    const double G = dot0 * dot1 / dist2;
    Hit h;
    C.st.rays_shadow++;
    const bool isect = intersect_scene(S, Ray{x, light_dir}, h, C.st);
    if (isect) {
      const bool cond = std::fabs(std::sqrt(dist2) - h.t) < 1e-3;
      if (cond) {
        if (h.entity >= 0 && h.entity != L.entity) return V3();
        return mul(emission, color * PS_INV_PI) * G / pdf;
      }
    }
  }
  return V3();
}

// BRDF_Diffuse, brdf.h:408-700
static V3 brdf_diffuse(Ctx& C, const PathKey& pk, const Ray& ray, const Hit& h, const Mat& m, const V3& n_orient,
                       double rr, int depth, uint32_t branch, const V3& T) {
  const Scene& S = *C.S;
  const Ent& e = S.ents[h.entity];
  if (e.light && e.type == PRT_ENTITY_UVPLANE) {
    if (dot(h.normal, ray.dir) >= 0.0) return V3();
  }
  if (e.light && (!S.nee || depth == 0)) {
    if (S.nee && depth == 0) C.direct_hit = true;  // *hit_direct_light = 1, brdf.h:443-452
    if (C.ad && h.entity == C.ad->light) adj_emission(C, T, V3(1.0 / rr, 1.0 / rr, 1.0 / rr));
    return m.emission / rr / (1.0 - 0.0);
  }
  if (e.light) return V3();
  V3 w, u, v;
  orthonormal_basis(n_orient, w, u, v);
  const double r1 = PS_TWOPI * rnd(pk, DIM_DIFF_U1);
  const double r2 = rnd(pk, DIM_DIFF_U2), r2s = std::sqrt(r2);
  const V3 dir = normalize((u * std::cos(r1) * r2s + v * std::sin(r1) * r2s + w * std::sqrt(1.0 - r2)));
  const V3 weight = m.color / rr / (1.0 - 0.0);
  const V3 incoming = radiance(C, Ray{h.position, dir}, depth + 1, branch, mul(T, weight));
  V3 direct;
  if (S.nee && !S.lights.empty()) {
    const int N = (int)S.lights.size();
    for (int i = 0; i < 3; ++i) {
      // LightList::randomLight, light.h:132-155 (returns probability 1)
      const double next = rnd(pk, DIM_NEE0 + 3 * i);
      int index = (int)((next) * (double)(N - 1) + 0.5);
      if (index >= N) index = N - 1;
      if (index < 0) index = 0;
      const V3 dl = direct_sample(C, pk, i, S.lights[index], h.position, n_orient, m.color) / 1.0;
      direct = direct + dl;
      if (C.ad && S.lights[index].entity == C.ad->light)
        adj_emission(C, T, safe_div(dl / 3 / rr, S.mats[S.ents[S.lights[index].entity].mat].emission));
    }
    direct = direct / 3;
    direct = direct / rr / (1.0 - 0.0);
  }
  const V3 R = S.env_light + m.emission + direct + mul(weight, incoming);
  if (C.ad) {
    adj_albedo(C, h, T, R, m);
    if (h.entity == C.ad->light) adj_emission(C, T, V3(1, 1, 1));
  }
  return R;
}

// BRDF_Specular, brdf.h:702-761
static V3 brdf_specular(Ctx& C, const Ray& ray, const Hit& h, const Mat& m, double rr, int depth, uint32_t branch,
                        const V3& T) {
  const Scene& S = *C.S;
  const V3 dir = normalize(ray.dir - h.normal * 2.0 * dot(h.normal, ray.dir));
  const V3 weight = m.color / rr / (1.0 - 0.0);
  V3 incoming = radiance(C, Ray{h.position, dir}, depth + 1, branch, mul(T, weight));
  if (S.nee) {
    int pe = -1;
    incoming = incoming + pickup(C, h.position, dir, &pe);
    if (C.ad && pe >= 0 && pe == C.ad->light) adj_emission(C, mul(T, weight), V3(1, 1, 1));
  }
  const V3 R = S.env_light + m.emission + mul(weight, incoming);
  if (C.ad) {
    adj_albedo(C, h, T, R, m);
    if (h.entity == C.ad->light) adj_emission(C, T, V3(1, 1, 1));
  }
  return R;
}

// BRDF_Refraction, brdf.h:763-953
static V3 brdf_refraction(Ctx& C, const PathKey& pk, const Ray& ray, const Hit& h, const Mat& m,
                          const V3& n_orient, double rr, int depth, uint32_t branch, const V3& T) {
  const Scene& S = *C.S;
  const V3 refl_dir = normalize(ray.dir - h.normal * 2.0 * dot(h.normal, ray.dir));
  const bool into = dot(h.normal, n_orient) > 0.0;
  V3 direct, direct_t;
  if (S.nee) direct = pickup(C, h.position, refl_dir);
  const double nc = 1.0;
  const double nt = m.ior;
  const double nnt = into ? nc / nt : nt / nc;
  const double ddn = dot(ray.dir, n_orient);
  const double cos2t = 1.0 - nnt * nnt * (1.0 - ddn * ddn);
  V3 incoming, weight;
  if (cos2t < 0.0) {
    weight = m.color / rr / (1.0 - 0.0);
    incoming = radiance(C, Ray{h.position, refl_dir}, depth + 1, branch, mul(T, weight));
    incoming = incoming + direct;
    return S.env_light + m.emission + mul(weight, incoming);
  }
  const V3 tdir = normalize(ray.dir * nnt - h.normal * (into ? 1.0 : -1.0) * (ddn * nnt + std::sqrt(std::fabs(cos2t))));
  const double a = nt - nc, b = nt + nc;
  const double R0 = (a * a) / (b * b);
  const double c = 1.0 - (into ? -ddn : dot(tdir, -1.0 * n_orient));
  const double Re = R0 + (1.0 - R0) * std::pow(c, 5.0);
  const double nnt2 = std::pow(into ? nc / nt : nt / nc, 2.0);
  const double Tr = (1.0 - Re) * nnt2;
  const double probability = 0.25 + 0.5 * Re;
  if (S.nee) direct_t = pickup(C, h.position, tdir);
  if (depth > 2) {
    if (rnd(pk, DIM_GLASS) < probability) {
      weight = m.color / (probability * rr) / (1.0 - 0.0);
      incoming = radiance(C, Ray{h.position, refl_dir}, depth + 1, branch, mul(T, weight) * Re) * Re;
      incoming = incoming + direct;
    } else {
      weight = m.color / ((1.0 - probability) * rr) / (1.0 - 0.0);
      incoming = radiance(C, Ray{h.position, tdir}, depth + 1, branch, mul(T, weight) * Tr) * Tr;
      incoming = incoming + direct_t;
    }
  } else {
    weight = m.color / (rr) / (1.0 - 0.0);
    incoming = radiance(C, Ray{h.position, refl_dir}, depth + 1, branch * 2, mul(T, weight) * Re) * Re +
               radiance(C, Ray{h.position, tdir}, depth + 1, branch * 2 + 1, mul(T, weight) * Tr) * Tr;
    incoming = incoming + direct + direct_t;
  }
  return S.env_light + m.emission + mul(weight, incoming);
}

// radiance, radiance.h:25-556 (RGB build, no media, no SSS)
static V3 radiance(Ctx& C, const Ray& ray, int depth, uint32_t branch, const V3& T) {
  const Scene& S = *C.S;
  if (depth >= S.max_vertices) { C.st.truncated++; return V3(); }
  if (depth == 0) C.st.rays_primary++; else C.st.rays_ext++;
  Hit h;
  if (!intersect_scene(S, ray, h, C.st)) return S.background;
  const Mat m = hit_material(S, h);
  const V3 n_orient = dot(h.normal, ray.dir) < 0.0 ? h.normal : (-1.0 * h.normal);
  double rr = rgb_max(m.emission + m.color);
  if (depth > S.depth_limit) rr *= std::pow(0.5, depth - S.depth_limit);
  PathKey pk = vertex_key(C.key, C.pixel, C.sample, depth, branch);
  pk.mlt = C.mlt;
  pk.mlt_depths = C.mlt_depths;
  if (depth > S.depth) {
    if (rnd(pk, DIM_RR) >= rr) {
      if (C.ad && h.entity == C.ad->light) adj_emission(C, T, V3(1, 1, 1));
      return m.emission;
    }
  } else {
    rr = 1.0;
  }
  switch (m.type) {
    case PRT_REFLECTION_DIFFUSE:
      if (rnd(pk, DIM_MATERIAL) < m.roughness) return brdf_diffuse(C, pk, ray, h, m, n_orient, rr, depth, branch, T);
      return brdf_specular(C, ray, h, m, rr, depth, branch, T);
    case PRT_REFLECTION_SPECULAR:
      if (rnd(pk, DIM_MATERIAL) < 1.0 - m.roughness) return brdf_specular(C, ray, h, m, rr, depth, branch, T);
      return brdf_diffuse(C, pk, ray, h, m, n_orient, rr, depth, branch, T);
    case PRT_REFLECTION_REFRACTION:
      return brdf_refraction(C, pk, ray, h, m, n_orient, rr, depth, branch, T);
  }
  return V3();
}

// CameraScreenPram, render.h:100-146, with SceneEnv's camera handling
// (scene_env.h:1881-1886, 1920)
static void setup_camera(Camera& cam, const prt_camera& c) {
  cam.position = from(c.camera_position);
  cam.dir = normalize(from(c.camera_dir));
  cam.up = from(c.camera_up);
  double wsw = c.world_screen_width, wsh = c.world_screen_height;
  if (c.camera_angle > 0.0) {
    wsw = 2.0 * c.world_screen_dist * std::tan(c.camera_angle * PS_RAD / 2.0);
    wsh = wsw;
  }
  double cam_phi = std::atan2(cam.dir.y, cam.dir.x);
  if (cam_phi < 0.0) cam_phi += 2.0 * PS_PI;
  if (std::fabs(cam_phi - PS_PI) < 1.0e-6 || cam_phi < 1.0e-6 || std::fabs(cam_phi - 2.0 * PS_PI) < 1.0e-6) {
    cam.dir.y += 1.0e-6;
    const double len = std::sqrt(cam.dir.x * cam.dir.x + cam.dir.y * cam.dir.y + cam.dir.z * cam.dir.z);
    cam.dir.x /= len; cam.dir.y /= len; cam.dir.z /= len;
  }
  cam.width = c.image_width;
  cam.height = c.image_height;
  cam.screen_width = wsw * c.image_width / c.image_height;
  cam.screen_height = wsh;
  cam.screen_dist = c.world_screen_dist;
  cam.screen_x = normalize(cross(cam.dir, cam.up)) * cam.screen_width;
  cam.screen_y = normalize(cross(cam.screen_x, cam.dir)) * cam.screen_height;
  cam.center = cam.position + cam.dir * cam.screen_dist;
  cam.iw = 1.0 / cam.width;
  cam.ih = 1.0 / cam.height;
}

// primary ray, render.h:1248-1269 (no tent filter, no lens)
static Ray camera_ray(const Camera& cam, int x, int y, int sx, int sy, int ss) {
  const double rate = 1.0 / ss;
  const double r1 = (sx + 0.5 + 0.0) * rate;
  const double r2 = (sy + 0.5 + 0.0) * rate;
  const V3 sp = cam.center + cam.screen_x * ((r1 + x) * cam.iw - 0.5) + cam.screen_y * ((r2 + y) * cam.ih - 0.5);
  return Ray{cam.position, normalize(sp - cam.position)};
}

static double luminance(const V3& c) { return dot(V3(0.2126, 0.7152, 0.0722), c); }  // mlt.h:171-174

}  // namespace orc

using namespace orc;

struct orc_scene {
  Scene S;
};

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox4x32_10(ctr, key, out);
}

orc_scene* orc_scene_create(const prt_scene_desc* d) {
  if (!d) { g_err = "null desc"; return nullptr; }
  orc_scene* sc = new orc_scene();
  Scene& S = sc->S;
  for (int i = 0; i < d->n_materials; ++i) {
    const prt_material& pm = d->materials[i];
    Mat m;
    m.color = from(pm.color);
    m.emission = from(pm.emission);
    m.roughness = pm.roughness;
    m.ior = pm.refractive_index;
    m.type = pm.reflection_type;
    if (m.type < 0 || m.type > 2) { g_err = "bad reflection_type"; delete sc; return nullptr; }
    S.mats.push_back(m);
  }
  S.ents.resize(d->n_entities);
  for (int i = 0; i < d->n_entities; ++i) {
    const prt_entity& pe = d->entities[i];
    Ent& e = S.ents[i];
    e.type = pe.type;
    e.mat = pe.material;
    e.ninv = pe.normal_vector_inverse < 0 ? -1 : 1;
    if (e.mat < 0 || e.mat >= (int)S.mats.size()) { g_err = "bad material index"; delete sc; return nullptr; }
    if (pe.type == PRT_ENTITY_SPHERE) {
      e.position = from(pe.p0);
      e.radius = pe.radius;
      e.area = (PS_FORPI * e.radius * e.radius);  // Sphere::CalcArea, sphere.h:23-27
    } else if (pe.type == PRT_ENTITY_UVPLANE) {
      // UVPlane ctor + CalcArea, plane.h:143-166
      e.org = from(pe.p0);
      const V3 U = from(pe.p1), V = from(pe.p2);
      e.u_length = length(U);
      e.v_length = length(V);
      e.u_axis = normalize(U);
      e.v_axis = normalize(V);
      e.normal = normalize(cross(e.u_axis, e.v_axis));
      e.area = length(e.u_axis * e.u_length) * length(e.v_axis * e.v_length) *
               (1.0 - std::pow(dot(e.u_axis, e.v_axis), 2.0));
    } else if (pe.type == PRT_ENTITY_POLYGON) {
      Mesh& m = e.mesh;
      if (pe.n_faces <= 0 || pe.n_vertices <= 0 || !pe.vertices || !pe.indices) {
        g_err = "empty mesh"; delete sc; return nullptr;
      }
      m.nf = (int)pe.n_faces;
      m.vert.assign(pe.vertices, pe.vertices + 3 * (size_t)pe.n_vertices);
      m.idx.assign(pe.indices, pe.indices + 3 * (size_t)pe.n_faces);
      for (int32_t k : m.idx)
        if (k < 0 || k >= pe.n_vertices) { g_err = "index out of range"; delete sc; return nullptr; }
      if (pe.face_albedo) m.albedo.assign(pe.face_albedo, pe.face_albedo + 3 * (size_t)pe.n_faces);
      mesh_build(m);
    } else {
      g_err = "unsupported entity type"; delete sc; return nullptr;
    }
    e.light = S.mats[e.mat].is_light();
  }
  // light list (scene_env.h:1567-1586, 1692-1719; light.h:95-99): power =
  // area*|emission|, descending
  for (int i = 0; i < (int)S.ents.size(); ++i) {
    const Ent& e = S.ents[i];
    if (!e.light) continue;
    if (e.type == PRT_ENTITY_POLYGON) {
      if (d->sampling.next_event_estimation) {
        g_err = "POLYGON lights cannot be sampled by NEE"; delete sc; return nullptr;
      }
      continue;
    }
    S.lights.push_back(Light{i, e.area * length(S.mats[e.mat].emission)});
  }
  std::stable_sort(S.lights.begin(), S.lights.end(), [](const Light& a, const Light& b) { return a.power > b.power; });
  setup_camera(S.cam, d->camera);
  S.samples = d->sampling.samples;
  S.supersamples = d->sampling.supersamples;
  S.depth = d->sampling.depth;
  S.depth_limit = d->sampling.depth_limit;
  S.nee = d->sampling.next_event_estimation;
  S.luminance_cutoff = d->sampling.luminance_cutoff;
  S.sensor_response = d->sampling.sensor_response;
  S.background = from(d->sampling.background_color);
  S.env_light = from(d->sampling.environment_light);
  if (S.cam.width <= 0 || S.cam.height <= 0 || S.samples <= 0 || S.supersamples <= 0) {
    g_err = "bad image size or sampling"; delete sc; return nullptr;
  }
  return sc;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

int orc_set_max_vertices(orc_scene* s, int cap) {
  if (!s || cap <= 0) return PRT_E_INVALID;
  s->S.max_vertices = cap;
  return PRT_OK;
}

int orc_light_order(const orc_scene* s, int32_t* entities, int cap) {
  const int n = (int)s->S.lights.size();
  for (int i = 0; i < n && i < cap; ++i) entities[i] = s->S.lights[i].entity;
  return n;
}

// Structural facts of a mesh entity's QBVH: node count, leaf slots, and a
// check that every face appears in exactly one leaf.
int orc_qbvh_info(const orc_scene* s, int entity, int64_t* n_nodes, int64_t* n_leaf_refs, int* each_face_once) {
  if (!s || entity < 0 || entity >= (int)s->S.ents.size()) return PRT_E_INVALID;
  const Ent& e = s->S.ents[entity];
  if (e.type != PRT_ENTITY_POLYGON) return PRT_E_INVALID;
  const Mesh& m = e.mesh;
  *n_nodes = (int64_t)m.qnodes.size();
  std::vector<int> cnt(m.nf, 0);
  int64_t refs = 0;
  for (int o : m.leaf_objects)
    if (o >= 0) { cnt[o]++; refs++; }
  if (m.root_is_leaf) { refs = m.nf; std::fill(cnt.begin(), cnt.end(), 1); }
  *n_leaf_refs = refs;
  int ok = 1;
  for (int c : cnt) if (c != 1) ok = 0;
  *each_face_once = ok;
  return PRT_OK;
}

// Closest hit for a batch of rays (mode 0: reference QBVH, 1: brute force).
int orc_trace_closest(const orc_scene* s, const prt_ray* rays, int64_t n, prt_hit* hits, int mode,
                      uint64_t* node_visits, uint64_t* tri_tests, int nthreads) {
  if (!s || (!rays && n > 0) || (!hits && n > 0)) return PRT_E_INVALID;
  uint64_t nodes = 0, tris = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) reduction(+ : nodes, tris)
#endif
  for (int64_t i = 0; i < n; ++i) {
    Stats st;
    Hit h;
    Ray r{from(rays[i].org), from(rays[i].dir)};
    bool ok = intersect_scene(s->S, r, h, st, mode);
    hits[i].t = ok ? h.t : PS_INF;
    hits[i].u = ok ? h.u : -PS_INF;
    hits[i].v = ok ? h.v : -PS_INF;
    hits[i].entity = ok ? h.entity : -1;
    hits[i].prim = ok ? h.prim : -1;
    nodes += st.nodes;
    tris += st.tris;
  }
  if (node_visits) *node_visits = nodes;
  if (tri_tests) *tri_tests = tris;
  return PRT_OK;
}

// Camera rays for the sub-pixel grid, as Run() forms them (for tests).
int orc_camera_rays(const orc_scene* s, prt_ray* out) {
  const Scene& S = s->S;
  const int ss = S.supersamples;
  int64_t k = 0;
  for (int y = 0; y < S.cam.height; ++y)
    for (int x = 0; x < S.cam.width; ++x)
      for (int sy = 0; sy < ss; ++sy)
        for (int sx = 0; sx < ss; ++sx) {
          Ray r = camera_ray(S.cam, x, y, sx, sy, ss);
          out[k].org[0] = r.org.x; out[k].org[1] = r.org.y; out[k].org[2] = r.org.z;
          out[k].dir[0] = r.dir.x; out[k].dir[1] = r.dir.y; out[k].dir[2] = r.dir.z;
          ++k;
        }
  return PRT_OK;
}

// One radiance() evaluation (for tests): path of camera sample `sample` of
// pixel (x, y) in Run()'s numbering.
int orc_radiance_sample(const orc_scene* s, uint64_t seed, int x, int y, int sample, double out[3]) {
  const Scene& S = s->S;
  const int ss = S.supersamples;
  const int sub = sample % (ss * ss);
  Ctx C;
  C.S = &S;
  C.key[0] = (uint32_t)seed; C.key[1] = (uint32_t)(seed >> 32);
  C.pixel = (uint32_t)(y * S.cam.width + x);
  C.sample = (uint32_t)sample;
  V3 v = radiance(C, camera_ray(S.cam, x, y, sub % ss, sub / ss, ss), 0, 1, V3(1, 1, 1));
  out[0] = v.x; out[1] = v.y; out[2] = v.z;
  return PRT_OK;
}

// radiance() of one given ray entering at path depth `depth` (for the
// closed-form tests: a ray aimed at a glass interface, a mirror or a light at
// a chosen depth).  Evaluates n independent paths (sample indices 0..n-1 of
// pixel 0) and returns their mean and the unbiased variance of that mean.
int orc_radiance_ray(const orc_scene* s, uint64_t seed, const double org[3], const double dir[3], int depth,
                     int64_t n, double mean[3], double var_of_mean[3], uint64_t* stats) {
  if (!s || !org || !dir || !mean || n <= 0 || depth < 0) return PRT_E_INVALID;
  const Scene& S = s->S;
  double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  Stats total;
  for (int64_t i = 0; i < n; ++i) {
    Ctx C;
    C.S = &S;
    C.key[0] = (uint32_t)seed; C.key[1] = (uint32_t)(seed >> 32);
    C.pixel = 0;
    C.sample = (uint32_t)i;
    const V3 v = radiance(C, Ray{from(org), from(dir)}, depth, 1, V3(1, 1, 1));
    const double c[3] = {v.x, v.y, v.z};
    for (int k = 0; k < 3; ++k) { s1[k] += c[k]; s2[k] += c[k] * c[k]; }
    total.add(C.st);
  }
  for (int k = 0; k < 3; ++k) {
    mean[k] = s1[k] / (double)n;
    if (var_of_mean)
      var_of_mean[k] = n > 1 ? std::max(0.0, (s2[k] - s1[k] * s1[k] / (double)n) / (double)(n - 1)) / (double)n : 0.0;
  }
  if (stats) {
    stats[0] = total.rays_primary; stats[1] = total.rays_ext; stats[2] = total.rays_shadow;
    stats[3] = total.rays_pickup; stats[4] = total.nodes; stats[5] = total.tris; stats[6] = total.truncated;
  }
  return PRT_OK;
}

// Render::Run (render.h:1125-1455) over rows [row_begin, row_end) of the
// reference's y numbering.  out_rgb / out_var: W*H*3 doubles in the image's
// row order (image_index = (H-1-y)*W + x); rows outside the range are left
// untouched.  out_var receives the unbiased variance of each pixel's mean.
// stats: [primary, extension, shadow, pickup, nodes, tris, truncated].
int orc_render(const orc_scene* s, uint64_t seed, int samples_override, int row_begin, int row_end,
               double* out_rgb, double* out_var, uint64_t* stats, int nthreads) {
  if (!s || !out_rgb) return PRT_E_INVALID;
  const Scene& S = s->S;
  const int W = S.cam.width, H = S.cam.height;
  if (row_begin < 0) row_begin = 0;
  if (row_end <= 0 || row_end > H) row_end = H;
  const int samples = samples_override > 0 ? samples_override : S.samples;
  const int ss = S.supersamples;
  const int spp = samples * ss * ss;
  const double smp = 1.0 / samples / (ss * ss);
  std::vector<double> sq((size_t)W * H * 3, 0.0);
  for (int y = row_begin; y < row_end; ++y)
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < 3; ++c) out_rgb[((size_t)(H - y - 1) * W + x) * 3 + c] = 0.0;
  Stats total;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
  for (int smp_i = 0; smp_i < samples; ++smp_i) {
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
      Stats local;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
      for (int y = row_begin; y < row_end; ++y) {
        for (int x = 0; x < W; ++x) {
          const size_t image_index = (size_t)(H - y - 1) * W + x;
          for (int sy = 0; sy < ss; ++sy) {
            for (int sx = 0; sx < ss; ++sx) {
              Ctx C;
              C.S = &S;
              C.key[0] = (uint32_t)seed; C.key[1] = (uint32_t)(seed >> 32);
              C.pixel = (uint32_t)(y * W + x);
              C.sample = (uint32_t)(smp_i * ss * ss + sy * ss + sx);
              const Ray ray = camera_ray(S.cam, x, y, sx, sy, ss);
              V3 v = radiance(C, ray, 0, 1, V3(1, 1, 1)) * smp;
              // luminance_cutoff is a float in SceneEnv (scene_env.h:221)
              const float cutoff = (float)S.luminance_cutoff;
              if (cutoff > 0.0f) {
                const float lv = (float)luminance(v);
                if (lv > cutoff) {
                  const float f = cutoff / lv;
                  v = v * f;
                }
              }
              out_rgb[image_index * 3 + 0] += v.x;
              out_rgb[image_index * 3 + 1] += v.y;
              out_rgb[image_index * 3 + 2] += v.z;
              sq[image_index * 3 + 0] += v.x * v.x;
              sq[image_index * 3 + 1] += v.y * v.y;
              sq[image_index * 3 + 2] += v.z * v.z;
              local.add(C.st);
            }
          }
        }
      }
#ifdef _OPENMP
#pragma omp critical
#endif
      total.add(local);
    }
  }
  const double sr = S.sensor_response;
  for (int y = row_begin; y < row_end; ++y)
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < 3; ++c) {
        const size_t k = ((size_t)(H - y - 1) * W + x) * 3 + c;
        const double s1 = out_rgb[k], s2 = sq[k];
        if (out_var) out_var[k] = spp > 1 ? std::max(0.0, (spp * s2 - s1 * s1) / (spp - 1)) * sr * sr : 0.0;
        out_rgb[k] = s1 * sr;  // render.h:1405-1409
      }
  if (stats) {
    stats[0] = total.rays_primary; stats[1] = total.rays_ext; stats[2] = total.rays_shadow;
    stats[3] = total.rays_pickup; stats[4] = total.nodes; stats[5] = total.tris; stats[6] = total.truncated;
  }
  return PRT_OK;
}

// CPU adjoint of Render::Run's image with respect to the per-face albedo of
// POLYGON entity `entity` and the emission of light entity `light`, for the
// adjoint image dimage (W*H*3 doubles, image row order): the FP64 recursion
// carries the path throughput T, and a vertex's return value is the
// radiance gathered from it onwards (no luminance clamp, no glass).
int orc_adjoint(const orc_scene* s, uint64_t seed, int samples_override, const double* dimage, int entity,
                double* grad_alb, int light, double* grad_e, int nthreads) {
  if (!s || !dimage) return PRT_E_INVALID;
  const Scene& S = s->S;
  for (const Mat& m : S.mats)
    if (m.type == PRT_REFLECTION_REFRACTION) return PRT_E_UNSUPPORTED;
  if (S.luminance_cutoff > 0.0) return PRT_E_UNSUPPORTED;
  const int W = S.cam.width, H = S.cam.height;
  const int samples = samples_override > 0 ? samples_override : S.samples;
  const int ss = S.supersamples;
  const double smp = 1.0 / samples / (ss * ss);
  const size_t nf = (entity >= 0 && entity < (int)S.ents.size() && S.ents[entity].type == PRT_ENTITY_POLYGON)
                        ? (size_t)S.ents[entity].mesh.nf : 0;
  std::vector<double> total(nf * 3, 0.0);
  double ge[3] = {0, 0, 0};
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    Adj ad;
    ad.entity = nf ? entity : -1;
    ad.light = light;
    ad.grad_alb.assign(nf * 3, 0.0);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int y = 0; y < H; ++y) {
      for (int x = 0; x < W; ++x) {
        const size_t ii = (size_t)(H - y - 1) * W + x;
        const double sc = smp * S.sensor_response;
        ad.g = V3(dimage[3 * ii] * sc, dimage[3 * ii + 1] * sc, dimage[3 * ii + 2] * sc);
        for (int k = 0; k < samples * ss * ss; ++k) {
          const int sub = k % (ss * ss);
          Ctx C;
          C.S = &S;
          C.key[0] = (uint32_t)seed; C.key[1] = (uint32_t)(seed >> 32);
          C.pixel = (uint32_t)(y * W + x);
          C.sample = (uint32_t)k;
          C.ad = &ad;
          radiance(C, camera_ray(S.cam, x, y, sub % ss, sub / ss, ss), 0, 1, V3(1, 1, 1));
        }
      }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    {
      for (size_t i = 0; i < total.size(); ++i) total[i] += ad.grad_alb[i];
      for (int c = 0; c < 3; ++c) ge[c] += ad.grad_e[c];
    }
  }
  if (grad_alb) std::memcpy(grad_alb, total.data(), total.size() * sizeof(double));
  if (grad_e) std::memcpy(grad_e, ge, sizeof(ge));
  return PRT_OK;
}

int orc_set_face_albedo(orc_scene* s, int entity, const float* albedo) {
  if (!s || !albedo || entity < 0 || entity >= (int)s->S.ents.size()) return PRT_E_INVALID;
  Ent& e = s->S.ents[entity];
  if (e.type != PRT_ENTITY_POLYGON || e.mesh.albedo.empty()) return PRT_E_INVALID;
  std::memcpy(e.mesh.albedo.data(), albedo, e.mesh.albedo.size() * sizeof(float));
  return PRT_OK;
}

int orc_set_emission(orc_scene* s, int entity, const double emission[3]) {
  if (!s || !emission || entity < 0 || entity >= (int)s->S.ents.size()) return PRT_E_INVALID;
  Scene& S = s->S;
  Mat& m = S.mats[S.ents[entity].mat];
  m.emission = from(emission);
  for (Light& L : S.lights) L.power = S.ents[L.entity].area * length(S.mats[S.ents[L.entity].mat].emission);
  std::stable_sort(S.lights.begin(), S.lights.end(), [](const Light& a, const Light& b) { return a.power > b.power; });
  return PRT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- Metropolis
// Primary-sample-space MLT as the HIP path runs it (prender_amd/csrc/prt_mlt.*):
// Render::RunMLT (render.h:360-780), KelemenMLT (mlt.h:70-170) and
// generate_new_path (render.h:195-356) with many chains side by side, all
// randomness from Philox streams keyed by (chain, step, tag).
namespace orc {

static const uint32_t MLT_TAG_LARGE = 0x4C415247u, MLT_TAG_MUTATE = 0x4D555441u, MLT_TAG_FRESH = 0x46524553u,
                      MLT_TAG_ACCEPT = 0x41434350u, MLT_TAG_INIT = 0x494E4954u, MLT_TAG_SEED = 0x53454544u;
static const double MLT_FIX_ONE = 1048576.0, MLT_LUM_ONE = 16777216.0;
static const uint32_t ERPT_TAG_CHAINS = 0x4E434841u, ERPT_TAG_PERTURB = 0x50455254u, ERPT_TAG_ACCEPT = 0x45414343u;

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static void mlt_key(uint64_t seed, uint64_t phase, uint64_t idx, uint32_t k[2]) {
  const uint64_t h = splitmix64(splitmix64(seed ^ (phase * 0xA0761D6478BD642Full)) + idx);
  k[0] = (uint32_t)h;
  k[1] = (uint32_t)(h >> 32);
}
static void philox_at(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, const uint32_t k[2], uint32_t out[4]) {
  const uint32_t c[4] = {c0, c1, c2, c3};
  philox4x32_10(c, k, out);
}
static uint32_t word_of(double x) { return x >= 1.0 ? 0u : (uint32_t)(x * 4294967296.0); }
// Mutate0, mlt.h:42-54
static uint32_t mutate0(uint32_t w, double r) {
  const double x = u01_from(w);
  const double s1 = 1.0 / 512.0, s2 = 1.0 / 16.0;
  const double dx = s1 / (s1 / s2 + (double)std::fabs((float)(2.0 * r - 1.0))) - s1 / (s1 / s2 + 1.0);
  if (r < 0.5) {
    const double x1 = x + dx;
    return word_of((x1 < 1.0) ? x1 : x1 - 1.0);
  }
  const double x1 = x - dx;
  return word_of((x1 < 0.0) ? x1 + 1.0 : x1);
}
// generate_new_path's camera (render.h:262-300) with the tent filter
static Ray mlt_camera_ray(const Camera& cam, const uint32_t* u, uint32_t& pix) {
  const int x = std::min((int)(u01_from(u[0]) * (double)cam.width), cam.width - 1);
  const int y = std::min((int)(u01_from(u[1]) * (double)cam.height), cam.height - 1);
  const int sx = u01_from(u[2]) < 0.5 ? 0 : 1;
  const int sy = u01_from(u[3]) < 0.5 ? 0 : 1;
  const double rr1 = 2.0 * u01_from(u[4]), rr2 = 2.0 * u01_from(u[5]);
  const double dx = rr1 < 1.0 ? std::sqrt(rr1) - 1.0 : 1.0 - std::sqrt(2.0 - rr1);
  const double dy = rr2 < 1.0 ? std::sqrt(rr2) - 1.0 : 1.0 - std::sqrt(2.0 - rr2);
  const double r1 = (sx + 0.5 + dx) * 0.5;
  const double r2 = (sy + 0.5 + dy) * 0.5;
  const V3 sp = cam.center + cam.screen_x * ((r1 + x) * cam.iw - 0.5) + cam.screen_y * ((r2 + y) * cam.ih - 0.5);
  pix = (uint32_t)((cam.height - y - 1) * cam.width + x);
  return Ray{cam.position, normalize(sp - cam.position)};
}
static void mlt_splat(std::vector<uint64_t>& img, uint32_t pix, double w, const V3& f) {
  const double c[3] = {f.x, f.y, f.z};
  for (int k = 0; k < 3; ++k) {
    const double v = w * c[k];
    if (v > 0.0) img[3 * (size_t)pix + k] += (uint64_t)(v * MLT_FIX_ONE + 0.5);
  }
}

}  // namespace orc

extern "C" {

// Kelemen's Mutate0 on one primary-sample word (for the tests).
uint32_t orc_mlt_mutate0(uint32_t w, double r) { return orc::mutate0(w, r); }

// Metropolis render (prt_render_mlt's algorithm).  out_rgb: W*H*3 doubles in
// image row order.  stats: [seed paths, chains, steps per chain, proposals,
// accepts, luminance total (fixed point)].
int orc_render_mlt(const orc_scene* s, uint64_t seed, double mutation, double pre_sample, double p_large,
                   int64_t mutations_per_chain, int samples_override, int chains, int sample_depths,
                   double* out_rgb, uint64_t* stats, int nthreads) {
  if (!s || !out_rgb) return PRT_E_INVALID;
  const Scene& S = s->S;
  const int W = S.cam.width, H = S.cam.height;
  const int64_t n_pix = (int64_t)W * H;
  const int samples = samples_override > 0 ? samples_override : S.samples;
  if (mutation <= 0.0) mutation = 32.0;
  if (pre_sample <= 0.0) pre_sample = 0.4;
  p_large = p_large > 0.0 ? std::min(p_large, 1.0) : 0.5;
  const int depths = sample_depths > 0 ? sample_depths : 8;
  if (depths > 32 || samples <= 0) return PRT_E_INVALID;
  int64_t N = (int64_t)(n_pix * pre_sample) * samples;
  if (N < 1) N = 1;
  if (N > ((int64_t)1 << 24)) return PRT_E_INVALID;
  const double total_steps = (double)samples * (double)n_pix * mutation;
  const int64_t C = chains > 0 ? chains : (int64_t)std::min(std::max(total_steps / 64.0, 1024.0), (double)(1 << 20));
  if (C > (1 << 20)) return PRT_E_INVALID;
  const int64_t M = mutations_per_chain > 0 ? mutations_per_chain
                                            : std::max<int64_t>(1, (int64_t)std::ceil(total_steps / (double)C));
  const int stride = MLT_CAMERA_WORDS + 4 * depths;
  uint32_t kc[2];
  mlt_key(seed, 3, 0, kc);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  (void)nthreads;
#endif
  auto seed_words = [&](uint32_t sidx, uint32_t* u) {
    for (int blk = 0; blk < stride / 4; ++blk) philox_at(sidx, 0u, MLT_TAG_SEED, (uint32_t)blk, kc, u + 4 * blk);
  };
  auto path = [&](const uint32_t key[2], uint32_t chain, const uint32_t* u, uint32_t& pix) {
    Ctx Cx;
    Cx.S = &S;
    Cx.key[0] = key[0]; Cx.key[1] = key[1];
    Cx.pixel = chain;
    Cx.sample = 0;
    Cx.mlt = u;
    Cx.mlt_depths = depths;
    const Ray ray = mlt_camera_ray(S.cam, u, pix);
    return radiance(Cx, ray, 0, 1, V3(1, 1, 1));
  };
  // seed paths, in batches of C (render.h:479-503)
  std::vector<V3> seed_f((size_t)N);
  std::vector<uint64_t> seed_lum((size_t)N);
  std::vector<uint32_t> seed_pix((size_t)N);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
  for (int64_t si = 0; si < N; ++si) {
    uint32_t key[2];
    mlt_key(seed, 1, (uint64_t)(si / C), key);
    std::vector<uint32_t> u((size_t)stride);
    seed_words((uint32_t)si, u.data());
    uint32_t pix;
    const V3 f = path(key, (uint32_t)(si % C), u.data(), pix);
    seed_f[(size_t)si] = f;
    const double l = luminance(f) * MLT_LUM_ONE + 0.5;
    seed_lum[(size_t)si] = l < 1099511627776.0 ? (uint64_t)l : 1099511627776ull;
    seed_pix[(size_t)si] = pix;
  }
  std::vector<uint64_t> cdf((size_t)N);
  uint64_t run = 0;
  for (int64_t i = 0; i < N; ++i) cdf[(size_t)i] = (run += seed_lum[(size_t)i]);
  const uint64_t total = cdf[(size_t)N - 1];
  std::vector<uint64_t> img((size_t)n_pix * 3, 0);
  uint64_t accepts = 0;
  int64_t steps_run = 0;
  if (total > 0) {
    const double b = (double)total / MLT_LUM_ONE / (double)N;
    const double weight = 4.0 * (double)n_pix, inv_m = 1.0 / (double)M;
    std::vector<uint32_t> cur((size_t)C * stride), prop((size_t)C * stride);
    std::vector<V3> f_old((size_t)C);
    std::vector<uint32_t> pix_old((size_t)C);
    for (int64_t c = 0; c < C; ++c) {  // start paths (mlt.h:201-220)
      uint32_t r4[4];
      philox_at((uint32_t)c, 0u, MLT_TAG_INIT, 0u, kc, r4);
      uint64_t r = (uint64_t)(u01_from(r4[0]) * (double)total);
      if (r >= total) r = total - 1;
      const size_t sidx = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
      seed_words((uint32_t)sidx, &cur[(size_t)c * stride]);
      f_old[(size_t)c] = seed_f[sidx];
      pix_old[(size_t)c] = seed_pix[sidx];
    }
    for (int64_t it = 0; it < M; ++it) {
      uint32_t key[2];
      mlt_key(seed, 2, (uint64_t)it, key);
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
      {
        std::vector<uint64_t> local((size_t)n_pix * 3, 0);
        uint64_t acc_local = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t c = 0; c < C; ++c) {
          uint32_t* pu = &prop[(size_t)c * stride];
          const uint32_t* cu = &cur[(size_t)c * stride];
          uint32_t r4[4];
          philox_at((uint32_t)c, (uint32_t)it, MLT_TAG_LARGE, 0u, kc, r4);
          const bool large = u01_from(r4[0]) < p_large;
          for (int blk = 0; blk < stride / 4; ++blk) {
            if (large) {
              philox_at((uint32_t)c, (uint32_t)it, MLT_TAG_FRESH, (uint32_t)blk, kc, pu + 4 * blk);
            } else {
              uint32_t r[4];
              philox_at((uint32_t)c, (uint32_t)it, MLT_TAG_MUTATE, (uint32_t)blk, kc, r);
              for (int k = 0; k < 4; ++k) pu[4 * blk + k] = mutate0(cu[4 * blk + k], u01_from(r[k]));
            }
          }
          uint32_t p_new;
          const V3 fn = path(key, (uint32_t)c, pu, p_new);
          const V3 fo = f_old[(size_t)c];
          // render.h:574-630
          const double l_new = luminance(fn), l_old = luminance(fo);
          const double a = l_old > 1.0e-16 ? std::min(1.0, l_new / l_old) : 1.0;
          const double lg = large ? 1.0 : 0.0;
          const double w_new = (a + lg) / (l_new / b + p_large) * inv_m;
          const double w_old = (1.0 - a) / (l_old / b + p_large) * inv_m;
          mlt_splat(local, pix_old[(size_t)c], w_old * weight, fo);
          mlt_splat(local, p_new, w_new * weight, fn);
          philox_at((uint32_t)c, (uint32_t)it, MLT_TAG_ACCEPT, 0u, kc, r4);
          if (u01_from(r4[0]) < a) {
            std::memcpy(&cur[(size_t)c * stride], pu, sizeof(uint32_t) * stride);
            f_old[(size_t)c] = fn;
            pix_old[(size_t)c] = p_new;
            ++acc_local;
          }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
          for (size_t i = 0; i < img.size(); ++i) img[i] += local[i];
          accepts += acc_local;
        }
      }
    }
    steps_run = M;
  }
  const double scale = S.sensor_response * 0.25 / (double)C;
  for (size_t i = 0; i < img.size(); ++i) out_rgb[i] = (double)img[i] * (1.0 / MLT_FIX_ONE) * scale;
  if (stats) {
    stats[0] = (uint64_t)N;
    stats[1] = (uint64_t)C;
    stats[2] = (uint64_t)steps_run;
    stats[3] = (uint64_t)C * (uint64_t)steps_run;
    stats[4] = accepts;
    stats[5] = total;
  }
  return PRT_OK;
}

// Energy redistribution path tracing as prt_render_erpt runs it: RunERPT's
// structure (render.h:786-1123) with the published chain (perturbations of
// the current path, deposits in its colour), in the same batches and Philox
// streams as the HIP path.  stats: [initial paths, chains, proposals, accepts].
int orc_render_erpt(const orc_scene* s, uint64_t seed, double mutation, double mutate_distance, int samples_override,
                    int sample_depths, double* out_rgb, uint64_t* stats, int nthreads) {
  if (!s || !out_rgb) return PRT_E_INVALID;
  const Scene& S = s->S;
  const int W = S.cam.width, H = S.cam.height;
  const int64_t n_pix = (int64_t)W * H;
  const int samples = samples_override > 0 ? samples_override : S.samples;
  if (mutation <= 0.0) mutation = 32.0;
  const int M = mutation < 1.0 ? 1 : (int)mutation;
  const double d = mutate_distance > 0.0 ? mutate_distance : 0.05;
  const int depths = sample_depths > 0 ? sample_depths : 8;
  if (depths > 32 || samples <= 0) return PRT_E_INVALID;
  const int ss2 = S.supersamples * S.supersamples;
  const int64_t n_init = n_pix * ss2;
  const int stride = MLT_CAMERA_WORDS + 4 * depths;
  const int64_t cap = (int64_t)1 << 20;
  const int64_t B = std::min(cap, std::max(n_init, n_pix));
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  (void)nthreads;
#endif
  uint32_t kA[2], kB[2], kC[2];
  mlt_key(seed, 4, 0, kA);
  mlt_key(seed, 5, 0, kB);
  mlt_key(seed, 6, 0, kC);
  uint64_t launch = 0;
  auto fresh_words = [&](const uint32_t k[2], uint32_t g, uint32_t* u) {
    for (int blk = 0; blk < stride / 4; ++blk) philox_at(g, 0u, MLT_TAG_SEED, (uint32_t)blk, k, u + 4 * blk);
  };
  auto path = [&](const uint32_t key[2], uint32_t slot, const uint32_t* u, uint32_t& pix, bool* direct) {
    Ctx Cx;
    Cx.S = &S;
    Cx.key[0] = key[0]; Cx.key[1] = key[1];
    Cx.pixel = slot;
    Cx.sample = 0;
    Cx.mlt = u;
    Cx.mlt_depths = depths;
    const Ray ray = mlt_camera_ray(S.cam, u, pix);
    const V3 f = radiance(Cx, ray, 0, 1, V3(1, 1, 1));
    if (direct) *direct = Cx.direct_hit;
    return f;
  };
  // deposition energy (render.h:831-849): sums of the paths' radiance in
  // 2^-24 fixed point, as the device sums its 32.32 accumulators >> 8
  uint64_t sums[3] = {0, 0, 0};
  for (int64_t first = 0; first < n_pix; first += cap) {
    const int64_t n = std::min(cap, n_pix - first);
    uint32_t key[2];
    mlt_key(seed, 7, launch++, key);
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
      uint64_t loc[3] = {0, 0, 0};
      std::vector<uint32_t> u((size_t)stride);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
      for (int64_t i = 0; i < n; ++i) {
        fresh_words(kA, (uint32_t)(first + i), u.data());
        uint32_t pix;
        const V3 f = path(key, (uint32_t)i, u.data(), pix, nullptr);
        const double c[3] = {f.x, f.y, f.z};
        for (int k = 0; k < 3; ++k) loc[k] += ((uint64_t)(c[k] * 4294967296.0 + 0.5)) >> 8;
      }
#ifdef _OPENMP
#pragma omp critical
#endif
      for (int k = 0; k < 3; ++k) sums[k] += loc[k];
    }
  }
  const V3 mean((double)sums[0] / MLT_LUM_ONE / (double)n_pix, (double)sums[1] / MLT_LUM_ONE / (double)n_pix,
                (double)sums[2] / MLT_LUM_ONE / (double)n_pix);
  const double ed = luminance(mean) / (double)M;
  const double dep = ed / (double)samples;
  std::vector<uint64_t> img((size_t)n_pix * 3, 0);
  uint64_t chain_base = 0, proposals = 0, accepts = 0;
  std::vector<uint32_t> pu((size_t)B * stride);
  std::vector<V3> pf((size_t)B);
  std::vector<uint32_t> ppix((size_t)B);
  std::vector<uint64_t> first_chain((size_t)B);
  for (int64_t first = 0; first < n_init; first += cap) {
    const int64_t n = std::min(cap, n_init - first);
    uint32_t key[2];
    mlt_key(seed, 7, launch++, key);
    std::vector<uint64_t> count((size_t)n, 0);
    std::vector<char> direct_hit((size_t)n, 0);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; ++i) {
      uint32_t* u = &pu[(size_t)i * stride];
      fresh_words(kB, (uint32_t)(first + i), u);
      bool direct = false;
      pf[(size_t)i] = path(key, (uint32_t)i, u, ppix[(size_t)i], &direct);
      const double l = luminance(pf[(size_t)i]);
      if (direct) {
        direct_hit[(size_t)i] = 1;
      } else if (l > 0.0 && ed > 0.0) {
        uint32_t r[4];
        philox_at((uint32_t)(first + i), 0u, ERPT_TAG_CHAINS, 0u, kB, r);
        const double c = std::floor(u01_from(r[0]) + l / ((double)M * ed));
        count[(size_t)i] = c > 0.0 ? (uint64_t)c : 0ull;
      }
    }
    // direct hits: F / samples at their pixel (render.h:927-931)
    for (int64_t i = 0; i < n; ++i)
      if (direct_hit[(size_t)i]) mlt_splat(img, ppix[(size_t)i], 1.0 / samples, pf[(size_t)i]);
    uint64_t run_total = 0;
    for (int64_t i = 0; i < n; ++i) first_chain[(size_t)i] = (run_total += count[(size_t)i]);
    const uint64_t n_chains = run_total;
    for (uint64_t j0 = 0; j0 < n_chains; j0 += (uint64_t)B) {
      const int64_t nc = (int64_t)std::min<uint64_t>((uint64_t)B, n_chains - j0);
      const uint32_t base = (uint32_t)(chain_base + j0);
      std::vector<uint32_t> cur((size_t)nc * stride), pix_y((size_t)nc), pix_now((size_t)nc), run((size_t)nc, 0);
      std::vector<V3> f_y((size_t)nc);
      for (int64_t j = 0; j < nc; ++j) {
        const uint64_t J = j0 + (uint64_t)j;
        const size_t p = (size_t)(std::upper_bound(first_chain.begin(), first_chain.begin() + n, J) -
                                  first_chain.begin());
        std::memcpy(&cur[(size_t)j * stride], &pu[p * stride], sizeof(uint32_t) * stride);
        f_y[(size_t)j] = pf[p];
        pix_y[(size_t)j] = ppix[p];
        const int64_t lp = (first + (int64_t)p) / ss2;  // loop pixel (render.h:971)
        pix_now[(size_t)j] = (uint32_t)((H - (int)(lp / W) - 1) * W + (int)(lp % W));
      }
      for (int step = 0; step < M; ++step) {
        uint32_t skey[2];
        mlt_key(seed, 7, launch++, skey);
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
        {
          std::vector<uint64_t> local((size_t)n_pix * 3, 0);
          std::vector<uint32_t> prop((size_t)stride);
          uint64_t acc_local = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
          for (int64_t j = 0; j < nc; ++j) {
            uint32_t* cu = &cur[(size_t)j * stride];
            for (int blk = 0; blk < stride / 4; ++blk) {
              uint32_t r[4];
              philox_at(base + (uint32_t)j, (uint32_t)step, ERPT_TAG_PERTURB, (uint32_t)blk, kC, r);
              for (int k = 0; k < 4; ++k) {
                double x = u01_from(cu[4 * blk + k]) + d * (2.0 * u01_from(r[k]) - 1.0);  // erpt.h:29-34
                if (x > 1.0) x -= 1.0;
                if (x < 0.0) x += 1.0;
                prop[(size_t)(4 * blk + k)] = word_of(x);
              }
            }
            uint32_t pz;
            const V3 fz = path(skey, (uint32_t)j, prop.data(), pz, nullptr);
            const double lz = luminance(fz), ly = luminance(f_y[(size_t)j]);
            const double q = std::min(1.0, ly > 1.0e-16 ? lz / ly : 0.0);
            uint32_t r4[4];
            philox_at(base + (uint32_t)j, (uint32_t)step, ERPT_TAG_ACCEPT, 0u, kC, r4);
            if (q > u01_from(r4[0])) {
              std::memcpy(cu, prop.data(), sizeof(uint32_t) * stride);
              f_y[(size_t)j] = fz;
              pix_y[(size_t)j] = pz;
              ++acc_local;
            }
            if (pix_now[(size_t)j] == pix_y[(size_t)j]) {
              ++run[(size_t)j];
            } else {
              pix_now[(size_t)j] = pix_y[(size_t)j];
              run[(size_t)j] = 0;
            }
            const double l = luminance(f_y[(size_t)j]);
            if (run[(size_t)j] < 20 && l > 0.0) mlt_splat(local, pix_y[(size_t)j], dep / l, f_y[(size_t)j]);
          }
#ifdef _OPENMP
#pragma omp critical
#endif
          {
            for (size_t i = 0; i < img.size(); ++i) img[i] += local[i];
            accepts += acc_local;
          }
        }
      }
      proposals += (uint64_t)nc * (uint64_t)M;
    }
    chain_base += n_chains;
  }
  const double scale = S.sensor_response * 0.25;
  for (size_t i = 0; i < img.size(); ++i) out_rgb[i] = (double)img[i] * (1.0 / MLT_FIX_ONE) * scale;
  if (stats) {
    stats[0] = (uint64_t)n_init;
    stats[1] = chain_base;
    stats[2] = proposals;
    stats[3] = accepts;
  }
  return PRT_OK;
}

}  // extern "C"
