// prt_api.cpp — the C ABI of include/prt.h: scene upload, closest-hit
// queries and the wavefront render loop on one HIP stream.
//
// Scene semantics follow SceneEnv::Load (proj/scene_env.h:357-1990): the
// light list holds the light-material spheres and uvplanes in entity order,
// sorted by power = area * |emission| descending (light.h:95-99,
// scene_env.h:1567-1719); the camera is CameraScreenPram (render.h:100-146)
// after SceneEnv's CAMERA_ANGLE handling and normalisation
// (scene_env.h:1881-1886, 1920).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prt.h"
#include "bvh.h"
#include "prt_mlt.h"
#include "prt_device.h"
#include "prt_internal.h"
#include "prt_kernels.h"

// 256-thread blocks per CU of the traversal grids
#ifndef PRT_TRAV_BLOCKS_PER_CU
#define PRT_TRAV_BLOCKS_PER_CU 8
#endif
// camera samples per depth-synchronous batch: 2^PRT_BATCH_LOG2 (half with
// glass).  2^24 holds a 512x512x64 frame whole: one drain of the depth tail
// per frame instead of eight (cfg2 216 -> 294 M samples/s against 2^21), for
// about 7.5 GB of queues, accumulators and hit records.
#ifndef PRT_BATCH_LOG2
#define PRT_BATCH_LOG2 24
#endif
// Per-scene renders of NEE scenes: where a depth's shadow rays are traced.
// 0: on the main stream after the depth's NEE; 1: on a second stream, while
// the next depth's extension rays are traced; 2: NEE and shadow rays both on
// the second stream, after the shade that wrote the NEE input; 3: the NEE on
// the main stream into one of two shadow queues, and its shadow rays on the
// second stream once the next depth's extension rays are done, so that they
// share the chip with that depth's shade and NEE (queue-streaming kernels)
// rather than with another traversal.
#ifndef PRT_OVERLAP_SHADOW
#define PRT_OVERLAP_SHADOW 3
#endif

// Shadow-ray blocks per CU.  The traversal kernels' LDS stacks allow 4
// resident blocks per CU; 3 leave room for the shade and NEE blocks that run
// beside the shadow rays in the alternating schedule (cfg2 348 -> 353 M
// samples/s; 2 let the shadow rays run into the next closest-hit launch).
#ifndef PRT_SHADOW_BLOCKS_PER_CU
#define PRT_SHADOW_BLOCKS_PER_CU 3
#endif
// the traversal-stack overflow areas hold one slot per thread of the
// closest-hit grid; a larger shadow grid would share slots between threads
static_assert(PRT_SHADOW_BLOCKS_PER_CU <= PRT_TRAV_BLOCKS_PER_CU,
              "shadow traversal grid must not exceed the closest-hit grid");

namespace prt {

thread_local std::string g_err;

int dev_alloc(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return PRT_OK;
  b.release();
  if (bytes == 0) return PRT_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();  // not left for the next launch's error check
    return fail(PRT_E_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
  }
  b.bytes = bytes;
  return PRT_OK;
}

namespace {
constexpr int kMaxPoolDevices = 64;
std::mutex g_pool_mu;
DevicePool* g_pools[kMaxPoolDevices] = {};
}  // namespace

DevicePool* acquire_device_pool(int device) {
  if (device < 0 || device >= kMaxPoolDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (!g_pools[device]) g_pools[device] = new (std::nothrow) DevicePool();
  if (g_pools[device]) g_pools[device]->scenes++;
  return g_pools[device];
}

void release_device_pool(int device) {
  if (device < 0 || device >= kMaxPoolDevices) return;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  DevicePool* p = g_pools[device];
  if (!p || --p->scenes > 0) return;
  // the last handle on the device is gone: free the memory (the pool object,
  // whose lock another thread could still name, stays)
  std::lock_guard<std::recursive_mutex> l2(p->m);
  p->ws.release();
  release_stream_workspace(p->sws);
  p->sws = nullptr;
}

void Workspace::release() {
  for (auto& a : q)
    for (auto& b : a) b.release();
  hit_t.release();
  hit_id.release();
  for (auto& b : sq) b.release();
  for (auto& b : sq2) b.release();
  for (auto& b : desc) b.release();
  slabs.release();
  accum.release();
  accum2.release();
  counts.release();
  fetch.release();
  ctr.release();
  if (h_counts) (void)hipHostFree(h_counts);
  h_counts = nullptr;
  batch_samples = queue_cap = shadow_cap = 0;
}

static const size_t kRayColBytes[RAYQ_COLS] = {8, 8, 8, 8, 8, 8, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};
static const size_t kShadowColBytes[SHADOWQ_COLS] = {8, 8, 8, 8, 8, 8, 8, 4, 4, 4, 4, 4};
size_t ray_entry_bytes() {
  size_t b = 0;
  for (size_t v : kRayColBytes) b += v;
  return b;
}
size_t shadow_entry_bytes() {
  size_t b = 0;
  for (size_t v : kShadowColBytes) b += v;
  return b;
}

int alloc_ray_queue(DevBuf* cols, int64_t cap) {
  const size_t* dsz = kRayColBytes;
  for (int k = 0; k < RAYQ_COLS; ++k) {
    const int rc = dev_alloc(cols[k], (size_t)cap * dsz[k]);
    if (rc) return rc;
  }
  return PRT_OK;
}

int alloc_shadow_queue(DevBuf* cols, int64_t cap) {
  const size_t* ssz = kShadowColBytes;
  for (int k = 0; k < SHADOWQ_COLS; ++k) {
    const int rc = dev_alloc(cols[k], (size_t)cap * ssz[k]);
    if (rc) return rc;
  }
  return PRT_OK;
}

RayQ ray_queue(DevBuf* b, int64_t cap) {
  RayQ q;
  q.ox = (double*)b[0].p; q.oy = (double*)b[1].p; q.oz = (double*)b[2].p;
  q.dx = (double*)b[3].p; q.dy = (double*)b[4].p; q.dz = (double*)b[5].p;
  q.tr = (float*)b[6].p; q.tg = (float*)b[7].p; q.tb = (float*)b[8].p;
  q.pr = (float*)b[9].p; q.pg = (float*)b[10].p; q.pb = (float*)b[11].p;
  q.slot = (uint32_t*)b[12].p;
  q.meta = (uint32_t*)b[13].p;
  q.nx = (float*)b[14].p; q.ny = (float*)b[15].p; q.nz = (float*)b[16].p;
  q.cap = (uint32_t)cap;
  return q;
}

ShadowQ shadow_queue(DevBuf* b, int64_t cap) {
  ShadowQ q;
  q.ox = (double*)b[0].p; q.oy = (double*)b[1].p; q.oz = (double*)b[2].p;
  q.dx = (double*)b[3].p; q.dy = (double*)b[4].p; q.dz = (double*)b[5].p;
  q.tmax = (double*)b[6].p;
  q.cr = (float*)b[7].p; q.cg = (float*)b[8].p; q.cb = (float*)b[9].p;
  q.slot = (uint32_t*)b[10].p;
  q.light = (int32_t*)b[11].p;
  q.cap = (uint32_t)cap;
  return q;
}

}  // namespace prt

using namespace prt;

namespace {

// Host Vector3d with the reference's operation order (vector3d.h).
struct V3 {
  double x, y, z;
};
V3 mk(double a, double b, double c) { return V3{a, b, c}; }
V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
V3 operator*(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
double dot(V3 a, V3 b) { return (a.x * b.x + a.z * b.z) + a.y * b.y; }
double length(V3 a) { return std::sqrt((a.x * a.x + a.z * a.z) + a.y * a.y); }
V3 normalize(V3 a) { return a * (1.0 / length(a)); }
V3 cross(V3 a, V3 b) { return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)); }
V3 from(const double* p) { return mk(p[0], p[1], p[2]); }
void put(double* d, V3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

}  // namespace

extern "C" {

int prt_abi_version(void) { return PRT_ABI_VERSION; }
const char* prt_last_error(void) { return g_err.c_str(); }

int prt_device_info(char* name, int name_len, int* cu_count, int* arch_major, int* arch_minor) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PRT_E_NODEVICE, hipGetErrorString(e));
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    std::snprintf(name, name_len, "%s", p.gcnArchName);
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (arch_major) *arch_major = p.major;
  if (arch_minor) *arch_minor = p.minor;
  return PRT_OK;
}

void prt_scene_destroy(prt_scene* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->stream2) (void)hipStreamSynchronize(s->stream2);
  s->nodes.release(); s->tris.release(); s->tri_nrm.release(); s->tri_alb.release();
  s->ents.release(); s->mats.release(); s->analytic.release(); s->lights.release(); s->ovf.release();
  if (s->pool) release_device_pool(s->device);
  s->pool = nullptr;
  s->out_rgb.release();
  s->out_var.release();
  s->d_slot_of_tri.release();
  s->grad_buf.release();
  s->tri_alb_rr.release();
  s->mlt.release();
  s->timer_events.release();
  for (hipEvent_t e : s->ev_nee)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : s->ev_sh)
    if (e) (void)hipEventDestroy(e);
  if (s->ev_shadow) (void)hipEventDestroy(s->ev_shadow);
  if (s->ev_shade) (void)hipEventDestroy(s->ev_shade);
  if (s->stream2) (void)hipStreamDestroy(s->stream2);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

static int scene_build(prt_scene* s, const prt_scene_desc* d) {
  if (!d) return fail(PRT_E_INVALID, "null scene description");
  const prt_camera& cam = d->camera;
  const prt_sampling& smp = d->sampling;
  if (cam.image_width <= 0 || cam.image_height <= 0) return fail(PRT_E_INVALID, "image size must be positive");
  if (smp.samples <= 0 || smp.supersamples <= 0) return fail(PRT_E_INVALID, "samples and supersamples must be positive");
  if ((int64_t)smp.samples * smp.supersamples * smp.supersamples > (1 << 20))
    return fail(PRT_E_INVALID, "samples per pixel above 2^20");
  if (d->n_materials <= 0 || !d->materials) return fail(PRT_E_INVALID, "no materials");
  if ((uint32_t)d->n_materials > TRI_SHADE_MAT_MASK + 1u) return fail(PRT_E_INVALID, "more than 2^24 materials");
  if (d->n_entities < 0 || (d->n_entities > 0 && !d->entities)) return fail(PRT_E_INVALID, "bad entity list");

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PRT_E_NODEVICE, hipGetErrorString(e));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(PRT_E_NODEVICE, std::string("libprt is built for gfx950, device is ") + prop.gcnArchName);
  s->device = dev;
  s->cu_count = prop.multiProcessorCount;
  s->pool = acquire_device_pool(dev);
  if (!s->pool) return fail(PRT_E_NOMEM, "no render memory pool for this device");
  HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&s->ev_nee[0], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&s->ev_nee[1], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&s->ev_shadow, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&s->ev_shade, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&s->ev_sh[0], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&s->ev_sh[1], hipEventDisableTiming));

  // materials (Material::isLight, material.h:195-210)
  s->h_mats.resize(d->n_materials);
  for (int i = 0; i < d->n_materials; ++i) {
    const prt_material& pm = d->materials[i];
    DevMaterial& m = s->h_mats[i];
    std::memcpy(m.color, pm.color, sizeof(m.color));
    std::memcpy(m.emission, pm.emission, sizeof(m.emission));
    m.roughness = pm.roughness;
    m.ior = pm.refractive_index;
    m.type = pm.reflection_type;
    if (m.type < PRT_REFLECTION_DIFFUSE || m.type > PRT_REFLECTION_REFRACTION)
      return fail(PRT_E_UNSUPPORTED, "reflection_type outside diffuse/specular/refraction");
    m.is_light = (m.type == PRT_REFLECTION_DIFFUSE && length(from(pm.color)) < 0.001 &&
                  length(from(pm.emission)) > 0.001) ? 1 : 0;
    if (m.type == PRT_REFLECTION_REFRACTION) s->has_glass = true;
  }

  // entities, triangles
  std::vector<int32_t> analytic;
  std::vector<float> v9;         // per input triangle: v0 v1 v2
  std::vector<int32_t> t_ent, t_face;
  std::vector<float> t_nrm, t_alb;
  bool any_albedo = false;
  s->h_ents.resize(d->n_entities);
  s->ent_first_tri.assign(d->n_entities, -1);
  s->ent_n_tris.assign(d->n_entities, 0);
  for (int i = 0; i < d->n_entities; ++i) {
    const prt_entity& pe = d->entities[i];
    DevEntity& de = s->h_ents[i];
    std::memset(&de, 0, sizeof(de));
    de.type = pe.type;
    de.material = pe.material;
    de.ninv = pe.normal_vector_inverse < 0 ? -1 : 1;
    if (pe.material < 0 || pe.material >= d->n_materials) return fail(PRT_E_INVALID, "entity material out of range");
    de.light = s->h_mats[pe.material].is_light;
    if (pe.type == PRT_ENTITY_SPHERE) {
      if (!(pe.radius > 0)) return fail(PRT_E_INVALID, "sphere radius must be positive");
      std::memcpy(de.p0, pe.p0, sizeof(de.p0));
      de.radius = pe.radius;
      de.area = (12.5663706143591729538505735331180115 * pe.radius * pe.radius);  // sphere.h:23-27
      analytic.push_back(i);
    } else if (pe.type == PRT_ENTITY_UVPLANE) {
      // UVPlane ctor + CalcArea, plane.h:143-166
      const V3 U = from(pe.p1), V = from(pe.p2);
      de.u_len = length(U);
      de.v_len = length(V);
      if (!(de.u_len > 0) || !(de.v_len > 0)) return fail(PRT_E_INVALID, "uvplane axes must be non-zero");
      const V3 ua = normalize(U), va = normalize(V);
      put(de.u_axis, ua);
      put(de.v_axis, va);
      put(de.normal, normalize(cross(ua, va)));
      std::memcpy(de.p0, pe.p0, sizeof(de.p0));
      de.area = length(ua * de.u_len) * length(va * de.v_len) * (1.0 - std::pow(dot(ua, va), 2.0));
      analytic.push_back(i);
    } else if (pe.type == PRT_ENTITY_POLYGON) {
      if (pe.n_faces <= 0 || pe.n_vertices <= 0 || !pe.vertices || !pe.indices)
        return fail(PRT_E_INVALID, "polygon entity without vertices or faces");
      if (de.light && smp.next_event_estimation)
        return fail(PRT_E_UNSUPPORTED, "POLYGON lights cannot be sampled by NEE (directLight_radiance.h:329)");
      de.has_albedo = pe.face_albedo ? 1 : 0;
      any_albedo |= (pe.face_albedo != nullptr);
      s->ent_first_tri[i] = (int64_t)t_ent.size();
      s->ent_n_tris[i] = pe.n_faces;
      for (int64_t f = 0; f < pe.n_faces; ++f) {
        const int32_t* ix = pe.indices + 3 * f;
        for (int k = 0; k < 3; ++k)
          if (ix[k] < 0 || ix[k] >= pe.n_vertices) return fail(PRT_E_INVALID, "triangle index out of range");
        const float* a = pe.vertices + 3 * (size_t)ix[0];
        const float* b = pe.vertices + 3 * (size_t)ix[1];
        const float* c = pe.vertices + 3 * (size_t)ix[2];
        v9.insert(v9.end(), {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]});
        t_ent.push_back(i);
        t_face.push_back((int32_t)f);
        // f32 face normal, ObjLoader.cpp:577-602 + 688-698
        const float dx1 = b[0] - a[0], dy1 = b[1] - a[1], dz1 = b[2] - a[2];
        const float dx2 = c[0] - a[0], dy2 = c[1] - a[1], dz2 = c[2] - a[2];
        float nx = dy1 * dz2 - dz1 * dy2;
        float ny = dz1 * dx2 - dx1 * dz2;
        float nz = dx1 * dy2 - dy1 * dx2;
        const float len = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (len != 0.0f) { nx /= len; ny /= len; nz /= len; }
        t_nrm.insert(t_nrm.end(), {nx, ny, nz});
        if (pe.face_albedo) {
          const float* al = pe.face_albedo + 3 * f;
          t_alb.insert(t_alb.end(), {al[0], al[1], al[2]});
        } else {
          t_alb.insert(t_alb.end(), {0.0f, 0.0f, 0.0f});
        }
      }
    } else {
      return fail(PRT_E_UNSUPPORTED, "entity type outside sphere/uvplane/polygon");
    }
  }
  s->n_tris = (int64_t)t_ent.size();

  // lights: light-material spheres and uvplanes, power-descending (stable)
  struct L { int32_t ent; double power; };
  std::vector<L> ls;
  for (int i = 0; i < d->n_entities; ++i) {
    const DevEntity& de = s->h_ents[i];
    if (!de.light || de.type == PRT_ENTITY_POLYGON) continue;
    ls.push_back(L{i, de.area * length(from(s->h_mats[de.material].emission))});
  }
  std::stable_sort(ls.begin(), ls.end(), [](const L& a, const L& b) { return a.power > b.power; });
  s->h_lights.clear();
  for (const L& l : ls) s->h_lights.push_back(l.ent);

  // BVH over all triangles, leaf-ordered triangle records
  BvhOutput bo;
  if (s->n_tris > 0) {
    BvhInput bi{v9.data(), s->n_tris, 0};
    const char* err = nullptr;
    if (!prt_build_bvh(bi, bo, &err)) return fail(PRT_E_UNSUPPORTED, err ? err : "bvh build failed");
  }
  s->bvh_depth = bo.depth;
  s->n_top = bo.n_top;
  s->bvh_ms = bo.build_ms;
  std::vector<float> tri4((size_t)s->n_tris * 12);
  std::vector<float> nrm((size_t)s->n_tris * 3), alb(any_albedo ? (size_t)s->n_tris * 3 : 0);
  s->slot_of_tri.assign(s->n_tris, 0);
  for (int64_t slot = 0; slot < s->n_tris; ++slot) {
    const uint32_t t = bo.order[slot];
    s->slot_of_tri[t] = (uint32_t)slot;
    float* r = &tri4[12 * slot];
    const float* v = &v9[9 * (size_t)t];
    int32_t ent = t_ent[t], face = t_face[t];
    const DevEntity& de = s->h_ents[ent];
    const uint32_t shade = ((uint32_t)de.material & TRI_SHADE_MAT_MASK) | (de.ninv < 0 ? TRI_SHADE_NINV : 0u) |
                           (de.light ? TRI_SHADE_LIGHT : 0u) | (de.has_albedo ? TRI_SHADE_ALBEDO : 0u);
    r[0] = v[0]; r[1] = v[1]; r[2] = v[2]; std::memcpy(&r[3], &ent, 4);
    r[4] = v[3]; r[5] = v[4]; r[6] = v[5]; std::memcpy(&r[7], &face, 4);
    r[8] = v[6]; r[9] = v[7]; r[10] = v[8]; std::memcpy(&r[11], &shade, 4);
    for (int k = 0; k < 3; ++k) nrm[3 * slot + k] = t_nrm[3 * (size_t)t + k];
    if (any_albedo)
      for (int k = 0; k < 3; ++k) alb[3 * slot + k] = t_alb[3 * (size_t)t + k];
  }

  hipStream_t st = s->stream;
  int rc;
  if ((rc = dev_upload(s->nodes, bo.nodes, st))) return rc;
  if ((rc = dev_upload(s->tris, tri4, st))) return rc;
  if ((rc = dev_upload(s->tri_nrm, nrm, st))) return rc;
  if (any_albedo && (rc = dev_upload(s->tri_alb, alb, st))) return rc;
  if ((rc = dev_upload(s->ents, s->h_ents, st))) return rc;
  if ((rc = dev_upload(s->mats, s->h_mats, st))) return rc;
  if ((rc = dev_upload(s->analytic, analytic, st))) return rc;
  if ((rc = dev_upload(s->lights, s->h_lights, st))) return rc;
  if (any_albedo && (rc = dev_upload(s->d_slot_of_tri, s->slot_of_tri, st))) return rc;
  // traversal stack overflow: one area per thread of the largest grid
  s->max_trav_blocks = std::max(1, s->cu_count) * PRT_TRAV_BLOCKS_PER_CU;
  const int64_t ovf_threads = (int64_t)s->max_trav_blocks * 256;
  const int ovf_words = ovf_words_for(s->bvh_depth);
  if ((rc = dev_alloc(s->ovf, (size_t)PRT_OVF_REGIONS * ovf_threads * ovf_words * sizeof(uint32_t))))
    return rc;
  HIPCHK(hipStreamSynchronize(st));

  // Params: camera (CameraScreenPram, render.h:100-146; scene_env.h:1881-1886, 1920)
  Params& P = s->P;
  std::memset(&P, 0, sizeof(P));
  V3 cpos = from(cam.camera_position), cdir = normalize(from(cam.camera_dir)), cup = from(cam.camera_up);
  double wsw = cam.world_screen_width, wsh = cam.world_screen_height;
  if (cam.camera_angle > 0.0) {
    wsw = 2.0 * cam.world_screen_dist * std::tan(cam.camera_angle * (3.14159265358979323846264338327950288 / 180.0) / 2.0);
    wsh = wsw;
  }
  const double PI = 3.14159265358979323846264338327950288;
  double cam_phi = std::atan2(cdir.y, cdir.x);
  if (cam_phi < 0.0) cam_phi += 2.0 * PI;
  if (std::fabs(cam_phi - PI) < 1.0e-6 || cam_phi < 1.0e-6 || std::fabs(cam_phi - 2.0 * PI) < 1.0e-6) {
    cdir.y += 1.0e-6;
    const double len = std::sqrt(cdir.x * cdir.x + cdir.y * cdir.y + cdir.z * cdir.z);
    cdir.x /= len; cdir.y /= len; cdir.z /= len;
  }
  const double screen_width = wsw * cam.image_width / cam.image_height;
  const double screen_height = wsh;
  const V3 sx = normalize(cross(cdir, cup)) * screen_width;
  const V3 sy = normalize(cross(sx, cdir)) * screen_height;
  put(P.cam_pos, cpos);
  put(P.screen_x, sx);
  put(P.screen_y, sy);
  put(P.center, cpos + cdir * cam.world_screen_dist);
  P.iw = 1.0 / cam.image_width;
  P.ih = 1.0 / cam.image_height;
  P.W = cam.image_width;
  P.H = cam.image_height;
  P.SS = smp.supersamples;
  P.samples = smp.samples;
  P.spp = smp.samples * smp.supersamples * smp.supersamples;
  P.depth = smp.depth;
  P.depth_limit = smp.depth_limit;
  P.nee = smp.next_event_estimation ? 1 : 0;
  P.max_vertices = 128;
  P.n_lights = (int32_t)s->h_lights.size();
  P.n_materials = (int32_t)s->h_mats.size();
  P.n_analytic = (int32_t)analytic.size();
  P.n_nodes = (int32_t)bo.nodes.size();
  P.smp = 1.0 / smp.samples / (smp.supersamples * smp.supersamples);
  P.sensor_response = smp.sensor_response;
  P.cutoff = (float)smp.luminance_cutoff;  // a float in SceneEnv (scene_env.h:221)
  std::memcpy(P.background, smp.background_color, sizeof(P.background));
  std::memcpy(P.env_light, smp.environment_light, sizeof(P.env_light));
  P.nodes = (const DevNode*)s->nodes.p;
  P.tris = (const float*)s->tris.p;
  P.tri_nrm = (const float*)s->tri_nrm.p;
  P.tri_alb = any_albedo ? (const float*)s->tri_alb.p : nullptr;
  P.ents = (const DevEntity*)s->ents.p;
  P.mats = (const DevMaterial*)s->mats.p;
  P.analytic = (const int32_t*)s->analytic.p;
  P.lights = (const int32_t*)s->lights.p;
  P.stack_ovf = (uint32_t*)s->ovf.p;
  P.ovf_threads = (int32_t)ovf_threads;
  P.ovf_words = ovf_words;
  P.n_top = P.n_nodes > 0 ? std::min(bo.n_top, PRT_TOP_NODES) : 0;
  P.n_tris = (int32_t)s->n_tris;
  return PRT_OK;
}

prt_scene* prt_scene_create(const prt_scene_desc* desc) {
  prt_scene* s = new (std::nothrow) prt_scene();
  if (!s) { g_err = "out of host memory"; return nullptr; }
  int rc;
  try {
    rc = scene_build(s, desc);
  } catch (const std::bad_alloc&) {
    rc = fail(PRT_E_NOMEM, "out of host memory while building the scene");
  }
  if (rc != PRT_OK) {
    std::string keep = g_err;
    prt_scene_destroy(s);
    g_err = keep;
    return nullptr;
  }
  return s;
}

int prt_scene_info(const prt_scene* s, int64_t* n_nodes, int64_t* n_tris, int64_t* n_lights, int64_t* bvh_bytes) {
  if (!s) return fail(PRT_E_INVALID, "null scene");
  if (n_nodes) *n_nodes = s->P.n_nodes;
  if (n_tris) *n_tris = s->n_tris;
  if (n_lights) *n_lights = (int64_t)s->h_lights.size();
  if (bvh_bytes) *bvh_bytes = (int64_t)s->P.n_nodes * (int64_t)sizeof(DevNode) + s->n_tris * 48;
  return PRT_OK;
}

extern "C++" void prt::fill_stats(prt_stats* st, const uint64_t* c) {
  st->node_visits += c[CTR_NODES] + c[CTR_SNODES];
  st->tri_tests += c[CTR_TRIS] + c[CTR_STRIS];
  st->shadow_node_visits += c[CTR_SNODES];
  st->shadow_tri_tests += c[CTR_STRIS];
  st->pickup_hits += c[CTR_PICKUP];
  st->truncated_paths += c[CTR_TRUNC];
}

static int trace_device(prt_scene* s, const prt_ray* d_rays, int64_t n, prt_hit* d_hits, prt_stats* stats) {
  HIPCHK(hipSetDevice(s->device));
  std::lock_guard<std::recursive_mutex> pool_lock(s->pool->m);
  int rc;
  if ((rc = dev_alloc(s->pool->ws.ctr, CTR_COUNT * sizeof(uint64_t)))) return rc;
  HIPCHK(launch_fill_u64((uint64_t*)s->pool->ws.ctr.p, 0ull, CTR_COUNT, s->stream));
  ScopedEvent e0, e1;
  HIPCHK(hipEventCreate(&e0.e));
  HIPCHK(hipEventCreate(&e1.e));
  HIPCHK(hipEventRecord(e0.e, s->stream));
  if (n > 0)
    HIPCHK(launch_trace(s->P, (const prt_ray_dev*)d_rays, n, (prt_hit_dev*)d_hits, (uint64_t*)s->pool->ws.ctr.p,
                        s->max_trav_blocks, s->stream));
  HIPCHK(hipEventRecord(e1.e, s->stream));
  uint64_t c[CTR_COUNT];
  HIPCHK(hipMemcpyAsync(c, s->pool->ws.ctr.p, sizeof(c), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0.e, e1.e));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    fill_stats(stats, c);
    stats->extension_rays = (uint64_t)n;
    stats->ms_total = ms;
    stats->ms_extend = ms;
    stats->launches_extend = 1;
    stats->bytes_extend = 128 * stats->node_visits + 48 * stats->tri_tests + 48 * (uint64_t)n;
    stats->bytes_algorithmic = stats->bytes_extend;
  }
  return PRT_OK;
}

int prt_trace_closest_device(prt_scene* s, const prt_ray* d_rays, int64_t n, prt_hit* d_hits, prt_stats* stats) {
  if (!s || n < 0 || (n > 0 && (!d_rays || !d_hits))) return fail(PRT_E_INVALID, "bad arguments");
  return trace_device(s, d_rays, n, d_hits, stats);
}

// The same traversal as prt_trace_closest_device with nothing allocated in
// the call: the counters are the library's per-GPU scratch array (cleared by
// a fill kernel, so every call starts from zero whatever the last call left)
// and no timing events are created.  Calls that overlap on one GPU share
// those counters (their hits stay exact; their counts mix).
int prt_trace_closest_device_noalloc(prt_scene* s, const prt_ray* d_rays, int64_t n, prt_hit* d_hits,
                                     prt_stats* stats) {
  if (!s || n < 0 || (n > 0 && (!d_rays || !d_hits))) return fail(PRT_E_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->device));
  uint64_t* ctr = nullptr;
  HIPCHK(noalloc_scratch(s->device, &ctr));
  HIPCHK(launch_fill_u64(ctr, 0, CTR_COUNT, s->stream));
  if (n > 0)
    HIPCHK(launch_trace(s->P, (const prt_ray_dev*)d_rays, n, (prt_hit_dev*)d_hits, ctr, s->max_trav_blocks,
                        s->stream));
  uint64_t c[CTR_COUNT];
  HIPCHK(hipMemcpyAsync(c, ctr, sizeof(c), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    fill_stats(stats, c);
    stats->extension_rays = (uint64_t)n;
    stats->launches_extend = n > 0 ? 1 : 0;
    stats->bytes_extend = 128 * stats->node_visits + 48 * stats->tri_tests + 48 * (uint64_t)n;
    stats->bytes_algorithmic = stats->bytes_extend;
  }
  return PRT_OK;
}

int prt_trace_closest(prt_scene* s, const prt_ray* rays, int64_t n, prt_hit* hits, prt_stats* stats) {
  if (!s || n < 0 || (n > 0 && (!rays || !hits))) return fail(PRT_E_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->device));
  void *dr = nullptr, *dh = nullptr;
  if (n > 0) {
    HIPCHK(hipMalloc(&dr, (size_t)n * sizeof(prt_ray)));
    hipError_t e = hipMalloc(&dh, (size_t)n * sizeof(prt_hit));
    if (e != hipSuccess) { (void)hipFree(dr); return fail(PRT_E_NOMEM, hipGetErrorString(e)); }
    e = hipMemcpyAsync(dr, rays, (size_t)n * sizeof(prt_ray), hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) { (void)hipFree(dr); (void)hipFree(dh); return fail(PRT_E_HIP, hipGetErrorString(e)); }
  }
  int rc = trace_device(s, (const prt_ray*)dr, n, (prt_hit*)dh, stats);
  if (rc == PRT_OK && n > 0) {
    hipError_t e = hipMemcpyAsync(hits, dh, (size_t)n * sizeof(prt_hit), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(PRT_E_HIP, hipGetErrorString(e));
  }
  if (dr) (void)hipFree(dr);
  if (dh) (void)hipFree(dh);
  return rc;
}

// ---------------------------------------------------------------- render
extern "C++" int prt::ensure_workspace(prt_scene* s, int64_t batch_samples, bool two_accum) {
  Workspace& w = s->pool->ws;
  // A slab writer over input chunks [c0, c1) owns fan * 64 * (c1 - c0)
  // entries from fan * 64 * c0 on, so a launch's outputs span fan * 64 *
  // (input chunks) <= fan * (input rays + 64 * PRT_SLAB_WAVES).  With glass a
  // vertex at depth <= 2 spawns two children (fan 2), deeper ones one, so
  // fan * input rays <= 8x the samples (2 * 4x at depth 2); without glass 1x.
  // Shadow rays: three per ray of the NEE input.
  const int64_t grow = s->has_glass ? 8 : 1;
  const int64_t W = PRT_SLAB_WAVES;
  int64_t qcap = grow * batch_samples + 2 * 64 * W + 64;
  int64_t scap = 3 * (grow * batch_samples + 64 * W);
  if (qcap >= ((int64_t)1 << 32) || scap >= ((int64_t)1 << 32)) return fail(PRT_E_INVALID, "batch too large");
  // the second shadow queue of the alternating schedule (NEE scenes only)
  bool two_sq = PRT_OVERLAP_SHADOW == 3 && s->P.nee && s->P.n_lights > 0;
  if (w.batch_samples >= batch_samples && w.queue_cap >= qcap && w.shadow_cap >= scap && w.h_counts &&
      (!two_accum || w.accum2.p) && (!two_sq || w.sq2[0].p))
    return PRT_OK;
  // The device's workspace serves every scene: it only grows (to the largest
  // sizes any call needed, keeping the optional buffers once they exist), so
  // scenes of different sizes in turn do not reallocate it back and forth.
  batch_samples = std::max(batch_samples, w.batch_samples);
  qcap = std::max(qcap, w.queue_cap);
  scap = std::max(scap, w.shadow_cap);
  two_sq = two_sq || w.sq2[0].p != nullptr;
  two_accum = two_accum || w.accum2.p != nullptr;
  // recorded sizes only once every buffer holds them: a failed allocation
  // has released its buffer, and the next call must not take the old sizes
  // for granted
  w.batch_samples = w.queue_cap = w.shadow_cap = 0;
  int rc;
  for (int qi = 0; qi < 2; ++qi)
    if ((rc = alloc_ray_queue(w.q[qi], qcap))) return rc;
  if ((rc = dev_alloc(w.hit_t, (size_t)qcap * 8))) return rc;
  if ((rc = dev_alloc(w.hit_id, (size_t)qcap * 4))) return rc;
  if ((rc = alloc_shadow_queue(w.sq, scap))) return rc;
  if (two_sq && (rc = alloc_shadow_queue(w.sq2, scap))) return rc;
  // chunks: a slab's last chunk may be partly filled
  for (int k = 0; k < (two_sq ? 4 : 3); ++k)
    if ((rc = dev_alloc(w.desc[k], (size_t)((k < 2 ? qcap : scap) / 64 + W + 1) * sizeof(uint2)))) return rc;
  if ((rc = dev_alloc(w.slabs, (size_t)3 * W * sizeof(uint2)))) return rc;
  if ((rc = dev_alloc(w.accum, (size_t)batch_samples * 3 * sizeof(uint64_t)))) return rc;
  if (two_accum && (rc = dev_alloc(w.accum2, (size_t)batch_samples * 3 * sizeof(uint64_t)))) return rc;
  if ((rc = dev_alloc(w.counts, 16 * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(w.fetch, (size_t)FETCH_LAUNCHES * PRT_FETCH_COUNTERS * PRT_FETCH_STRIDE * sizeof(uint32_t))))
    return rc;
  if ((rc = dev_alloc(w.ctr, CTR_COUNT * sizeof(uint64_t)))) return rc;
  if (!w.h_counts) HIPCHK(hipHostMalloc((void**)&w.h_counts, 16 * sizeof(uint32_t), hipHostMallocDefault));
  w.batch_samples = batch_samples;
  w.queue_cap = qcap;
  w.shadow_cap = scap;
  return PRT_OK;
}

// Settings of one render call derived from the scene and the options.
struct RenderSetup {
  Params P;
  int64_t n_pix = 0, bp = 0;
};

static int setup_render(prt_scene* s, const prt_render_opts* o, RenderSetup& R) {
  prt_render_opts opts{};
  if (o) opts = *o;
  Params P = s->P;
  if (opts.samples_override > 0) {
    P.samples = opts.samples_override;
    P.spp = P.samples * P.SS * P.SS;
    P.smp = 1.0 / P.samples / (P.SS * P.SS);
  }
  if (P.spp > (1 << 20)) return fail(PRT_E_INVALID, "samples per pixel above 2^20");
  P.max_vertices = opts.max_depth > 0 ? std::min(opts.max_depth, 255) : 128;
  P.key0 = (uint32_t)opts.seed;
  P.key1 = (uint32_t)(opts.seed >> 32);
  R.n_pix = (int64_t)P.W * P.H;
  // 2^PRT_BATCH_LOG2 camera samples per batch; half with glass, whose queues
  // are sized for up to 8 path segments per sample
  const int64_t batch_samples = (int64_t)1 << (s->has_glass ? PRT_BATCH_LOG2 - 1 : PRT_BATCH_LOG2);
  int64_t bp = opts.batch_pixels > 0 ? opts.batch_pixels : std::max<int64_t>(1, batch_samples / P.spp);
  bp = std::min(bp, R.n_pix);
  // queues hold up to 16x (glass splits into slabs) and 24x (shadow rays)
  // the batch, indexed by uint32
  while (bp > 1 && bp * P.spp * 32 >= ((int64_t)1 << 32)) bp /= 2;
  R.bp = bp;
  R.P = P;
  return PRT_OK;
}

// Depths enqueued between host reads of the ray count (a depth with an
// empty queue costs only near-empty launches).
#ifndef PRT_DEPTHS_PER_SYNC
#define PRT_DEPTHS_PER_SYNC 4
#endif

// One batch through the wavefront: generate, then extend / shade / NEE /
// shadow per path depth until no ray is left.  Sample radiance accumulates
// into `accum` (fixed point, 3 per camera sample of the batch).  Each
// slab-writing kernel is followed by k_slab_index, which lists the chunks
// of the queue it wrote for the kernels that read it.  The host reads the
// ray count once per PRT_DEPTHS_PER_SYNC depths.  Without the adjoint (which
// reads the partial sums of a sample in path order) a depth's shadow rays are
// traced on the scene's second stream as PRT_OVERLAP_SHADOW selects; a NEE
// that rewrites a shadow queue waits for the shadow launch that read it, and
// the batch's end waits for every shadow launch.
static int run_batch_on(prt_scene* s, const Params& P, const Batch& B, uint64_t* accum, const Adjoint& A, Timer& tm,
                        CallTotals& tot) {
  Workspace& w = s->pool->ws;
  hipStream_t st = s->stream;
  const int overlap = A.on ? 0 : PRT_OVERLAP_SHADOW;
  hipStream_t st_sh = overlap ? s->stream2 : st;
  uint32_t* counts = (uint32_t*)w.counts.p;
  uint32_t* fetch = (uint32_t*)w.fetch.p;
  uint2* slabs = (uint2*)w.slabs.p;
  uint64_t* ctr = (uint64_t*)w.ctr.p;
  RayQ q[2] = {ray_queue(w.q[0], w.queue_cap), ray_queue(w.q[1], w.queue_cap)};
  ShadowQ sq = shadow_queue(w.sq, w.shadow_cap);
  const int tb = s->max_trav_blocks;
  // the shadow grid on the second stream leaves room on each CU; on the main
  // stream (no overlap) it is the closest-hit grid
  const int tb_sh = overlap ? std::max(1, s->cu_count) * PRT_SHADOW_BLOCKS_PER_CU : tb;
  const bool nee = P.nee && P.n_lights > 0;
  const uint32_t W = PRT_SLAB_WAVES;
  const size_t fetch_launch = (size_t)PRT_FETCH_COUNTERS * PRT_FETCH_STRIDE;
  // children per shaded vertex: glass splits in two up to depth 2
  auto fan = [&](int depth) { return (uint32_t)(s->has_glass && depth <= 2 ? 2 : 1); };
  auto chunks = [&](int k) { return ChunkList{(const uint2*)w.desc[k].p, counts + k}; };
  auto index = [&](int k, uint32_t n_slabs, const uint2* sl, hipStream_t on) {
    const uint32_t max_chunks = (uint32_t)(w.desc[k].bytes / sizeof(uint2));
    return launch_slab_index(sl, n_slabs, (uint2*)w.desc[k].p, max_chunks, counts + k, counts + 4 + k, ctr,
                             k < 2 ? CTR_QUEUED : -1, on);
  };
  int fi = 0;  // traversal launches so far; each takes fresh claim counters
  auto next_fetch = [&](uint32_t** f) -> int {
    if (fi >= FETCH_LAUNCHES) return fail(PRT_E_INVALID, "too many wavefront launches in one batch");
    if (fi % FETCH_GROUP == 0) HIPCHK(launch_fill_u32(fetch + fi * fetch_launch, 0u, FETCH_GROUP * fetch_launch, st));
    *f = fetch + fi++ * fetch_launch;
    return PRT_OK;
  };
  int rc;
  uint32_t* f;
  // NEE of the rays in q[k] and the any-hit tests of its shadow rays
  auto nee_and_shadow = [&](int k) -> int {
    if (overlap == 2) {
      // both on the second stream, after the shade that wrote q[k]; the
      // claim counters are cleared on the main stream before the hand-over
      if ((rc = next_fetch(&f))) return rc;
      HIPCHK(hipEventRecord(s->ev_shade, st));
      HIPCHK(hipStreamWaitEvent(st_sh, s->ev_shade, 0));
      if ((rc = tm.begin_side(st_sh))) return rc;
      HIPCHK(launch_nee(P, B, q[k], chunks(k), sq, slabs + 2 * W, ctr, st_sh));
      HIPCHK(index(2, W, slabs + 2 * W, st_sh));
      if ((rc = tm.end_side(st_sh, T_NEE))) return rc;
      HIPCHK(hipEventRecord(s->ev_nee[k], st_sh));
      if ((rc = tm.begin_side(st_sh))) return rc;
      HIPCHK(launch_shadow(P, B, sq, chunks(2), f, accum, A, ctr, tb_sh, st_sh));
      if ((rc = tm.end_side(st_sh, T_SHADOW))) return rc;
      HIPCHK(hipEventRecord(s->ev_shadow, st_sh));
      return PRT_OK;
    }
    if (overlap) HIPCHK(hipStreamWaitEvent(st, s->ev_shadow, 0));  // the last shadow launch has read sq
    HIPCHK(launch_nee(P, B, q[k], chunks(k), sq, slabs + 2 * W, ctr, st));
    HIPCHK(index(2, W, slabs + 2 * W, st));
    if ((rc = tm.mark(T_NEE))) return rc;
    if ((rc = next_fetch(&f))) return rc;
    if (overlap) {
      HIPCHK(hipEventRecord(s->ev_shade, st));
      HIPCHK(hipStreamWaitEvent(st_sh, s->ev_shade, 0));
      if ((rc = tm.begin_side(st_sh))) return rc;
    }
    HIPCHK(launch_shadow(P, B, sq, chunks(2), f, accum, A, ctr, tb_sh, st_sh));
    if (overlap) {
      if ((rc = tm.end_side(st_sh, T_SHADOW))) return rc;
      HIPCHK(hipEventRecord(s->ev_shadow, st_sh));
    } else if ((rc = tm.mark(T_SHADOW))) {
      return rc;
    }
    return PRT_OK;
  };
  // Alternating schedule (overlap 3): the NEE of q[k] on the main stream
  // into shadow queue sh_next, whose shadow rays are traced on the second
  // stream after the next depth's closest-hit launch (shadow_alt).
  const bool alt = overlap == 3 && nee;
  const ShadowQ sqs[2] = {sq, alt ? shadow_queue(w.sq2, w.shadow_cap) : sq};
  int sh_next = 0, sh_pending = -1;
  bool sh_used[2] = {false, false};
  auto nee_alt = [&](int k) -> int {
    const int b = sh_next;
    sh_next ^= 1;
    if (sh_used[b]) HIPCHK(hipStreamWaitEvent(st, s->ev_sh[b], 0));  // its last shadow launch has read it
    HIPCHK(launch_nee(P, B, q[k], chunks(k), sqs[b], slabs + 2 * W, ctr, st));
    HIPCHK(index(2 + b, W, slabs + 2 * W, st));
    if ((rc = tm.mark(T_NEE))) return rc;
    sh_pending = b;
    return PRT_OK;
  };
  // the pending shadow queue's rays, after everything enqueued so far on the
  // main stream (on the main stream itself when `side` is false)
  auto shadow_alt = [&](bool side) -> int {
    if (sh_pending < 0) return PRT_OK;
    const int b = sh_pending;
    sh_pending = -1;
    if ((rc = next_fetch(&f))) return rc;  // a fill clearing its counters goes on st first
    hipStream_t on = side ? st_sh : st;
    if (side) {
      HIPCHK(hipEventRecord(s->ev_shade, st));
      HIPCHK(hipStreamWaitEvent(on, s->ev_shade, 0));
      if ((rc = tm.begin_side(on))) return rc;
    }
    HIPCHK(launch_shadow(P, B, sqs[b], chunks(2 + b), f, accum, A, ctr, side ? tb_sh : tb, on));
    if (side) {
      if ((rc = tm.end_side(on, T_SHADOW))) return rc;
      HIPCHK(hipEventRecord(s->ev_sh[b], on));
      sh_used[b] = true;
    } else if ((rc = tm.mark(T_SHADOW))) {
      return rc;
    }
    return PRT_OK;
  };
  const int64_t ns = B.n_pixels * P.spp;
  const int64_t np = B.n_pixels * P.SS * P.SS;
  HIPCHK(launch_fill_u64(accum, 0ull, 3 * ns, st));
  HIPCHK(launch_fill_u32(counts, 0u, 16, st));
  // the primary rays: one slab of np rays at the start of q[0]
  HIPCHK(launch_fill_u32((uint32_t*)slabs, 0u, 1, st));
  HIPCHK(launch_fill_u32((uint32_t*)slabs + 1, (uint32_t)np, 1, st));
  if (P.mlt_u) HIPCHK(launch_mlt_gen(P, (uint32_t)np, q[0], st));
  else HIPCHK(launch_gen_primary(P, B, q[0], st));
  HIPCHK(index(0, 1, slabs, st));
  if ((rc = tm.mark(T_OTHER))) return rc;
  if ((rc = next_fetch(&f))) return rc;
  HIPCHK(launch_extend(P, q[0], chunks(0), np, f, (double*)w.hit_t.p, (int32_t*)w.hit_id.p, ctr, tb, st));
  if ((rc = tm.mark(T_EXTEND0))) return rc;
  tot.primary_rays += (uint64_t)np;
  tot.launches++;
  // Shade q[kin] (primary: every camera sample against its sub-pixel's
  // primary hit) into q[kout], then the NEE of the children.
  auto shade_nee = [&](int primary, int kin, int kout, uint32_t fn) -> int {
    HIPCHK(launch_shade(P, B, primary, q[kin], chunks(kin), (const double*)w.hit_t.p, (const int32_t*)w.hit_id.p,
                        q[kout], fn, slabs + kout * W, accum, A, ctr, st));
    HIPCHK(index(kout, W, slabs + kout * W, st));
    if ((rc = tm.mark(T_SHADE))) return rc;
    return nee ? (alt ? nee_alt(kout) : nee_and_shadow(kout)) : PRT_OK;
  };
  if ((rc = shade_nee(1, 0, 1, fan(0)))) return rc;
  int cur = 1;
  for (int depth = 1;; ++depth) {  // depth of the rays in q[cur]
    if ((depth - 1) % PRT_DEPTHS_PER_SYNC == 0) {
      HIPCHK(hipMemcpyAsync(w.h_counts, counts, 16 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if (w.h_counts[4 + cur] == 0) break;
    }
    const int nxt = cur ^ 1;
    if ((rc = next_fetch(&f))) return rc;
    if ((rc = tm.mark(T_OTHER))) return rc;
    HIPCHK(launch_extend(P, q[cur], chunks(cur), (int64_t)tb * 256, f, (double*)w.hit_t.p, (int32_t*)w.hit_id.p, ctr,
                         tb, st));
    if ((rc = tm.mark(T_EXTEND))) return rc;
    tot.launches++;
    // the shadow rays of q[cur]'s NEE, alongside this depth's shade and NEE
    if (alt && (rc = shadow_alt(true))) return rc;
    // the NEE two depths back read q[nxt] and its chunk list
    if (overlap == 2 && nee) HIPCHK(hipStreamWaitEvent(st, s->ev_nee[nxt], 0));
    if ((rc = shade_nee(0, cur, nxt, fan(depth)))) return rc;
    cur = nxt;
  }
  if (alt) {
    if ((rc = shadow_alt(false))) return rc;
    for (int b = 0; b < 2; ++b)
      if (sh_used[b]) HIPCHK(hipStreamWaitEvent(st, s->ev_sh[b], 0));
  } else if (overlap) {
    HIPCHK(hipStreamWaitEvent(st, s->ev_shadow, 0));
  }
  return PRT_OK;
}

extern "C++" int prt::run_batch(prt_scene* s, const Params& P, const Batch& B, uint64_t* accum, const Adjoint& A, Timer& tm,
                   CallTotals& tot) {
  const int rc = run_batch_on(s, P, B, accum, A, tm, tot);
  // a failed batch may have left shadow work queued on the second stream
  if (rc != PRT_OK) (void)hipStreamSynchronize(s->stream2);
  return rc;
}

static int finish_stats(prt_scene* s, const RenderSetup& R, Timer& tm, const CallTotals& tot, prt_stats* stats,
                        uint64_t* c) {
  hipStream_t st = s->stream;
  HIPCHK(hipMemcpyAsync(c, s->pool->ws.ctr.p, CTR_COUNT * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (c[CTR_DROP] != 0) return fail(PRT_E_INVALID, "ray queue overflow (" + std::to_string(c[CTR_DROP]) + " rays)");
  // every queued ray must have been traced exactly once
  if (c[CTR_TRACED] != c[CTR_QUEUED] || c[CTR_STRACED] != c[CTR_SHADOW])
    return fail(PRT_E_INTERNAL, "traversal finished " + std::to_string(c[CTR_TRACED]) + " of " +
                                    std::to_string(c[CTR_QUEUED]) + " rays and " +
                                    std::to_string(c[CTR_STRACED]) + " of " + std::to_string(c[CTR_SHADOW]) +
                                    " shadow rays");
  double cat[T_NCAT] = {0, 0, 0, 0, 0, 0};
  double trav = 0.0;
  const double wall = tm.collect(cat, &trav, 1u << T_EXTEND | 1u << T_EXTEND0 | 1u << T_SHADOW);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    fill_stats(stats, c);
    stats->camera_samples = (uint64_t)R.n_pix * R.P.spp;
    stats->primary_rays = tot.primary_rays;
    stats->extension_rays = c[CTR_QUEUED] - tot.primary_rays;
    stats->shadow_rays = c[CTR_SHADOW];
    stats->ms_extend = cat[T_EXTEND] + cat[T_EXTEND0];
    stats->ms_shade = cat[T_SHADE];
    stats->ms_shadow = cat[T_SHADOW];
    stats->ms_other = cat[T_OTHER];
    stats->ms_extend_primary = cat[T_EXTEND0];
    stats->ms_nee = cat[T_NEE];
    stats->ms_total = wall;
    stats->ms_traversal = trav;
    stats->launches_extend = tot.launches;
    // per ray query: 128 B per node record, 48 B per triangle record, 48 B
    // of ray in + hit out
    stats->bytes_extend = 128 * c[CTR_NODES] + 48 * c[CTR_TRIS] + 48 * c[CTR_QUEUED];
    stats->bytes_shadow = 128 * c[CTR_SNODES] + 48 * c[CTR_STRIS] + 48 * c[CTR_SHADOW];
    stats->bytes_algorithmic = stats->bytes_extend + stats->bytes_shadow;
  }
  return PRT_OK;
}

// Progress images of a render with prt_dump_opts (Render::Run's periodic
// OutputBmp, render.h:1390-1401): the device image is copied to the host
// between batches once the interval has passed.
struct ImageDumper {
  const prt_dump_opts* d = nullptr;
  std::chrono::steady_clock::time_point last;
  int32_t index = 0;
  std::vector<float> host;
  std::vector<double> host64;

  bool on() const { return d != nullptr && d->seconds >= 0.0; }
  bool due() const {
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - last).count();
    return d->seconds == 0.0 || el >= d->seconds;
  }
  int dump(prt_scene* s, const void* d_rgb, int f64, double fraction) {
    const int W = s->P.W, H = s->P.H;
    const size_t n = (size_t)W * H * 3;
    host.resize(n);
    HIPCHK(hipStreamSynchronize(s->stream));
    if (f64) {
      host64.resize(n);
      HIPCHK(hipMemcpy(host64.data(), d_rgb, n * sizeof(double), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) host[i] = (float)host64[i];
    } else {
      HIPCHK(hipMemcpy(host.data(), d_rgb, n * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (d->callback) d->callback(d->user, host.data(), W, H, index, fraction);
    if (d->directory) {
      char name[32];
      std::snprintf(name, sizeof(name), "output_W%06d.bmp", (int)index);
      const std::string path = std::string(d->directory) + "/" + name;
      const int rc = prt_write_bmp(path.c_str(), host.data(), W, H, d->gamma_offset);
      if (rc != PRT_OK) return rc;
    }
    ++index;
    last = std::chrono::steady_clock::now();
    return PRT_OK;
  }
};

static int render_device(prt_scene* s, const prt_render_opts* o, float* d_rgb, float* d_var, prt_stats* stats,
                         ImageDumper* dumper = nullptr) {
  HIPCHK(hipSetDevice(s->device));
  std::lock_guard<std::recursive_mutex> pool_lock(s->pool->m);
  // Scenes with glass stream through the batched renderer's tile queue: a
  // depth-synchronous batch of them drains through long, thin tails after
  // the glass splits, which the stream keeps filled with new tiles.  Same
  // image bits either way (fixed-point sums, counter-based random numbers).
  if (o && o->reserved_ != 0) return fail(PRT_E_INVALID, "prt_render_opts.reserved_ must be 0");
  const int engine = o ? o->engine : 0;
  const bool stream_ok = !o || (o->image_f64 == 0 && o->samples_override <= 0 && o->batch_pixels <= 0);
  if (engine == 2 && !stream_ok)
    return fail(PRT_E_INVALID, "the streaming engine writes float images without samples_override / batch_pixels");
  const bool dumping = dumper && dumper->on();
  if (dumping && (engine == 2 || engine == 3))
    return fail(PRT_E_INVALID, "progress images need the depth-synchronous engine");
  if (engine < 0 || engine > 3) return fail(PRT_E_INVALID, "engine must be 0..3");
  // Scenes without next-event estimation render fastest in one persistent
  // launch that follows each sample's whole path (cfg1: 1.64 against 2.71 ms
  // per frame); with NEE its lanes trace their shadow rays one after another
  // and the queued engines win.
  const bool nee_scene = s->P.nee && s->P.n_lights > 0;
  const bool persist_ok = stream_ok && (uint64_t)s->P.W * s->P.H * s->P.spp <= ((uint64_t)1 << 28);
  if (engine == 3 && !persist_ok)
    return fail(PRT_E_INVALID, "the persistent engine writes float images of at most 2^28 samples, without "
                               "samples_override / batch_pixels");
  if (engine == 3 || (engine == 0 && !nee_scene && persist_ok && !dumping)) {
    prt_batch_opts b{};
    b.seed = o ? o->seed : 0;
    b.max_depth = o ? o->max_depth : 0;
    b.persistent = 1;
    b.trav_blocks_per_cu = 2;  // the compiler's own register budget (fastest without NEE)
    prt_scene* const one[1] = {s};
    float* const rgb[1] = {d_rgb};
    float* const var[1] = {d_var};
    return prt_render_batch_device(one, 1, &b, rgb, var, stats);
  }
  if (engine == 2 || (engine == 0 && s->has_glass && stream_ok && !dumping)) {
    prt_batch_opts b{};
    b.seed = o ? o->seed : 0;
    b.max_depth = o ? o->max_depth : 0;
    b.timing = o ? o->timing : 0;
    b.use_graph = -1;            // plain launches (faster than graph replay; timed renders need them)
    // samples in flight: a whole frame of up to 2^21, else the batched
    // renderer's automatic size (the most, up to 2^24, whose workspace fits
    // its share of the device memory)
    const int64_t frame = (int64_t)s->P.W * s->P.H * s->P.spp;
    b.tile_samples = 4096;
    b.inflight_tiles = frame <= ((int64_t)1 << 21) ? (int32_t)std::max<int64_t>(1, (frame + 4095) / 4096) : 0;
    prt_scene* const one[1] = {s};
    float* const rgb[1] = {d_rgb};
    float* const var[1] = {d_var};
    return prt_render_batch_device(one, 1, &b, rgb, var, stats);
  }
  RenderSetup R;
  int rc;
  if ((rc = setup_render(s, o, R))) return rc;
  if ((rc = ensure_workspace(s, R.bp * R.P.spp, false))) return rc;
  hipStream_t st = s->stream;
  HIPCHK(launch_fill_u64((uint64_t*)s->pool->ws.ctr.p, 0ull, CTR_COUNT, st));
  Timer tm;
  if ((rc = tm.begin(o && o->timing != 0, st, &s->timer_events))) return rc;
  CallTotals tot;
  Adjoint none{};
  const int f64 = o && o->image_f64;
  if (dumping) {
    // pixels not yet rendered read as 0 in the progress images
    const int64_t n = R.n_pix * 3;
    if (f64)
      HIPCHK(launch_fill_u64((uint64_t*)d_rgb, 0ull, n, st));
    else
      HIPCHK(launch_fill_f32(d_rgb, 0.0f, n, st));
    dumper->last = std::chrono::steady_clock::now();
  }
  for (int64_t pb = 0; pb < R.n_pix; pb += R.bp) {
    Batch B{pb, std::min(R.bp, R.n_pix - pb)};
    if ((rc = run_batch(s, R.P, B, (uint64_t*)s->pool->ws.accum.p, none, tm, tot))) return rc;
    HIPCHK(launch_resolve(R.P, B, (const uint64_t*)s->pool->ws.accum.p, d_rgb, d_var, f64, st));
    if ((rc = tm.mark(T_OTHER))) return rc;
    const int64_t done = pb + B.n_pixels;
    if (dumping && done < R.n_pix && dumper->due() &&
        (rc = dumper->dump(s, d_rgb, f64, (double)done / (double)R.n_pix)))
      return rc;
  }
  uint64_t c[CTR_COUNT];
  return finish_stats(s, R, tm, tot, stats, c);
}

int prt_render_device(prt_scene* s, const prt_render_opts* opts, float* d_rgb, float* d_var, prt_stats* stats) {
  if (!s || !d_rgb) return fail(PRT_E_INVALID, "bad arguments");
  return render_device(s, opts, d_rgb, d_var, stats);
}

int prt_render(prt_scene* s, const prt_render_opts* opts, float* out_rgb, float* out_var, prt_stats* stats) {
  if (!s || !out_rgb) return fail(PRT_E_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->device));
  const size_t bytes = (size_t)s->P.W * s->P.H * 3 * sizeof(float);
  int rc;
  if ((rc = dev_alloc(s->out_rgb, bytes))) return rc;
  if (out_var && (rc = dev_alloc(s->out_var, bytes))) return rc;
  prt_render_opts o32{};
  if (opts) o32 = *opts;
  o32.image_f64 = 0;  // host outputs are float
  rc = render_device(s, &o32, (float*)s->out_rgb.p, out_var ? (float*)s->out_var.p : nullptr, stats);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out_rgb, s->out_rgb.p, bytes, hipMemcpyDeviceToHost, s->stream));
  if (out_var) HIPCHK(hipMemcpyAsync(out_var, s->out_var.p, bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return PRT_OK;
}

int prt_render_dump(prt_scene* s, const prt_render_opts* opts, const prt_dump_opts* dump, float* out_rgb,
                    float* out_var, prt_stats* stats) {
  if (!s || !out_rgb) return fail(PRT_E_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->device));
  const size_t bytes = (size_t)s->P.W * s->P.H * 3 * sizeof(float);
  int rc;
  if ((rc = dev_alloc(s->out_rgb, bytes))) return rc;
  if (out_var && (rc = dev_alloc(s->out_var, bytes))) return rc;
  prt_render_opts o32{};
  if (opts) o32 = *opts;
  o32.image_f64 = 0;  // host outputs are float
  ImageDumper dumper;
  dumper.d = dump;
  rc = render_device(s, &o32, (float*)s->out_rgb.p, out_var ? (float*)s->out_var.p : nullptr, stats, &dumper);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out_rgb, s->out_rgb.p, bytes, hipMemcpyDeviceToHost, s->stream));
  if (out_var) HIPCHK(hipMemcpyAsync(out_var, s->out_var.p, bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return PRT_OK;
}

// ---------------------------------------------------------------- differentiable parameters
int prt_set_face_albedo_device(prt_scene* s, int entity, const float* d_albedo) {
  if (!s || !d_albedo || entity < 0 || entity >= (int)s->h_ents.size()) return fail(PRT_E_INVALID, "bad arguments");
  const DevEntity& e = s->h_ents[entity];
  if (e.type != PRT_ENTITY_POLYGON || !e.has_albedo)
    return fail(PRT_E_INVALID, "entity is not a POLYGON created with per-face albedo");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(launch_scatter_albedo((float*)s->tri_alb.p, (const uint32_t*)s->d_slot_of_tri.p, s->ent_first_tri[entity],
                               s->ent_n_tris[entity], d_albedo, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return PRT_OK;
}

static void sort_lights(prt_scene* s) {
  struct L { int32_t ent; double power; };
  std::vector<L> ls;
  for (int i = 0; i < (int)s->h_ents.size(); ++i) {
    const DevEntity& de = s->h_ents[i];
    if (!de.light || de.type == PRT_ENTITY_POLYGON) continue;
    ls.push_back(L{i, de.area * length(from(s->h_mats[de.material].emission))});
  }
  std::stable_sort(ls.begin(), ls.end(), [](const L& a, const L& b) { return a.power > b.power; });
  s->h_lights.clear();
  for (const L& l : ls) s->h_lights.push_back(l.ent);
}

int prt_hold_rr_albedo(prt_scene* s, int hold) {
  if (!s) return fail(PRT_E_INVALID, "null scene");
  HIPCHK(hipSetDevice(s->device));
  if (!hold) {
    s->P.tri_alb_rr = nullptr;
    return PRT_OK;
  }
  if (!s->tri_alb.p) return fail(PRT_E_INVALID, "scene has no per-face albedo");
  const size_t bytes = (size_t)s->n_tris * 3 * sizeof(float);
  int rc;
  if ((rc = dev_alloc(s->tri_alb_rr, bytes))) return rc;
  HIPCHK(hipMemcpyAsync(s->tri_alb_rr.p, s->tri_alb.p, bytes, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->P.tri_alb_rr = (const float*)s->tri_alb_rr.p;
  return PRT_OK;
}

int prt_set_emission(prt_scene* s, int entity, const double emission[3]) {
  if (!s || !emission || entity < 0 || entity >= (int)s->h_ents.size()) return fail(PRT_E_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->device));
  DevEntity& e = s->h_ents[entity];
  DevMaterial& m = s->h_mats[e.material];
  std::memcpy(m.emission, emission, sizeof(m.emission));
  const int was_light = m.is_light;
  m.is_light = (m.type == PRT_REFLECTION_DIFFUSE && length(from(m.color)) < 0.001 && length(from(m.emission)) > 0.001);
  if (m.is_light != was_light) return fail(PRT_E_UNSUPPORTED, "changing whether an entity is a light");
  // the light list order follows power (light.h:99)
  sort_lights(s);
  int rc;
  if ((rc = dev_upload(s->mats, s->h_mats, s->stream))) return rc;
  if ((rc = dev_upload(s->lights, s->h_lights, s->stream))) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return PRT_OK;
}

// Path-replay adjoint: per batch, a primal pass records every sample's
// radiance; a replay with the same random numbers re-traces the paths and
// scatters d(loss)/d(albedo) and d(loss)/d(emission).  Memory is O(samples
// of a batch); nothing per path vertex is stored.
int prt_adjoint_device(prt_scene* s, const prt_render_opts* o, const float* d_dimage, int entity,
                       float* d_albedo_grad, int light, double* emission_grad, prt_stats* stats) {
  if (!s || !d_dimage) return fail(PRT_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> pool_lock(s->pool->m);
  if (s->has_glass)
    return fail(PRT_E_UNSUPPORTED, "the adjoint needs one path per sample (no glass splits)");
  if (s->P.cutoff > 0.0f) return fail(PRT_E_UNSUPPORTED, "the adjoint needs luminance_cutoff <= 0");
  const bool want_alb = d_albedo_grad != nullptr;
  if (want_alb) {
    if (entity < 0 || entity >= (int)s->h_ents.size() || s->h_ents[entity].type != PRT_ENTITY_POLYGON ||
        !s->h_ents[entity].has_albedo)
      return fail(PRT_E_INVALID, "albedo gradient needs a POLYGON entity created with per-face albedo");
  }
  if (light >= (int)s->h_ents.size()) return fail(PRT_E_INVALID, "bad light entity");
  HIPCHK(hipSetDevice(s->device));
  RenderSetup R;
  int rc;
  if ((rc = setup_render(s, o, R))) return rc;
  if ((rc = ensure_workspace(s, R.bp * R.P.spp, true))) return rc;
  hipStream_t st = s->stream;
  const int64_t nf = want_alb ? s->ent_n_tris[entity] : 0;
  if ((rc = dev_alloc(s->grad_buf, (size_t)(3 * nf + 3) * sizeof(double)))) return rc;
  double* g_alb = (double*)s->grad_buf.p;
  double* g_e = g_alb + 3 * nf;
  HIPCHK(launch_fill_u64((uint64_t*)g_alb, 0ull, 3 * nf + 3, st));
  HIPCHK(launch_fill_u64((uint64_t*)s->pool->ws.ctr.p, 0ull, CTR_COUNT, st));
  Timer tm;
  if ((rc = tm.begin(o && o->timing != 0, st, &s->timer_events))) return rc;
  CallTotals tot;
  Adjoint none{};
  Adjoint A{};
  A.on = 1;
  A.entity = want_alb ? entity : -1;
  A.light = light;
  A.dimage = d_dimage;
  A.ltot = (const uint64_t*)s->pool->ws.accum2.p;
  A.grad_alb = g_alb;
  A.grad_e = g_e;
  A.scale = R.P.sensor_response * R.P.smp;
  for (int64_t pb = 0; pb < R.n_pix; pb += R.bp) {
    Batch B{pb, std::min(R.bp, R.n_pix - pb)};
    if ((rc = run_batch(s, R.P, B, (uint64_t*)s->pool->ws.accum2.p, none, tm, tot))) return rc;
    if ((rc = run_batch(s, R.P, B, (uint64_t*)s->pool->ws.accum.p, A, tm, tot))) return rc;
    HIPCHK(launch_count_mismatch((const uint64_t*)s->pool->ws.accum.p, (const uint64_t*)s->pool->ws.accum2.p,
                                 3 * B.n_pixels * R.P.spp, (uint64_t*)s->pool->ws.ctr.p, st));
  }
  if (want_alb) HIPCHK(launch_grad_to_f32(g_alb, d_albedo_grad, 3 * nf, st));
  double ge[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(ge, g_e, sizeof(ge), hipMemcpyDeviceToHost, st));
  uint64_t c[CTR_COUNT];
  if ((rc = finish_stats(s, R, tm, tot, stats, c))) return rc;
  if (c[CTR_REPLAY_MISMATCH] != 0)
    return fail(PRT_E_INVALID, "adjoint replay diverged from the primal pass on " +
                                   std::to_string(c[CTR_REPLAY_MISMATCH]) + " sample channels");
  if (emission_grad) std::memcpy(emission_grad, ge, sizeof(ge));
  return PRT_OK;
}

}  // extern "C"
