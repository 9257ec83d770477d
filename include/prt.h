/*
 * prt.h — C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * This is the drop-in boundary for prender's render path
 *   main() -> Render::Run() -> radiance() -> intersect_scene() -> QBVH / Moller-Trumbore
 * (proj/main.cpp:199, proj/render.h:1125-1455, proj/radiance.h:25-556,
 *  proj/scene.cpp:83-214, proj/QBVH.cpp:17-159, proj/polygon.h:78-129).
 *
 * The reference has no FFI; its seams are C++ functions that read a global
 * SceneEnv.  Here the scene is an explicit handle built from plain structs whose
 * fields carry the SceneEnv / Material / Entity keywords of the in-scope subset
 * (proj/scene_env.h:262-350 defaults, 1392-1520 MATERIAL, 1544-1850 OBJECT).
 *
 * Conventions (all entry points):
 *   - every call returns int status: PRT_OK (0) or a negative PRT_E_* code,
 *     except prt_scene_create (NULL on failure) and prt_scene_destroy (void);
 *   - details of the last failure on the calling thread: prt_last_error();
 *   - the caller keeps ownership of every array it passes; arrays are copied
 *     to device memory inside the call;
 *   - a scene handle is not thread-safe and owns two HIP streams (render work,
 *     and shadow rays traced alongside it); different handles may be used
 *     from different threads;
 *   - the render memory (ray and shadow queues, sample accumulators, the
 *     batched renderer's workspace) is one pool per device, shared by every
 *     handle on it and sized for the largest call so far: render, trace,
 *     batch, adjoint and Metropolis calls on one device take turns on it
 *     (a call holds the device's pool for its duration; calls return after
 *     the device is done), and the pool is freed with the device's last
 *     handle.
 */
#ifndef PRT_H
#define PRT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRT_ABI_VERSION 1

/* status codes */
#define PRT_OK              0
#define PRT_E_INVALID      -1   /* bad argument / malformed scene description */
#define PRT_E_HIP          -2   /* HIP runtime error (message in prt_last_error) */
#define PRT_E_NOMEM        -3   /* device or host allocation failed */
#define PRT_E_NODEVICE     -4   /* no gfx950 device visible */
#define PRT_E_UNSUPPORTED  -5   /* feature outside the in-scope subset */
#define PRT_E_INTERNAL     -6   /* a render failed its own consistency check (message in prt_last_error) */

/* Material::reflection_type subset (proj/material.h:19-29). */
#define PRT_REFLECTION_DIFFUSE     0
#define PRT_REFLECTION_SPECULAR    1
#define PRT_REFLECTION_REFRACTION  2

/* EntityType numbering kept from proj/entity.h:17-24. */
#define PRT_ENTITY_SPHERE   1
#define PRT_ENTITY_UVPLANE  3
#define PRT_ENTITY_POLYGON  5

/* Material (proj/material.h:54-240, in-scope fields only).
 * Keyword semantics (proj/scene_env.h:1437-1491): "reflection diffuse" sets
 * roughness 1, "reflection specular" roughness 0, "reflection reflection" is
 * glass with IOR default 1.5.  A DIFFUSE material with roughness 0 is a mirror
 * (proj/radiance.h:514-529). */
typedef struct prt_material {
    double color[3];          /* Material::color (reflectance) */
    double emission[3];       /* Material::emission */
    double roughness;         /* Material::roughness */
    double refractive_index;  /* Material::refractive_index (REFRACTIVE_FORMULA0) */
    int32_t reflection_type;  /* PRT_REFLECTION_* */
    int32_t pad_;
} prt_material;

/* Entity (proj/entity.h:29-70 plus the per-type members).
 *   SPHERE  (proj/sphere.h:9-21):   p0 = position, radius
 *   UVPLANE (proj/plane.h:129-166): p0 = org, p1 = U, p2 = V (unnormalised; their
 *                                   lengths are u_length, v_length)
 *   POLYGON (proj/polygon.h:708-969 over ObjLoader's Obj): float32 xyz
 *           vertices, int32 0-based triangle indices, optional per-face RGB
 *           albedo (float32, n_faces*3) that replaces the material colour the
 *           way per-vertex colours do in proj/polygon.h:214-225.
 * normal_vector_inverse is Entity::normal_vector_inverse (1 or -1). */
typedef struct prt_entity {
    int32_t type;                  /* PRT_ENTITY_* */
    int32_t material;              /* index into prt_scene_desc.materials */
    int32_t normal_vector_inverse; /* 1 or -1 */
    int32_t n_vertices;
    double p0[3];
    double p1[3];
    double p2[3];
    double radius;
    const float* vertices;         /* n_vertices * 3 */
    const int32_t* indices;        /* n_faces * 3 */
    const float* face_albedo;      /* n_faces * 3 or NULL */
    int64_t n_faces;
} prt_entity;

/* Camera: SceneEnv camera fields (proj/scene_env.h:326-343) consumed by
 * CameraScreenPram (proj/render.h:70-158).  When camera_angle > 0 it is the
 * CAMERA_ANGLE keyword in degrees and overrides world_screen_width/height as
 * in proj/scene_env.h:1881-1886. */
typedef struct prt_camera {
    double camera_position[3];
    double camera_dir[3];          /* normalised by the library, as SceneEnv::Load does */
    double camera_up[3];
    double world_screen_width;
    double world_screen_height;
    double world_screen_dist;
    double camera_angle;           /* degrees; <= 0 means "not given" */
    int32_t image_width;
    int32_t image_height;
} prt_camera;

/* Sampling keywords (proj/scene_env.h:221, 329-349). */
typedef struct prt_sampling {
    int32_t samples;               /* SAMPLING */
    int32_t supersamples;          /* SUPERSAMPLING (fixed SSxSS sub-pixel grid) */
    int32_t depth;                 /* DEPTH: RR starts above it */
    int32_t depth_limit;           /* DEPTH_LIMIT: RR probability halves per level above it */
    int32_t next_event_estimation; /* NEXTEVENTESTIMATION */
    int32_t pad_;
    double luminance_cutoff;       /* LUMINANCE_CUTOFF (<= 0 disables the clamp) */
    double sensor_response;
    double background_color[3];    /* BackgroundColor */
    double environment_light[3];   /* Environment_Light */
} prt_sampling;

typedef struct prt_scene_desc {
    prt_camera camera;
    prt_sampling sampling;
    const prt_material* materials;
    int32_t n_materials;
    int32_t n_entities;
    const prt_entity* entities;
} prt_scene_desc;

/* Ray (proj/ray.h:9-14): origin and direction in double.  Directions are used
 * as given (the reference normalises before tracing). */
typedef struct prt_ray {
    double org[3];
    double dir[3];
} prt_ray;

/* Closest hit as intersect_scene() reports it (proj/scene.cpp:83-214):
 * t = IntersectionPos::distance measured from the nudged origin
 * org + 1e-6*dir; entity = Intersection::object_id (entity index, -1 = miss);
 * prim = triangle face index inside a POLYGON entity (-1 otherwise);
 * u, v = Moller-Trumbore barycentrics for triangles. */
typedef struct prt_hit {
    double t;
    double u;
    double v;
    int32_t entity;
    int32_t prim;
} prt_hit;

/* Counters of one render / trace call. */
typedef struct prt_stats {
    uint64_t camera_samples;   /* W*H*spp */
    uint64_t primary_rays;     /* camera rays traversed (unique sub-pixel rays) */
    uint64_t extension_rays;   /* bounce rays traversed */
    uint64_t shadow_rays;      /* NEE shadow rays traversed */
    uint64_t pickup_hits;      /* specular-pickup contributions (no extra ray) */
    uint64_t node_visits;      /* BVH nodes fetched over all traversals */
    uint64_t tri_tests;        /* triangle tests over all traversals */
    uint64_t truncated_paths;  /* paths stopped by the safety depth cap */
    uint64_t bytes_algorithmic;/* 128*nodes + 48*tris + 48*rays */
    double ms_total;           /* device time of the whole call */
    double ms_extend;          /* closest-hit kernels */
    double ms_shade;           /* shading kernels */
    double ms_shadow;          /* shadow kernels */
    double ms_other;           /* generate / resolve / clears */
    int64_t launches_extend;
    uint64_t shadow_node_visits;  /* part of node_visits spent on shadow rays */
    uint64_t shadow_tri_tests;    /* part of tri_tests spent on shadow rays */
    uint64_t bytes_extend;        /* algorithmic bytes of the closest-hit kernels */
    uint64_t bytes_shadow;        /* algorithmic bytes of the shadow kernels */
    double ms_extend_primary;     /* part of ms_extend spent on camera rays */
    double ms_nee;                /* next-event-estimation kernels (light sampling) */
    uint64_t mixed_scene_passes;  /* batched renders: extra passes of waves whose rays
                                     belong to more than one scene */
    double ms_traversal;          /* device time during which a closest-hit or shadow kernel
                                     ran (shadow kernels overlap the next depth's closest-hit
                                     kernel on a second stream, so this can be less than
                                     ms_extend + ms_shadow) */
    uint64_t mlt_proposals;       /* Metropolis / ERPT renders: proposals evaluated */
    uint64_t mlt_accepts;         /* Metropolis / ERPT renders: proposals accepted */
} prt_stats;

typedef struct prt_render_opts {
    uint64_t seed;             /* RNG key */
    int32_t samples_override;  /* > 0 replaces prt_sampling.samples */
    int32_t batch_pixels;      /* pixels per wavefront batch (0 = auto) */
    int32_t max_depth;         /* safety cap on path vertices (0 = 128) */
    int32_t reserved_;         /* must be 0 (per-scene renders launch their depth loop directly;
                                  prt_batch_opts.use_graph selects graph replay for batches) */
    int32_t timing;            /* 1 = per-kernel device timing into prt_stats */
    int32_t image_f64;         /* 1 = out_rgb / out_var are double arrays (device entry points) */
    int32_t engine;            /* 0 = auto: scenes without next-event estimation render in
                                  one persistent launch, scenes with glass stream through
                                  the batched renderer's tile queue, others run
                                  depth-synchronous batches; 1 = depth-synchronous
                                  batches; 2 = streaming; 3 = persistent (2 and 3: float
                                  images, no samples_override / batch_pixels; 3: at most
                                  2^28 camera samples).  The image is bit-identical
                                  whichever runs. */
} prt_render_opts;

/* ---- version / device ---------------------------------------------------- */
int prt_abi_version(void);
const char* prt_last_error(void);
/* Fills name (>= 64 bytes) and the CU count of the current device. */
int prt_device_info(char* name, int name_len, int* cu_count, int* arch_major, int* arch_minor);

/* ---- scene (replaces SceneEnv::Load + CreateEntList: scene_env.h:357,
 *      scene.cpp:38; QBVH::Construct QBVH.cpp:161) ------------------------- */
typedef struct prt_scene prt_scene;
prt_scene* prt_scene_create(const prt_scene_desc* desc);
void prt_scene_destroy(prt_scene* scene);
/* Number of BVH nodes / leaf triangles of the flattened device BVH. */
int prt_scene_info(const prt_scene* scene, int64_t* n_nodes, int64_t* n_tris,
                   int64_t* n_lights, int64_t* bvh_bytes);

/* ---- closest-hit query (replaces intersect_scene, scene.cpp:448) --------- */
/* Host arrays. */
int prt_trace_closest(prt_scene* scene, const prt_ray* rays, int64_t n,
                      prt_hit* hits, prt_stats* stats);
/* Device arrays (pointers from hipMalloc / torch); runs on the scene's stream
 * and returns after the kernel completes. */
int prt_trace_closest_device(prt_scene* scene, const prt_ray* d_rays, int64_t n,
                             prt_hit* d_hits, prt_stats* stats);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same query and the same hits as
 * prt_trace_closest_device, with no hipMalloc, hipFree or event creation in
 * the call.  Its traversal counters live in a fixed 256-byte device array of
 * the library (one per GPU, held until the process ends), cleared on every
 * call; no argument size changes what it needs, so no call is refused for
 * size.  Calls that overlap on one GPU (two threads or two streams) share
 * that array: their hits are exact, the counts in *stats are not.  stats may
 * be NULL; no timings are reported.  Returns after the kernel completes. */
int prt_trace_closest_device_noalloc(prt_scene* scene, const prt_ray* d_rays, int64_t n,
                                     prt_hit* d_hits, prt_stats* stats);

/* ---- render (replaces Render::Run, render.h:1125-1455) ------------------- */
/* out_rgb: W*H*3 float, reference row order (row 0 = top, render.h:1234).
 * out_var: W*H*3 float or NULL — unbiased variance of each pixel's mean. */
int prt_render(prt_scene* scene, const prt_render_opts* opts,
               float* out_rgb, float* out_var, prt_stats* stats);
/* Same, writing device buffers. */
int prt_render_device(prt_scene* scene, const prt_render_opts* opts,
                      float* d_rgb, float* d_var, prt_stats* stats);

/* Progress images during a render: Render::Run's periodic OutputBmp(index)
 * every imageDumpTime seconds (render.h:1390-1401; IMAGEDUMP keyword and
 * environment variable, default 60, scene_env.h:305, 402, 1905).  Between
 * batches of pixels, once `seconds` of wall time have passed since the render
 * began or since the last dump (0: after every batch; < 0: never), the image
 * so far is copied to the host and handed to `callback` (if not NULL) and
 * written as <directory>/output_W%06d.bmp (if `directory` is not NULL: the
 * reference's <dir of OUTPUT>/image folder; index from 0; `gamma_offset` as
 * in prt_write_bmp).  The image so far holds the finished pixels (all their
 * samples, sensor_response applied) and zeros elsewhere, in the reference's
 * row order; the reference's holds every pixel with the sample passes done
 * so far.  Dumps come only between batches, whose size
 * prt_render_opts.batch_pixels sets (the automatic size is 2^24 camera
 * samples: a whole 512x512x64 frame is one batch), and a render with dumps
 * runs the depth-synchronous engine (the same image bits). */
typedef struct prt_dump_opts {
    double seconds;
    const char* directory;
    double gamma_offset;
    void (*callback)(void* user, const float* rgb, int32_t width, int32_t height, int32_t index,
                     double fraction_done);
    void* user;
} prt_dump_opts;
/* prt_render with progress images; out_rgb / out_var as in prt_render. */
int prt_render_dump(prt_scene* scene, const prt_render_opts* opts, const prt_dump_opts* dump,
                    float* out_rgb, float* out_var, prt_stats* stats);

/* ---- Metropolis light transport (replaces Render::RunMLT, render.h:360-780;
 *      KelemenMLT, mlt.h:70-170; generate_new_path, render.h:195-356) ------- */
/* Primary-sample-space MLT (Kelemen et al. 2002) over the same radiance():
 * seed paths with fresh primary samples estimate the normalisation b (mean
 * luminance) and give each Markov chain its start, drawn in proportion to
 * luminance; then every chain proposes a large step (all primary samples
 * fresh, probability p_large) or a small step (every primary sample moved by
 * mlt.h:42-54's Mutate0), accepts with min(1, L_new/L_old) and splats both
 * paths with Kelemen's weights.  Camera: a pixel from two primary samples, a
 * 2x2 sub-pixel choice and the tent filter (render.h:262-300).  Primary
 * samples drive Russian roulette, the BSDF choice and direction of the first
 * sample_depths vertices; light sampling takes fresh numbers per proposal, as
 * the reference's Next01() does (directLight_radiance.h:343-359,
 * light.h:146-151).  The reference runs `samples` chains one after another,
 * each of W*H*mutation steps; here `chains` chains run side by side, each a
 * wavefront batch slot, for samples*W*H*mutation/chains steps. */
typedef struct prt_mlt_opts {
    uint64_t seed;               /* RNG key */
    double mutation;             /* steps per pixel per chain of the reference (SceneEnv::mutation; <= 0: 32) */
    double pre_sample;           /* seed paths per pixel per chain (SceneEnv::pre_sample; <= 0: 0.4) */
    double p_large;              /* large-step probability (render.h:551; <= 0: 0.5) */
    int64_t mutations_per_chain; /* > 0 overrides samples*W*H*mutation / chains */
    int32_t samples_override;    /* > 0 replaces prt_sampling.samples (the reference's chain count) */
    int32_t chains;              /* chains run side by side (0: auto, at most 2^20) */
    int32_t sample_depths;       /* path vertices whose random numbers are primary samples (0: 8, <= 32) */
    int32_t max_depth;           /* safety cap on path vertices (0 = 128) */
    int32_t image_f64;           /* 1 = d_rgb is a double array (device entry point) */
    int32_t pad_;
} prt_mlt_opts;
/* out_rgb: W*H*3 float in the reference's row order, already scaled by the
 * MLT sensor response (0.25 * SENSOR_RESPONSE, scene_env.h:1945-1956). */
int prt_render_mlt(prt_scene* scene, const prt_mlt_opts* opts, float* out_rgb, prt_stats* stats);
int prt_render_mlt_device(prt_scene* scene, const prt_mlt_opts* opts, float* d_rgb, prt_stats* stats);

/* ---- energy redistribution path tracing (replaces Render::RunERPT,
 *      render.h:786-1123; ERPTSampler, erpt.h:14-88) ------------------------ */
/* W*H paths estimate the deposition energy ed = L(mean radiance) / mutation
 * (render.h:831-849); then SUPERSAMPLING^2 initial paths per pixel: a path
 * that returns a light directly with NEE on is splatted as F / samples, any
 * other starts floor(u + L / (mutation * ed)) chains of `mutation` steps, each
 * step a perturbation of the chain's primary samples by up to
 * +-mutate_distance, accepted with min(1, L'/L), and a deposit of
 * ed / samples worth of the current path's colour at its pixel (skipped after
 * 20 deposits in a row on one pixel).  The image is scaled by
 * 0.25 * SENSOR_RESPONSE (scene_env.h:1970).  Chains run side by side as
 * wavefront batches of up to 2^20. */
typedef struct prt_erpt_opts {
    uint64_t seed;               /* RNG key */
    double mutation;             /* chain steps (SceneEnv::mutation; <= 0: 32) */
    double mutate_distance;      /* primary-sample perturbation (erpt.h:107; <= 0: 0.05) */
    int32_t samples_override;    /* > 0 replaces prt_sampling.samples (the deposit divisor) */
    int32_t sample_depths;       /* path vertices whose random numbers are primary samples (0: 8, <= 32) */
    int32_t max_depth;           /* safety cap on path vertices (0 = 128) */
    int32_t image_f64;           /* 1 = d_rgb is a double array (device entry point) */
} prt_erpt_opts;
int prt_render_erpt(prt_scene* scene, const prt_erpt_opts* opts, float* out_rgb, prt_stats* stats);
int prt_render_erpt_device(prt_scene* scene, const prt_erpt_opts* opts, float* d_rgb, prt_stats* stats);

/* ---- batched rendering of many scenes ------------------------------------- */
/* No reference counterpart: the reference renders one scene per process
 * (main.cpp:173-200).  All scenes of a batch go through one shared work queue
 * of pixel tiles (a tile = consecutive pixels of one scene with all their
 * samples); a fixed pool of tile slots keeps rays of many tiles, of whatever
 * scenes, in flight, and the wavefront iterates, two iterations per host
 * step, until the queue is drained.  Each scene's image is
 * bit-identical to prt_render_device of that scene with the same seed. */
typedef struct prt_batch_opts {
    uint64_t seed;
    int32_t tile_samples;       /* camera samples per tile, rounded up to a power of two
                                   and to the largest spp (0 = 4096) */
    int32_t inflight_tiles;     /* tile slots (0 = automatic: the most samples in flight, up to 2^24
                                   and up to the batch's camera samples, whose workspace fits 3/8
                                   of the device memory; more in flight is faster) */
    int32_t stack_lds;          /* LDS traversal stack entries: 8, 12, 16 or 32 (0 = 12; 12 runs
                                   the traversal kernels at 5 waves per SIMD) */
    int32_t trav_blocks_per_cu; /* persistent traversal blocks per CU, <= 16 (0 = 8) */
    int32_t use_graph;          /* 1: the iterations replayed as a captured hipGraph; 0 (or -1):
                                   plain launches, the faster form measured on MI355X */
    int32_t tile_stride;        /* render only every k-th tile (tuning trials; 0 = all) */
    int32_t max_depth;          /* safety cap on path vertices (0 = 128) */
    int32_t timing;             /* 1 = per-kernel device timing into prt_stats (plain launches only) */
    int32_t persistent;         /* 0 = the tile queue above.
                                   1 = one persistent launch renders every camera sample of the batch:
                                   waves claim 64-sample chunks from one shared counter and follow each
                                   sample's whole path (shading, NEE, shadow rays, glass branches) with
                                   no ray queues and no waits between waves; then one launch folds the
                                   pixels.  trav_blocks_per_cu (2..4) sets its waves per SIMD;
                                   tile_samples, inflight_tiles, use_graph and tile_stride do not
                                   apply; at most 2^28 camera samples (a fixed-point accumulator each).
                                   2 = split: the scenes without next-event estimation in one
                                   persistent launch, the others through the tile queue (with these
                                   options), one after the other.
                                   The images are the same bit for bit in every mode. */
    int32_t persist_waves;      /* waves per SIMD the persistent kernel's registers are budgeted for
                                   (2..4) in modes 1 and 2; 0 = mode 1: trav_blocks_per_cu, mode 2: 2 */
} prt_batch_opts;
/* d_rgb[i] / d_var[i]: W_i*H_i*3 float device buffers of scene i (d_var or
 * its entries may be NULL).  All scenes must live on one device; the call
 * runs on scenes[0]'s stream, which also owns the batch workspace. */
int prt_render_batch_device(prt_scene* const* scenes, int n_scenes,
                            const prt_batch_opts* opts, float* const* d_rgb,
                            float* const* d_var, prt_stats* stats);
/* Measures candidate tile sizes (1024..16384 samples), LDS stack depths,
 * traversal grids and the launch form (plain or graph replay) by coordinate
 * descent on a subsample of the batch's tiles (every k-th tile, k from
 * base->tile_stride or chosen for ~16M samples per trial), then the
 * persistent engine and the split form, each at 2, 3 and 4 waves per SIMD, on
 * the whole batch (when it fits and base->tile_stride asks for no
 * subsample), and writes the fastest
 * options to *best (tile_stride copied from base; inflight_tiles from base,
 * 0 = the automatic size).  The trial renders write the output buffers. */
int prt_tune_batch(prt_scene* const* scenes, int n_scenes, const prt_batch_opts* base,
                   float* const* d_rgb, float* const* d_var, prt_batch_opts* best,
                   double* best_ns_per_sample);

/* ---- differentiable rendering (per-face albedo + light emission) --------- */
/* Replace the per-face albedo of POLYGON entity `entity` (float n_faces*3,
 * device pointer) and scale factor of the emission of light entity `light`. */
int prt_set_face_albedo_device(prt_scene* scene, int entity, const float* d_albedo);
int prt_set_emission(prt_scene* scene, int entity, const double emission[3]);
/* hold = 1: Russian-roulette probabilities keep using the per-face albedo as
 * it is now, whatever later prt_set_face_albedo_device calls set (the
 * primal decisions, for finite-difference checks under common random
 * numbers); hold = 0: they follow the current albedo again. */
int prt_hold_rr_albedo(prt_scene* scene, int hold);
/* Path-replay adjoint: given dL/dimage (W*H*3 device floats), accumulate
 * dL/dalbedo into d_albedo_grad (n_faces*3 device floats of `entity`) and
 * dL/demission (3 doubles, host) for light `light`. */
int prt_adjoint_device(prt_scene* scene, const prt_render_opts* opts,
                       const float* d_dimage, int entity, float* d_albedo_grad,
                       int light, double* d_emission_grad, prt_stats* stats);

/* ---- scene files (replaces SceneEnv::Load + CreateEntList for the in-scope
 *      keyword subset: scene_env.h:357-1990, Expr.cpp, ObjLoader.cpp) ------- */
/* Parses a scene file: IMAGE, SCREEN, SAMPLING, SUPERSAMPLING, DEPTH,
 * DEPTH_LIMIT, CAMERA_POS / _DIR / _UPVEC / _ANGLE / _MATRIX, TARGET_POS,
 * BACKGROUND, ENVLIGHT, nextEventEstimation, luminance_cutoff,
 * sensor_response, VAR, OUTPUT, gamma_offset; OBJECT sphere / uvplane /
 * objfile with emission, color, roughness, refractive_index, reflection
 * diffuse|specular|reflection, normal, rotation / translate / scale; the
 * expression syntax of every numeric field and the environment overrides
 * applied after the file.  Keywords that would change the image outside the
 * subset fail with PRT_E_UNSUPPORTED (message in prt_last_error).  OBJ paths
 * are tried as given, then relative to the scene file.  NULL on failure. */
typedef struct prt_scene_file prt_scene_file;
prt_scene_file* prt_scene_file_load(const char* path);
/* The parsed description (valid until prt_scene_file_free); feed it to
 * prt_scene_create. */
const prt_scene_desc* prt_scene_file_desc(const prt_scene_file* file);
/* OUTPUT (default "output.ppm") and gamma_offset of the file. */
const char* prt_scene_file_output(const prt_scene_file* file, double* gamma_offset);
/* The file's Markov-chain switches (metropolisTransport,
 * energyRedistributionPathTracing, mutation, pre_sample and the METROPOLIS /
 * ERPT / MUTATION / PRESAMPLE overrides, scene_env.h:642-672, 1908-1911,
 * 1946-1968): returns 2 when the file asks for RunERPT (tested first, as
 * main.cpp:189-196 does), 1 for RunMLT, else 0, and fills *opts (if not
 * NULL) with its mutation and pre_sample, every other field zero (library
 * defaults); an ERPT render takes opts->mutation as its chain length. */
int prt_scene_file_mlt(const prt_scene_file* file, prt_mlt_opts* opts);
/* The file's IMAGEDUMP seconds (default 60) after the IMAGEDUMP environment
 * override, read with atoi as scene_env.h:402 and 1905 do (for
 * prt_dump_opts.seconds); -1 for a NULL file. */
int prt_scene_file_image_dump(const prt_scene_file* file);
void prt_scene_file_free(prt_scene_file* file);
/* One numeric field as the scene-file reader evaluates it (Expr.cpp
 * expr_eval: + - * / % ^, parentheses, pi, e, sqrt ... datan2, floor, ceil,
 * "%NAME%" environment references); *status = PRT_OK or an error. */
double prt_eval_expression(const char* text, int* status);

/* ---- image output (ppm.h:14-35, render.h:1458-1667, bitmap.cpp, hdr.h) --- */
/* rgb: W*H*3 floats in the reference's storage order (row 0 = top).
 * 8-bit values int(pow(clamp(x,0,1), gamma) * 255 + 0.5); gamma <= 0 means
 * 1/2.2. */
int prt_image_to_rgb8(const float* rgb, int width, int height, double gamma, uint8_t* out);
/* ASCII P3 file as Render::Output writes it. */
int prt_write_ppm(const char* path, const float* rgb, int width, int height);
/* 24-bit BMP as Render::OutputBmp writes it; |gamma_offset| > 0.01 replaces
 * the 1/2.2 exponent. */
int prt_write_bmp(const char* path, const float* rgb, int width, int height, double gamma_offset);
/* Radiance RGBE file as Render::OutputHdr writes it (render.h:1647-1667,
 * hdr.h:212-246): linear values, no gamma. */
int prt_write_hdr(const char* path, const float* rgb, int width, int height);

#ifdef __cplusplus
}
#endif

#endif /* PRT_H */
