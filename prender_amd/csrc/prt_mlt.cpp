// prt_mlt.cpp — host side of Metropolis renders (prt_render_mlt*), the
// MI355X form of Render::RunMLT (render.h:360-780).
//
// The reference runs `samples` Markov chains one after another on a CPU
// thread each: W*H*pre_sample seed paths, a start path drawn from them in
// proportion to luminance, then W*H*mutation Kelemen steps.  Here `chains`
// chains advance side by side, one wavefront batch slot each, so that every
// step is one batch of `chains` paths through the same kernels as a
// path-traced render; the seed paths (the reference's total,
// samples * W*H*pre_sample) are traced first in batches of the same size and
// shared by all chains, which also gives the normalisation b from all of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/prt.h"
#include "prt_internal.h"
#include "prt_kernels.h"
#include "prt_mlt.h"

namespace prt {
namespace {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Philox key of one random stream of a render: phase 1 = seed-path batch
// `idx`, 2 = chain step `idx`, 3 = chain control.
void mlt_key(uint64_t seed, uint64_t phase, uint64_t idx, uint32_t& k0, uint32_t& k1) {
  const uint64_t h = splitmix64(splitmix64(seed ^ (phase * 0xA0761D6478BD642Full)) + idx);
  k0 = (uint32_t)h;
  k1 = (uint32_t)(h >> 32);
}

struct MltPlan {
  int64_t chains = 0, steps = 0, seeds = 0;
  int depths = 8;
  double p_large = 0.5;
};

int plan_mlt(const Params& P, const prt_mlt_opts& o, MltPlan& m) {
  const int samples = o.samples_override > 0 ? o.samples_override : P.samples;
  const double mutation = o.mutation > 0.0 ? o.mutation : 32.0;      // scene_env.h:1949
  const double pre_sample = o.pre_sample > 0.0 ? o.pre_sample : 0.4;  // scene_env.h:1948
  if (samples <= 0) return fail(PRT_E_INVALID, "samples must be positive");
  const int64_t n_pix = (int64_t)P.W * P.H;
  // SeedPathMax = (int)(W*H*pre_sample) per chain of the reference (render.h:400)
  m.seeds = (int64_t)(n_pix * pre_sample) * samples;
  if (m.seeds < 1) m.seeds = 1;
  if (m.seeds > ((int64_t)1 << 24)) return fail(PRT_E_INVALID, "more than 2^24 seed paths (lower pre_sample or samples)");
  const double total = (double)samples * (double)n_pix * mutation;  // M = W*H*mutation (render.h:432)
  if (o.chains > 0) m.chains = o.chains;
  else m.chains = (int64_t)std::min(std::max(total / 64.0, 1024.0), (double)(1 << 20));
  if (m.chains > (1 << 20)) return fail(PRT_E_INVALID, "more than 2^20 chains");
  m.steps = o.mutations_per_chain > 0 ? o.mutations_per_chain
                                      : std::max<int64_t>(1, (int64_t)std::ceil(total / (double)m.chains));
  if (m.steps > 0x7FFFFFFF) return fail(PRT_E_INVALID, "too many steps per chain");
  m.depths = o.sample_depths > 0 ? o.sample_depths : 8;
  if (m.depths > 32) return fail(PRT_E_INVALID, "sample_depths above 32");
  m.p_large = o.p_large > 0.0 ? std::min(o.p_large, 1.0) : 0.5;  // render.h:551
  return PRT_OK;
}

}  // namespace

static int render_mlt(prt_scene* s, const prt_mlt_opts* o, void* d_out, int as_f64, prt_stats* stats) {
  prt_mlt_opts opts{};
  if (o) opts = *o;
  HIPCHK(hipSetDevice(s->device));
  std::lock_guard<std::recursive_mutex> pool_lock(s->pool->m);
  Params P = s->P;
  MltPlan m;
  int rc;
  if ((rc = plan_mlt(P, opts, m))) return rc;
  const int64_t n_pix = (int64_t)P.W * P.H;
  const uint32_t C = (uint32_t)m.chains, N = (uint32_t)m.seeds;
  const uint32_t stride = MLT_CAMERA_WORDS + 4u * (uint32_t)m.depths;
  // one path per batch slot: no sub-pixel sharing, no per-sample scaling
  P.SS = 1;
  P.samples = 1;
  P.spp = 1;
  P.smp = 1.0;
  P.max_vertices = opts.max_depth > 0 ? std::min(opts.max_depth, 255) : 128;
  P.mlt_stride = (int32_t)stride;
  P.mlt_depths = m.depths;

  if ((rc = ensure_workspace(s, C, false))) return rc;
  MltWorkspace& w = s->mlt;
  if ((rc = dev_alloc(w.cur, (size_t)C * stride * 4)) || (rc = dev_alloc(w.prop, (size_t)C * stride * 4)) ||
      (rc = dev_alloc(w.pix_new, (size_t)C * 4)) || (rc = dev_alloc(w.pix_old, (size_t)C * 4)) ||
      (rc = dev_alloc(w.f_old, (size_t)C * 24)) || (rc = dev_alloc(w.img, (size_t)n_pix * 24)) ||
      (rc = dev_alloc(w.seed_f, (size_t)N * 24)) || (rc = dev_alloc(w.seed_lum, (size_t)N * 8)) ||
      (rc = dev_alloc(w.seed_pix, (size_t)N * 4)) || (rc = dev_alloc(w.seed_cdf, (size_t)N * 8)) ||
      (rc = dev_alloc(w.tile_sums, (size_t)((N + MLT_SCAN_TILE - 1) / MLT_SCAN_TILE) * 8)))
    return rc;
  MltChains CH{(uint32_t*)w.cur.p, (uint32_t*)w.prop.p, (uint32_t*)w.pix_new.p, (uint32_t*)w.pix_old.p,
               (double*)w.f_old.p, (uint64_t*)w.img.p, C, stride, 0u, 0u};
  mlt_key(opts.seed, 3, 0, CH.k0, CH.k1);
  MltSeeds SD{(double*)w.seed_f.p, (uint64_t*)w.seed_lum.p, (uint32_t*)w.seed_pix.p, (uint64_t*)w.seed_cdf.p,
              (uint64_t*)w.tile_sums.p, N};
  P.mlt_u = CH.prop;
  P.mlt_pix = CH.pix_new;

  hipStream_t st = s->stream;
  uint64_t* ctr = (uint64_t*)s->pool->ws.ctr.p;
  uint64_t* accum = (uint64_t*)s->pool->ws.accum.p;
  HIPCHK(launch_fill_u64(ctr, 0ull, CTR_COUNT, st));
  HIPCHK(launch_fill_u64(CH.img, 0ull, 3 * n_pix, st));
  Timer tm;
  if ((rc = tm.begin(false, st))) return rc;
  CallTotals tot;
  const Adjoint none{};
  auto t0 = std::chrono::steady_clock::now();

  // seed paths: large steps from fresh primary samples (render.h:479-503)
  for (uint32_t first = 0, batch = 0; first < N; first += C, ++batch) {
    const uint32_t n = std::min(C, N - first);
    HIPCHK(launch_mlt_mutate(CH, 0, first, n, 0u, m.p_large, st));
    Params Pb = P;
    mlt_key(opts.seed, 1, batch, Pb.key0, Pb.key1);
    if ((rc = run_batch(s, Pb, Batch{0, n}, accum, none, tm, tot))) return rc;
    HIPCHK(launch_mlt_seed_record(accum, CH.pix_new, first, n, SD, st));
  }
  HIPCHK(launch_mlt_scan(SD, st));
  uint64_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, SD.cdf + (N - 1), sizeof(total), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));

  int64_t steps_run = 0;
  if (total > 0) {  // else every seed is black: the reference skips the chain (render.h:521-524)
    const double b = (double)total / MLT_LUM_ONE / (double)N;  // render.h:550
    HIPCHK(launch_mlt_init(CH, SD, st));
    const double weight = 4.0 * (double)n_pix;  // generate_new_path pdf: W*H, x4 for the tent filter
    const double inv_m = 1.0 / (double)m.steps;
    for (int64_t it = 0; it < m.steps; ++it) {
      HIPCHK(launch_mlt_mutate(CH, 1, 0u, C, (uint32_t)it, m.p_large, st));
      Params Pi = P;
      mlt_key(opts.seed, 2, (uint64_t)it, Pi.key0, Pi.key1);
      if ((rc = run_batch(s, Pi, Batch{0, C}, accum, none, tm, tot))) return rc;
      HIPCHK(launch_mlt_step(CH, accum, (uint32_t)it, m.p_large, b, weight, inv_m, ctr, st));
    }
    steps_run = m.steps;
  }
  // image * sensor_response / samples, with the MLT mode's 0.25 factor
  // (scene_env.h:1955) and the chains in the reference's place of samples
  HIPCHK(launch_mlt_resolve(CH.img, 3 * n_pix, P.sensor_response * 0.25 / (double)C, d_out, as_f64, st));
  uint64_t c[CTR_COUNT];
  HIPCHK(hipMemcpyAsync(c, ctr, sizeof(c), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (c[CTR_DROP] != 0) return fail(PRT_E_INVALID, "ray queue overflow (" + std::to_string(c[CTR_DROP]) + " rays)");
  if (c[CTR_TRACED] != c[CTR_QUEUED] || c[CTR_STRACED] != c[CTR_SHADOW])
    return fail(PRT_E_INTERNAL, "traversal finished " + std::to_string(c[CTR_TRACED]) + " of " +
                                    std::to_string(c[CTR_QUEUED]) + " rays");
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    fill_stats(stats, c);
    stats->camera_samples = (uint64_t)N + (uint64_t)C * (uint64_t)steps_run;
    stats->primary_rays = tot.primary_rays;
    stats->extension_rays = c[CTR_QUEUED] - tot.primary_rays;
    stats->shadow_rays = c[CTR_SHADOW];
    stats->launches_extend = tot.launches;
    stats->ms_total = ms;
    stats->mlt_proposals = (uint64_t)C * (uint64_t)steps_run;
    stats->mlt_accepts = c[CTR_MLT_ACCEPT];
    stats->bytes_extend = 128 * c[CTR_NODES] + 48 * c[CTR_TRIS] + 48 * c[CTR_QUEUED];
    stats->bytes_shadow = 128 * c[CTR_SNODES] + 48 * c[CTR_STRIS] + 48 * c[CTR_SHADOW];
    stats->bytes_algorithmic = stats->bytes_extend + stats->bytes_shadow;
  }
  return PRT_OK;
}

// Render::RunERPT (render.h:786-1123) as wavefront batches: the deposition
// energy paths and the initial paths are batches of fresh primary samples;
// an initial path's chains are listed by an integer scan of their counts and
// run side by side, one batch per step.
static int render_erpt(prt_scene* s, const prt_erpt_opts* o, void* d_out, int as_f64, prt_stats* stats) {
  prt_erpt_opts opts{};
  if (o) opts = *o;
  HIPCHK(hipSetDevice(s->device));
  std::lock_guard<std::recursive_mutex> pool_lock(s->pool->m);
  Params P = s->P;
  const int samples = opts.samples_override > 0 ? opts.samples_override : P.samples;
  if (samples <= 0) return fail(PRT_E_INVALID, "samples must be positive");
  const double mutation = opts.mutation > 0.0 ? opts.mutation : 32.0;  // scene_env.h:1962
  const int M = mutation < 1.0 ? 1 : (int)mutation;                   // render.h:809
  const double d = opts.mutate_distance > 0.0 ? opts.mutate_distance : 0.05;  // erpt.h:107
  const int depths = opts.sample_depths > 0 ? opts.sample_depths : 8;
  if (depths > 32) return fail(PRT_E_INVALID, "sample_depths above 32");
  const int ss2 = P.SS * P.SS;
  const int64_t n_pix = (int64_t)P.W * P.H, n_init = n_pix * ss2;
  if (n_init >= ((int64_t)1 << 32)) return fail(PRT_E_INVALID, "image too large for ERPT");
  const uint32_t stride = MLT_CAMERA_WORDS + 4u * (uint32_t)depths;
  const uint32_t cap = 1u << 20;  // paths per wavefront batch
  const uint32_t B = (uint32_t)std::min<int64_t>(cap, std::max(n_init, n_pix));
  P.SS = 1;
  P.samples = 1;
  P.spp = 1;
  P.smp = 1.0;
  P.max_vertices = opts.max_depth > 0 ? std::min(opts.max_depth, 255) : 128;
  P.mlt_stride = (int32_t)stride;
  P.mlt_depths = depths;
  int rc;
  if ((rc = ensure_workspace(s, B, false))) return rc;
  MltWorkspace& w = s->mlt;
  if ((rc = dev_alloc(w.path_u, (size_t)B * stride * 4)) || (rc = dev_alloc(w.path_f, (size_t)B * 24)) ||
      (rc = dev_alloc(w.path_lum, (size_t)B * 8)) || (rc = dev_alloc(w.path_pix, (size_t)B * 4)) ||
      (rc = dev_alloc(w.path_direct, (size_t)B * 4)) || (rc = dev_alloc(w.path_count, (size_t)B * 8)) ||
      (rc = dev_alloc(w.path_first, (size_t)B * 8)) ||
      (rc = dev_alloc(w.tile_sums, (size_t)((B + MLT_SCAN_TILE - 1) / MLT_SCAN_TILE) * 8)) ||
      (rc = dev_alloc(w.cur, (size_t)B * stride * 4)) || (rc = dev_alloc(w.prop, (size_t)B * stride * 4)) ||
      (rc = dev_alloc(w.pix_new, (size_t)B * 4)) || (rc = dev_alloc(w.pix_old, (size_t)B * 4)) ||
      (rc = dev_alloc(w.pix_now, (size_t)B * 4)) || (rc = dev_alloc(w.run, (size_t)B * 4)) ||
      (rc = dev_alloc(w.f_old, (size_t)B * 24)) || (rc = dev_alloc(w.img, (size_t)n_pix * 24)) ||
      (rc = dev_alloc(w.sums, 3 * sizeof(uint64_t))))
    return rc;
  uint32_t kA[2], kB[2], kC[2];
  mlt_key(opts.seed, 4, 0, kA[0], kA[1]);  // deposition-energy paths
  mlt_key(opts.seed, 5, 0, kB[0], kB[1]);  // initial paths and their chain counts
  mlt_key(opts.seed, 6, 0, kC[0], kC[1]);  // chain steps
  // fresh primary samples of a batch of paths go to path_u through the
  // Metropolis mutate kernel's seed mode
  MltChains fresh{nullptr, (uint32_t*)w.path_u.p, nullptr, nullptr, nullptr, nullptr, B, stride, 0u, 0u};
  ErptPaths E{(double*)w.path_f.p, (double*)w.path_lum.p, (uint32_t*)w.path_pix.p, (uint32_t*)w.path_direct.p,
              (uint64_t*)w.path_count.p, (uint64_t*)w.path_first.p, (uint64_t*)w.tile_sums.p, 0u};
  ErptChains CH{(uint32_t*)w.cur.p, (uint32_t*)w.prop.p, (uint32_t*)w.pix_new.p, (uint32_t*)w.pix_old.p,
                (uint32_t*)w.pix_now.p, (uint32_t*)w.run.p, (double*)w.f_old.p, (uint64_t*)w.img.p, 0u, stride,
                kC[0], kC[1], 0u, 0.0};

  hipStream_t st = s->stream;
  uint64_t* ctr = (uint64_t*)s->pool->ws.ctr.p;
  uint64_t* accum = (uint64_t*)s->pool->ws.accum.p;
  uint64_t* sums = (uint64_t*)w.sums.p;
  HIPCHK(launch_fill_u64(ctr, 0ull, CTR_COUNT, st));
  HIPCHK(launch_fill_u64(CH.img, 0ull, 3 * n_pix, st));
  HIPCHK(launch_fill_u64(sums, 0ull, 3, st));
  Timer tm;
  if ((rc = tm.begin(false, st))) return rc;
  CallTotals tot;
  const Adjoint none{};
  uint64_t launch = 0;  // path batches traced so far: each takes its own Philox key
  auto trace = [&](const uint32_t* u, uint32_t* pix, uint32_t* flags, uint32_t n) {
    Params Pb = P;
    Pb.mlt_u = u;
    Pb.mlt_pix = pix;
    Pb.mlt_flags = flags;
    mlt_key(opts.seed, 7, launch++, Pb.key0, Pb.key1);
    return run_batch(s, Pb, Batch{0, n}, accum, none, tm, tot);
  };
  auto t0 = std::chrono::steady_clock::now();

  // deposition energy: one path per pixel (render.h:831-849)
  fresh.k0 = kA[0];
  fresh.k1 = kA[1];
  for (uint32_t first = 0; first < (uint32_t)n_pix; first += cap) {
    const uint32_t n = (uint32_t)std::min<int64_t>(cap, n_pix - first);
    HIPCHK(launch_mlt_mutate(fresh, 0, first, n, 0u, 0.5, st));
    if ((rc = trace(fresh.prop, E.pix, nullptr, n))) return rc;
    HIPCHK(launch_erpt_sum(accum, n, sums, st));
  }
  uint64_t hs[3];
  HIPCHK(hipMemcpyAsync(hs, sums, sizeof(hs), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double mean[3] = {(double)hs[0] / MLT_LUM_ONE / (double)n_pix, (double)hs[1] / MLT_LUM_ONE / (double)n_pix,
                          (double)hs[2] / MLT_LUM_ONE / (double)n_pix};
  const double ed = ((0.2126 * mean[0] + 0.0722 * mean[2]) + 0.7152 * mean[1]) / (double)M;
  CH.dep = ed / (double)samples;

  // initial paths, SS^2 per pixel in row-major pixel order (render.h:897-1010)
  fresh.k0 = kB[0];
  fresh.k1 = kB[1];
  uint64_t chain_base = 0, proposals = 0;
  for (uint32_t first = 0; first < (uint32_t)n_init; first += cap) {
    const uint32_t n = (uint32_t)std::min<int64_t>(cap, n_init - first);
    HIPCHK(launch_mlt_mutate(fresh, 0, first, n, 0u, 0.5, st));
    if ((rc = trace(fresh.prop, E.pix, E.direct, n))) return rc;
    E.n = n;
    HIPCHK(launch_erpt_paths(accum, E, first, CH.img, 1.0 / samples, ed, (double)M, kB[0], kB[1], st));
    MltSeeds scan{nullptr, E.count, nullptr, E.first, E.tile_sums, n};
    HIPCHK(launch_mlt_scan(scan, st));
    uint64_t n_chains = 0;
    HIPCHK(hipMemcpyAsync(&n_chains, E.first + (n - 1), sizeof(n_chains), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (chain_base + n_chains >= ((uint64_t)1 << 32)) return fail(PRT_E_INVALID, "more than 2^32 ERPT chains");
    for (uint64_t j0 = 0; j0 < n_chains; j0 += B) {  // sub-batches fit the workspace
      CH.n = (uint32_t)std::min<uint64_t>(B, n_chains - j0);
      CH.base = (uint32_t)(chain_base + j0);
      HIPCHK(launch_erpt_chain_init(CH, E, fresh.prop, j0, first, ss2, P.W, P.H, st));
      for (int step = 0; step < M; ++step) {
        HIPCHK(launch_erpt_perturb(CH, (uint32_t)step, d, st));
        if ((rc = trace(CH.prop, CH.pix_new, nullptr, CH.n))) return rc;
        HIPCHK(launch_erpt_step(CH, accum, (uint32_t)step, ctr, st));
      }
      proposals += (uint64_t)CH.n * (uint64_t)M;
    }
    chain_base += n_chains;
  }
  // the reference multiplies only image[0] by sensor_response (render.h:1075-1078);
  // every pixel is scaled here, with ERPT's 0.25 factor (scene_env.h:1970)
  HIPCHK(launch_mlt_resolve(CH.img, 3 * n_pix, P.sensor_response * 0.25, d_out, as_f64, st));
  uint64_t c[CTR_COUNT];
  HIPCHK(hipMemcpyAsync(c, ctr, sizeof(c), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (c[CTR_DROP] != 0) return fail(PRT_E_INVALID, "ray queue overflow (" + std::to_string(c[CTR_DROP]) + " rays)");
  if (c[CTR_TRACED] != c[CTR_QUEUED] || c[CTR_STRACED] != c[CTR_SHADOW])
    return fail(PRT_E_INTERNAL, "traversal finished " + std::to_string(c[CTR_TRACED]) + " of " +
                                    std::to_string(c[CTR_QUEUED]) + " rays");
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    fill_stats(stats, c);
    stats->camera_samples = (uint64_t)n_pix + (uint64_t)n_init + proposals;
    stats->primary_rays = tot.primary_rays;
    stats->extension_rays = c[CTR_QUEUED] - tot.primary_rays;
    stats->shadow_rays = c[CTR_SHADOW];
    stats->launches_extend = tot.launches;
    stats->ms_total = ms;
    stats->mlt_proposals = proposals;
    stats->mlt_accepts = c[CTR_MLT_ACCEPT];
    stats->bytes_extend = 128 * c[CTR_NODES] + 48 * c[CTR_TRIS] + 48 * c[CTR_QUEUED];
    stats->bytes_shadow = 128 * c[CTR_SNODES] + 48 * c[CTR_STRIS] + 48 * c[CTR_SHADOW];
    stats->bytes_algorithmic = stats->bytes_extend + stats->bytes_shadow;
  }
  return PRT_OK;
}

}  // namespace prt

using namespace prt;

extern "C" {

int prt_render_mlt_device(prt_scene* s, const prt_mlt_opts* opts, float* d_rgb, prt_stats* stats) {
  if (!s || !d_rgb) return fail(PRT_E_INVALID, "bad arguments");
  return render_mlt(s, opts, d_rgb, opts && opts->image_f64, stats);
}

int prt_render_mlt(prt_scene* s, const prt_mlt_opts* opts, float* out_rgb, prt_stats* stats) {
  if (!s || !out_rgb) return fail(PRT_E_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->device));
  const size_t bytes = (size_t)s->P.W * s->P.H * 3 * sizeof(float);
  int rc;
  if ((rc = dev_alloc(s->out_rgb, bytes))) return rc;
  if ((rc = render_mlt(s, opts, s->out_rgb.p, 0, stats))) return rc;
  HIPCHK(hipMemcpyAsync(out_rgb, s->out_rgb.p, bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return PRT_OK;
}

int prt_render_erpt_device(prt_scene* s, const prt_erpt_opts* opts, float* d_rgb, prt_stats* stats) {
  if (!s || !d_rgb) return fail(PRT_E_INVALID, "bad arguments");
  return render_erpt(s, opts, d_rgb, opts && opts->image_f64, stats);
}

int prt_render_erpt(prt_scene* s, const prt_erpt_opts* opts, float* out_rgb, prt_stats* stats) {
  if (!s || !out_rgb) return fail(PRT_E_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->device));
  const size_t bytes = (size_t)s->P.W * s->P.H * 3 * sizeof(float);
  int rc;
  if ((rc = dev_alloc(s->out_rgb, bytes))) return rc;
  if ((rc = render_erpt(s, opts, s->out_rgb.p, 0, stats))) return rc;
  HIPCHK(hipMemcpyAsync(out_rgb, s->out_rgb.p, bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return PRT_OK;
}

}  // extern "C"
