// prt_mlt.hip — kernels of the primary-sample-space Metropolis renderer
// (Render::RunMLT, render.h:360-780; KelemenMLT, mlt.h:70-170).  The paths
// themselves go through the wavefront kernels of prt_kernels.hip; these
// kernels mutate primary samples, make the proposals' camera rays, pick the
// chains' start paths and run the acceptance / splatting step.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "prt_math.h"
#include "prt_mlt.h"
#include "prt_shade.h"

namespace {

__device__ __forceinline__ uint32_t word_of(double x) {
  // x in [0, 1]; 1 wraps to 0 (the primary sample space is a torus)
  return x >= 1.0 ? 0u : (uint32_t)(x * 4294967296.0);
}

// Mutate0 (mlt.h:42-54), LuxRender's small step: an exponentially
// distributed offset between s1 / (s1/s2 + 1) and s2, up or down with equal
// probability, wrapped into [0, 1).  fabsf keeps the reference's float
// absolute value.
__device__ __forceinline__ uint32_t mutate0(uint32_t w, double r) {
  const double x = u01(w);
  const double s1 = 1.0 / 512.0, s2 = 1.0 / 16.0;
  const double dx = s1 / (s1 / s2 + (double)fabsf((float)(2.0 * r - 1.0))) - s1 / (s1 / s2 + 1.0);
  if (r < 0.5) {
    const double x1 = x + dx;
    return word_of((x1 < 1.0) ? x1 : x1 - 1.0);
  }
  const double x1 = x - dx;
  return word_of((x1 < 0.0) ? x1 + 1.0 : x1);
}

__device__ __forceinline__ bool large_step(const MltChains& C, uint32_t chain, uint32_t iter, double p_large) {
  uint32_t b[4];
  philox(chain, iter, MLT_TAG_LARGE, 0u, C.k0, C.k1, b);
  return u01(b[0]) < p_large;
}

__device__ __forceinline__ void seed_words(const MltChains& C, uint32_t seed, uint32_t blk, uint32_t w[4]) {
  philox(seed, 0u, MLT_TAG_SEED, blk, C.k0, C.k1, w);
}

// mlt.h:171-174 with the Vector3d dot pairing
__device__ __forceinline__ double lum3(const double f[3]) { return (0.2126 * f[0] + 0.0722 * f[2]) + 0.7152 * f[1]; }

__device__ __forceinline__ void splat(uint64_t* img, uint32_t pix, double w, const double f[3]) {
  for (int k = 0; k < 3; ++k) {
    const double v = w * f[k];
    if (v > 0.0) atomicAdd((unsigned long long*)&img[3 * (size_t)pix + k], (unsigned long long)(v * MLT_FIX_ONE + 0.5));
  }
}

// One thread per (chain, block of 4 words).
__global__ __launch_bounds__(256) void k_mlt_mutate(MltChains C, int mode, uint32_t first, uint32_t n, uint32_t iter,
                                                    double p_large) {
  const uint32_t blocks = C.stride / 4u;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * blocks) return;
  const uint32_t chain = (uint32_t)(t / blocks), blk = (uint32_t)(t % blocks);
  uint4* out = reinterpret_cast<uint4*>(C.prop + (size_t)chain * C.stride) + blk;
  uint32_t w[4];
  if (mode == 0) {
    seed_words(C, first + chain, blk, w);
  } else if (large_step(C, chain, iter, p_large)) {
    // large step: fresh primary samples (mlt.h:116-121)
    philox(chain, iter, MLT_TAG_FRESH, blk, C.k0, C.k1, w);
  } else {
    // small step: Mutate0 of every word (mlt.h:123-135; mutating the words a
    // path does not use changes nothing it evaluates, and the lazy catch-up
    // there applies the same mutations when a word is next used)
    const uint4 v = reinterpret_cast<const uint4*>(C.cur + (size_t)chain * C.stride)[blk];
    uint32_t r[4];
    philox(chain, iter, MLT_TAG_MUTATE, blk, C.k0, C.k1, r);
    w[0] = mutate0(v.x, u01(r[0]));
    w[1] = mutate0(v.y, u01(r[1]));
    w[2] = mutate0(v.z, u01(r[2]));
    w[3] = mutate0(v.w, u01(r[3]));
  }
  *out = make_uint4(w[0], w[1], w[2], w[3]);
}

// generate_new_path's camera part (render.h:262-300) with the tent filter the
// MLT mode switches on (scene_env.h:1954): pixel (u0 W, u1 H), sub-pixel
// (u2 < 1/2, u3 < 1/2) of a 2x2 grid, tent offsets from u4, u5.
__global__ __launch_bounds__(256) void k_mlt_gen(Params P, uint32_t n, RayQ q) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* u = P.mlt_u + (size_t)i * (uint32_t)P.mlt_stride;
  const int x = min((int)(u01(u[0]) * (double)P.W), P.W - 1);
  const int y = min((int)(u01(u[1]) * (double)P.H), P.H - 1);
  const int sx = u01(u[2]) < 0.5 ? 0 : 1;
  const int sy = u01(u[3]) < 0.5 ? 0 : 1;
  const double rr1 = 2.0 * u01(u[4]), rr2 = 2.0 * u01(u[5]);
  const double dx = rr1 < 1.0 ? sqrt(rr1) - 1.0 : 1.0 - sqrt(2.0 - rr1);
  const double dy = rr2 < 1.0 ? sqrt(rr2) - 1.0 : 1.0 - sqrt(2.0 - rr2);
  const double r1 = (sx + 0.5 + dx) * 0.5;
  const double r2 = (sy + 0.5 + dy) * 0.5;
  const d3 sp = ld3(P.center) + ld3(P.screen_x) * ((r1 + x) * P.iw - 0.5) + ld3(P.screen_y) * ((r2 + y) * P.ih - 0.5);
  const d3 o = ld3(P.cam_pos);
  const d3 d = normalize(sp - o);
  q.ox[i] = o.x; q.oy[i] = o.y; q.oz[i] = o.z;
  q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
  P.mlt_pix[i] = (uint32_t)((P.H - y - 1) * P.W + x);  // render.h:585-586
}

__global__ __launch_bounds__(256) void k_mlt_seed_record(const uint64_t* __restrict__ accum,
                                                         const uint32_t* __restrict__ pix_new, uint32_t first,
                                                         uint32_t n, MltSeeds S) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f[3] = {accum_get(accum + 3 * (size_t)i), accum_get(accum + 3 * (size_t)i + 1),
                       accum_get(accum + 3 * (size_t)i + 2)};
  const size_t j = (size_t)first + i;
  S.f[3 * j] = f[0];
  S.f[3 * j + 1] = f[1];
  S.f[3 * j + 2] = f[2];
  // luminance in fixed point, capped so that 2^24 seeds cannot overflow
  const double l = lum3(f) * MLT_LUM_ONE + 0.5;
  S.lum[j] = l < 1099511627776.0 ? (uint64_t)l : 1099511627776ull;
  S.pix[j] = pix_new[i];
}

// Scan of MLT_SCAN_TILE-entry tiles: inclusive prefix within the tile, tile
// totals to tile_sums.
__global__ __launch_bounds__(256) void k_mlt_scan_tiles(MltSeeds S) {
  __shared__ uint64_t s_part[256];
  const uint32_t base = blockIdx.x * MLT_SCAN_TILE + threadIdx.x * 4u;
  uint64_t v[4];
  uint64_t run = 0;
  for (int k = 0; k < 4; ++k) {
    const uint32_t j = base + (uint32_t)k;
    run += j < S.n ? S.lum[j] : 0ull;
    v[k] = run;
  }
  s_part[threadIdx.x] = run;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint64_t add = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0ull;
    __syncthreads();
    s_part[threadIdx.x] += add;
    __syncthreads();
  }
  const uint64_t before = threadIdx.x ? s_part[threadIdx.x - 1] : 0ull;
  for (int k = 0; k < 4; ++k) {
    const uint32_t j = base + (uint32_t)k;
    if (j < S.n) S.cdf[j] = v[k] + before;
  }
  if (threadIdx.x == 255) S.tile_sums[blockIdx.x] = s_part[255];
}

// One block: exclusive prefix of the tile totals, in place.
__global__ __launch_bounds__(256) void k_mlt_scan_sums(MltSeeds S, uint32_t n_tiles) {
  __shared__ uint64_t s_part[256];
  uint64_t carry = 0;
  for (uint32_t b0 = 0; b0 < n_tiles; b0 += 256) {
    const uint32_t j = b0 + threadIdx.x;
    const uint64_t x = j < n_tiles ? S.tile_sums[j] : 0ull;
    s_part[threadIdx.x] = x;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
      const uint64_t add = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0ull;
      __syncthreads();
      s_part[threadIdx.x] += add;
      __syncthreads();
    }
    if (j < n_tiles) S.tile_sums[j] = carry + s_part[threadIdx.x] - x;
    const uint64_t total = s_part[255];
    __syncthreads();
    carry += total;
  }
}

__global__ __launch_bounds__(256) void k_mlt_scan_add(MltSeeds S) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= S.n) return;
  S.cdf[j] += S.tile_sums[j / MLT_SCAN_TILE];
}

// Start path of each chain: the first seed whose inclusive CDF exceeds
// u * total (mlt.h:201-220 picks the first with r <= accumulated, the same
// rule up to the boundary); then its primary samples, radiance and pixel.
__global__ __launch_bounds__(256) void k_mlt_init(MltChains C, MltSeeds S) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C.n) return;
  const uint64_t total = S.cdf[S.n - 1];
  uint32_t b[4];
  philox(c, 0u, MLT_TAG_INIT, 0u, C.k0, C.k1, b);
  uint64_t r = (uint64_t)(u01(b[0]) * (double)total);
  if (r >= total) r = total - 1;
  uint32_t lo = 0, hi = S.n - 1;  // smallest i with cdf[i] > r
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2u;
    if (S.cdf[mid] > r) hi = mid;
    else lo = mid + 1u;
  }
  const uint32_t seed = lo;
  uint4* cur = reinterpret_cast<uint4*>(C.cur + (size_t)c * C.stride);
  for (uint32_t blk = 0; blk < C.stride / 4u; ++blk) {
    uint32_t w[4];
    seed_words(C, seed, blk, w);
    cur[blk] = make_uint4(w[0], w[1], w[2], w[3]);
  }
  C.f_old[3 * (size_t)c] = S.f[3 * (size_t)seed];
  C.f_old[3 * (size_t)c + 1] = S.f[3 * (size_t)seed + 1];
  C.f_old[3 * (size_t)c + 2] = S.f[3 * (size_t)seed + 2];
  C.pix_old[c] = S.pix[seed];
}

// render.h:574-630: a = min(1, L_new / L_old) (1 when L_old is ~0); the
// proposal gets (a + large) / (L_new / b + p_large) / M, the current path
// (1 - a) / (L_old / b + p_large) / M, each times the path weight 4*W*H and
// its radiance; then accept with probability a.
__global__ __launch_bounds__(256) void k_mlt_step(MltChains C, const uint64_t* __restrict__ accum, uint32_t iter,
                                                  double p_large, double b, double weight, double inv_m,
                                                  uint64_t* ctr) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  bool acc = false;
  if (c < C.n) {
    const double fn[3] = {accum_get(accum + 3 * (size_t)c), accum_get(accum + 3 * (size_t)c + 1),
                          accum_get(accum + 3 * (size_t)c + 2)};
    const double fo[3] = {C.f_old[3 * (size_t)c], C.f_old[3 * (size_t)c + 1], C.f_old[3 * (size_t)c + 2]};
    const double l_new = lum3(fn), l_old = lum3(fo);
    const double a = l_old > 1.0e-16 ? fmin(1.0, l_new / l_old) : 1.0;
    const double large = large_step(C, c, iter, p_large) ? 1.0 : 0.0;
    const double w_new = (a + large) / (l_new / b + p_large) * inv_m;
    const double w_old = (1.0 - a) / (l_old / b + p_large) * inv_m;
    const uint32_t p_new = C.pix_new[c], p_old = C.pix_old[c];
    splat(C.img, p_old, w_old * weight, fo);
    splat(C.img, p_new, w_new * weight, fn);
    uint32_t r[4];
    philox(c, iter, MLT_TAG_ACCEPT, 0u, C.k0, C.k1, r);
    if (u01(r[0]) < a) {
      acc = true;
      const uint4* src = reinterpret_cast<const uint4*>(C.prop + (size_t)c * C.stride);
      uint4* dst = reinterpret_cast<uint4*>(C.cur + (size_t)c * C.stride);
      for (uint32_t blk = 0; blk < C.stride / 4u; ++blk) dst[blk] = src[blk];
      C.f_old[3 * (size_t)c] = fn[0];
      C.f_old[3 * (size_t)c + 1] = fn[1];
      C.f_old[3 * (size_t)c + 2] = fn[2];
      C.pix_old[c] = p_new;
    }
  }
  const uint64_t n_acc = __popcll(__ballot(acc));
  if (lane_id() == 0 && n_acc) atomicAdd((unsigned long long*)&ctr[CTR_MLT_ACCEPT], (unsigned long long)n_acc);
}

__global__ __launch_bounds__(256) void k_mlt_resolve(const uint64_t* __restrict__ img, int64_t n, double scale,
                                                     void* out, int as_f64) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = (double)img[i] * MLT_FIX_INV * scale;
    if (as_f64) static_cast<double*>(out)[i] = v;
    else static_cast<float*>(out)[i] = (float)v;
  }
}

// ---------------------------------------------------------------- ERPT
// Block sums of the batch's path radiance (accumulators >> 8: 2^-24 fixed
// point), one atomic per block and channel.
__global__ __launch_bounds__(256) void k_erpt_sum(const uint64_t* __restrict__ accum, uint32_t n, uint64_t* sums) {
  __shared__ uint64_t s_sum[3][4];
  uint64_t v[3] = {0, 0, 0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int k = 0; k < 3; ++k) v[k] += accum[3 * (size_t)i + k] >> 8;
  for (int k = 0; k < 3; ++k) {
    uint64_t x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane_id() == 0) s_sum[k][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const uint64_t t = s_sum[threadIdx.x][0] + s_sum[threadIdx.x][1] + s_sum[threadIdx.x][2] + s_sum[threadIdx.x][3];
    if (t) atomicAdd((unsigned long long*)&sums[threadIdx.x], (unsigned long long)t);
  }
}

__global__ __launch_bounds__(256) void k_erpt_paths(const uint64_t* __restrict__ accum, ErptPaths E, uint32_t first,
                                                    uint64_t* img, double inv_samples, double ed, double mutation,
                                                    uint32_t k0, uint32_t k1) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E.n) return;
  const double f[3] = {accum_get(accum + 3 * (size_t)i), accum_get(accum + 3 * (size_t)i + 1),
                       accum_get(accum + 3 * (size_t)i + 2)};
  const double l = lum3(f);
  E.f[3 * (size_t)i] = f[0];
  E.f[3 * (size_t)i + 1] = f[1];
  E.f[3 * (size_t)i + 2] = f[2];
  E.lum[i] = l;
  uint64_t chains = 0;
  if (E.direct[i]) {
    splat(img, E.pix[i], inv_samples, f);  // render.h:927-931
  } else if (l > 0.0 && ed > 0.0) {
    uint32_t r[4];
    philox(first + i, 0u, ERPT_TAG_CHAINS, 0u, k0, k1, r);
    const double c = floor(u01(r[0]) + l / (mutation * ed));  // render.h:938
    chains = c > 0.0 ? (uint64_t)c : 0ull;
  }
  E.count[i] = chains;
}

__global__ __launch_bounds__(256) void k_erpt_chain_init(ErptChains C, ErptPaths E, const uint32_t* __restrict__ path_u,
                                                         uint64_t j0, uint32_t first, int ss2, int width, int height) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= C.n) return;
  const uint64_t J = j0 + j;
  uint32_t lo = 0, hi = E.n - 1;  // the path whose chain range holds J
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2u;
    if (E.first[mid] > J) hi = mid;
    else lo = mid + 1u;
  }
  const uint32_t p = lo;
  const uint4* src = reinterpret_cast<const uint4*>(path_u + (size_t)p * C.stride);
  uint4* dst = reinterpret_cast<uint4*>(C.cur + (size_t)j * C.stride);
  for (uint32_t blk = 0; blk < C.stride / 4u; ++blk) dst[blk] = src[blk];
  C.f_y[3 * (size_t)j] = E.f[3 * (size_t)p];
  C.f_y[3 * (size_t)j + 1] = E.f[3 * (size_t)p + 1];
  C.f_y[3 * (size_t)j + 2] = E.f[3 * (size_t)p + 2];
  C.pix_y[j] = E.pix[p];
  // the loop pixel of initial path g = (y*W + x)*SS^2 + i (render.h:897-917)
  const uint32_t lp = (first + p) / (uint32_t)ss2;
  const int x = (int)(lp % (uint32_t)width), y = (int)(lp / (uint32_t)width);
  C.pix_now[j] = (uint32_t)((height - y - 1) * width + x);
  C.run[j] = 0u;
}

__global__ __launch_bounds__(256) void k_erpt_perturb(ErptChains C, uint32_t step, double d) {
  const uint32_t blocks = C.stride / 4u;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)C.n * blocks) return;
  const uint32_t j = (uint32_t)(t / blocks), blk = (uint32_t)(t % blocks);
  const uint4 v = reinterpret_cast<const uint4*>(C.cur + (size_t)j * C.stride)[blk];
  uint32_t r[4];
  philox(C.base + j, step, ERPT_TAG_PERTURB, blk, C.k0, C.k1, r);
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  uint32_t o[4];
  for (int k = 0; k < 4; ++k) {
    double x = u01(w[k]) + d * (2.0 * u01(r[k]) - 1.0);  // erpt.h:29-34
    if (x > 1.0) x -= 1.0;
    if (x < 0.0) x += 1.0;
    o[k] = word_of(x);
  }
  reinterpret_cast<uint4*>(C.prop + (size_t)j * C.stride)[blk] = make_uint4(o[0], o[1], o[2], o[3]);
}

__global__ __launch_bounds__(256) void k_erpt_step(ErptChains C, const uint64_t* __restrict__ accum, uint32_t step,
                                                   uint64_t* ctr) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  bool acc = false;
  if (j < C.n) {
    const double fz[3] = {accum_get(accum + 3 * (size_t)j), accum_get(accum + 3 * (size_t)j + 1),
                          accum_get(accum + 3 * (size_t)j + 2)};
    double fy[3] = {C.f_y[3 * (size_t)j], C.f_y[3 * (size_t)j + 1], C.f_y[3 * (size_t)j + 2]};
    const double lz = lum3(fz), ly = lum3(fy);
    const double q = fmin(1.0, ly > 1.0e-16 ? lz / ly : 0.0);  // render.h:978-980
    uint32_t r[4];
    philox(C.base + j, step, ERPT_TAG_ACCEPT, 0u, C.k0, C.k1, r);
    uint32_t py = C.pix_y[j];
    if (q > u01(r[0])) {
      acc = true;
      const uint4* src = reinterpret_cast<const uint4*>(C.prop + (size_t)j * C.stride);
      uint4* dst = reinterpret_cast<uint4*>(C.cur + (size_t)j * C.stride);
      for (uint32_t blk = 0; blk < C.stride / 4u; ++blk) dst[blk] = src[blk];
      for (int k = 0; k < 3; ++k) {
        fy[k] = fz[k];
        C.f_y[3 * (size_t)j + k] = fz[k];
      }
      py = C.pix_new[j];
      C.pix_y[j] = py;
    }
    uint32_t run = C.run[j];
    if (C.pix_now[j] == py) {
      ++run;
    } else {
      C.pix_now[j] = py;
      run = 0;
    }
    C.run[j] = run;
    const double l = lum3(fy);
    if (run < ERPT_MAX_STACK && l > 0.0) splat(C.img, py, C.dep / l, fy);
  }
  const uint64_t n_acc = __popcll(__ballot(acc));
  if (lane_id() == 0 && n_acc) atomicAdd((unsigned long long*)&ctr[CTR_MLT_ACCEPT], (unsigned long long)n_acc);
}

inline unsigned grid_of(uint64_t n, unsigned block = 256) { return (unsigned)((n + block - 1) / block); }

}  // namespace

hipError_t launch_mlt_mutate(const MltChains& C, int mode, uint32_t first, uint32_t n, uint32_t iter,
                             double p_large, hipStream_t s) {
  if (C.stride % 4u != 0 || n > C.n) return hipErrorInvalidValue;
  const uint64_t threads = (uint64_t)n * (C.stride / 4u);
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlt_mutate, dim3(grid_of(threads)), dim3(256), 0, s, C, mode, first, n, iter, p_large);
  return hipGetLastError();
}

hipError_t launch_mlt_gen(const Params& P, uint32_t n, const RayQ& q, hipStream_t s) {
  if (!P.mlt_u || !P.mlt_pix || n > q.cap) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlt_gen, dim3(grid_of(n)), dim3(256), 0, s, P, n, q);
  return hipGetLastError();
}

hipError_t launch_mlt_seed_record(const uint64_t* accum, const uint32_t* pix_new, uint32_t first, uint32_t n,
                                  const MltSeeds& S, hipStream_t s) {
  if ((uint64_t)first + n > S.n) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlt_seed_record, dim3(grid_of(n)), dim3(256), 0, s, accum, pix_new, first, n, S);
  return hipGetLastError();
}

hipError_t launch_mlt_scan(const MltSeeds& S, hipStream_t s) {
  if (S.n == 0) return hipErrorInvalidValue;
  const uint32_t tiles = (S.n + MLT_SCAN_TILE - 1) / MLT_SCAN_TILE;
  hipLaunchKernelGGL(k_mlt_scan_tiles, dim3(tiles), dim3(256), 0, s, S);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlt_scan_sums, dim3(1), dim3(256), 0, s, S, tiles);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlt_scan_add, dim3(grid_of(S.n)), dim3(256), 0, s, S);
  return hipGetLastError();
}

hipError_t launch_mlt_init(const MltChains& C, const MltSeeds& S, hipStream_t s) {
  if (S.n == 0 || C.stride % 4u != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mlt_init, dim3(grid_of(C.n)), dim3(256), 0, s, C, S);
  return hipGetLastError();
}

hipError_t launch_mlt_step(const MltChains& C, const uint64_t* accum, uint32_t iter, double p_large, double b,
                           double weight, double inv_m, uint64_t* ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_mlt_step, dim3(grid_of(C.n)), dim3(256), 0, s, C, accum, iter, p_large, b, weight, inv_m,
                     ctr);
  return hipGetLastError();
}

hipError_t launch_mlt_resolve(const uint64_t* img, int64_t n, double scale, void* out, int as_f64, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_mlt_resolve, dim3(g), dim3(256), 0, s, img, n, scale, out, as_f64);
  return hipGetLastError();
}

hipError_t launch_erpt_sum(const uint64_t* accum, uint32_t n, uint64_t* sums, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<uint64_t>(grid_of(n), 2048);
  hipLaunchKernelGGL(k_erpt_sum, dim3(g), dim3(256), 0, s, accum, n, sums);
  return hipGetLastError();
}

hipError_t launch_erpt_paths(const uint64_t* accum, const ErptPaths& E, uint32_t first, uint64_t* img,
                             double inv_samples, double ed, double mutation, uint32_t k0, uint32_t k1,
                             hipStream_t s) {
  if (E.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_erpt_paths, dim3(grid_of(E.n)), dim3(256), 0, s, accum, E, first, img, inv_samples, ed,
                     mutation, k0, k1);
  return hipGetLastError();
}

hipError_t launch_erpt_chain_init(const ErptChains& C, const ErptPaths& E, const uint32_t* path_u, uint64_t j0,
                                  uint32_t first, int ss2, int width, int height, hipStream_t s) {
  if (C.n == 0) return hipSuccess;
  if (E.n == 0 || C.stride % 4u != 0 || ss2 <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_erpt_chain_init, dim3(grid_of(C.n)), dim3(256), 0, s, C, E, path_u, j0, first, ss2, width,
                     height);
  return hipGetLastError();
}

hipError_t launch_erpt_perturb(const ErptChains& C, uint32_t step, double d, hipStream_t s) {
  if (C.stride % 4u != 0) return hipErrorInvalidValue;
  const uint64_t threads = (uint64_t)C.n * (C.stride / 4u);
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_erpt_perturb, dim3(grid_of(threads)), dim3(256), 0, s, C, step, d);
  return hipGetLastError();
}

hipError_t launch_erpt_step(const ErptChains& C, const uint64_t* accum, uint32_t step, uint64_t* ctr,
                            hipStream_t s) {
  if (C.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_erpt_step, dim3(grid_of(C.n)), dim3(256), 0, s, C, accum, step, ctr);
  return hipGetLastError();
}
