// prt_mlt.h — device side of the primary-sample-space Metropolis renderer
// (Kelemen et al. 2002 as Render::RunMLT runs it, render.h:360-780).
//
// Each Markov chain is one slot of a wavefront batch: per iteration the
// chain's proposal (its primary samples mutated, or all fresh for a large
// step) is traced through the same extend / shade / NEE / shadow kernels as a
// camera sample, and k_mlt_step then splats the current and the proposed path
// with Kelemen's weights and accepts or rejects.  All random decisions of the
// chain logic come from Philox streams keyed by (chain, iteration, tag), so a
// render is a pure function of the seed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prt_device.h"
#include "prt_kernels.h"

// Counter tags of the chain-control Philox streams (third counter word).
enum : uint32_t {
  MLT_TAG_LARGE = 0x4C415247u,   // large-step choice per (chain, iteration)
  MLT_TAG_MUTATE = 0x4D555441u,  // Mutate0 numbers per (chain, iteration, word block)
  MLT_TAG_FRESH = 0x46524553u,   // large-step primary samples per (chain, iteration, word block)
  MLT_TAG_ACCEPT = 0x41434350u,  // acceptance number per (chain, iteration)
  MLT_TAG_INIT = 0x494E4954u,    // start-path choice per chain
  MLT_TAG_SEED = 0x53454544u     // primary samples of seed path s per word block
};

// Fixed-point scale of the splatted image (order-independent integer sums).
#define MLT_FIX_ONE 1048576.0
#define MLT_FIX_INV 9.5367431640625e-07
// Fixed-point scale of seed-path luminance in the start-path CDF.
#define MLT_LUM_ONE 16777216.0

// Chain state of one render.
struct MltChains {
  uint32_t* cur;        // [n][stride] primary samples of each chain's current path
  uint32_t* prop;       // [n][stride] of its proposal (Params::mlt_u)
  uint32_t* pix_new;    // [n] image element of the proposal (Params::mlt_pix)
  uint32_t* pix_old;    // [n] image element of the current path
  double* f_old;        // [n][3] radiance of the current path
  uint64_t* img;        // [W*H*3] splatted image, fixed point (MLT_FIX_ONE)
  uint32_t n;           // chains
  uint32_t stride;      // words per chain
  uint32_t k0, k1;      // Philox key of the chain-control streams
};

// Seed paths of one render.
struct MltSeeds {
  double* f;            // [N][3] radiance
  uint64_t* lum;        // [N] luminance, fixed point (MLT_LUM_ONE)
  uint32_t* pix;        // [N] image element
  uint64_t* cdf;        // [N] inclusive prefix sums of lum
  uint64_t* tile_sums;  // [ceil(N / MLT_SCAN_TILE)] scratch of the scan
  uint32_t n;
};
#define MLT_SCAN_TILE 1024u

// mode 0: the n chains of the batch are seed paths first..first+n-1, all
// primary samples fresh (MLT_TAG_SEED); mode 1: iteration `iter` of every
// chain, a large step with probability p_large, else Mutate0 of every word.
hipError_t launch_mlt_mutate(const MltChains& C, int mode, uint32_t first, uint32_t n, uint32_t iter,
                             double p_large, hipStream_t s);
// Camera rays of the n proposals into q (entries 0..n-1) and their image
// elements into P.mlt_pix.
hipError_t launch_mlt_gen(const Params& P, uint32_t n, const RayQ& q, hipStream_t s);
// Radiance, luminance and image element of seed paths first..first+n-1
// from the batch's accumulators.
hipError_t launch_mlt_seed_record(const uint64_t* accum, const uint32_t* pix_new, uint32_t first, uint32_t n,
                                  const MltSeeds& S, hipStream_t s);
// Inclusive prefix sums of S.lum into S.cdf (exact: integers).
hipError_t launch_mlt_scan(const MltSeeds& S, hipStream_t s);
// Start path of every chain: seed path drawn in proportion to luminance
// (mlt.h:201-220), its primary samples, radiance and image element.
hipError_t launch_mlt_init(const MltChains& C, const MltSeeds& S, hipStream_t s);
// Acceptance and splatting of iteration `iter` (render.h:574-630): b is the
// mean seed luminance, weight = 4*W*H (tent filter pdf), inv_m = 1/steps.
hipError_t launch_mlt_step(const MltChains& C, const uint64_t* accum, uint32_t iter, double p_large, double b,
                           double weight, double inv_m, uint64_t* ctr, hipStream_t s);
// ---- energy redistribution path tracing (Render::RunERPT, render.h:786-1123)
enum : uint32_t {
  ERPT_TAG_CHAINS = 0x4E434841u,   // chain count of initial path g
  ERPT_TAG_PERTURB = 0x50455254u,  // perturbation numbers per (chain, step, word block)
  ERPT_TAG_ACCEPT = 0x45414343u    // acceptance number per (chain, step)
};
#define ERPT_MAX_STACK 20  // render.h:969

// Initial paths of one batch: radiance, luminance, image element, chain count.
struct ErptPaths {
  double* f;         // [n][3]
  double* lum;       // [n]
  uint32_t* pix;     // [n] (Params::mlt_pix of the batch)
  uint32_t* direct;  // [n] (Params::mlt_flags of the batch)
  uint64_t* count;   // [n] chains started from the path
  uint64_t* first;   // [n] inclusive prefix of count
  uint64_t* tile_sums;
  uint32_t n;
};

// One sub-batch of chains.  A chain proposes a perturbation of its current
// path's primary samples, accepts with min(1, L_z / L_y) and after every step
// deposits ed / samples worth of its current path's colour at that path's
// pixel (Cline et al. 2005, as render.h:955-1008 lays the loop out).
struct ErptChains {
  uint32_t* cur;      // [n][stride] current path's primary samples
  uint32_t* prop;     // [n][stride] proposal (Params::mlt_u)
  uint32_t* pix_new;  // [n] proposal's image element (Params::mlt_pix)
  uint32_t* pix_y;    // [n] current path's image element
  uint32_t* pix_now;  // [n] element of the current run of deposits
  uint32_t* run;      // [n] length of that run (stack_num)
  double* f_y;        // [n][3] current path's radiance
  uint64_t* img;      // [W*H*3] image, fixed point (MLT_FIX_ONE)
  uint32_t n, stride, k0, k1;
  uint32_t base;      // global index of chain 0 of the sub-batch (Philox counter)
  double dep;         // ed / samples: the luminance of one deposit
};

// Sum of the batch's path radiance, fixed point 2^-24, into sums[3].
hipError_t launch_erpt_sum(const uint64_t* accum, uint32_t n, uint64_t* sums, hipStream_t s);
// Initial paths first..first+n-1 (global index, row-major over pixels, SS^2
// per pixel): a path that returned a light directly is splatted as
// F/samples; any other path with luminance > 0 starts
// floor(u + L / (mutation * ed)) chains (render.h:927-938).
hipError_t launch_erpt_paths(const uint64_t* accum, const ErptPaths& E, uint32_t first, uint64_t* img,
                             double inv_samples, double ed, double mutation, uint32_t k0, uint32_t k1,
                             hipStream_t s);
// Chains j0..j0+C.n-1 of the batch's path list: parent path, start state;
// the run element starts at the parent's loop pixel (render.h:971).
hipError_t launch_erpt_chain_init(const ErptChains& C, const ErptPaths& E, const uint32_t* path_u, uint64_t j0,
                                  uint32_t first, int ss2, int width, int height, hipStream_t s);
// Proposals: every primary sample moved by d * (2u - 1), wrapped (erpt.h:29-34).
hipError_t launch_erpt_perturb(const ErptChains& C, uint32_t step, double d, hipStream_t s);
// Acceptance min(1, L_z / L_y) and the deposit, skipped once 20 deposits in a
// row fell on one pixel (render.h:975-1007).
hipError_t launch_erpt_step(const ErptChains& C, const uint64_t* accum, uint32_t step, uint64_t* ctr,
                            hipStream_t s);

// Final image: img * scale, float or double.
hipError_t launch_mlt_resolve(const uint64_t* img, int64_t n, double scale, void* out, int as_f64, hipStream_t s);
