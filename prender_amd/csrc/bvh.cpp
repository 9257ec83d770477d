// bvh.cpp — binned-SAH BVH build and 4-wide collapse (host, multi-threaded).
#include "bvh.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace {

struct Box {
  float lo[3], hi[3];
  void clear() {
    for (int k = 0; k < 3; ++k) { lo[k] = std::numeric_limits<float>::infinity(); hi[k] = -lo[k]; }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
  }
  void grow(const float* p) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
  }
  float half_area() const {
    if (!(hi[0] >= lo[0])) return 0.0f;
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dx * dz + dy * dz;
  }
};

struct BNode {
  Box box;
  int32_t left = -1, right = -1;
  uint32_t first = 0, count = 0;  // leaf when count > 0
};

constexpr int kBins = 32;
// 1: 4-wide nodes by the cost-optimal collapse (DpCollapser), 0: by opening
// the largest child until four (Collapser)
#ifndef PRT_BVH_DP_COLLAPSE
#define PRT_BVH_DP_COLLAPSE 1
#endif
// SAH costs of a binary node step and of a triangle test
#ifndef PRT_SAH_COST_TRI
#define PRT_SAH_COST_TRI 1.0f
#endif
constexpr float kCostNode = 1.0f;
constexpr float kCostTri = PRT_SAH_COST_TRI;
constexpr int64_t kParallelMin = 1 << 15;

struct Builder {
  std::vector<Box> tbox;
  std::vector<float> cen;  // 3 per tri
  std::vector<uint32_t> idx;
  std::vector<BNode> nodes;
  std::atomic<int32_t> next{0};
  int spawn_depth = 0;

  int32_t alloc(int n) { return next.fetch_add(n); }

  void build(int32_t ni, uint32_t b, uint32_t e, int level) {
    BNode& node = nodes[ni];
    Box bb, cb;
    bb.clear();
    cb.clear();
    for (uint32_t i = b; i < e; ++i) {
      bb.grow(tbox[idx[i]]);
      cb.grow(&cen[3 * (size_t)idx[i]]);
    }
    node.box = bb;
    const uint32_t n = e - b;
    if (n <= 2) { node.first = b; node.count = n; return; }
    // binned SAH over the centroid bounds
    float best_cost = std::numeric_limits<float>::infinity();
    int best_axis = -1, best_bin = -1;
    for (int k = 0; k < 3; ++k) {
      const float ext = cb.hi[k] - cb.lo[k];
      if (!(ext > 0.0f)) continue;
      const float scale = kBins / ext;
      Box bins[kBins];
      uint32_t cnt[kBins] = {0};
      for (int j = 0; j < kBins; ++j) bins[j].clear();
      for (uint32_t i = b; i < e; ++i) {
        int j = (int)((cen[3 * (size_t)idx[i] + k] - cb.lo[k]) * scale);
        j = std::min(std::max(j, 0), kBins - 1);
        cnt[j]++;
        bins[j].grow(tbox[idx[i]]);
      }
      float right_area[kBins];
      uint32_t right_cnt[kBins];
      Box acc;
      acc.clear();
      uint32_t c = 0;
      for (int j = kBins - 1; j > 0; --j) {
        acc.grow(bins[j]);
        c += cnt[j];
        right_area[j] = acc.half_area();
        right_cnt[j] = c;
      }
      acc.clear();
      c = 0;
      for (int j = 0; j < kBins - 1; ++j) {
        acc.grow(bins[j]);
        c += cnt[j];
        if (c == 0 || right_cnt[j + 1] == 0) continue;
        const float cost = acc.half_area() * c + right_area[j + 1] * right_cnt[j + 1];
        if (cost < best_cost) { best_cost = cost; best_axis = k; best_bin = j; }
      }
    }
    const float parent_area = std::max(bb.half_area(), 1e-30f);
    const float split_cost = kCostNode + kCostTri * best_cost / parent_area;
    const float leaf_cost = kCostTri * n;
    uint32_t mid;
    if (best_axis < 0 || (n <= PRT_MAX_LEAF && leaf_cost <= split_cost)) {
      if (n <= PRT_MAX_LEAF) { node.first = b; node.count = n; return; }
      // degenerate centroids: split by index
      mid = b + n / 2;
    } else {
      const int k = best_axis;
      const float scale = kBins / (cb.hi[k] - cb.lo[k]);
      auto left_of = [&](uint32_t t) {
        int j = (int)((cen[3 * (size_t)t + k] - cb.lo[k]) * scale);
        j = std::min(std::max(j, 0), kBins - 1);
        return j <= best_bin;
      };
      mid = (uint32_t)(std::partition(idx.begin() + b, idx.begin() + e, left_of) - idx.begin());
      if (mid == b || mid == e) mid = b + n / 2;
    }
    const int32_t c0 = alloc(2);
    node.left = c0;
    node.right = c0 + 1;
    if (level < spawn_depth && n >= kParallelMin) {
      auto fut = std::async(std::launch::async, [=] { build(c0, b, mid, level + 1); });
      build(c0 + 1, mid, e, level + 1);
      fut.get();
    } else {
      build(c0, b, mid, level + 1);
      build(c0 + 1, mid, e, level + 1);
    }
  }
};

struct Collapser {
  const std::vector<BNode>& bn;
  std::vector<DevNode>& out;
  int max_depth = 0;

  int32_t emit(int32_t bi, int depth) {
    max_depth = std::max(max_depth, depth);
    const int32_t me = (int32_t)out.size();
    out.emplace_back();
    // open the internal child with the largest area until 4 children
    int32_t ch[4] = {bn[bi].left, bn[bi].right, -1, -1};
    int nc = 2;
    while (nc < 4) {
      int pick = -1;
      float best = -1.0f;
      for (int i = 0; i < nc; ++i) {
        const BNode& c = bn[ch[i]];
        if (c.count == 0 && c.box.half_area() > best) { best = c.box.half_area(); pick = i; }
      }
      if (pick < 0) break;
      const int32_t open = ch[pick];
      ch[pick] = bn[open].left;
      ch[nc++] = bn[open].right;
    }
    int32_t code[4];
    for (int i = 0; i < 4; ++i) {
      if (i >= nc) { code[i] = 0; continue; }
      const BNode& c = bn[ch[i]];
      if (c.count > 0) code[i] = prt_leaf_code(c.first, c.count);
      else code[i] = emit(ch[i], depth + 1);
    }
    DevNode& d = out[me];
    const float inf = std::numeric_limits<float>::infinity();
    for (int i = 0; i < 4; ++i) {
      if (i < nc) {
        const Box& b = bn[ch[i]].box;
        d.lo_x[i] = b.lo[0]; d.lo_y[i] = b.lo[1]; d.lo_z[i] = b.lo[2];
        d.hi_x[i] = b.hi[0]; d.hi_y[i] = b.hi[1]; d.hi_z[i] = b.hi[2];
      } else {  // empty slot: a box at +inf that no ray enters
        d.lo_x[i] = d.lo_y[i] = d.lo_z[i] = inf;
        d.hi_x[i] = d.hi_y[i] = d.hi_z[i] = inf;
      }
      d.child[i] = code[i];
      d.pad[i] = 0;
    }
    return me;
  }
};

// Cost-optimal collapse of the binary tree into 4-wide nodes (dynamic
// programming over the binary tree, in the manner of Ylitie, Karras and
// Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed
// Wide BVHs", HPG 2017, for width 4): every binary node n gets the least
// surface-area cost of its subtree as one root of a wide node's child list,
// either a leaf (when its triangles, a contiguous range, number at most
// PRT_MAX_LEAF) or a wide node whose up-to-4 children are the cheapest cut of
// n's subtree.  Cost = half area x (PRT_WIDE_COST_NODE per wide-node visit,
// PRT_WIDE_COST_TRI per triangle test).
#ifndef PRT_WIDE_COST_NODE
#define PRT_WIDE_COST_NODE 1.0f
#endif
#ifndef PRT_WIDE_COST_TRI
#define PRT_WIDE_COST_TRI 1.0f
#endif
struct DpCollapser {
  const std::vector<BNode>& bn;
  std::vector<DevNode>& out;
  int max_depth = 0;
  std::vector<float> cost;                 // best cost of n as one root
  std::vector<uint8_t> leaf;               // that best is a leaf
  std::vector<std::array<float, 5>> dcut;  // least cost of covering n's subtree with at most j roots
  std::vector<std::array<int8_t, 5>> kcut; // its split: 0 = n itself, k = k roots from the left child
  std::vector<uint32_t> first, count;

  DpCollapser(const std::vector<BNode>& b, std::vector<DevNode>& o)
      : bn(b), out(o), cost(b.size()), leaf(b.size()), dcut(b.size()), kcut(b.size()), first(b.size()),
        count(b.size()) {}

  void eval(int32_t n) {
    const BNode& b = bn[n];
    const float a = b.box.half_area();
    if (b.count > 0) {
      first[n] = b.first;
      count[n] = b.count;
      cost[n] = a * PRT_WIDE_COST_TRI * (float)b.count;
      leaf[n] = 1;
    } else {
      eval(b.left);
      eval(b.right);
      first[n] = first[b.left];
      count[n] = count[b.left] + count[b.right];
      float open = std::numeric_limits<float>::infinity();
      for (int k = 1; k <= 3; ++k) open = std::min(open, dcut[b.left][k] + dcut[b.right][4 - k]);
      const float wide = a * PRT_WIDE_COST_NODE + open;
      const float as_leaf = count[n] <= (uint32_t)PRT_MAX_LEAF ? a * PRT_WIDE_COST_TRI * (float)count[n]
                                                               : std::numeric_limits<float>::infinity();
      leaf[n] = as_leaf <= wide;
      cost[n] = leaf[n] ? as_leaf : wide;
    }
    dcut[n][0] = std::numeric_limits<float>::infinity();
    kcut[n][0] = 0;
    for (int j = 1; j <= 4; ++j) {
      dcut[n][j] = cost[n];
      kcut[n][j] = 0;
      if (b.count > 0 || j == 1) continue;
      for (int k = 1; k < j; ++k) {
        const float c = dcut[b.left][k] + dcut[b.right][j - k];
        if (c < dcut[n][j]) {
          dcut[n][j] = c;
          kcut[n][j] = (int8_t)k;
        }
      }
    }
  }
  void collect(int32_t n, int j, std::vector<int32_t>& roots) const {
    const int k = kcut[n][j];
    if (k == 0) {
      roots.push_back(n);
      return;
    }
    collect(bn[n].left, k, roots);
    collect(bn[n].right, j - k, roots);
  }
  // emits binary node n (internal) as a wide node
  int32_t emit(int32_t n, int depth) {
    max_depth = std::max(max_depth, depth);
    const int32_t me = (int32_t)out.size();
    out.emplace_back();
    const BNode& b = bn[n];
    int kbest = 1;
    float best = std::numeric_limits<float>::infinity();
    for (int k = 1; k <= 3; ++k) {
      const float c = dcut[b.left][k] + dcut[b.right][4 - k];
      if (c < best) { best = c; kbest = k; }
    }
    std::vector<int32_t> roots;
    collect(b.left, kbest, roots);
    collect(b.right, 4 - kbest, roots);
    int32_t code[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < roots.size(); ++i) {
      const int32_t r = roots[i];
      code[i] = leaf[r] ? prt_leaf_code(first[r], count[r]) : emit(r, depth + 1);
    }
    DevNode& d = out[me];
    const float inf = std::numeric_limits<float>::infinity();
    for (int i = 0; i < 4; ++i) {
      if (i < (int)roots.size()) {
        const Box& x = bn[roots[i]].box;
        d.lo_x[i] = x.lo[0]; d.lo_y[i] = x.lo[1]; d.lo_z[i] = x.lo[2];
        d.hi_x[i] = x.hi[0]; d.hi_y[i] = x.hi[1]; d.hi_z[i] = x.hi[2];
      } else {
        d.lo_x[i] = d.lo_y[i] = d.lo_z[i] = inf;
        d.hi_x[i] = d.hi_y[i] = d.hi_z[i] = inf;
      }
      d.child[i] = code[i];
      d.pad[i] = 0;
    }
    return me;
  }
};

// Renumbers the nodes so that those of the top PRT_TOP_LEVELS levels come
// first, breadth first (the traversal kernels keep these in LDS); the other
// nodes keep their depth-first order.  Internal child references (>= 0) are
// remapped; leaf codes (< 0) are unchanged.
void number_top_levels(BvhOutput& out) {
  const size_t n = out.nodes.size();
  std::vector<int32_t> top;
  top.push_back(0);
  for (size_t head = 0, level_end = 1, level = 1; head < top.size() && level < PRT_TOP_LEVELS;) {
    const DevNode& d = out.nodes[(size_t)top[head]];
    for (int i = 0; i < 4; ++i)
      if (d.child[i] > 0 && d.lo_x[i] != std::numeric_limits<float>::infinity()) top.push_back(d.child[i]);
    if (++head == level_end) {
      level_end = top.size();
      ++level;
    }
  }
  std::vector<int32_t> perm(n, -1);  // old index -> new index
  for (size_t k = 0; k < top.size(); ++k) perm[(size_t)top[k]] = (int32_t)k;
  int32_t next = (int32_t)top.size();
  for (size_t i = 0; i < n; ++i)
    if (perm[i] < 0) perm[i] = next++;
  std::vector<DevNode> nodes(n);
  for (size_t i = 0; i < n; ++i) {
    DevNode d = out.nodes[i];
    for (int c = 0; c < 4; ++c)
      if (d.child[c] >= 0) d.child[c] = perm[(size_t)d.child[c]];
    nodes[(size_t)perm[i]] = d;
  }
  out.nodes.swap(nodes);
  out.n_top = (int)top.size();
}

}  // namespace

bool prt_build_bvh(const BvhInput& in, BvhOutput& out, const char** err) {
  const auto t0 = std::chrono::steady_clock::now();
  out.nodes.clear();
  out.order.clear();
  out.depth = 0;
  const int64_t n = in.n_tris;
  if (n <= 0) return true;
  if (n >= (int64_t)1 << 27) { *err = "too many triangles for the leaf encoding (max 2^27)"; return false; }
  Builder B;
  B.tbox.resize(n);
  B.cen.resize(3 * n);
  B.idx.resize(n);
  float max_abs = 1.0f;
  for (int64_t i = 0; i < 9 * n; ++i) max_abs = std::max(max_abs, std::fabs(in.v0v1v2[i]));
  // conservative margin for the float slab test: well above its rounding
  // error for origins within ~100x the scene extent
  const float pad = max_abs * 1.0e-5f;
  out.pad = pad;
  for (int64_t t = 0; t < n; ++t) {
    Box b;
    b.clear();
    for (int v = 0; v < 3; ++v) b.grow(&in.v0v1v2[9 * t + 3 * v]);
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = std::nextafter(b.lo[k] - pad, -std::numeric_limits<float>::infinity());
      b.hi[k] = std::nextafter(b.hi[k] + pad, std::numeric_limits<float>::infinity());
      B.cen[3 * t + k] = 0.5f * (b.lo[k] + b.hi[k]);
    }
    B.tbox[t] = b;
    B.idx[t] = (uint32_t)t;
  }
  B.nodes.resize(2 * (size_t)n + 1);
  B.next = 1;
  int threads = in.n_threads > 0 ? in.n_threads : (int)std::thread::hardware_concurrency();
  B.spawn_depth = 0;
  while ((1 << B.spawn_depth) < threads && B.spawn_depth < 6) ++B.spawn_depth;
  B.build(0, 0, (uint32_t)n, 0);
  B.nodes.resize(B.next.load());
  out.order = B.idx;
  if (B.nodes[0].count > 0) {
    // a single leaf: wrap it in one 4-wide node
    DevNode d;
    const float inf = std::numeric_limits<float>::infinity();
    for (int i = 0; i < 4; ++i) {
      d.lo_x[i] = d.lo_y[i] = d.lo_z[i] = inf;
      d.hi_x[i] = d.hi_y[i] = d.hi_z[i] = inf;
      d.child[i] = 0;
      d.pad[i] = 0;
    }
    const Box& b = B.nodes[0].box;
    d.lo_x[0] = b.lo[0]; d.lo_y[0] = b.lo[1]; d.lo_z[0] = b.lo[2];
    d.hi_x[0] = b.hi[0]; d.hi_y[0] = b.hi[1]; d.hi_z[0] = b.hi[2];
    d.child[0] = prt_leaf_code(0, (uint32_t)n);
    out.nodes.push_back(d);
    out.depth = 1;
  } else {
    out.nodes.reserve(B.nodes.size() / 2 + 1);
#if PRT_BVH_DP_COLLAPSE
    DpCollapser C(B.nodes, out.nodes);
    C.eval(0);
    C.emit(0, 1);
#else
    Collapser C{B.nodes, out.nodes};
    C.emit(0, 1);
#endif
    out.depth = C.max_depth;
  }
  number_top_levels(out);
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (3 * out.depth + 1 > PRT_STACK_TOTAL) {
    *err = "BVH too deep for the traversal stack";
    return false;
  }
  return true;
}
