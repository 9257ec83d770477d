// bvh.h — host builder of the device BVH: binned-SAH binary tree collapsed
// into 4-wide 128-byte DevNode records.
//
// Replaces the reference's per-mesh QBVH::Construct (QBVH.cpp:161-290) over a
// full-sort SAH BVH (BVH.cpp:194-348) with one tree over the triangles of all
// POLYGON entities.  The closest hit it yields is the reference's: every
// triangle sits in exactly one leaf whose box contains it with a conservative
// margin, so the closest triangle with t > 1e-6 over all leaves the ray
// enters is the closest over all triangles.
#pragma once
#include <cstdint>
#include <vector>

#include "prt_device.h"

struct BvhInput {
  const float* v0v1v2;  // n_tris * 9 floats (v0.xyz v1.xyz v2.xyz)
  int64_t n_tris;
  int n_threads;        // 0 = hardware concurrency
};

struct BvhOutput {
  std::vector<DevNode> nodes;
  std::vector<uint32_t> order;  // leaf-order slot -> input triangle index
  int depth = 0;                // levels of 4-wide nodes
  int n_top = 0;                // nodes 0..n_top-1: the top PRT_TOP_LEVELS levels, breadth first
  float pad = 0.0f;             // box margin used
  double build_ms = 0.0;
};

// Returns false (with a message) when the tree would exceed the traversal
// stack bound (4-wide depth > 63).
bool prt_build_bvh(const BvhInput& in, BvhOutput& out, const char** err);
