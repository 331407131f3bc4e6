// Acceleration-structure builders.  Both produce the same 4-wide node layout
// (BvhNode4) over a primitive permutation; the device traversal is agnostic
// of which builder ran.  Replaces the Embree BVH build (embree.cpp:18-43).
#pragma once

#include <cstdint>
#include <vector>

#include "scene_dev.h"

namespace rad {

struct BvhBuildResult {
  std::vector<BvhNode4> nodes;
  std::vector<uint32_t> order;  // leaf-order position -> input primitive index
  uint32_t depth = 0;
  uint32_t stack_need = 0;  // traversal stack entries the tree can need (kTraceStackEntries)
  float sah_cost = 0;
  double build_ms = 0;
};

// Host binned-SAH build.  `bounds` holds 6 floats per primitive
// (lo.xyz, hi.xyz).  Leaves hold at most `max_leaf` primitives.
void BuildBvh4Host(const float* bounds, uint32_t count, int max_leaf, BvhBuildResult& out);

// SAH cost of a finished tree (traversal 1, intersection 1, relative to root area).
float BvhSahCost(const std::vector<BvhNode4>& nodes);
uint32_t BvhDepth(const std::vector<BvhNode4>& nodes);
// Worst-case traversal stack entries of the tree (see kTraceStackEntries).
uint32_t BvhStackNeed(const std::vector<BvhNode4>& nodes);

}  // namespace rad
