// Device BVH builder: Morton-ordered primitives, level-synchronous binned
// SAH splits, collapse to the 4-wide node layout of the traversal.
#pragma once

#include <cstdint>

#include "scene_host.h"

namespace rad {

struct DeviceBvhStats {
  uint32_t node_count = 0;  // 4-wide nodes
  uint32_t depth = 0;       // 4-wide levels
  uint32_t stack_need = 0;  // traversal stack entries the tree can need (kTraceStackEntries)
  uint32_t levels = 0;      // binary build levels
  float sah_cost = 0;
  double build_ms = 0;
};

// bounds: 6 floats per primitive (lo.xyz, hi.xyz); records: 3 float4 per
// primitive (host memory).  Writes the nodes and the primitive records in
// leaf order to device buffers.
void BuildBvh4Device(const float* bounds, const float4* records, uint32_t n, int maxLeaf, hipStream_t st,
                     DeviceBuffer& nodesOut, DeviceBuffer& primsOut, DeviceBvhStats& stats);

// The traversal's 64-byte copy of m nodes (device buffers); throws if a
// child box cannot be bounded.
void QuantizeBvh4(const DeviceBuffer& nodes, uint32_t m, hipStream_t st, DeviceBuffer& out);

}  // namespace rad
