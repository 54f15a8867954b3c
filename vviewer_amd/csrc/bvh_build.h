/*
 * bvh_build.h - host builder for the two-level wide BVH.
 *
 * The reference delegates acceleration structures to the Vulkan driver
 * (vkCmdBuildAccelerationStructuresKHR, VulkanAccelerationStructure.cpp:137-140,
 * 258-259): one BLAS per unique mesh in object space and one TLAS over the
 * instances (VulkanScene.cpp:306-381).  This builder keeps that two-level shape:
 * a binned-SAH binary tree per mesh and over the instances, collapsed into
 * 8-wide nodes with quantised child boxes (layout.h).
 */
#pragma once
#include <cstdint>
#include <vector>

#include "layout.h"

namespace vvpt {

struct Box3 {
    float lo[3] = {1e30f, 1e30f, 1e30f};
    float hi[3] = {-1e30f, -1e30f, -1e30f};
    void grow(const float *p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = p[a] < lo[a] ? p[a] : lo[a];
            hi[a] = p[a] > hi[a] ? p[a] : hi[a];
        }
    }
    void grow(const Box3 &b) {
        if (!b.valid()) return; /* an empty box's corners (+1e30, -1e30) would span everything */
        grow(b.lo);
        grow(b.hi);
    }
    float area() const {
        float d[3];
        for (int a = 0; a < 3; ++a) d[a] = hi[a] > lo[a] ? hi[a] - lo[a] : 0.0f;
        return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
    bool valid() const { return lo[0] <= hi[0]; }
};

struct WideBuildResult {
    uint32_t root = 0;   /* index of the root node in the caller's node array */
    uint32_t depth = 0;  /* number of wide-node levels */
};

/* Build a wide BVH over `boxes` and append it to `nodes`.  Leaf references are
 * appended to `leaf_refs` (primitive indices in leaf order); leaf_base values in
 * the emitted nodes are absolute indices into `leaf_refs`.  max_leaf (1..3) is
 * the largest number of primitives one leaf child may hold. */
WideBuildResult build_wide_bvh(const std::vector<Box3> &boxes, int max_leaf, std::vector<WideNode> &nodes,
                               std::vector<uint32_t> &leaf_refs);

}  // namespace vvpt
