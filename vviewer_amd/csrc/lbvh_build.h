/*
 * lbvh_build.h - entry points of the LBVH builder (lbvh.h): the host build
 * (lbvh_host.cpp, the bit-exact reference for the device build) and the gfx950
 * build (lbvh.hip).
 */
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "lbvh.h"

namespace vvpt {

struct LbvhOutput {
    /* always filled */
    std::vector<WideNode> nodes;     /* node 0 is the root; child_base / leaf_base index these arrays */
    std::vector<uint32_t> leaf_refs; /* primitive indices in leaf order */
    uint32_t depth = 0;              /* wide-node levels */
    double ms = 0.0;                 /* build time (device: events around the kernels) */
    /* filled when requested */
    std::vector<uint32_t> codes;     /* Morton codes in sorted order */
    std::vector<uint32_t> order;     /* primitive at each sorted position */
    std::vector<uint32_t> inner;     /* 2 (n - 1): left, right child reference of each internal node */
};

/* Morton frame from the centroid bounds (1024 / extent per axis, 0 for a flat axis). */
MortonFrame lbvh_frame_from_bounds(const float lo[3], const float hi[3]);

/* boxes: n boxes of 6 floats (lo xyz, hi xyz).  max_leaf 1..3. */
void lbvh_build_host(const float *boxes, uint32_t n, uint32_t max_leaf, bool keep_debug, LbvhOutput &out);

/* Device build on the current device and `stream` (a hipStream_t); boxes are host memory.  Returns an empty string
 * on success, else the failing call. */
std::string lbvh_build_device(const float *boxes, uint32_t n, uint32_t max_leaf, bool keep_debug, void *stream,
                              LbvhOutput &out);

}  // namespace vvpt
