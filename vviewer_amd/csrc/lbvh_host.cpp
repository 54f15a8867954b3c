/*
 * lbvh_host.cpp - the LBVH build on the host, step for step the same as the
 * gfx950 build in lbvh.hip (same per-item functions from lbvh.h, same
 * breadth-first allocation), so that the two produce identical bytes.
 */
#include <algorithm>
#include <chrono>
#include <cstring>

#include "lbvh.h"
#include "lbvh_build.h"

namespace vvpt {

namespace {

MortonFrame frame_from_bounds(const float lo[3], const float hi[3]) {
    MortonFrame f;
    for (int a = 0; a < 3; ++a) {
        f.lo[a] = lo[a];
        const float extent = hi[a] - lo[a];
        f.scale[a] = extent > 0.0f ? 1024.0f / extent : 0.0f;
    }
    return f;
}

WideNode empty_node() {
    WideNode w;
    std::memset(&w, 0, sizeof(w));
    w.e_imask = 127u | (127u << 8) | (127u << 16);
    return w;
}

}  // namespace

MortonFrame lbvh_frame_host(const float *boxes, uint32_t n) {
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = 0; i < n; ++i) {
        float c[3];
        centroid_of(boxes + 6 * (size_t)i, c);
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], c[a]);
            hi[a] = std::max(hi[a], c[a]);
        }
    }
    return frame_from_bounds(lo, hi);
}

MortonFrame lbvh_frame_from_bounds(const float lo[3], const float hi[3]) { return frame_from_bounds(lo, hi); }

void lbvh_build_host(const float *boxes, uint32_t n, uint32_t max_leaf, bool keep_debug, LbvhOutput &out) {
    const auto t0 = std::chrono::steady_clock::now();
    out = LbvhOutput();
    max_leaf = std::max(1u, std::min(3u, max_leaf));
    if (n == 0) {
        out.nodes.push_back(empty_node());
        out.depth = 1;
        return;
    }
    /* 1. Morton codes */
    const MortonFrame f = lbvh_frame_host(boxes, n);
    std::vector<uint32_t> code(n), order(n);
    for (uint32_t i = 0; i < n; ++i) {
        code[i] = morton_of_box(boxes + 6 * (size_t)i, f);
        order[i] = i;
    }
    /* 2. stable sort by code */
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    std::vector<uint32_t> sorted(n);
    for (uint32_t i = 0; i < n; ++i) sorted[i] = code[order[i]];
    /* 3. hierarchy; 4. boxes in post-order (an explicit stack: the tree can be deep) */
    std::vector<LbvhInner> inner(n > 1 ? n - 1 : 0);
    for (uint32_t i = 0; i + 1 < n; ++i) inner[i] = lbvh_inner(sorted.data(), (int)n, (int)i);
    std::vector<HBox> ibox(inner.size());
    LbvhView v{inner.data(), ibox.data(), boxes, order.data(), n, max_leaf};
    if (n > 1) {
        std::vector<std::pair<uint32_t, bool>> st = {{0u, false}};
        while (!st.empty()) {
            auto [k, ready] = st.back();
            st.pop_back();
            if (ready) {
                HBox b = lbvh_ref_box(v, inner[k].left);
                hbox_grow(b, lbvh_ref_box(v, inner[k].right));
                ibox[k] = b;
                continue;
            }
            st.push_back({k, true});
            if (!(inner[k].left & kLeafBit)) st.push_back({inner[k].left, false});
            if (!(inner[k].right & kLeafBit)) st.push_back({inner[k].right, false});
        }
    }
    /* 5. breadth-first collapse */
    std::vector<uint32_t> level = {n > 1 ? 0u : kLeafBit};
    uint32_t level_base = 0, leaf_cursor = 0;
    out.nodes.resize(1);
    while (!level.empty()) {
        const uint32_t m = (uint32_t)level.size();
        std::vector<uint32_t> ni(m), np(m);
        uint32_t tot_i = 0, tot_p = 0;
        for (uint32_t k = 0; k < m; ++k) {
            uint32_t a, b;
            lbvh_collapse_count(v, level[k], a, b);
            ni[k] = tot_i;
            np[k] = tot_p;
            tot_i += a;
            tot_p += b;
        }
        const uint32_t next_base = level_base + m;
        out.nodes.resize(next_base + tot_i);
        out.leaf_refs.resize(leaf_cursor + tot_p);
        std::vector<uint32_t> next(tot_i);
        for (uint32_t k = 0; k < m; ++k)
            out.nodes[level_base + k] = lbvh_collapse_emit(v, level[k], next_base + ni[k], leaf_cursor + np[k],
                                                           out.leaf_refs.data(), next.data() + ni[k]);
        level.swap(next);
        level_base = next_base;
        leaf_cursor += tot_p;
        ++out.depth;
    }
    if (keep_debug) {
        out.codes = sorted;
        out.order = order;
        out.inner.resize(2 * inner.size());
        for (size_t k = 0; k < inner.size(); ++k) {
            out.inner[2 * k] = inner[k].left;
            out.inner[2 * k + 1] = inner[k].right;
        }
    }
    out.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace vvpt
