/*
 * lbvh.h - linear BVH (Karras, "Maximizing Parallelism in the Construction of
 * BVHs, Octrees, and k-d Trees", HPG 2012) collapsed breadth-first into the
 * 8-wide compressed nodes of layout.h.  The per-item steps live here and are
 * compiled for the host builder (lbvh_host.cpp) and for the gfx950 kernels
 * (lbvh.hip), so both produce the same Morton codes, the same binary topology
 * and the same wide-node bytes.
 *
 * The reference has no builder of its own: it hands triangles to the Vulkan
 * driver (vkCmdBuildAccelerationStructuresKHR, VulkanAccelerationStructure.cpp:
 * 137-140, 258-259).  This builder and the binned-SAH one (bvh_build.cpp) are
 * this project's replacements for that driver step.
 *
 * Steps:
 *   1. centroid bounds; Morton code of each box centroid, 10 bits per axis
 *      (x in bit 3k+2, y in 3k+1, z in 3k);
 *   2. stable sort of the codes (ties keep primitive order);
 *   3. internal node i of n-1 (Karras): direction, range [first, last], split;
 *      child references carry kLeafBit for a leaf (a sorted position);
 *   4. boxes bottom-up;
 *   5. breadth-first collapse: a wide node opens its largest-area internal
 *      child until it has 8 children; a binary subtree of at most max_leaf
 *      primitives is one leaf child.  Each level's wide nodes get their
 *      internal children as consecutive blocks in (node, slot) order, and
 *      their leaf primitives the same way, by prefix sums over the level.
 */
#pragma once
#include <stdint.h>

#include "wide_node.h"

namespace vvpt {

constexpr uint32_t kLeafBit = 0x80000000u;

/* Spread the low 10 bits of v to every third bit. */
VVPT_HD inline uint32_t morton_spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

/* Quantisation frame of the centroids: lo and 1024 / extent per axis (0 for a flat axis). */
struct MortonFrame {
    float lo[3];
    float scale[3];
};

VVPT_HD inline void centroid_of(const float *box6, float c[3]) {
    VVPT_NO_CONTRACT
    for (int a = 0; a < 3; ++a) c[a] = (box6[a] + box6[3 + a]) * 0.5f;
}

VVPT_HD inline uint32_t morton_of_box(const float *box6, const MortonFrame &f) {
    VVPT_NO_CONTRACT
    float c[3];
    centroid_of(box6, c);
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        float t = (c[a] - f.lo[a]) * f.scale[a];
        t = t > 0.0f ? t : 0.0f;
        q[a] = t >= 1023.0f ? 1023u : (uint32_t)t;
    }
    return (morton_spread10(q[0]) << 2) | (morton_spread10(q[1]) << 1) | morton_spread10(q[2]);
}

/* Length of the common prefix of the keys (code, position) at sorted positions i and j; -1 outside [0, n). */
VVPT_HD inline int lbvh_delta(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = codes[i], b = codes[j];
    if (a != b) return __builtin_clz(a ^ b);
    return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
}

struct LbvhInner {
    uint32_t left, right; /* child references: kLeafBit | position, or an internal node index */
    uint32_t first, last; /* sorted positions covered */
};

/* Internal node i (0 <= i < n - 1) of the Karras hierarchy over n >= 2 sorted codes. */
VVPT_HD inline LbvhInner lbvh_inner(const uint32_t *codes, int n, int i) {
    const int d = lbvh_delta(codes, n, i, i + 1) > lbvh_delta(codes, n, i, i - 1) ? 1 : -1;
    const int dmin = lbvh_delta(codes, n, i, i - d);
    int lmax = 2;
    while (lbvh_delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lbvh_delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = lbvh_delta(codes, n, i, j);
    int s = 0;
    for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) { /* ceil(l/2), ceil(l/4), ..., 1 */
        if (lbvh_delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    LbvhInner r;
    r.first = (uint32_t)(i < j ? i : j);
    r.last = (uint32_t)(i < j ? j : i);
    r.left = (r.first == (uint32_t)gamma) ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
    r.right = (r.last == (uint32_t)gamma + 1) ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    return r;
}

/* Read-only view of a finished binary LBVH for the collapse. */
struct LbvhView {
    const LbvhInner *inner; /* n - 1 internal nodes */
    const HBox *ibox;       /* their boxes */
    const float *boxes;     /* primitive boxes, 6 floats each, in primitive order */
    const uint32_t *order;  /* primitive at each sorted position */
    uint32_t n;
    uint32_t max_leaf;      /* 1..3 */
};

VVPT_HD inline HBox lbvh_ref_box(const LbvhView &v, uint32_t ref) {
    HBox b;
    if (ref & kLeafBit) {
        const float *p = v.boxes + 6 * (size_t)v.order[ref & ~kLeafBit];
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = p[a];
            b.hi[a] = p[3 + a];
        }
        return b;
    }
    return v.ibox[ref];
}

VVPT_HD inline uint32_t lbvh_ref_count(const LbvhView &v, uint32_t ref) {
    if (ref & kLeafBit) return 1u;
    return v.inner[ref].last - v.inner[ref].first + 1u;
}

VVPT_HD inline uint32_t lbvh_ref_first(const LbvhView &v, uint32_t ref) {
    return (ref & kLeafBit) ? (ref & ~kLeafBit) : v.inner[ref].first;
}

/* Children of the wide node made from binary reference `root`: open the internal child of largest surface area
 * (ties: the earlier child) until 8 children or none can open; a subtree of at most max_leaf primitives is a leaf
 * child.  Fills ch / cb / cnt (cnt = 0 for a child that becomes a wide node).  Returns the child count. */
VVPT_HD inline int lbvh_open(const LbvhView &v, uint32_t root, uint32_t ch[8], HBox cb[8], uint32_t cnt[8]) {
    int n = 0;
    const uint32_t rc = lbvh_ref_count(v, root);
    if ((root & kLeafBit) || rc <= v.max_leaf) {
        ch[0] = root;
        n = 1;
    } else {
        ch[0] = v.inner[root].left;
        ch[1] = v.inner[root].right;
        n = 2;
    }
    for (int i = 0; i < n; ++i) cb[i] = lbvh_ref_box(v, ch[i]);
    while (n < 8) {
        int pick = -1;
        float best = -1.0f;
        for (int i = 0; i < n; ++i) {
            if ((ch[i] & kLeafBit) || lbvh_ref_count(v, ch[i]) <= v.max_leaf) continue;
            const float a = hbox_area(cb[i]);
            if (a > best) {
                best = a;
                pick = i;
            }
        }
        if (pick < 0) break;
        const uint32_t c = ch[pick];
        ch[pick] = v.inner[c].left;
        ch[n] = v.inner[c].right;
        cb[pick] = lbvh_ref_box(v, ch[pick]);
        cb[n] = lbvh_ref_box(v, ch[n]);
        ++n;
    }
    for (int i = 0; i < n; ++i) {
        const uint32_t c = lbvh_ref_count(v, ch[i]);
        cnt[i] = ((ch[i] & kLeafBit) || c <= v.max_leaf) ? c : 0u;
    }
    return n;
}

/* Internal-child and leaf-primitive counts of one wide node of a level. */
VVPT_HD inline void lbvh_collapse_count(const LbvhView &v, uint32_t root, uint32_t &n_internal, uint32_t &n_prims) {
    uint32_t ch[8], cnt[8];
    HBox cb[8];
    const int n = lbvh_open(v, root, ch, cb, cnt);
    n_internal = 0;
    n_prims = 0;
    for (int i = 0; i < n; ++i) {
        n_internal += cnt[i] == 0;
        n_prims += cnt[i];
    }
}

/* Emit one wide node of a level: the node, its leaf primitive references (from leaf_base) and the binary
 * references of its internal children (next[0 .. n_internal), in slot order). */
VVPT_HD inline WideNode lbvh_collapse_emit(const LbvhView &v, uint32_t root, uint32_t child_base, uint32_t leaf_base,
                                           uint32_t *leaf_refs, uint32_t *next) {
    uint32_t ch[8], cnt[8], off[8];
    HBox cb[8], nb = hbox_empty();
    const int n = lbvh_open(v, root, ch, cb, cnt);
    for (int i = 0; i < n; ++i) hbox_grow(nb, cb[i]);
    int child_at[8];
    wide_assign_slots(cb, n, nb, child_at);
    const WideNode w = wide_encode(cb, cnt, n, child_at, child_base, leaf_base, off);
    uint32_t rank = 0;
    for (int s = 0; s < 8; ++s) {
        const int i = child_at[s];
        if (i < 0) continue;
        if (cnt[i] == 0) {
            next[rank++] = ch[i];
        } else {
            const uint32_t f = lbvh_ref_first(v, ch[i]);
            for (uint32_t k = 0; k < cnt[i]; ++k) leaf_refs[leaf_base + off[i] + k] = v.order[f + k];
        }
    }
    return w;
}

}  // namespace vvpt
