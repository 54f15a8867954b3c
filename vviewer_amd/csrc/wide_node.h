/*
 * wide_node.h - per-node steps of the wide-BVH collapse, compiled for the host
 * builders and for the gfx950 LBVH kernels alike, so that both produce the
 * same bytes: greedy opening of binary children by surface area, octant-ordered
 * slot assignment and conservative 8-bit quantisation of the child boxes into a
 * compressed 80-byte node (layout.h).  Floating-point contraction is off in
 * every function: a fused multiply-add on one side only would change a
 * comparison and with it the tree.
 */
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "layout.h"

#if defined(__HIPCC__)
#define VVPT_HD __host__ __device__
#else
#define VVPT_HD
#endif

#if defined(__clang__)
#define VVPT_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define VVPT_NO_CONTRACT
#endif

namespace vvpt {

struct HBox {
    float lo[3], hi[3];
};

VVPT_HD inline HBox hbox_empty() {
    HBox b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = 1e30f;
        b.hi[a] = -1e30f;
    }
    return b;
}

VVPT_HD inline void hbox_grow(HBox &b, const HBox &o) {
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = o.lo[a] < b.lo[a] ? o.lo[a] : b.lo[a];
        b.hi[a] = o.hi[a] > b.hi[a] ? o.hi[a] : b.hi[a];
    }
}

/* Surface area, 2 (dx dy + dy dz + dz dx), rounded in that order. */
VVPT_HD inline float hbox_area(const HBox &b) {
    VVPT_NO_CONTRACT
    float d[3];
    for (int a = 0; a < 3; ++a) d[a] = b.hi[a] > b.lo[a] ? b.hi[a] - b.lo[a] : 0.0f;
    const float s = d[0] * d[1] + d[1] * d[2];
    return 2.0f * (s + d[2] * d[0]);
}

/* Smallest exponent e with p + 255 * 2^e >= hi (checked in float). */
VVPT_HD inline int wide_pick_exponent(float p, float hi) {
    VVPT_NO_CONTRACT
    const float extent = hi - p;
    int e = -126;
    if (extent > 0.0f) {
        int ex;
        frexpf(extent / 255.0f, &ex);
        e = ex - 1 > -126 ? ex - 1 : -126;
    }
    for (;;) {
        const float s = ldexpf(1.0f, e);
        if (p + 255.0f * s >= hi || e >= 127) break;
        ++e;
    }
    return e;
}

/* Octant-ordered slot assignment.  cost[i][s] = offset of child i's centre from the parent centre projected on the
 * octant direction s (bit0: -x, bit1: -y, bit2: -z); pairs are taken greedily by the smallest cost, ties to the lower
 * child then the lower slot.  Fills child_at[s] (child index or -1). */
VVPT_HD inline void wide_assign_slots(const HBox *cb, int n, const HBox &nb, int child_at[8]) {
    VVPT_NO_CONTRACT
    float pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (nb.lo[a] + nb.hi[a]);
    float cost[8][8];
    for (int i = 0; i < n; ++i) {
        float d[3];
        for (int a = 0; a < 3; ++a) d[a] = 0.5f * (cb[i].lo[a] + cb[i].hi[a]) - pc[a];
        for (int s = 0; s < 8; ++s) {
            const float dx = (s & 1) ? -d[0] : d[0], dy = (s & 2) ? -d[1] : d[1], dz = (s & 4) ? -d[2] : d[2];
            cost[i][s] = (dx + dy) + dz;
        }
    }
    bool child_done[8] = {false, false, false, false, false, false, false, false};
    bool slot_used[8] = {false, false, false, false, false, false, false, false};
    for (int s = 0; s < 8; ++s) child_at[s] = -1;
    for (int k = 0; k < n; ++k) {
        float best = 1e30f;
        int bi = -1, bs = -1;
        for (int i = 0; i < n; ++i) {
            if (child_done[i]) continue;
            for (int s = 0; s < 8; ++s) {
                if (slot_used[s]) continue;
                if (cost[i][s] < best || bi < 0) {
                    best = cost[i][s];
                    bi = i;
                    bs = s;
                }
            }
        }
        child_done[bi] = true;
        slot_used[bs] = true;
        child_at[bs] = bi;
    }
}

/* Encode one wide node.  Children i = 0..n-1 have boxes cb[i] and leaf_count[i] (0 = internal child).  Slots come
 * from wide_assign_slots.  Internal children get consecutive node indices from child_base in slot order; leaf
 * children get consecutive primitive offsets from leaf_base in slot order (leaf_off[i] receives child i's offset,
 * relative to leaf_base).  Returns the node. */
VVPT_HD inline WideNode wide_encode(const HBox *cb, const uint32_t *leaf_count, int n, const int child_at[8],
                                    uint32_t child_base, uint32_t leaf_base, uint32_t *leaf_off) {
    VVPT_NO_CONTRACT
    HBox nb = hbox_empty();
    for (int i = 0; i < n; ++i) hbox_grow(nb, cb[i]);
    WideNode w;
    memset(&w, 0, sizeof(w));
    const float p[3] = {nb.lo[0], nb.lo[1], nb.lo[2]};
    int e[3];
    float sc[3];
    for (int a = 0; a < 3; ++a) {
        e[a] = wide_pick_exponent(p[a], nb.hi[a]);
        sc[a] = ldexpf(1.0f, e[a]);
    }
    w.px = p[0];
    w.py = p[1];
    w.pz = p[2];
    uint32_t imask = 0;
    for (int s = 0; s < 8; ++s)
        if (child_at[s] >= 0 && leaf_count[child_at[s]] == 0) imask |= 1u << s;
    w.e_imask = (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) | ((uint32_t)(e[2] + 127) << 16) | (imask << 24);
    w.child_base = child_base;
    w.leaf_base = leaf_base;
    uint8_t meta[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint8_t qlo[3][8], qhi[3][8];
    memset(qlo, 0, sizeof(qlo));
    memset(qhi, 0, sizeof(qhi));
    uint32_t off = 0;
    for (int s = 0; s < 8; ++s) {
        const int i = child_at[s];
        if (i < 0) continue;
        for (int a = 0; a < 3; ++a) {
            float lo = floorf((cb[i].lo[a] - p[a]) / sc[a]);
            float hi = ceilf((cb[i].hi[a] - p[a]) / sc[a]);
            lo = fminf(fmaxf(lo, 0.0f), 255.0f);
            hi = fminf(fmaxf(hi, 0.0f), 255.0f);
            while (lo > 0.0f && p[a] + lo * sc[a] > cb[i].lo[a]) lo -= 1.0f;
            while (hi < 255.0f && p[a] + hi * sc[a] < cb[i].hi[a]) hi += 1.0f;
            qlo[a][s] = (uint8_t)lo;
            qhi[a][s] = (uint8_t)hi;
        }
        const uint32_t cnt = leaf_count[i];
        if (cnt == 0) {
            meta[s] = (uint8_t)(0x20 | (24 + s));
        } else {
            const uint32_t unary = cnt == 1 ? 1u : (cnt == 2 ? 3u : 7u);
            meta[s] = (uint8_t)((unary << 5) | off);
            if (leaf_off) leaf_off[i] = off;
            off += cnt;
        }
    }
    const auto pack = [](const uint8_t *b, uint32_t *o) {
        o[0] = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
        o[1] = (uint32_t)b[4] | ((uint32_t)b[5] << 8) | ((uint32_t)b[6] << 16) | ((uint32_t)b[7] << 24);
    };
    pack(meta, w.meta);
    pack(qlo[0], w.qlo_x);
    pack(qlo[1], w.qlo_y);
    pack(qlo[2], w.qlo_z);
    pack(qhi[0], w.qhi_x);
    pack(qhi[1], w.qhi_y);
    pack(qhi[2], w.qhi_z);
    return w;
}

}  // namespace vvpt
