/*
 * bvh_build.cpp - binned-SAH binary build, collapse to 8-wide nodes,
 * octant-ordered child slots and conservative 8-bit box quantisation.
 */
#include "bvh_build.h"
#include "wide_node.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace vvpt {
namespace {

struct BNode {
    Box3 box;
    uint32_t left = 0, right = 0; /* internal: child node indices */
    uint32_t first = 0, count = 0; /* leaf: range in the primitive order */
    bool leaf() const { return count > 0; }
};

struct Builder {
    const std::vector<Box3> &boxes;
    int max_leaf;
    std::vector<float> cx, cy, cz;
    std::vector<uint32_t> order;
    std::vector<BNode> nodes;

    Builder(const std::vector<Box3> &b, int ml) : boxes(b), max_leaf(ml) {}

    float cen(uint32_t p, int a) const { return a == 0 ? cx[p] : (a == 1 ? cy[p] : cz[p]); }

    void build() {
        uint32_t n = (uint32_t)boxes.size();
        cx.resize(n);
        cy.resize(n);
        cz.resize(n);
        order.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            cx[i] = 0.5f * (boxes[i].lo[0] + boxes[i].hi[0]);
            cy[i] = 0.5f * (boxes[i].lo[1] + boxes[i].hi[1]);
            cz[i] = 0.5f * (boxes[i].lo[2] + boxes[i].hi[2]);
            order[i] = i;
        }
        nodes.reserve(2 * (size_t)n + 1);
        nodes.emplace_back();
        struct Task { uint32_t node, first, count; };
        std::vector<Task> stack = {{0, 0, n}};
        constexpr int kBins = 32;
        while (!stack.empty()) {
            Task t = stack.back();
            stack.pop_back();
            Box3 box, cbox;
            for (uint32_t i = t.first; i < t.first + t.count; ++i) {
                uint32_t p = order[i];
                box.grow(boxes[p]);
                float c[3] = {cx[p], cy[p], cz[p]};
                cbox.grow(c);
            }
            nodes[t.node].box = box;
            if ((int)t.count <= 1) {
                nodes[t.node].first = t.first;
                nodes[t.node].count = t.count;
                continue;
            }
            /* binned SAH over the three axes */
            float best_cost = 1e30f;
            int best_axis = -1, best_split = 0;
            for (int a = 0; a < 3; ++a) {
                float lo = cbox.lo[a], hi = cbox.hi[a];
                if (!(hi > lo)) continue;
                float scale = kBins / (hi - lo);
                Box3 bb[kBins];
                uint32_t bc[kBins] = {0};
                for (uint32_t i = t.first; i < t.first + t.count; ++i) {
                    uint32_t p = order[i];
                    int b = std::min(kBins - 1, (int)((cen(p, a) - lo) * scale));
                    bc[b]++;
                    bb[b].grow(boxes[p]);
                }
                float rarea[kBins];
                uint32_t rcount[kBins];
                Box3 acc;
                uint32_t cnt = 0;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.grow(bb[b]);
                    cnt += bc[b];
                    rarea[b] = acc.valid() ? acc.area() : 0.0f;
                    rcount[b] = cnt;
                }
                Box3 lacc;
                uint32_t lcnt = 0;
                for (int b = 0; b < kBins - 1; ++b) {
                    lacc.grow(bb[b]);
                    lcnt += bc[b];
                    if (lcnt == 0 || rcount[b + 1] == 0) continue;
                    float c = lacc.area() * lcnt + rarea[b + 1] * rcount[b + 1];
                    if (c < best_cost) {
                        best_cost = c;
                        best_axis = a;
                        best_split = b;
                    }
                }
            }
            float leaf_cost = box.area() * t.count;
            bool can_leaf = (int)t.count <= max_leaf;
            uint32_t mid;
            if (best_axis < 0) {
                if (can_leaf) {
                    nodes[t.node].first = t.first;
                    nodes[t.node].count = t.count;
                    continue;
                }
                mid = t.first + t.count / 2; /* coincident centroids: split by index */
            } else {
                /* SAH with unit traversal cost relative to one primitive test */
                if (can_leaf && leaf_cost <= box.area() * 1.0f + best_cost) {
                    nodes[t.node].first = t.first;
                    nodes[t.node].count = t.count;
                    continue;
                }
                float lo = cbox.lo[best_axis];
                float scale = kBins / (cbox.hi[best_axis] - lo);
                auto it = std::partition(order.begin() + t.first, order.begin() + t.first + t.count, [&](uint32_t p) {
                    return std::min(kBins - 1, (int)((cen(p, best_axis) - lo) * scale)) <= best_split;
                });
                mid = (uint32_t)(it - order.begin());
                if (mid == t.first || mid == t.first + t.count) mid = t.first + t.count / 2;
            }
            uint32_t l = (uint32_t)nodes.size();
            nodes.emplace_back();
            nodes.emplace_back();
            nodes[t.node].left = l;
            nodes[t.node].right = l + 1;
            nodes[t.node].count = 0;
            stack.push_back({l, t.first, mid - t.first});
            stack.push_back({l + 1, mid, t.first + t.count - mid});
        }
    }
};

inline HBox to_hbox(const Box3 &b) {
    HBox h;
    for (int a = 0; a < 3; ++a) {
        h.lo[a] = b.lo[a];
        h.hi[a] = b.hi[a];
    }
    return h;
}

/* Depth-first collapse of the binary SAH tree into wide nodes: a node's internal children are allocated as one
 * consecutive block when the node is emitted, then emitted in slot order. */
struct Collapser {
    const Builder &B;
    std::vector<WideNode> &out;
    std::vector<uint32_t> &leaf_refs;
    uint32_t max_depth = 0;

    Collapser(const Builder &b, std::vector<WideNode> &o, std::vector<uint32_t> &l) : B(b), out(o), leaf_refs(l) {}

    /* Fill wide node `wi` from binary node `bi`; depth is the level of `wi`. */
    void emit(uint32_t wi, uint32_t bi, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        const BNode &bn = B.nodes[bi];
        uint32_t ch[8];
        int n = 0;
        if (bn.leaf()) {
            ch[n++] = bi;
        } else {
            ch[n++] = bn.left;
            ch[n++] = bn.right;
        }
        while (n < 8) { /* open the internal child of largest surface area */
            int pick = -1;
            float best = -1.0f;
            for (int i = 0; i < n; ++i) {
                const BNode &c = B.nodes[ch[i]];
                if (c.leaf()) continue;
                const float a = hbox_area(to_hbox(c.box));
                if (a > best) {
                    best = a;
                    pick = i;
                }
            }
            if (pick < 0) break;
            const uint32_t c = ch[pick];
            ch[pick] = B.nodes[c].left;
            ch[n++] = B.nodes[c].right;
        }
        HBox cb[8], nb = hbox_empty();
        uint32_t cnt[8], off[8];
        uint32_t n_internal = 0;
        for (int i = 0; i < n; ++i) {
            cb[i] = to_hbox(B.nodes[ch[i]].box);
            hbox_grow(nb, cb[i]);
            cnt[i] = B.nodes[ch[i]].count;
            n_internal += cnt[i] == 0;
        }
        int child_at[8];
        wide_assign_slots(cb, n, nb, child_at);
        const uint32_t child_base = (uint32_t)out.size();
        out.resize(out.size() + n_internal);
        const uint32_t leaf_base = (uint32_t)leaf_refs.size();
        out[wi] = wide_encode(cb, cnt, n, child_at, child_base, leaf_base, off);
        uint32_t nleaf = 0;
        for (int i = 0; i < n; ++i) nleaf += cnt[i];
        leaf_refs.resize(leaf_base + nleaf);
        uint32_t rank = 0;
        std::vector<std::pair<uint32_t, uint32_t>> internal; /* (wide index, binary index) */
        for (int s = 0; s < 8; ++s) {
            const int i = child_at[s];
            if (i < 0) continue;
            if (cnt[i] == 0) {
                internal.push_back({child_base + rank++, ch[i]});
            } else {
                const BNode &c = B.nodes[ch[i]];
                for (uint32_t k = 0; k < c.count; ++k) leaf_refs[leaf_base + off[i] + k] = B.order[c.first + k];
            }
        }
        for (auto &ib : internal) emit(ib.first, ib.second, depth + 1);
    }
};

}  // namespace

WideBuildResult build_wide_bvh(const std::vector<Box3> &boxes, int max_leaf, std::vector<WideNode> &nodes,
                               std::vector<uint32_t> &leaf_refs) {
    WideBuildResult r;
    max_leaf = std::max(1, std::min(3, max_leaf));
    Builder b(boxes, max_leaf);
    if (boxes.empty()) {
        /* an empty node: no children, every ray misses */
        WideNode n;
        std::memset(&n, 0, sizeof(n));
        n.e_imask = 127u | (127u << 8) | (127u << 16);
        r.root = (uint32_t)nodes.size();
        nodes.push_back(n);
        r.depth = 1;
        return r;
    }
    b.build();
    /* binary leaves must fit one wide-node leaf child */
    Collapser c(b, nodes, leaf_refs);
    r.root = (uint32_t)nodes.size();
    nodes.emplace_back();
    c.emit(r.root, 0, 1);
    r.depth = c.max_depth;
    return r;
}

}  // namespace vvpt
