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
 * consecutive block when the node is emitted, then emitted in slot order.
 *
 * Which binary nodes become wide nodes, leaves or neither is the SAH-optimal choice of Ylitie, Karras and Laine
 * ("Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017, section 3.1), by dynamic
 * programming over the binary tree: C(n, i) is the least cost of the subtree of n spread over at most i child slots
 * of its wide parent,
 *   C(n, 1) = min(leaf: A(n) N(n) c_tri if N(n) <= 3,  wide node: A(n) c_node + D(n, 8))
 *   C(n, i) = min(C(n, i - 1), D(n, i)),   D(n, j) = min over k of C(left, k) + C(right, j - k).
 * c_tri / c_node = 0.5: the 1080p icosphere-field frame measured 1.3 % faster than with 0.26 (the ratio of a
 * triangle test's instructions to a wide node test's), and 0.15 and 1.0 were slower; a triangle test costs more than
 * its instructions because its 48-byte record is a dependent load. */
struct Collapser {
    const Builder &B;
    std::vector<WideNode> &out;
    std::vector<uint32_t> &leaf_refs;
    uint32_t max_depth = 0;
    int max_leaf;
    static constexpr int kSlots = 8;
    static constexpr float kCostNode = 1.0f, kCostTri = 0.5f;
    static constexpr float kInf = 3.0e38f;
    struct Plan {
        float c[kSlots + 1];     /* C(n, i), i = 1..8 */
        int8_t split[kSlots + 1]; /* i >= 2: slots of the left child in D(n, i), or 0 when C(n, i) = C(n, i-1) */
        int8_t dsplit8;           /* slots of the left child in D(n, 8) (the node as a wide node) */
        bool leaf;                /* C(n, 1) is the leaf */
        uint32_t first, count;    /* the subtree's primitives: order[first, first + count) */
    };
    std::vector<Plan> plan;

    Collapser(const Builder &b, std::vector<WideNode> &o, std::vector<uint32_t> &l, int ml)
        : B(b), out(o), leaf_refs(l), max_leaf(ml) {}

    /* D(n, j) and its best split, for an internal binary node */
    float dist(const BNode &bn, int j, int &best_k) const {
        const Plan &L = plan[bn.left], &R = plan[bn.right];
        float best = kInf;
        best_k = 0;
        for (int k = 1; k < j; ++k) {
            const float v = L.c[k] + R.c[j - k];
            if (v < best) {
                best = v;
                best_k = k;
            }
        }
        return best;
    }

    void make_plan() {
        const size_t nn = B.nodes.size();
        plan.assign(nn, Plan());
        /* children are appended after their parent, so reverse index order visits children first */
        for (size_t q = nn; q-- > 0;) {
            const BNode &bn = B.nodes[q];
            Plan &P = plan[q];
            const float area = hbox_area(to_hbox(bn.box));
            if (bn.leaf()) {
                P.first = bn.first;
                P.count = bn.count;
                P.leaf = true;
                P.dsplit8 = 0;
                const float cl = area * (float)bn.count * kCostTri;
                for (int i = 1; i <= kSlots; ++i) {
                    P.c[i] = cl;
                    P.split[i] = 0;
                }
                continue;
            }
            const Plan &L = plan[bn.left], &R = plan[bn.right];
            P.first = std::min(L.first, R.first);
            P.count = L.count + R.count;
            int k8 = 0;
            const float cw = area * kCostNode + dist(bn, kSlots, k8);
            const float cl = (int)P.count <= max_leaf ? area * (float)P.count * kCostTri : kInf;
            P.dsplit8 = (int8_t)k8;
            P.leaf = cl <= cw;
            P.c[1] = P.leaf ? cl : cw;
            P.split[1] = 0;
            for (int i = 2; i <= kSlots; ++i) {
                int k = 0;
                const float d = dist(bn, i, k);
                if (d < P.c[i - 1]) {
                    P.c[i] = d;
                    P.split[i] = (int8_t)k;
                } else {
                    P.c[i] = P.c[i - 1];
                    P.split[i] = 0;
                }
            }
        }
    }

    /* the binary nodes that stand for subtree n in at most i slots */
    void expand(uint32_t n, int i, uint32_t *ch, int &cnt) const {
        const BNode &bn = B.nodes[n];
        while (i > 1 && plan[n].split[i] == 0) --i;
        if (i <= 1 || bn.leaf()) {
            ch[cnt++] = n;
            return;
        }
        const int k = plan[n].split[i];
        expand(bn.left, k, ch, cnt);
        expand(bn.right, i - k, ch, cnt);
    }

    /* Fill wide node `wi` from binary node `bi`; depth is the level of `wi`. */
    void emit(uint32_t wi, uint32_t bi, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        const BNode &bn = B.nodes[bi];
        uint32_t ch[kSlots];
        int n = 0;
        if (bn.leaf()) {
            ch[n++] = bi;
        } else {
            const int k = plan[bi].dsplit8;
            expand(bn.left, k, ch, n);
            expand(bn.right, kSlots - k, ch, n);
        }
        HBox cb[kSlots], nb = hbox_empty();
        uint32_t cnt[kSlots], off[kSlots];
        uint32_t n_internal = 0;
        for (int i = 0; i < n; ++i) {
            cb[i] = to_hbox(B.nodes[ch[i]].box);
            hbox_grow(nb, cb[i]);
            cnt[i] = plan[ch[i]].leaf ? plan[ch[i]].count : 0u; /* 0: a wide child node */
            n_internal += cnt[i] == 0;
        }
        int child_at[kSlots];
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
        for (int s = 0; s < kSlots; ++s) {
            const int i = child_at[s];
            if (i < 0) continue;
            if (cnt[i] == 0) {
                internal.push_back({child_base + rank++, ch[i]});
            } else {
                const Plan &c = plan[ch[i]];
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
    Collapser c(b, nodes, leaf_refs, max_leaf);
    c.make_plan();
    r.root = (uint32_t)nodes.size();
    nodes.emplace_back();
    c.emit(r.root, 0, 1);
    r.depth = c.max_depth;
    return r;
}

}  // namespace vvpt
