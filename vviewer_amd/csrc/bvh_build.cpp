/*
 * bvh_build.cpp - binned-SAH binary build, collapse to 8-wide nodes,
 * octant-ordered child slots and conservative 8-bit box quantisation.
 */
#include "bvh_build.h"

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

/* Smallest e with 255 * 2^e covering `extent` from p (checked in float). */
int pick_exponent(float p, float hi) {
    float extent = hi - p;
    int e = -126;
    if (extent > 0.0f) {
        int ex;
        std::frexp(extent / 255.0f, &ex);
        e = std::max(-126, ex - 1);
    }
    for (;;) {
        float s = std::ldexp(1.0f, e);
        if (p + 255.0f * s >= hi || e >= 127) break;
        ++e;
    }
    return e;
}

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
        std::vector<uint32_t> ch;
        if (bn.leaf()) {
            ch.push_back(bi);
        } else {
            ch.push_back(bn.left);
            ch.push_back(bn.right);
        }
        while (ch.size() < 8) {
            int pick = -1;
            float best = -1.0f;
            for (size_t i = 0; i < ch.size(); ++i) {
                const BNode &c = B.nodes[ch[i]];
                if (c.leaf()) continue;
                float a = c.box.area();
                if (a > best) {
                    best = a;
                    pick = (int)i;
                }
            }
            if (pick < 0) break;
            uint32_t c = ch[pick];
            ch[pick] = B.nodes[c].left;
            ch.push_back(B.nodes[c].right);
        }
        /* octant-ordered slot assignment (greedy on the centroid-offset cost) */
        Box3 nb;
        for (uint32_t c : ch) nb.grow(B.nodes[c].box);
        float pc[3];
        for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (nb.lo[a] + nb.hi[a]);
        float cost[8][8];
        for (size_t i = 0; i < ch.size(); ++i) {
            const Box3 &b = B.nodes[ch[i]].box;
            float d[3];
            for (int a = 0; a < 3; ++a) d[a] = 0.5f * (b.lo[a] + b.hi[a]) - pc[a];
            for (int s = 0; s < 8; ++s) {
                float dx = (s & 1) ? -1.0f : 1.0f, dy = (s & 2) ? -1.0f : 1.0f, dz = (s & 4) ? -1.0f : 1.0f;
                cost[i][s] = d[0] * dx + d[1] * dy + d[2] * dz;
            }
        }
        int slot_of[8];
        bool child_done[8] = {false}, slot_used[8] = {false};
        for (size_t k = 0; k < ch.size(); ++k) {
            float best = 1e30f;
            int bi2 = -1, bs = -1;
            for (size_t i = 0; i < ch.size(); ++i) {
                if (child_done[i]) continue;
                for (int s = 0; s < 8; ++s) {
                    if (slot_used[s]) continue;
                    if (cost[i][s] < best) {
                        best = cost[i][s];
                        bi2 = (int)i;
                        bs = s;
                    }
                }
            }
            child_done[bi2] = true;
            slot_used[bs] = true;
            slot_of[bi2] = bs;
        }
        int child_at[8];
        for (int s = 0; s < 8; ++s) child_at[s] = -1;
        for (size_t i = 0; i < ch.size(); ++i) child_at[slot_of[i]] = (int)i;

        WideNode n;
        std::memset(&n, 0, sizeof(n));
        /* quantisation frame */
        float p[3] = {nb.lo[0], nb.lo[1], nb.lo[2]};
        int e[3];
        float sc[3];
        for (int a = 0; a < 3; ++a) {
            e[a] = pick_exponent(p[a], nb.hi[a]);
            sc[a] = std::ldexp(1.0f, e[a]);
        }
        n.px = p[0];
        n.py = p[1];
        n.pz = p[2];
        uint32_t imask = 0;
        uint32_t n_internal = 0;
        for (int s = 0; s < 8; ++s)
            if (child_at[s] >= 0 && !B.nodes[ch[child_at[s]]].leaf()) {
                imask |= 1u << s;
                ++n_internal;
            }
        n.e_imask = (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) | ((uint32_t)(e[2] + 127) << 16) | (imask << 24);
        uint32_t child_base = (uint32_t)out.size();
        out.resize(out.size() + n_internal);
        n.child_base = child_base;
        n.leaf_base = (uint32_t)leaf_refs.size();
        uint32_t leaf_off = 0;
        uint8_t meta[8] = {0}, qlo[3][8] = {{0}}, qhi[3][8] = {{0}};
        std::vector<std::pair<uint32_t, uint32_t>> internal; /* (wide index, binary index) */
        uint32_t internal_rank = 0;
        for (int s = 0; s < 8; ++s) {
            if (child_at[s] < 0) continue;
            const BNode &c = B.nodes[ch[child_at[s]]];
            for (int a = 0; a < 3; ++a) {
                float lo = std::floor((c.box.lo[a] - p[a]) / sc[a]);
                float hi = std::ceil((c.box.hi[a] - p[a]) / sc[a]);
                lo = std::min(std::max(lo, 0.0f), 255.0f);
                hi = std::min(std::max(hi, 0.0f), 255.0f);
                while (lo > 0.0f && p[a] + lo * sc[a] > c.box.lo[a]) lo -= 1.0f;
                while (hi < 255.0f && p[a] + hi * sc[a] < c.box.hi[a]) hi += 1.0f;
                qlo[a][s] = (uint8_t)lo;
                qhi[a][s] = (uint8_t)hi;
            }
            if (!c.leaf()) {
                meta[s] = (uint8_t)(0x20 | (24 + s));
                internal.push_back({child_base + internal_rank, ch[child_at[s]]});
                ++internal_rank;
            } else {
                uint32_t cnt = c.count;
                uint32_t unary = cnt == 1 ? 1u : (cnt == 2 ? 3u : 7u);
                meta[s] = (uint8_t)((unary << 5) | leaf_off);
                for (uint32_t k = 0; k < cnt; ++k) leaf_refs.push_back(B.order[c.first + k]);
                leaf_off += cnt;
            }
        }
        auto pack = [](const uint8_t *b, uint32_t *w) {
            w[0] = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
            w[1] = b[4] | (b[5] << 8) | (b[6] << 16) | ((uint32_t)b[7] << 24);
        };
        pack(meta, n.meta);
        pack(qlo[0], n.qlo_x);
        pack(qlo[1], n.qlo_y);
        pack(qlo[2], n.qlo_z);
        pack(qhi[0], n.qhi_x);
        pack(qhi[1], n.qhi_y);
        pack(qhi[2], n.qhi_z);
        out[wi] = n;
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
