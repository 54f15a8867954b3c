/*
 * bvh_check.cpp - host-side checks of the wide-BVH builders (product code in
 * vviewer_amd/csrc/bvh_build.cpp, binned SAH, and vviewer_amd/csrc/lbvh_host.cpp,
 * LBVH), loaded by tests/test_bvh_build.py.
 *
 * bvh_check_build() builds over a triangle soup and verifies the encoding the
 * gfx950 traversal relies on: every child/leaf index in range, every primitive
 * referenced exactly once, leaf counts 1..3 with offsets below 24, internal
 * children contiguous at child_base, dequantised child boxes containing their
 * primitives, octant slot order consistent with the meta encoding.
 * bvh_check_closest() decodes the nodes exactly like the kernel (p + q * 2^e)
 * and answers closest-hit queries by an exhaustive node walk, so tests can
 * compare with a brute-force loop.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../vviewer_amd/csrc/bvh_build.h"
#include "../../vviewer_amd/csrc/lbvh_build.h"

using namespace vvpt;

namespace {

struct Built {
    std::vector<WideNode> nodes;
    std::vector<uint32_t> refs;
    WideBuildResult r;
};

inline uint8_t byte_of(const uint32_t *w, int slot) { return (uint8_t)((w[slot >> 2] >> (8 * (slot & 3))) & 0xff); }

void child_box(const WideNode &n, int slot, float *lo, float *hi) {
    const float p[3] = {n.px, n.py, n.pz};
    const uint32_t *qlo[3] = {n.qlo_x, n.qlo_y, n.qlo_z};
    const uint32_t *qhi[3] = {n.qhi_x, n.qhi_y, n.qhi_z};
    for (int a = 0; a < 3; ++a) {
        int e = (int)((n.e_imask >> (8 * a)) & 0xff) - 127;
        float s = std::ldexp(1.0f, e);
        lo[a] = p[a] + (float)byte_of(qlo[a], slot) * s;
        hi[a] = p[a] + (float)byte_of(qhi[a], slot) * s;
    }
}

bool tri_hit(const float *o, const float *d, const float *v0, const float *v1, const float *v2, float &t) {
    float e1[3], e2[3], s[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = v1[k] - v0[k];
        e2[k] = v2[k] - v0[k];
        s[k] = o[k] - v0[k];
    }
    p[0] = std::fma(d[1], e2[2], -(d[2] * e2[1]));
    p[1] = std::fma(d[2], e2[0], -(d[0] * e2[2]));
    p[2] = std::fma(d[0], e2[1], -(d[1] * e2[0]));
    float det = std::fma(e1[0], p[0], std::fma(e1[1], p[1], e1[2] * p[2]));
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    float u = std::fma(s[0], p[0], std::fma(s[1], p[1], s[2] * p[2])) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    q[0] = std::fma(s[1], e1[2], -(s[2] * e1[1]));
    q[1] = std::fma(s[2], e1[0], -(s[0] * e1[2]));
    q[2] = std::fma(s[0], e1[1], -(s[1] * e1[0]));
    float v = std::fma(d[0], q[0], std::fma(d[1], q[1], d[2] * q[2])) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    t = std::fma(e2[0], q[0], std::fma(e2[1], q[1], e2[2] * q[2])) * inv;
    return true;
}

bool box_hit(const float *o, const float *d, const float *lo, const float *hi, float tmax) {
    float tn = 0.0f, tf = tmax;
    for (int a = 0; a < 3; ++a) {
        float inv = d[a] == 0.0f ? std::copysign(1e30f, d[a]) : 1.0f / d[a];
        float t0 = (lo[a] - o[a]) * inv, t1 = (hi[a] - o[a]) * inv;
        if (t0 > t1) std::swap(t0, t1);
        tn = std::max(tn, t0);
        tf = std::min(tf, t1);
    }
    return tn <= tf * 1.000001f;
}

}  // namespace

extern "C" {

/* Returns 0 on success or a negative code naming the first violated invariant;
 * info[0..3] = node count, depth, max stack need, leaf refs. */
int bvh_check_build(const float *tris, int ntris, int max_leaf, int builder, uint32_t *info, void **handle) {
    Built *b = new Built();
    std::vector<Box3> boxes(ntris);
    for (int i = 0; i < ntris; ++i)
        for (int k = 0; k < 3; ++k) boxes[i].grow(tris + 9 * i + 3 * k);
    if (builder == 0) {
        b->r = build_wide_bvh(boxes, max_leaf, b->nodes, b->refs);
    } else { /* LBVH on the host */
        std::vector<float> b6(6 * (size_t)ntris);
        for (int i = 0; i < ntris; ++i)
            for (int a = 0; a < 3; ++a) {
                b6[6 * i + a] = boxes[i].lo[a];
                b6[6 * i + 3 + a] = boxes[i].hi[a];
            }
        LbvhOutput o;
        lbvh_build_host(b6.data(), (uint32_t)ntris, (uint32_t)max_leaf, false, o);
        b->nodes = o.nodes;
        b->refs = o.leaf_refs;
        b->r.root = 0;
        b->r.depth = o.depth;
    }
    info[0] = (uint32_t)b->nodes.size();
    info[1] = b->r.depth;
    info[2] = 0;
    info[3] = (uint32_t)b->refs.size();
    *handle = b;
    if (ntris == 0) return 0;
    if ((int)b->refs.size() != ntris) return -1;
    std::vector<int> seen(ntris, 0);
    for (uint32_t r : b->refs) {
        if ((int)r >= ntris) return -2;
        seen[r]++;
    }
    for (int i = 0; i < ntris; ++i)
        if (seen[i] != 1) return -3;
    /* walk from the root, checking encoding and containment */
    std::vector<int> visited(b->nodes.size(), 0);
    uint32_t maxd = 0;
    int err = 0;
    /* returns the bounds of every primitive below `node` (lo/hi) */
    struct Rec {
        Built *b;
        const float *tris;
        int max_leaf;
        std::vector<int> &visited;
        uint32_t &maxd;
        int &err;
        void run(uint32_t node, uint32_t depth, float *plo, float *phi) {
            for (int a = 0; a < 3; ++a) { plo[a] = 1e30f; phi[a] = -1e30f; }
            if (err) return;
            if (node >= b->nodes.size()) { err = -4; return; }
            if (visited[node]++) { err = -5; return; }
            maxd = std::max(maxd, depth);
            const WideNode &n = b->nodes[node];
            uint32_t imask = n.e_imask >> 24;
            uint32_t rank = 0, leaf_total = 0;
            for (int s = 0; s < 8; ++s) {
                uint8_t meta = byte_of(n.meta, s);
                float lo[3], hi[3];
                child_box(n, s, lo, hi);
                bool internal = (imask >> s) & 1u;
                if (meta == 0) { if (internal) { err = -6; return; } continue; }
                float clo[3], chi[3];
                if (internal) {
                    if (meta != (uint8_t)(0x20 | (24 + s))) { err = -7; return; }
                    uint32_t child = n.child_base + rank++;
                    if (child >= b->nodes.size()) { err = -8; return; }
                    run(child, depth + 1, clo, chi);
                    if (err) return;
                } else {
                    uint32_t unary = meta >> 5, off = meta & 0x1f;
                    uint32_t cnt = unary == 1 ? 1 : (unary == 3 ? 2 : (unary == 7 ? 3 : 0));
                    if (cnt == 0 || (int)cnt > max_leaf) { err = -9; return; }
                    if (off + cnt > 24) { err = -10; return; }
                    leaf_total += cnt;
                    for (int a = 0; a < 3; ++a) { clo[a] = 1e30f; chi[a] = -1e30f; }
                    for (uint32_t k = 0; k < cnt; ++k) {
                        uint32_t li = n.leaf_base + off + k;
                        if (li >= b->refs.size()) { err = -11; return; }
                        const float *t = tris + 9 * b->refs[li];
                        for (int v = 0; v < 3; ++v)
                            for (int a = 0; a < 3; ++a) {
                                clo[a] = std::min(clo[a], t[3 * v + a]);
                                chi[a] = std::max(chi[a], t[3 * v + a]);
                            }
                    }
                }
                for (int a = 0; a < 3; ++a) {
                    if (clo[a] < lo[a] || chi[a] > hi[a]) { err = -12; return; }
                    plo[a] = std::min(plo[a], clo[a]);
                    phi[a] = std::max(phi[a], chi[a]);
                }
            }
            if (leaf_total > 24) err = -13;
        }
    } rec{b, tris, max_leaf, visited, maxd, err};
    float rlo[3], rhi[3];
    rec.run(b->r.root, 1, rlo, rhi);
    if (err) return err;
    info[1] = maxd;
    info[2] = maxd + 1;
    return 0;
}

/* Closest hit through the encoded nodes (exhaustive walk pruned by the dequantised boxes). */
void bvh_check_closest(void *handle, const float *tris, int nrays, const float *orig, const float *dir, float tmin,
                       float tmax, int32_t *prim, float *tout) {
    Built *b = (Built *)handle;
    for (int r = 0; r < nrays; ++r) {
        const float *o = orig + 3 * r, *d = dir + 3 * r;
        float best = tmax;
        int32_t bp = -1;
        std::vector<uint32_t> st = {b->r.root};
        while (!st.empty()) {
            const WideNode &n = b->nodes[st.back()];
            st.pop_back();
            uint32_t imask = n.e_imask >> 24, rank = 0;
            for (int s = 0; s < 8; ++s) {
                uint8_t meta = byte_of(n.meta, s);
                if (meta == 0) continue;
                bool internal = (imask >> s) & 1u;
                float lo[3], hi[3];
                child_box(n, s, lo, hi);
                uint32_t child = internal ? n.child_base + rank++ : 0;
                if (!box_hit(o, d, lo, hi, best)) continue;
                if (internal) {
                    st.push_back(child);
                } else {
                    uint32_t unary = meta >> 5, off = meta & 0x1f;
                    uint32_t cnt = unary == 1 ? 1 : (unary == 3 ? 2 : 3);
                    for (uint32_t k = 0; k < cnt; ++k) {
                        uint32_t p = b->refs[n.leaf_base + off + k];
                        const float *t = tris + 9 * p;
                        float tt;
                        if (tri_hit(o, d, t, t + 3, t + 6, tt) && tt >= tmin && tt <= best)
                            if (tt < best || bp < 0 || (int32_t)p < bp) {
                                best = tt;
                                bp = (int32_t)p;
                            }
                    }
                }
            }
        }
        prim[r] = bp;
        tout[r] = bp >= 0 ? best : 0.0f;
    }
}

void bvh_check_free(void *handle) { delete (Built *)handle; }

/* Box3 union with an empty box must leave the box as it was (an empty box's corners are +1e30 / -1e30). */
int bvh_check_empty_union(void) {
    Box3 a, e;
    const float p[3] = {1.0f, 2.0f, 3.0f}, q[3] = {2.0f, 4.0f, 5.0f};
    a.grow(p);
    a.grow(q);
    const Box3 before = a;
    a.grow(e);
    for (int k = 0; k < 3; ++k)
        if (a.lo[k] != before.lo[k] || a.hi[k] != before.hi[k]) return 0;
    Box3 f;
    f.grow(e);
    return f.valid() ? 0 : 1;
}

/* Surface-area cost of a built wide BVH relative to its root box: sum over children of area x (1 for a wide node,
 * 0.26 x primitives for a leaf), the measure the collapse minimises. */
double bvh_check_sah(void *handle) {
    Built *b = (Built *)handle;
    double total = 0.0, root = 0.0;
    for (size_t k = b->r.root; k < b->nodes.size(); ++k) {
        const WideNode &n = b->nodes[k];
        for (int s = 0; s < 8; ++s) {
            const uint8_t meta = byte_of(n.meta, s);
            if (!meta) continue;
            float lo[3], hi[3];
            child_box(n, s, lo, hi);
            const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
            const double area = 2.0 * (dx * dy + dy * dz + dz * dx);
            const uint32_t unary = meta >> 5;
            total += (meta & 0x1f) >= 24 ? area : area * 0.26 * (unary == 1 ? 1 : (unary == 3 ? 2 : 3));
        }
    }
    /* root area: union of the root's children */
    const WideNode &r = b->nodes[b->r.root];
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int s = 0; s < 8; ++s) {
        if (!byte_of(r.meta, s)) continue;
        float l[3], h[3];
        child_box(r, s, l, h);
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    }
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    root = 2.0 * (dx * dy + dy * dz + dz * dx);
    return total / root;
}

} /* extern "C" */
