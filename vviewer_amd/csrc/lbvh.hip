/*
 * lbvh.hip - the LBVH build (lbvh.h) on gfx950.
 *
 *   k_lbvh_bounds    centroid bounds: wave reduction, one atomic min/max per wave on order-preserving integer images
 *   k_lbvh_morton    Morton code per box
 *   k_rs_*           stable LSD radix sort of (code, index), 8-bit digits, 4 passes: per-tile digit histograms,
 *                    one exclusive scan (digit-major), then a scatter that ranks keys inside the tile with wave
 *                    ballots (lanes whose digit matches, lower lanes first) and per-wave digit counts in LDS
 *   k_lbvh_inner     Karras internal nodes and parent links
 *   k_lbvh_boxes     boxes bottom-up: one thread per leaf climbs; the second arrival at a node merges its children
 *   k_lbvh_count /   breadth-first collapse, one thread per wide node of a level: counts, exclusive scans of the
 *   k_lbvh_emit      counts, then node emission (child and leaf blocks placed by the scans)
 * Every step is deterministic, so the output equals the host build (lbvh_host.cpp) byte for byte.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "lbvh_build.h"

namespace vvpt {
namespace {

constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr uint32_t kRsTile = kRsThreads * kRsItems;

__device__ inline uint32_t float_order(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float order_float(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline uint32_t wave_min(uint32_t v) {
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ inline uint32_t wave_max(uint32_t v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

/* bounds[0..2] = ordered min, bounds[3..5] = ordered max (initialised to ~0 and 0) */
__global__ __launch_bounds__(256) void k_lbvh_bounds(const float *boxes, uint32_t n, uint32_t *bounds) {
    uint32_t lo[3] = {~0u, ~0u, ~0u}, hi[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float c[3];
        centroid_of(boxes + 6 * (size_t)i, c);
        for (int a = 0; a < 3; ++a) {
            const uint32_t o = float_order(c[a]);
            lo[a] = min(lo[a], o);
            hi[a] = max(hi[a], o);
        }
    }
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) {
            atomicMin(&bounds[a], lo[a]);
            atomicMax(&bounds[3 + a], hi[a]);
        }
}

__global__ __launch_bounds__(256) void k_lbvh_morton(const float *boxes, uint32_t n, MortonFrame f, uint32_t *codes,
                                                     uint32_t *idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    codes[i] = morton_of_box(boxes + 6 * (size_t)i, f);
    idx[i] = i;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t *keys, uint32_t n, int shift, uint32_t ntiles,
                                                        uint32_t *hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        const uint32_t i = base + r * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

/* Exclusive scan of data[0 .. count) in place by one block; *total receives the sum. */
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t *data, uint32_t count, uint32_t *total) {
    __shared__ uint32_t s[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (count + 1023u) / 1024u;
    const uint32_t b = min(count, t * per), e = min(count, b + per);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += data[i];
    s[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? s[t - 1] : 0u;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
    if (t == 1023 && total) *total = s[1023];
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                           uint32_t *vout, uint32_t n, int shift, uint32_t ntiles,
                                                           const uint32_t *hist) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[kRsThreads / 64][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    run[t] = hist[t * ntiles + blockIdx.x];
    for (int k = 0; k < kRsThreads / 64; ++k) wc[k][t] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRsTile;
    const uint64_t lower = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int r = 0; r < kRsItems; ++r) {
        const uint32_t i = base + r * kRsThreads + t;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t same = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1u;
            const uint64_t bb = __ballot(on);
            same &= on ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(same & lower);
        if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int k = 0; k < w; ++k) pos += wc[k][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int k = 0; k < kRsThreads / 64; ++k) {
            add += wc[k][t];
            wc[k][t] = 0;
        }
        run[t] += add;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_lbvh_inner(const uint32_t *codes, uint32_t n, LbvhInner *inner,
                                                    uint32_t *leaf_parent, uint32_t *inner_parent) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= n) return;
    const LbvhInner r = lbvh_inner(codes, (int)n, (int)i);
    inner[i] = r;
    if (r.left & kLeafBit) leaf_parent[r.left & ~kLeafBit] = i; else inner_parent[r.left] = i;
    if (r.right & kLeafBit) leaf_parent[r.right & ~kLeafBit] = i; else inner_parent[r.right] = i;
}

__device__ inline float coherent_load(const float *p) {
    return __uint_as_float(__hip_atomic_load((const uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

/* One thread per leaf climbs towards the root; at each node the first arrival stops and the second, which then
 * sees both children final, writes the node's box. */
__global__ __launch_bounds__(256) void k_lbvh_boxes(const float *boxes, const uint32_t *order, uint32_t n,
                                                    const LbvhInner *inner, const uint32_t *leaf_parent,
                                                    const uint32_t *inner_parent, uint32_t *visits, HBox *ibox) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || n < 2) return;
    uint32_t node = leaf_parent[j];
    for (;;) {
        __threadfence();
        if (atomicAdd(&visits[node], 1u) == 0u) return;
        __threadfence();
        HBox b = hbox_empty();
        const uint32_t kids[2] = {inner[node].left, inner[node].right};
        for (int c = 0; c < 2; ++c) {
            HBox cb;
            if (kids[c] & kLeafBit) {
                const float *p = boxes + 6 * (size_t)order[kids[c] & ~kLeafBit];
                for (int a = 0; a < 3; ++a) {
                    cb.lo[a] = p[a];
                    cb.hi[a] = p[3 + a];
                }
            } else {
                const float *q = (const float *)&ibox[kids[c]];
                for (int a = 0; a < 3; ++a) {
                    cb.lo[a] = coherent_load(q + a);
                    cb.hi[a] = coherent_load(q + 3 + a);
                }
            }
            hbox_grow(b, cb);
        }
        ibox[node] = b;
        if (node == 0) return;
        node = inner_parent[node];
    }
}

__global__ __launch_bounds__(128) void k_lbvh_count(LbvhView v, const uint32_t *level, uint32_t m, uint32_t *ni,
                                                    uint32_t *np) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    uint32_t a, b;
    lbvh_collapse_count(v, level[k], a, b);
    ni[k] = a;
    np[k] = b;
}

__global__ __launch_bounds__(128) void k_lbvh_emit(LbvhView v, const uint32_t *level, uint32_t m, uint32_t level_base,
                                                   uint32_t next_base, uint32_t leaf_cursor, const uint32_t *ni,
                                                   const uint32_t *np, WideNode *nodes, uint32_t *leaf_refs,
                                                   uint32_t *next) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    nodes[level_base + k] =
        lbvh_collapse_emit(v, level[k], next_base + ni[k], leaf_cursor + np[k], leaf_refs, next + ni[k]);
}

struct DevMem {
    std::vector<void *> ptrs;
    ~DevMem() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T **p, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 16));
        if (e == hipSuccess) ptrs.push_back(q);
        *p = (T *)q;
        return e;
    }
};

#define LB_TRY(expr)                                                         \
    do {                                                                     \
        hipError_t _e = (expr);                                              \
        if (_e != hipSuccess) return std::string(#expr) + ": " + hipGetErrorString(_e); \
    } while (0)

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

std::string lbvh_build_device(const float *h_boxes, uint32_t n, uint32_t max_leaf, bool keep_debug, void *stream,
                              LbvhOutput &out) {
    out = LbvhOutput();
    max_leaf = std::max(1u, std::min(3u, max_leaf));
    if (n == 0) {
        lbvh_build_host(h_boxes, 0, max_leaf, keep_debug, out);
        return "";
    }
    hipStream_t s = (hipStream_t)stream;
    DevMem mem;
    float *boxes;
    uint32_t *codes, *idx, *codes2, *idx2, *hist, *bounds, *totals, *leaf_parent, *inner_parent, *visits, *leaf_refs;
    uint32_t *lvl_a, *lvl_b, *ni, *np;
    LbvhInner *inner;
    HBox *ibox;
    WideNode *nodes;
    const uint32_t ntiles = blocks_for(n, kRsTile);
    const uint32_t ninner = n > 1 ? n - 1 : 1;
    LB_TRY(mem.alloc(&boxes, 6 * (size_t)n));
    LB_TRY(mem.alloc(&codes, n));
    LB_TRY(mem.alloc(&idx, n));
    LB_TRY(mem.alloc(&codes2, n));
    LB_TRY(mem.alloc(&idx2, n));
    LB_TRY(mem.alloc(&hist, 256 * (size_t)ntiles));
    LB_TRY(mem.alloc(&bounds, 8));
    LB_TRY(mem.alloc(&totals, 4));
    LB_TRY(mem.alloc(&leaf_parent, n));
    LB_TRY(mem.alloc(&inner_parent, ninner));
    LB_TRY(mem.alloc(&visits, ninner));
    LB_TRY(mem.alloc(&inner, ninner));
    LB_TRY(mem.alloc(&ibox, ninner));
    LB_TRY(mem.alloc(&nodes, n));      /* a wide node holds at least two children: at most n - 1 nodes, 1 if n = 1 */
    LB_TRY(mem.alloc(&leaf_refs, n));
    LB_TRY(mem.alloc(&lvl_a, n));
    LB_TRY(mem.alloc(&lvl_b, n));
    LB_TRY(mem.alloc(&ni, n));
    LB_TRY(mem.alloc(&np, n));
    LB_TRY(hipMemcpyAsync(boxes, h_boxes, 6 * sizeof(float) * (size_t)n, hipMemcpyHostToDevice, s));
    const uint32_t init_bounds[8] = {~0u, ~0u, ~0u, 0u, 0u, 0u, 0u, 0u};
    LB_TRY(hipMemcpyAsync(bounds, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice, s));
    LB_TRY(hipMemsetAsync(visits, 0, sizeof(uint32_t) * ninner, s));
    LB_TRY(hipStreamSynchronize(s));

    hipEvent_t e0, e1;
    LB_TRY(hipEventCreate(&e0));
    LB_TRY(hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evg{e0, e1};
    LB_TRY(hipEventRecord(e0, s));
    /* 1. bounds and codes */
    hipLaunchKernelGGL(k_lbvh_bounds, dim3(std::min<uint32_t>(blocks_for(n, 256), 1024)), dim3(256), 0, s, boxes, n,
                       bounds);
    LB_TRY(hipGetLastError());
    uint32_t hb[8];
    LB_TRY(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, s));
    LB_TRY(hipStreamSynchronize(s));
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = order_float(hb[a]);
        hi[a] = order_float(hb[3 + a]);
    }
    const MortonFrame f = lbvh_frame_from_bounds(lo, hi);
    hipLaunchKernelGGL(k_lbvh_morton, dim3(blocks_for(n, 256)), dim3(256), 0, s, boxes, n, f, codes, idx);
    LB_TRY(hipGetLastError());
    /* 2. stable radix sort, 4 passes of 8 bits (ends in codes / idx) */
    uint32_t *kin = codes, *vin = idx, *kout = codes2, *vout = idx2;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, s, kin, n, shift, ntiles, hist);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, hist, 256 * ntiles, totals);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, s, kin, vin, kout, vout, n, shift, ntiles,
                           hist);
        LB_TRY(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    /* 3. hierarchy, 4. boxes */
    if (n > 1) {
        hipLaunchKernelGGL(k_lbvh_inner, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, kin, n, inner, leaf_parent,
                           inner_parent);
        hipLaunchKernelGGL(k_lbvh_boxes, dim3(blocks_for(n, 256)), dim3(256), 0, s, boxes, vin, n, inner, leaf_parent,
                           inner_parent, visits, ibox);
        LB_TRY(hipGetLastError());
    }
    /* 5. breadth-first collapse */
    LbvhView v{inner, ibox, boxes, vin, n, max_leaf};
    const uint32_t root = n > 1 ? 0u : kLeafBit;
    LB_TRY(hipMemcpyAsync(lvl_a, &root, 4, hipMemcpyHostToDevice, s));
    uint32_t m = 1, level_base = 0, leaf_cursor = 0, depth = 0;
    uint32_t *cur = lvl_a, *nxt = lvl_b;
    while (m > 0) {
        hipLaunchKernelGGL(k_lbvh_count, dim3(blocks_for(m, 128)), dim3(128), 0, s, v, cur, m, ni, np);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, ni, m, totals);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, np, m, totals + 1);
        LB_TRY(hipGetLastError());
        uint32_t tot[2];
        LB_TRY(hipMemcpyAsync(tot, totals, 8, hipMemcpyDeviceToHost, s));
        LB_TRY(hipStreamSynchronize(s));
        const uint32_t next_base = level_base + m;
        if ((uint64_t)next_base + tot[0] > n || (uint64_t)leaf_cursor + tot[1] > n)
            return "LBVH collapse exceeded its bounds";
        hipLaunchKernelGGL(k_lbvh_emit, dim3(blocks_for(m, 128)), dim3(128), 0, s, v, cur, m, level_base, next_base,
                           leaf_cursor, ni, np, nodes, leaf_refs, nxt);
        LB_TRY(hipGetLastError());
        std::swap(cur, nxt);
        level_base = next_base;
        leaf_cursor += tot[1];
        m = tot[0];
        ++depth;
    }
    LB_TRY(hipEventRecord(e1, s));
    LB_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    LB_TRY(hipEventElapsedTime(&ms, e0, e1));
    out.ms = ms;
    out.depth = depth;
    out.nodes.resize(level_base);
    out.leaf_refs.resize(leaf_cursor);
    LB_TRY(hipMemcpy(out.nodes.data(), nodes, sizeof(WideNode) * (size_t)level_base, hipMemcpyDeviceToHost));
    LB_TRY(hipMemcpy(out.leaf_refs.data(), leaf_refs, sizeof(uint32_t) * (size_t)leaf_cursor, hipMemcpyDeviceToHost));
    if (keep_debug) {
        out.codes.resize(n);
        out.order.resize(n);
        LB_TRY(hipMemcpy(out.codes.data(), kin, 4 * (size_t)n, hipMemcpyDeviceToHost));
        LB_TRY(hipMemcpy(out.order.data(), vin, 4 * (size_t)n, hipMemcpyDeviceToHost));
        if (n > 1) {
            std::vector<LbvhInner> in(n - 1);
            LB_TRY(hipMemcpy(in.data(), inner, sizeof(LbvhInner) * (size_t)(n - 1), hipMemcpyDeviceToHost));
            out.inner.resize(2 * (size_t)(n - 1));
            for (uint32_t k = 0; k + 1 < n; ++k) {
                out.inner[2 * k] = in[k].left;
                out.inner[2 * k + 1] = in[k].right;
            }
        }
    }
    return "";
}

}  // namespace vvpt
