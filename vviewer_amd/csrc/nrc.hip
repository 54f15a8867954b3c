/*
 * nrc.hip - the neural radiance cache's kernels on gfx950 and its host-side
 * model (parameters, Adam state, training steps).
 *
 *   k_nrc_forward<TRAIN, WAVES>  16 samples per wave, WAVES waves per block: hash-grid + SH encoding into an LDS tile, then the six
 *                         layers as v_mfma_f32_16x16x16_bf16 products against the weights held in LDS; TRAIN keeps
 *                         every layer's input and writes the relative-L2 loss gradient of the outputs
 *   k_nrc_bwd_dx          dL/dx_l = g_l W_l^T (masked by ReLU') per 16-sample tile, on MFMA
 *   k_nrc_bwd_dw          dW_l = x_l^T g_l: each wave reduces 64 samples on MFMA, then adds its 64 x N block
 *   k_nrc_bwd_grid        hash-grid gradients: trilinear weights times dL/dx_0, scattered with atomics
 *   k_nrc_adam            Adam on fp32 masters; refreshes the bf16 weights; clears the gradients
 * MFMA lane layout (16x16x16, wave64): lane l holds A[l%16][4(l/16)+i], B[4(l/16)+i][l%16] and
 * D[4(l/16)+i][l%16], i = 0..3.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "nrc.h"
#include "nrc_model.h"

namespace vvpt {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
/* Samples each wave of k_nrc_bwd_dw reduces on MFMA before adding its block of weight gradients with atomics: more
 * waves hide the latency of the gathered operands better than fewer atomics save (per 1080p 64-spp cached render:
 * 64 samples 56.7 ms of cache work, 256: 65.2, 1024: 103). */
constexpr uint32_t kDwSamples = 64;
/* Inference blocks of 4 waves (each block holds its own LDS copy of the weights).  Measured per 1080p 64-spp
 * cached render: 4-wave blocks 65 ms of cache work, 8-wave 68, 16-wave 73 (one weight copy for more waves, but
 * barriers across more waves). */
constexpr int kInferWaves = 4;
constexpr int kTile = 16;
constexpr int kBlockSamples = kWaves * kTile;

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

/* Trilinear lookup of one level: features (f0, f1); optionally scatters g0, g1 into a gradient grid instead. */
template <bool SCATTER>
__device__ __forceinline__ void grid_level(const NrcView &V, const float *p, int lv, const float *grid, float *ggrid,
                                           float g0, float g1, float &f0, float &f1) {
    const float N = V.res[lv];
    const uint32_t n = (uint32_t)N;
    uint32_t i0[3];
    float fr[3];
    for (int a = 0; a < 3; ++a) {
        const float x = p[a] * N;
        const uint32_t c = min((uint32_t)floorf(x), n - 1u);
        i0[a] = c;
        fr[a] = x - (float)c;
    }
    f0 = f1 = 0.0f;
    for (int c = 0; c < 8; ++c) {
        const uint32_t cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const float w = (cx ? fr[0] : 1.0f - fr[0]) * (cy ? fr[1] : 1.0f - fr[1]) * (cz ? fr[2] : 1.0f - fr[2]);
        const uint32_t idx = nrc_grid_index(i0[0] + cx, i0[1] + cy, i0[2] + cz, n);
        const size_t o = ((size_t)lv * kNrcT + idx) * kNrcFeat;
        if (SCATTER) {
            atomicAdd(ggrid + o, w * g0);
            atomicAdd(ggrid + o + 1, w * g1);
        } else {
            f0 += w * grid[o];
            f1 += w * grid[o + 1];
        }
    }
}

__device__ __forceinline__ void norm_pos(const NrcView &V, const float *x, float *p) {
    for (int a = 0; a < 3; ++a) p[a] = fminf(fmaxf((x[a] - V.lo[a]) * V.inv_ext[a], 0.0f), 1.0f);
}

template <bool TRAIN, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_nrc_forward(NrcView V, const float *in, uint32_t n,
                                                            const uint32_t *n_dev, float *pred, uint16_t *acts,
                                                            uint32_t npad, const float *targets, float *dout,
                                                            float *loss) {
    constexpr uint32_t kGroup = WAVES * kTile; /* samples per block pass */
    /* weights transposed per layer ([out][in]) with rows padded to kWRow, so a lane's four B-operand values
     * (consecutive inputs of one output) are one 8-byte LDS read, and the 16 lanes of a row group fall in
     * different banks */
    constexpr int kWRow = kNrcWidth + 4;
    __shared__ __align__(16) uint16_t W[(kNrcHidden * kNrcWidth + kNrcOut) * kWRow];
    __shared__ uint16_t act[WAVES][kTile][kNrcWidth];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (uint32_t e = tid; e < kNrcMlpParams; e += WAVES * 64) {
        const uint32_t l = min(e / (uint32_t)(kNrcWidth * kNrcWidth), (uint32_t)kNrcHidden);
        const uint32_t off = e - l * kNrcWidth * kNrcWidth, nout = l < (uint32_t)kNrcHidden ? kNrcWidth : kNrcOut;
        W[(l * kNrcWidth + off % nout) * kWRow + off / nout] = V.w[e];
    }
    if (n_dev) n = min(*n_dev, n); /* a count produced on the device; n is then the capacity */
    /* blocks stride over groups of kGroup samples (one pass each when the grid covers n) */
    for (uint32_t blk = blockIdx.x; blk * kGroup < n; blk += gridDim.x) {
    const uint32_t s0 = (blk * WAVES + wave) * kTile;
    __syncthreads();
    {
        /* encoding: lane = (sample s, level quarter q) */
        const int s = lane & 15, q = lane >> 4;
        const uint32_t gs = s0 + s;
        float x[kNrcIn];
        for (int k = 0; k < kNrcIn; ++k) x[k] = gs < n ? in[(size_t)gs * kNrcIn + k] : 0.0f;
        float p[3];
        norm_pos(V, x, p);
        for (int lv = q * 4; lv < q * 4 + 4; ++lv) {
            float f0, f1;
            grid_level<false>(V, p, lv, V.grid, nullptr, 0.f, 0.f, f0, f1);
            act[wave][s][2 * lv] = f32_to_bf16(f0);
            act[wave][s][2 * lv + 1] = f32_to_bf16(f1);
        }
        if (q == 0) {
            const float dl = sqrtf(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
            const float id = dl > 0.0f ? 1.0f / dl : 0.0f;
            float sh[16];
            nrc_sh16(x[3] * id, x[4] * id, x[5] * id, sh);
            for (int k = 0; k < 16; ++k) act[wave][s][32 + k] = f32_to_bf16(sh[k]);
            act[wave][s][48] = f32_to_bf16(x[6]);
            act[wave][s][49] = f32_to_bf16(x[7]);
            act[wave][s][50] = f32_to_bf16(x[8]);
            act[wave][s][51] = f32_to_bf16(x[9]);
            act[wave][s][52] = f32_to_bf16(1.0f);
            for (int k = 53; k < kNrcWidth; ++k) act[wave][s][k] = 0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < kNrcLayers; ++l) {
        const int nout = nrc_layer_out(l);
        const uint16_t *WT = W + (size_t)l * kNrcWidth * kWRow;
        if (TRAIN) { /* keep the layer input: 16 rows x 64 bf16 per wave, 16 values per lane */
            const int r = lane >> 2, c0 = (lane & 3) * 16;
            uint16_t *dst = acts + ((size_t)l * npad + s0 + r) * kNrcWidth + c0;
            for (int k = 0; k < 16; ++k) dst[k] = act[wave][r][c0 + k];
        }
        bf16x4 a[4];
        for (int kb = 0; kb < 4; ++kb)
            for (int i = 0; i < 4; ++i) a[kb][i] = (short)act[wave][lane & 15][kb * 16 + 4 * (lane >> 4) + i];
        f32x4 acc[4];
        for (int nb = 0; nb < nout / 16; ++nb) {
            acc[nb] = {0.f, 0.f, 0.f, 0.f};
            for (int kb = 0; kb < 4; ++kb) {
                const bf16x4 b = *reinterpret_cast<const bf16x4 *>(WT + (nb * 16 + (lane & 15)) * kWRow + kb * 16 +
                                                                  4 * (lane >> 4));
                acc[nb] = mfma16(a[kb], b, acc[nb]);
            }
        }
        __syncthreads();
        if (l < kNrcHidden) {
            for (int nb = 0; nb < 4; ++nb)
                for (int i = 0; i < 4; ++i)
                    act[wave][4 * (lane >> 4) + i][nb * 16 + (lane & 15)] = f32_to_bf16(fmaxf(acc[nb][i], 0.0f));
        } else {
            float lsum = 0.0f;
            const int c = lane & 15;
            for (int i = 0; i < 4; ++i) {
                const uint32_t gs = s0 + 4 * (lane >> 4) + i;
                const float y = acc[0][i];
                if (c < 3 && gs < n && pred) pred[(size_t)gs * 3 + c] = y;
                if (TRAIN) {
                    float g = 0.0f;
                    if (c < 3 && gs < n) {
                        const float t = targets[(size_t)gs * 3 + c];
                        const float den = y * y + 0.01f;
                        const float d = y - t;
                        lsum += d * d / den;
                        g = 2.0f * d / den / (3.0f * (float)n);
                    }
                    dout[(size_t)gs * kNrcOut + c] = g;
                }
            }
            if (TRAIN) {
                for (int o = 32; o >= 1; o >>= 1) lsum += __shfl_xor(lsum, o);
                if (lane == 0) atomicAdd(loss, lsum / (3.0f * (float)n));
            }
        }
        __syncthreads();
    }
    }
}

/* g_in = (g W_l^T) masked by x_l > 0 (layers l > 0); layer 0 writes the 32 grid columns unmasked. */
template <int nout>
__global__ __launch_bounds__(kWaves * 64) void k_nrc_bwd_dx(int l, const float *g, const uint16_t *w16,
                                                            const uint16_t *acts, uint32_t npad, float *gin) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t s0 = (blockIdx.x * kWaves + wave) * kTile;
    if (s0 >= npad) return;
    const uint16_t *Wl = w16 + nrc_layer_offset(l);
    const int kblocks = nout / 16;
    for (int nb = 0; nb < 4; ++nb) {
        if (l == 0 && nb >= 2) break; /* only the grid features need a gradient */
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < kblocks; ++kb) {
            bf16x4 a, b;
            for (int i = 0; i < 4; ++i) {
                const int k = kb * 16 + 4 * (lane >> 4) + i;
                a[i] = (short)f32_to_bf16(g[(size_t)(s0 + (lane & 15)) * nout + k]);
                b[i] = (short)Wl[(nb * 16 + (lane & 15)) * nout + k]; /* B[k][j] = W[j][k] */
            }
            acc = mfma16(a, b, acc);
        }
        for (int i = 0; i < 4; ++i) {
            const uint32_t s = s0 + 4 * (lane >> 4) + i;
            const int j = nb * 16 + (lane & 15);
            if (l == 0) {
                gin[(size_t)s * 32 + j] = acc[i];
            } else {
                const float x = bf16_to_f32(acts[((size_t)l * npad + s) * kNrcWidth + j]);
                gin[(size_t)s * kNrcWidth + j] = x > 0.0f ? acc[i] : 0.0f;
            }
        }
    }
}

/* dW_l += x_l^T g_l over 256 samples per wave. */
template <int nout>
__global__ __launch_bounds__(kWaves * 64) void k_nrc_bwd_dw(int l, const float *g, const uint16_t *acts,
                                                            uint32_t npad, float *dw) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t c0 = (blockIdx.x * kWaves + wave) * kDwSamples;
    if (c0 >= npad) return;
    constexpr int nblocks = nout / 16;
    f32x4 acc[4][4];
    for (int mb = 0; mb < 4; ++mb)
        for (int nb = 0; nb < 4; ++nb) acc[mb][nb] = {0.f, 0.f, 0.f, 0.f};
    const uint32_t c1 = min(npad, c0 + kDwSamples);
    for (uint32_t sb = c0; sb < c1; sb += 16) {
        bf16x4 b[4];
        for (int nb = 0; nb < nblocks; ++nb)
            for (int i = 0; i < 4; ++i)
                b[nb][i] = (short)f32_to_bf16(g[(size_t)(sb + 4 * (lane >> 4) + i) * nout + nb * 16 + (lane & 15)]);
        for (int mb = 0; mb < 4; ++mb) {
            bf16x4 a;
            for (int i = 0; i < 4; ++i)
                a[i] = (short)acts[((size_t)l * npad + sb + 4 * (lane >> 4) + i) * kNrcWidth + mb * 16 + (lane & 15)];
            for (int nb = 0; nb < nblocks; ++nb) acc[mb][nb] = mfma16(a, b[nb], acc[mb][nb]);
        }
    }
    float *dWl = dw + nrc_layer_offset(l);
    for (int mb = 0; mb < 4; ++mb)
        for (int nb = 0; nb < nblocks; ++nb)
            for (int i = 0; i < 4; ++i)
                atomicAdd(dWl + (size_t)(mb * 16 + 4 * (lane >> 4) + i) * nout + nb * 16 + (lane & 15), acc[mb][nb][i]);
}

__global__ __launch_bounds__(256) void k_nrc_bwd_grid(NrcView V, const float *in, const float *gin, uint32_t n,
                                                      float *ggrid) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kNrcLevels) return;
    const uint32_t s = t / kNrcLevels;
    const int lv = (int)(t % kNrcLevels);
    float p[3];
    norm_pos(V, in + (size_t)s * kNrcIn, p);
    float f0, f1;
    grid_level<true>(V, p, lv, nullptr, ggrid, gin[(size_t)s * 32 + 2 * lv], gin[(size_t)s * 32 + 2 * lv + 1], f0, f1);
}

__global__ __launch_bounds__(256) void k_nrc_adam(float *p, float *g, float *m, float *v, size_t count, float lr,
                                                  float b1, float b2, float eps, float bc1, float bc2, uint16_t *p16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        p[i] = pi;
        g[i] = 0.0f;
        if (p16) p16[i] = f32_to_bf16(pi);
    }
}

__global__ __launch_bounds__(256) void k_nrc_to_bf16(const float *p, uint16_t *p16, size_t count) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < count) p16[i] = f32_to_bf16(p[i]);
}

__global__ __launch_bounds__(256) void k_nrc_grid_init(float *grid, size_t count, uint32_t seed) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 0x9e3779b9u ^ seed;
        h ^= h >> 16;
        h *= 0x7feb352du;
        h ^= h >> 15;
        h *= 0x846ca68bu;
        h ^= h >> 16;
        grid[i] = ((float)(h >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * 1e-4f;
    }
}

inline uint32_t blocks_of(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

#define NRC_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) return std::string(#expr) + ": " + hipGetErrorString(_e);     \
    } while (0)

NrcModel::~NrcModel() {
    for (void *p : {w32, w16, wg, wm, wv, grid, gg, gm, gv, acts, ga, gb, gin, dout, loss, in, tgt, pred})
        if (p) (void)hipFree(p);
}

std::string NrcModel::alloc(void **p, size_t bytes) {
    NRC_TRY(hipMalloc(p, std::max<size_t>(bytes, 16)));
    NRC_TRY(hipMemset(*p, 0, std::max<size_t>(bytes, 16)));
    return "";
}

std::string NrcModel::init(hipStream_t s, uint32_t seed) {
    stream = s;
    std::string e;
    if (!w32) {
        const size_t G = kNrcGridParams;
        const size_t bytes_mlp = kNrcMlpParams * sizeof(float);
        for (void **p : {&w32, &wg, &wm, &wv})
            if (!(e = alloc(p, bytes_mlp)).empty()) return e;
        if (!(e = alloc(&w16, kNrcMlpParams * sizeof(uint16_t))).empty()) return e;
        for (void **p : {&grid, &gg, &gm, &gv})
            if (!(e = alloc(p, G * sizeof(float))).empty()) return e;
        if (!(e = alloc(&acts, (size_t)kNrcLayers * kMaxBatch * kNrcWidth * sizeof(uint16_t))).empty()) return e;
        for (void **p : {&ga, &gb})
            if (!(e = alloc(p, (size_t)kMaxBatch * kNrcWidth * sizeof(float))).empty()) return e;
        if (!(e = alloc(&gin, (size_t)kMaxBatch * 32 * sizeof(float))).empty()) return e;
        if (!(e = alloc(&dout, (size_t)kMaxBatch * kNrcOut * sizeof(float))).empty()) return e;
        if (!(e = alloc(&loss, sizeof(float))).empty()) return e;
        if (!(e = alloc(&in, (size_t)kMaxBatch * kNrcIn * sizeof(float))).empty()) return e;
        if (!(e = alloc(&tgt, (size_t)kMaxBatch * 3 * sizeof(float))).empty()) return e;
        if (!(e = alloc(&pred, (size_t)kMaxBatch * 3 * sizeof(float))).empty()) return e;
    }
    /* Xavier-uniform weights (host RNG, deterministic), grid entries uniform in +-1e-4 */
    std::vector<float> w(kNrcMlpParams);
    std::mt19937 rng(seed);
    for (int l = 0; l < kNrcLayers; ++l) {
        const int nout = nrc_layer_out(l);
        const int used_out = l < kNrcHidden ? kNrcWidth : 3;
        const float lim = std::sqrt(6.0f / (float)(kNrcWidth + used_out));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int j = 0; j < kNrcWidth; ++j)
            for (int k = 0; k < nout; ++k) w[nrc_layer_offset(l) + j * nout + k] = k < used_out ? U(rng) : 0.0f;
    }
    NRC_TRY(hipMemcpyAsync(w32, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    for (void *p : {wg, wm, wv}) NRC_TRY(hipMemsetAsync(p, 0, kNrcMlpParams * sizeof(float), stream));
    for (void *p : {gg, gm, gv}) NRC_TRY(hipMemsetAsync(p, 0, kNrcGridParams * sizeof(float), stream));
    hipLaunchKernelGGL(k_nrc_grid_init, dim3(4096), dim3(256), 0, stream, (float *)grid, kNrcGridParams, seed);
    hipLaunchKernelGGL(k_nrc_to_bf16, dim3(blocks_of(kNrcMlpParams, 256)), dim3(256), 0, stream, (const float *)w32,
                       (uint16_t *)w16, (size_t)kNrcMlpParams);
    NRC_TRY(hipGetLastError());
    NRC_TRY(hipStreamSynchronize(stream));
    step = 0;
    nrc_resolutions(res);
    return "";
}

void NrcModel::set_bounds(const float lo_[3], const float hi_[3]) {
    for (int a = 0; a < 3; ++a) {
        const float ext = hi_[a] - lo_[a];
        lo[a] = lo_[a];
        inv_ext[a] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
}

NrcView NrcModel::view() const {
    NrcView V;
    V.w = (const uint16_t *)w16;
    V.grid = (const float *)grid;
    for (int l = 0; l < kNrcLevels; ++l) V.res[l] = res[l];
    for (int a = 0; a < 3; ++a) {
        V.lo[a] = lo[a];
        V.inv_ext[a] = inv_ext[a];
    }
    return V;
}

std::string NrcModel::infer(const float *d_in, uint32_t n, float *d_pred) {
    if (n == 0) return "";
    hipLaunchKernelGGL((k_nrc_forward<false, kInferWaves>), dim3(blocks_of(n, kInferWaves * kTile)),
                       dim3(kInferWaves * 64), 0, stream, view(), d_in, n, (const uint32_t *)nullptr, d_pred, nullptr,
                       0u, nullptr, nullptr, nullptr);
    NRC_TRY(hipGetLastError());
    return "";
}

std::string NrcModel::infer_dev(const float *d_in, const uint32_t *d_n, uint32_t cap, float *d_pred, uint32_t blocks) {
    if (cap == 0) return "";
    hipLaunchKernelGGL((k_nrc_forward<false, kInferWaves>),
                       dim3(std::max(1u, std::min(blocks, blocks_of(cap, kInferWaves * kTile)))), dim3(kInferWaves * 64),
                       0, stream, view(), d_in, cap, d_n, d_pred, nullptr, 0u, nullptr, nullptr, nullptr);
    NRC_TRY(hipGetLastError());
    return "";
}

std::string NrcModel::gradients(const float *d_in, const float *d_tgt, uint32_t n, float *loss_out) {
    if (n == 0 || n > kMaxBatch) return "batch size out of range";
    const uint32_t npad = blocks_of(n, kBlockSamples) * kBlockSamples;
    NRC_TRY(hipMemsetAsync(loss, 0, sizeof(float), stream));
    hipLaunchKernelGGL((k_nrc_forward<true, kWaves>), dim3(npad / kBlockSamples), dim3(kWaves * 64), 0, stream, view(), d_in, n,
                       (const uint32_t *)nullptr, (float *)pred, (uint16_t *)acts, npad, d_tgt, (float *)dout,
                       (float *)loss);
    const float *g = (const float *)dout;
    float *bufs[2] = {(float *)ga, (float *)gb};
    for (int l = kNrcLayers - 1; l >= 0; --l) {
        float *gnext = l > 0 ? bufs[l & 1] : (float *)gin;
        if (l == kNrcLayers - 1) {
            hipLaunchKernelGGL((k_nrc_bwd_dw<kNrcOut>), dim3(blocks_of(npad, kWaves * kDwSamples)), dim3(kWaves * 64), 0,
                               stream, l, g, (const uint16_t *)acts, npad, (float *)wg);
            hipLaunchKernelGGL((k_nrc_bwd_dx<kNrcOut>), dim3(npad / kBlockSamples), dim3(kWaves * 64), 0, stream, l,
                               g, (const uint16_t *)w16, (const uint16_t *)acts, npad, gnext);
        } else {
            hipLaunchKernelGGL((k_nrc_bwd_dw<kNrcWidth>), dim3(blocks_of(npad, kWaves * kDwSamples)), dim3(kWaves * 64), 0,
                               stream, l, g, (const uint16_t *)acts, npad, (float *)wg);
            hipLaunchKernelGGL((k_nrc_bwd_dx<kNrcWidth>), dim3(npad / kBlockSamples), dim3(kWaves * 64), 0, stream,
                               l, g, (const uint16_t *)w16, (const uint16_t *)acts, npad, gnext);
        }
        g = gnext;
    }
    hipLaunchKernelGGL(k_nrc_bwd_grid, dim3(blocks_of((size_t)n * kNrcLevels, 256)), dim3(256), 0, stream, view(), d_in,
                       (const float *)gin, n, (float *)gg);
    NRC_TRY(hipGetLastError());
    if (loss_out) {
        NRC_TRY(hipMemcpyAsync(loss_out, loss, sizeof(float), hipMemcpyDeviceToHost, stream));
        NRC_TRY(hipStreamSynchronize(stream));
    }
    return "";
}

std::string NrcModel::train_step(const float *d_in, const float *d_tgt, uint32_t n, float *loss_out) {
    std::string e = gradients(d_in, d_tgt, n, loss_out);
    if (!e.empty()) return e;
    ++step;
    const float b1 = 0.9f, b2 = 0.99f;
    const float bc1 = 1.0f - std::pow(b1, (float)step), bc2 = 1.0f - std::pow(b2, (float)step);
    const float lr = lr0 * std::pow(lr_decay, (float)step);
    hipLaunchKernelGGL(k_nrc_adam, dim3(blocks_of(kNrcMlpParams, 256)), dim3(256), 0, stream, (float *)w32, (float *)wg,
                       (float *)wm, (float *)wv, (size_t)kNrcMlpParams, lr, b1, b2, 1e-10f, bc1, bc2, (uint16_t *)w16);
    hipLaunchKernelGGL(k_nrc_adam, dim3(8192), dim3(256), 0, stream, (float *)grid, (float *)gg, (float *)gm,
                       (float *)gv, kNrcGridParams, lr, b1, b2, 1e-10f, bc1, bc2, (uint16_t *)nullptr);
    NRC_TRY(hipGetLastError());
    return "";
}

}  // namespace vvpt
