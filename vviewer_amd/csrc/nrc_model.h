/*
 * nrc_model.h - host-side state of the neural radiance cache (nrc.h): device
 * parameters with their Adam moments and gradients, staging for training
 * batches, and the entry points that launch the kernels of nrc.hip.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "nrc.h"

namespace vvpt {

/* Inference blocks per compute unit for NrcModel::infer_dev (4-wave blocks). */
constexpr uint32_t kNrcInferBlocksPerCU = 4;

struct NrcModel {
    static constexpr uint32_t kMaxBatch = 65536; /* samples per training step (a multiple of 64) */
    hipStream_t stream = nullptr;
    /* MLP: fp32 masters, bf16 copies, gradients and Adam moments */
    void *w32 = nullptr, *w16 = nullptr, *wg = nullptr, *wm = nullptr, *wv = nullptr;
    /* hash grid: fp32 entries, gradients and Adam moments */
    void *grid = nullptr, *gg = nullptr, *gm = nullptr, *gv = nullptr;
    /* training scratch: layer inputs (bf16), layer gradients, grid-input gradients, output gradients, loss */
    void *acts = nullptr, *ga = nullptr, *gb = nullptr, *gin = nullptr, *dout = nullptr, *loss = nullptr;
    /* staging for host batches */
    void *in = nullptr, *tgt = nullptr, *pred = nullptr;
    float res[kNrcLevels] = {};
    float lo[3] = {0, 0, 0}, inv_ext[3] = {1, 1, 1};
    uint32_t step = 0;
    float lr0 = 1e-2f;       /* Adam learning rate */
    float lr_decay = 0.999f; /* exponential decay per step */

    ~NrcModel();
    std::string init(hipStream_t s, uint32_t seed);
    void set_bounds(const float lo[3], const float hi[3]);
    NrcView view() const;
    /* predictions for n device inputs (kNrcIn floats each) into d_pred (3 floats each) */
    std::string infer(const float *d_in, uint32_t n, float *d_pred);
    /* the same for a count *d_n (<= cap) known only on the device, over at most `blocks` persistent blocks */
    std::string infer_dev(const float *d_in, const uint32_t *d_n, uint32_t cap, float *d_pred, uint32_t blocks);
    /* forward + backward of one batch; gradients accumulate into wg / gg; optional loss to the host */
    std::string gradients(const float *d_in, const float *d_tgt, uint32_t n, float *loss_out);
    /* gradients + one Adam step */
    std::string train_step(const float *d_in, const float *d_tgt, uint32_t n, float *loss_out);

  private:
    std::string alloc(void **p, size_t bytes);
};

}  // namespace vvpt
