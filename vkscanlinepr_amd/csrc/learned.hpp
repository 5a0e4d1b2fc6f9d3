// learned.hpp -- the per-path count predictor of the learned single-pass allocation: model
// parameters, their BF16 device copy for inference, FP32 training on the GPU, save and load.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "kernels.hpp"

namespace spr {

struct MlpModel {
    std::vector<float> params;        // kMlpParams, FP32 master copy (layout in kernels.hpp)
    std::vector<float> mean, inv_std; // kMlpIn each
    float log_margin = 0.5f;          // capacity = 2^(prediction + log_margin) - 1 + slack
    int32_t slack = 8;
    bool ready = false;
};

// Offsets of the parameter blocks in the flat vector.
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpIn * kMlpHidden, kOffW2 = kOffB1 + kMlpHidden,
              kOffB2 = kOffW2 + kMlpHidden * kMlpHidden, kOffW3 = kOffB2 + kMlpHidden,
              kOffB3 = kOffW3 + kMlpHidden * kMlpOut;

struct TrainOptions {
    int epochs = 60;
    int batch = 512;
    int min_steps = 4000;             // epochs are added until at least this many Adam steps ran
    float lr = 3e-3f;
    uint64_t seed = 4;
    float val_frac = 0.1f;
    float overflow_quantile = 0.95f;  // the margin keeps this fraction of validation paths in their arena
};

struct TrainReport {
    int samples = 0, steps = 0;
    float train_loss = 0, val_loss = 0, first_loss = 0;
    float log_margin = 0;
    double ms = 0;
};

// Device copy of the model for k_mlp_infer (BF16 weights, FP32 biases and normalisation).
class MlpDevice {
public:
    ~MlpDevice();
    void upload(const MlpModel& m, hipStream_t s);
    const MlpDev& dev() const { return dev_; }
    bool valid() const { return valid_; }
    void release();

private:
    void* blob_ = nullptr;
    size_t bytes_ = 0;
    MlpDev dev_{};
    bool valid_ = false;
};

// Fits the model to n samples (raw features d_feat[n][16], targets d_tgt[n][2] in log2(1+count))
// with the FP32 training kernels; normalisation statistics come from the samples.
void mlp_train(MlpModel& m, const float* d_feat, const float* d_tgt, int n, const TrainOptions& o,
               TrainReport& rep, hipStream_t s);
// One forward/backward pass on the given batch (raw features, host arrays) with the model's
// current parameters: gradients of the mean squared error (flat layout) and the loss.
void mlp_gradients(const MlpModel& m, const float* X, const float* T, int B, float* grads, float* loss,
                   hipStream_t s);
void mlp_init(MlpModel& m, uint64_t seed);
void mlp_save(const MlpModel& m, const std::string& path);
void mlp_load(MlpModel& m, const std::string& path);
uint16_t f32_to_bf16(float f);

}  // namespace spr
