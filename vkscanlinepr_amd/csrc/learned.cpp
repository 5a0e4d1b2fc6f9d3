// learned.cpp -- training, device upload, save and load of the per-path count predictor.
#include "learned.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>

namespace spr {

namespace {

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct DevVec {
    T* p = nullptr;
    size_t n = 0;
    explicit DevVec(size_t count) : n(count) {
        hip_ok(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
    }
    ~DevVec() { if (p) (void)hipFree(p); }
    DevVec(const DevVec&) = delete;
    DevVec& operator=(const DevVec&) = delete;
};

// Forward + backward of one batch already on the device (normalised inputs X[B][16]).
struct Work {
    DevVec<float> h1, h2, y, dy, sq, dh2, dh1;
    explicit Work(int B)
        : h1((size_t)B * kMlpHidden), h2((size_t)B * kMlpHidden), y((size_t)B * kMlpOut),
          dy((size_t)B * kMlpOut), sq((size_t)B * kMlpOut), dh2((size_t)B * kMlpHidden),
          dh1((size_t)B * kMlpHidden) {}
};

void forward(const float* P, const float* X, int B, Work& w, hipStream_t s) {
    hip_ok(launch_dense_fwd(X, B, kMlpIn, P + kOffW1, P + kOffB1, kMlpHidden, 1, w.h1.p, s), "fwd1");
    hip_ok(launch_dense_fwd(w.h1.p, B, kMlpHidden, P + kOffW2, P + kOffB2, kMlpHidden, 1, w.h2.p, s), "fwd2");
    hip_ok(launch_dense_fwd(w.h2.p, B, kMlpHidden, P + kOffW3, P + kOffB3, kMlpOut, 0, w.y.p, s), "fwd3");
}

void backward(const float* P, const float* X, const float* T, int B, Work& w, float* G, hipStream_t s) {
    hip_ok(launch_loss_grad(w.y.p, T, B, kMlpOut, w.dy.p, w.sq.p, s), "loss");
    hip_ok(launch_dense_bwd_weight(w.h2.p, w.dy.p, B, kMlpHidden, kMlpOut, G + kOffW3, G + kOffB3, s), "bw3");
    hip_ok(launch_dense_bwd_data(w.dy.p, B, kMlpOut, P + kOffW3, kMlpHidden, w.h2.p, w.dh2.p, s), "bd3");
    hip_ok(launch_dense_bwd_weight(w.h1.p, w.dh2.p, B, kMlpHidden, kMlpHidden, G + kOffW2, G + kOffB2, s), "bw2");
    hip_ok(launch_dense_bwd_data(w.dh2.p, B, kMlpHidden, P + kOffW2, kMlpHidden, w.h1.p, w.dh1.p, s), "bd2");
    hip_ok(launch_dense_bwd_weight(X, w.dh1.p, B, kMlpIn, kMlpHidden, G + kOffW1, G + kOffB1, s), "bw1");
}

float mean_sq(const Work& w, int B, hipStream_t s) {
    std::vector<float> h((size_t)B * kMlpOut);
    hip_ok(hipMemcpyAsync(h.data(), w.sq.p, h.size() * 4, hipMemcpyDeviceToHost, s), "d2h");
    hip_ok(hipStreamSynchronize(s), "sync");
    double a = 0;
    for (float v : h) a += v;
    return (float)(a / std::max<size_t>(h.size(), 1));
}

}  // namespace

uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

void mlp_init(MlpModel& m, uint64_t seed) {
    m.params.assign(kMlpParams, 0.0f);
    std::mt19937_64 g(seed);
    auto fill = [&](int off, int fan_in, int count) {
        std::normal_distribution<float> d(0.0f, std::sqrt(2.0f / (float)fan_in));
        for (int i = 0; i < count; ++i) m.params[off + i] = d(g);
    };
    fill(kOffW1, kMlpIn, kMlpIn * kMlpHidden);
    fill(kOffW2, kMlpHidden, kMlpHidden * kMlpHidden);
    fill(kOffW3, kMlpHidden, kMlpHidden * kMlpOut);
    if (m.mean.size() != (size_t)kMlpIn) m.mean.assign(kMlpIn, 0.0f);
    if (m.inv_std.size() != (size_t)kMlpIn) m.inv_std.assign(kMlpIn, 1.0f);
}

MlpDevice::~MlpDevice() { release(); }

void MlpDevice::release() {
    if (blob_) { (void)hipFree(blob_); (void)hipGetLastError(); }
    blob_ = nullptr;
    bytes_ = 0;
    valid_ = false;
}

void MlpDevice::upload(const MlpModel& m, hipStream_t s) {
    if (!m.ready || m.params.size() != (size_t)kMlpParams) throw std::runtime_error("model not trained or loaded");
    // layout: w1 bf16[16*64] | w2 bf16[64*64] | w3 bf16[64*16] | b1[64] b2[64] b3[2 -> 4] mean[16] inv_std[16]
    const size_t n16 = (size_t)kMlpIn * kMlpHidden + (size_t)kMlpHidden * kMlpHidden + (size_t)kMlpHidden * 16;
    const size_t nf = kMlpHidden * 2 + 4 + kMlpIn * 2;
    std::vector<uint8_t> host(n16 * 2 + nf * 4);
    uint16_t* w = reinterpret_cast<uint16_t*>(host.data());
    for (int i = 0; i < kMlpIn * kMlpHidden; ++i) w[i] = f32_to_bf16(m.params[kOffW1 + i]);
    w += kMlpIn * kMlpHidden;
    for (int i = 0; i < kMlpHidden * kMlpHidden; ++i) w[i] = f32_to_bf16(m.params[kOffW2 + i]);
    w += kMlpHidden * kMlpHidden;
    for (int k = 0; k < kMlpHidden; ++k)
        for (int j = 0; j < 16; ++j) w[k * 16 + j] = j < kMlpOut ? f32_to_bf16(m.params[kOffW3 + k * kMlpOut + j]) : 0;
    float* f = reinterpret_cast<float*>(host.data() + n16 * 2);
    std::memcpy(f, &m.params[kOffB1], kMlpHidden * 4);
    std::memcpy(f + kMlpHidden, &m.params[kOffB2], kMlpHidden * 4);
    f[2 * kMlpHidden] = m.params[kOffB3];
    f[2 * kMlpHidden + 1] = m.params[kOffB3 + 1];
    f[2 * kMlpHidden + 2] = f[2 * kMlpHidden + 3] = 0.0f;
    std::memcpy(f + 2 * kMlpHidden + 4, m.mean.data(), kMlpIn * 4);
    std::memcpy(f + 2 * kMlpHidden + 4 + kMlpIn, m.inv_std.data(), kMlpIn * 4);
    if (bytes_ < host.size()) {
        release();
        hip_ok(hipMalloc(&blob_, host.size()), "hipMalloc");
        bytes_ = host.size();
    }
    hip_ok(hipMemcpyAsync(blob_, host.data(), host.size(), hipMemcpyHostToDevice, s), "h2d");
    hip_ok(hipStreamSynchronize(s), "sync");
    const uint16_t* d16 = reinterpret_cast<const uint16_t*>(blob_);
    const float* df = reinterpret_cast<const float*>(reinterpret_cast<const uint8_t*>(blob_) + n16 * 2);
    dev_.w1 = d16;
    dev_.w2 = d16 + kMlpIn * kMlpHidden;
    dev_.w3 = d16 + kMlpIn * kMlpHidden + kMlpHidden * kMlpHidden;
    dev_.b1 = df;
    dev_.b2 = df + kMlpHidden;
    dev_.b3 = df + 2 * kMlpHidden;
    dev_.mean = df + 2 * kMlpHidden + 4;
    dev_.inv_std = df + 2 * kMlpHidden + 4 + kMlpIn;
    valid_ = true;
}

void mlp_train(MlpModel& m, const float* d_feat, const float* d_tgt, int n, const TrainOptions& o,
               TrainReport& rep, hipStream_t s) {
    if (n < 16) throw std::runtime_error("too few training samples");
    const auto t0 = std::chrono::steady_clock::now();
    // normalisation from the samples
    std::vector<float> X((size_t)n * kMlpIn), T((size_t)n * kMlpOut);
    hip_ok(hipMemcpyAsync(X.data(), d_feat, X.size() * 4, hipMemcpyDeviceToHost, s), "d2h");
    hip_ok(hipMemcpyAsync(T.data(), d_tgt, T.size() * 4, hipMemcpyDeviceToHost, s), "d2h");
    hip_ok(hipStreamSynchronize(s), "sync");
    m.mean.assign(kMlpIn, 0.0f);
    m.inv_std.assign(kMlpIn, 1.0f);
    for (int k = 0; k < kMlpIn; ++k) {
        double a = 0, b = 0;
        for (int i = 0; i < n; ++i) a += X[(size_t)i * kMlpIn + k];
        a /= n;
        for (int i = 0; i < n; ++i) { const double d = X[(size_t)i * kMlpIn + k] - a; b += d * d; }
        m.mean[k] = (float)a;
        m.inv_std[k] = (float)(1.0 / std::max(std::sqrt(b / n), 1e-3));
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < kMlpIn; ++k)
            X[(size_t)i * kMlpIn + k] = (X[(size_t)i * kMlpIn + k] - m.mean[k]) * m.inv_std[k];
    // train / validation split
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 g(o.seed);
    std::shuffle(perm.begin(), perm.end(), g);
    const int n_val = std::max(1, std::min(n - 1, (int)(n * o.val_frac)));
    const int n_tr = n - n_val;
    DevVec<float> dX((size_t)n * kMlpIn), dT((size_t)n * kMlpOut);
    hip_ok(hipMemcpyAsync(dX.p, X.data(), X.size() * 4, hipMemcpyHostToDevice, s), "h2d");
    hip_ok(hipMemcpyAsync(dT.p, T.data(), T.size() * 4, hipMemcpyHostToDevice, s), "h2d");
    DevVec<int32_t> dperm(n);
    DevVec<float> P(kMlpParams), G(kMlpParams), M(kMlpParams), V(kMlpParams);
    if (!m.ready || m.params.size() != (size_t)kMlpParams) mlp_init(m, o.seed);
    hip_ok(hipMemcpyAsync(P.p, m.params.data(), kMlpParams * 4, hipMemcpyHostToDevice, s), "h2d");
    hip_ok(hipMemsetAsync(M.p, 0, kMlpParams * 4, s), "memset");
    hip_ok(hipMemsetAsync(V.p, 0, kMlpParams * 4, s), "memset");
    const int B = std::max(16, std::min(o.batch, n_tr));
    DevVec<float> xb((size_t)B * kMlpIn), tb((size_t)B * kMlpOut);
    Work w(std::max(B, n_val));
    DevVec<float> xv((size_t)n_val * kMlpIn), tv((size_t)n_val * kMlpOut);
    hip_ok(hipMemcpyAsync(dperm.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s), "h2d");
    hip_ok(launch_gather_rows(dX.p, dperm.p + n_tr, n_val, kMlpIn, xv.p, s), "gather");
    hip_ok(launch_gather_rows(dT.p, dperm.p + n_tr, n_val, kMlpOut, tv.p, s), "gather");
    std::vector<int32_t> order(perm.begin(), perm.begin() + n_tr);
    int step = 0;
    const float b1 = 0.9f, b2 = 0.999f;
    const int per_epoch = std::max(1, n_tr / B);
    const int epochs = std::max(o.epochs, (o.min_steps + per_epoch - 1) / per_epoch);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(order.begin(), order.end(), g);
        hip_ok(hipMemcpyAsync(dperm.p, order.data(), (size_t)n_tr * 4, hipMemcpyHostToDevice, s), "h2d");
        const float lr = o.lr * (e < epochs * 3 / 4 ? 1.0f : 0.2f);
        for (int b0 = 0; b0 + B <= n_tr; b0 += B) {
            hip_ok(launch_gather_rows(dX.p, dperm.p + b0, B, kMlpIn, xb.p, s), "gather");
            hip_ok(launch_gather_rows(dT.p, dperm.p + b0, B, kMlpOut, tb.p, s), "gather");
            forward(P.p, xb.p, B, w, s);
            backward(P.p, xb.p, tb.p, B, w, G.p, s);
            ++step;
            const float bc1 = 1.0f - std::pow(b1, (float)step), bc2 = 1.0f - std::pow(b2, (float)step);
            hip_ok(launch_adam(P.p, G.p, M.p, V.p, kMlpParams, lr, b1, b2, 1e-8f, bc1, bc2, s), "adam");
            if (step == 1) rep.first_loss = mean_sq(w, B, s);
        }
        // the order buffer is rewritten next epoch: wait for this epoch's gathers
        hip_ok(hipStreamSynchronize(s), "sync");
    }
    forward(P.p, xb.p, B, w, s);
    hip_ok(launch_loss_grad(w.y.p, tb.p, B, kMlpOut, w.dy.p, w.sq.p, s), "loss");
    rep.train_loss = mean_sq(w, B, s);
    // validation loss and the capacity margin from the validation residuals
    forward(P.p, xv.p, n_val, w, s);
    hip_ok(launch_loss_grad(w.y.p, tv.p, n_val, kMlpOut, w.dy.p, w.sq.p, s), "loss");
    rep.val_loss = mean_sq(w, n_val, s);
    std::vector<float> yv((size_t)n_val * kMlpOut), tvh((size_t)n_val * kMlpOut);
    hip_ok(hipMemcpyAsync(yv.data(), w.y.p, yv.size() * 4, hipMemcpyDeviceToHost, s), "d2h");
    hip_ok(hipMemcpyAsync(tvh.data(), tv.p, tvh.size() * 4, hipMemcpyDeviceToHost, s), "d2h");
    m.params.resize(kMlpParams);
    hip_ok(hipMemcpyAsync(m.params.data(), P.p, kMlpParams * 4, hipMemcpyDeviceToHost, s), "d2h");
    hip_ok(hipStreamSynchronize(s), "sync");
    std::vector<float> res;
    for (int i = 0; i < n_val; ++i)
        if (tvh[(size_t)i * 2] > 0.0f) res.push_back(tvh[(size_t)i * 2] - yv[(size_t)i * 2]);
    float margin = 0.5f;
    if (!res.empty()) {
        std::sort(res.begin(), res.end());
        const size_t q = std::min(res.size() - 1, (size_t)(o.overflow_quantile * (double)res.size()));
        margin = std::max(0.0f, res[q]);
    }
    m.log_margin = margin;
    m.ready = true;
    rep.samples = n;
    rep.steps = step;
    rep.log_margin = margin;
    rep.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void mlp_gradients(const MlpModel& m, const float* X, const float* T, int B, float* grads, float* loss,
                   hipStream_t s) {
    if (B < 1) throw std::runtime_error("empty batch");
    if (m.params.size() != (size_t)kMlpParams) throw std::runtime_error("model has no parameters");
    std::vector<float> xn((size_t)B * kMlpIn);
    for (int i = 0; i < B; ++i)
        for (int k = 0; k < kMlpIn; ++k) {
            const float mu = m.mean.empty() ? 0.0f : m.mean[k], is = m.inv_std.empty() ? 1.0f : m.inv_std[k];
            xn[(size_t)i * kMlpIn + k] = (X[(size_t)i * kMlpIn + k] - mu) * is;
        }
    DevVec<float> dX(xn.size()), dT((size_t)B * kMlpOut), P(kMlpParams), G(kMlpParams);
    Work w(B);
    hip_ok(hipMemcpyAsync(dX.p, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, s), "h2d");
    hip_ok(hipMemcpyAsync(dT.p, T, (size_t)B * kMlpOut * 4, hipMemcpyHostToDevice, s), "h2d");
    hip_ok(hipMemcpyAsync(P.p, m.params.data(), kMlpParams * 4, hipMemcpyHostToDevice, s), "h2d");
    forward(P.p, dX.p, B, w, s);
    backward(P.p, dX.p, dT.p, B, w, G.p, s);
    hip_ok(hipMemcpyAsync(grads, G.p, kMlpParams * 4, hipMemcpyDeviceToHost, s), "d2h");
    *loss = mean_sq(w, B, s);
}

void mlp_save(const MlpModel& m, const std::string& path) {
    if (!m.ready) throw std::runtime_error("model not trained or loaded");
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write \"" + path + "\"");
    const char magic[8] = {'S', 'P', 'R', 'M', 'L', 'P', '1', 0};
    const uint32_t hdr[4] = {(uint32_t)kMlpIn, (uint32_t)kMlpHidden, (uint32_t)kMlpOut, (uint32_t)kMlpParams};
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 4, f) == 4 &&
              std::fwrite(m.params.data(), 4, kMlpParams, f) == (size_t)kMlpParams &&
              std::fwrite(m.mean.data(), 4, kMlpIn, f) == (size_t)kMlpIn &&
              std::fwrite(m.inv_std.data(), 4, kMlpIn, f) == (size_t)kMlpIn &&
              std::fwrite(&m.log_margin, 4, 1, f) == 1 && std::fwrite(&m.slack, 4, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw std::runtime_error("write failed: \"" + path + "\"");
}

void mlp_load(MlpModel& m, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open \"" + path + "\"");
    char magic[8];
    uint32_t hdr[4];
    MlpModel t;
    t.params.resize(kMlpParams);
    t.mean.resize(kMlpIn);
    t.inv_std.resize(kMlpIn);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "SPRMLP1", 8) == 0 &&
              std::fread(hdr, 4, 4, f) == 4 && hdr[0] == (uint32_t)kMlpIn && hdr[1] == (uint32_t)kMlpHidden &&
              hdr[2] == (uint32_t)kMlpOut && hdr[3] == (uint32_t)kMlpParams &&
              std::fread(t.params.data(), 4, kMlpParams, f) == (size_t)kMlpParams &&
              std::fread(t.mean.data(), 4, kMlpIn, f) == (size_t)kMlpIn &&
              std::fread(t.inv_std.data(), 4, kMlpIn, f) == (size_t)kMlpIn &&
              std::fread(&t.log_margin, 4, 1, f) == 1 && std::fread(&t.slack, 4, 1, f) == 1;
    std::fclose(f);
    if (!ok) throw std::runtime_error("not a model file of this layout: \"" + path + "\"");
    t.ready = true;
    m = t;
}

}  // namespace spr
