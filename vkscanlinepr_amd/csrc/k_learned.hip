// k_learned.hip -- learned single-pass allocation (gfx950).
//
// The plain frame counts every curve's intersections, scans the counts and waits for the host
// to read the total before anything is written (make_intersection_0.comp -> naive_scan.comp ->
// host readback, scanline_rasterizer.cpp:317-395).  The learned frame sizes its buffers before
// the count instead, from predicted per-path counts, and never waits for the host:
//   1. k_path_features: 16 cheap features per path from its transformed control points (and a
//      proven upper bound of the path's intersections);
//   2. k_mlp_infer: a 16-64-64-2 MLP (BF16 weights, FP32 accumulation on the matrix cores)
//      predicts log2(1 + intersections) and log2(1 + records) per path;
//   3. k_caps: predicted capacities with a safety margin; their sum (one device scan) is the
//      allocation the host sizes the frame's buffers by;
//   4. k_path_overflow: the paths whose exact count (the count pass still runs, on the device)
//      exceeded their capacity.
// The entries keep the plain frame's layout (exact offsets from the device scan of the counts);
// a frame whose exact total exceeds its budget is culled on the device and rendered with the plain
// method, so the output never changes.
//
// The training kernels (k_dense_*, k_loss_grad, k_adam) are plain FP32 forward/backward passes
// of the same network, used to fit it on counts produced by the plain method.
#include "device_common.hpp"
#include "kernels.hpp"
#include "scan_common.hpp"

namespace spr {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kF = kMlpIn, kH = kMlpHidden;

__device__ inline short f2bf(float f) {  // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ inline float wave_sum_f(float v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline float wave_min_f(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_max_f(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// One wave per path; curves of path p are [pcurve[p], pcurve[p+1]).
__global__ __launch_bounds__(256) void k_path_features(const float2* __restrict__ tpos,
                                                       const uint32_t* __restrict__ cpos,
                                                       const uint32_t* __restrict__ ctype,
                                                       const uint32_t* __restrict__ pcurve,
                                                       const int32_t* __restrict__ visible,
                                                       const uint32_t* __restrict__ rule, uint32_t Np,
                                                       int w, int h, float* __restrict__ feat,
                                                       unsigned long long* __restrict__ bound_total,
                                                       const int2* __restrict__ pcanvas,
                                                       int32_t* __restrict__ pbound,
                                                       const int32_t* __restrict__ skip) {
    if (skip && *skip) return;  // the prediction of an earlier frame still holds
    __shared__ unsigned long long s_bound[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int w0 = w, h0 = h;
    unsigned long long bsum = 0;  // lane 0: this wave's share of the bound total
    // four paths per wave, one atomic per workgroup for the bound total
    for (int j = 0; j < 4; ++j) {
    const uint32_t p = (blockIdx.x * 4 + wave) * 4 + j;
    if (p >= Np) break;
    w = w0; h = h0;
    if (pcanvas) { w = pcanvas[p].x; h = pcanvas[p].y; }
    const uint32_t c0 = pcurve[p], c1 = pcurve[p + 1];
    float n_cub = 0, xmin = 3.4e38f, xmax = -3.4e38f, ymin = 3.4e38f, ymax = -3.4e38f;
    float plen = 0, sdx = 0, sdy = 0, cdx = 0, cdy = 0, turn = 0, inside = 0, npts = 0;
    for (uint32_t c = c0 + lane; c < c1; c += 64) {
        const uint32_t k = ctype[c] & 7u;  // points: line 2, quadratic and arc 3, cubic 4
        const uint32_t b = cpos[c];
        float2 q[4];  // static indices only (the line's points repeat its end point)
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) q[i] = tpos[b + (i < k ? i : k - 1)];
        n_cub += k > 2u ? 1.0f : 0.0f;
        float pdx = 0, pdy = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            if (i >= k) break;
            xmin = fminf(xmin, q[i].x); xmax = fmaxf(xmax, q[i].x);
            ymin = fminf(ymin, q[i].y); ymax = fmaxf(ymax, q[i].y);
            inside += (q[i].x >= 0 && q[i].x < (float)w && q[i].y >= 0 && q[i].y < (float)h) ? 1.0f : 0.0f;
            npts += 1.0f;
            if (i > 0) {
                const float dx = q[i].x - q[i - 1].x, dy = q[i].y - q[i - 1].y;
                plen += sqrtf(dx * dx + dy * dy);
                sdx += fabsf(dx); sdy += fabsf(dy);
                if (i > 1) turn += fabsf(atan2f(pdx * dy - pdy * dx, pdx * dx + pdy * dy));
                pdx = dx; pdy = dy;
            }
        }
        const float2 qe = k == 4u ? q[3] : (k == 3u ? q[2] : q[1]);
        cdx += fabsf(qe.x - q[0].x);
        cdy += fabsf(qe.y - q[0].y);
    }
    n_cub = wave_sum_f(n_cub); plen = wave_sum_f(plen); sdx = wave_sum_f(sdx); sdy = wave_sum_f(sdy);
    cdx = wave_sum_f(cdx); cdy = wave_sum_f(cdy); turn = wave_sum_f(turn);
    inside = wave_sum_f(inside); npts = wave_sum_f(npts);
    xmin = wave_min_f(xmin); xmax = wave_max_f(xmax); ymin = wave_min_f(ymin); ymax = wave_max_f(ymax);
    if (lane == 0) {
        const float bw = npts > 0 ? xmax - xmin : 0.0f, bh = npts > 0 ? ymax - ymin : 0.0f;
        const float cw = npts > 0 ? fmaxf(0.0f, fminf(xmax, (float)w) - fmaxf(xmin, 0.0f)) : 0.0f;
        const float ch = npts > 0 ? fmaxf(0.0f, fminf(ymax, (float)h) - fmaxf(ymin, 0.0f)) : 0.0f;
        float* f = feat + (size_t)p * kF;
        f[0] = log2f(1.0f + (float)(c1 - c0));
        f[1] = log2f(1.0f + n_cub);
        f[2] = log2f(1.0f + bw);
        f[3] = log2f(1.0f + bh);
        f[4] = log2f(1.0f + plen);
        f[5] = log2f(1.0f + 0.5f * sdx);
        f[6] = log2f(1.0f + 0.5f * sdy);
        f[7] = log2f(1.0f + 0.5f * cdx);
        f[8] = log2f(1.0f + 0.5f * cdy);
        f[9] = log2f(1.0f + turn * 0.15915494f);
        f[10] = log2f(1.0f + cw);
        f[11] = log2f(1.0f + ch);
        f[12] = visible_mask(visible[p]) ? 1.0f : 0.0f;
        f[13] = npts > 0 ? inside / npts : 0.0f;
        f[14] = (float)rule[p];
        f[15] = log2f(1.0f + 0.25f * cw * ch);
        // Upper bound of the path's intersections.  A Bezier curve (also a rational one with
        // positive weights: x(t) = c where the Bernstein coefficients w_i (x_i - c) change sign)
        // varies no more than its control polygon; its <= 5 pieces cover [0, 1] at most three
        // times over (the 4-cut order quirk runs one piece backwards); a piece from a to b gives at most
        // 1 + (|dx|/2 + 1) + (|dy|/2 + 1) entries.  The float terms carry a 0.1% margin.
        double b = 0.0;
        if (visible_mask(visible[p])) b = 1.5 * ((double)sdx + (double)sdy) * 1.001 + 20.0 * (double)(c1 - c0) + 2.0;
        if (!(b < 1e15)) b = 1e15;  // non-finite coordinates
        if (b > 0.0) bsum += (unsigned long long)b;
        if (pbound) pbound[p] = b < 1073741824.0 ? (int32_t)b : 1073741824;
    }
    }
    if (lane == 0) s_bound[wave] = bsum;
    __syncthreads();
    if (threadIdx.x == 0 && bound_total) {
        const unsigned long long t = s_bound[0] + s_bound[1] + s_bound[2] + s_bound[3];
        if (t) atomicAdd(bound_total, t);
    }
}

// MLP inference on the matrix cores: one wave per 16 paths, v_mfma_f32_16x16x16_bf16 tiles.
// Operand layout of the 16x16x16 BF16 MFMA: lane l holds A[l%16][4*(l/16)+q], B[4*(l/16)+q][l%16]
// (q = 0..3) and receives D[4*(l/16)+q][l%16].
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ feat, uint32_t Np, MlpDev m,
                                                   float* __restrict__ pred, const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ float act[4][16][kH + 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    const int li = lane & 15, lq = lane >> 4;  // row/col inside a tile, k group
    // layer 1: 16 inputs -> 64
    bf16x4 a;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = 4 * lq + q;
        const int64_t r = r0 + li;
        const float x = r < Np ? (feat[r * kF + k] - m.mean[k]) * m.inv_std[k] : 0.0f;
        a[q] = f2bf(x);
    }
#pragma unroll
    for (int t = 0; t < kH / 16; ++t) {
        bf16x4 b;
#pragma unroll
        for (int q = 0; q < 4; ++q) b[q] = (short)m.w1[(4 * lq + q) * kH + 16 * t + li];
        f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) act[wave][4 * lq + q][16 * t + li] = fmaxf(d[q] + m.b1[16 * t + li], 0.0f);
    }
    __syncthreads();
    // layer 2: 64 -> 64
    f32x4 d2[kH / 16];
#pragma unroll
    for (int t = 0; t < kH / 16; ++t) d2[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < kH / 16; ++ks) {
        bf16x4 a2;
#pragma unroll
        for (int q = 0; q < 4; ++q) a2[q] = f2bf(act[wave][li][16 * ks + 4 * lq + q]);
#pragma unroll
        for (int t = 0; t < kH / 16; ++t) {
            bf16x4 b;
#pragma unroll
            for (int q = 0; q < 4; ++q) b[q] = (short)m.w2[(16 * ks + 4 * lq + q) * kH + 16 * t + li];
            d2[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b, d2[t], 0, 0, 0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kH / 16; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) act[wave][4 * lq + q][16 * t + li] = fmaxf(d2[t][q] + m.b2[16 * t + li], 0.0f);
    __syncthreads();
    // layer 3: 64 -> 2 (the 16-wide tile is padded with zero columns)
    f32x4 d3 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < kH / 16; ++ks) {
        bf16x4 a3, b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a3[q] = f2bf(act[wave][li][16 * ks + 4 * lq + q]);
            b[q] = (short)m.w3[(16 * ks + 4 * lq + q) * 16 + li];
        }
        d3 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, b, d3, 0, 0, 0);
    }
    if (li < kMlpOut) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t r = r0 + 4 * lq + q;
            if (r < Np) pred[r * kMlpOut + li] = d3[q] + m.b3[li];
        }
    }
}

// Predicted capacities: cap = 2^(pred + margin) - 1 + slack (0 for culled paths).  capN goes
// to `arena` (scanned afterwards into arena bases); the record capacities are summed.
__global__ void k_caps(const float* __restrict__ pred, const int32_t* __restrict__ visible, uint32_t Np,
                       float log_margin, int32_t slack, int32_t* __restrict__ arena, int32_t* __restrict__ cap_n,
                       int32_t* __restrict__ cap_r_total, const int32_t* __restrict__ pbound,
                       const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t cr = 0;
    if (p < Np) {
        int32_t cn = 0;
        if (visible_mask(visible[p])) {
            const float en = fminf(pred[2 * p] + log_margin, 28.0f);
            const float er = fminf(pred[2 * p + 1] + log_margin, 28.0f);
            cn = (int32_t)ceilf(exp2f(en)) - 1 + slack;
            cr = (int32_t)ceilf(exp2f(er)) - 1 + slack;
            cn = cn < 0 ? 0 : cn;
            cr = cr < 0 ? 0 : cr;
            // never more than the path's proven bound (records: two per entry), so the arena
            // total stays within the bound total
            if (pbound) { cn = min(cn, pbound[p]); cr = min(cr, 2 * min(pbound[p], 1 << 29)); }
        }
        arena[p] = cn;
        cap_n[p] = cn;
    }
    for (int o = 32; o >= 1; o >>= 1) cr += __shfl_xor(cr, o);
    if ((threadIdx.x & 63) == 0 && cr) atomicAdd(cap_r_total, cr);
}

// Paths whose exact entry count exceeds their predicted capacity: one atomic per wave.
__global__ void k_path_overflow(const int32_t* __restrict__ seg, const int32_t* __restrict__ cap_n,
                                const int32_t* __restrict__ visible, uint32_t Np, int32_t* __restrict__ n_over,
                                const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool over = p < Np && visible_mask(visible[p]) && seg[p + 1] - seg[p] > cap_n[p];
    const unsigned long long m = __ballot(over);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(n_over, (int32_t)__popcll(m));
}

// sstart[p] = first sorted element of path p = bucket offset of the path's first bucket.
__global__ void k_sorted_starts(const int32_t* __restrict__ boff, const int32_t* __restrict__ bucket_base,
                                uint32_t Np, int32_t* __restrict__ sstart) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= Np) sstart[p] = boff[bucket_base[p]];
}

// Per-path labels for training: log2(1 + intersections), log2(1 + records).
__global__ void k_path_targets(const int32_t* __restrict__ n_per_path_or_seg, int seg_form,
                               const int2* __restrict__ brec, const int32_t* __restrict__ bucket_base,
                               uint32_t Np, float* __restrict__ targets, int32_t* __restrict__ exact) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Np) return;
    const int32_t n = seg_form ? n_per_path_or_seg[p + 1] - n_per_path_or_seg[p] : n_per_path_or_seg[p];
    int64_t r = 0;
    for (int32_t b = bucket_base[p]; b < bucket_base[p + 1]; ++b) {
        const int2 q = brec[b];
        r += q.y > q.x ? q.y - q.x : 0;
    }
    targets[2 * p] = log2f(1.0f + (float)n);
    targets[2 * p + 1] = log2f(1.0f + (float)r);
    if (exact) { exact[2 * p] = n; exact[2 * p + 1] = (int32_t)r; }
}

// ---- FP32 training kernels (row-major activations, W[k][n])
__global__ void k_dense_fwd(const float* __restrict__ X, int B, int K, const float* __restrict__ W,
                            const float* __restrict__ bias, int N, int relu, float* __restrict__ Y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * N) return;
    const int r = (int)(i / N), n = (int)(i % N);
    float acc = bias[n];
    for (int k = 0; k < K; ++k) acc += X[(int64_t)r * K + k] * W[k * N + n];
    Y[i] = relu ? fmaxf(acc, 0.0f) : acc;
}

__global__ void k_loss_grad(const float* __restrict__ Y, const float* __restrict__ T, int B, int N,
                            float* __restrict__ dY, float* __restrict__ sq) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * N) return;
    const float e = Y[i] - T[i];
    dY[i] = 2.0f * e / (float)((int64_t)B * N);
    sq[i] = e * e;
}

// dA = (A > 0) * dY W^T  (A is the layer's ReLU output that fed this layer)
__global__ void k_dense_bwd_data(const float* __restrict__ dY, int B, int N, const float* __restrict__ W, int K,
                                 const float* __restrict__ A, float* __restrict__ dA) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * K) return;
    const int r = (int)(i / K), k = (int)(i % K);
    float acc = 0.0f;
    for (int n = 0; n < N; ++n) acc += dY[(int64_t)r * N + n] * W[k * N + n];
    dA[i] = A[i] > 0.0f ? acc : 0.0f;
}

// dW[k][n] = sum_r X[r][k] dY[r][n] for k < K; output row K gives db[n] = sum_r dY[r][n].  One
// wave per output: the lanes take every 64th row of the batch and a wave reduction adds them up.
__global__ __launch_bounds__(256) void k_dense_bwd_weight(const float* __restrict__ X, const float* __restrict__ dY,
                                                          int B, int K, int N, float* __restrict__ dW,
                                                          float* __restrict__ db) {
    const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (o >= (K + 1) * N) return;  // wave-uniform
    const int k = o / N, n = o % N;
    float acc = 0.0f;
    for (int r = lane; r < B; r += 64) acc += (k < K ? X[(int64_t)r * K + k] : 1.0f) * dY[(int64_t)r * N + n];
    acc = wave_sum_f(acc);
    if (lane == 0) {
        if (k < K) dW[k * N + n] = acc;
        else db[n] = acc;
    }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void k_gather_rows(const float* __restrict__ src, const int32_t* __restrict__ idx, int B, int K,
                              float* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * K) return;
    const int r = (int)(i / K), k = (int)(i % K);
    dst[i] = src[(int64_t)idx[r] * K + k];
}

__global__ void k_doc_winding_base(const int32_t* __restrict__ winding, const int32_t* __restrict__ sstart,
                                   const int32_t* __restrict__ doc_first_path, uint32_t Np,
                                   int32_t* __restrict__ wbase) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < Np) wbase[p] = winding[sstart[doc_first_path[p]]];
}

// Winding number at each document's first sorted entry from the bands: that position is a band
// start (a path's first non-empty bucket starts a band) or the end, whose prefix is the total.
__global__ void k_doc_winding_base_bands(const int32_t* __restrict__ band_start, const int32_t* __restrict__ bpre,
                                         const int32_t* __restrict__ n_bands, const int32_t* __restrict__ sstart,
                                         const int32_t* __restrict__ doc_first_path, uint32_t Np,
                                         int32_t* __restrict__ wbase) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Np) return;
    const int32_t pos = sstart[doc_first_path[p]];
    int32_t lo = 0, hi = *n_bands;  // band_start[lo] <= pos < band_start[hi], band_start[n_bands] = end
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (band_start[mid] <= pos) lo = mid; else hi = mid;
    }
    wbase[p] = band_start[lo] == pos ? bpre[lo] : bpre[hi];
}

inline unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b < 1 ? 1 : (n + b - 1) / b); }

}  // namespace

hipError_t launch_path_features(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                                const uint32_t* pcurve, const int32_t* visible, const uint32_t* rule,
                                uint32_t Np, int w, int h, float* feat, hipStream_t s,
                                unsigned long long* bound_total, const int2* pcanvas, int32_t* pbound,
                                const int32_t* skip) {
    if (Np == 0) return hipSuccess;
    k_path_features<<<(Np + 15) / 16, 256, 0, s>>>(tpos, cpos, ctype, pcurve, visible, rule, Np, w, h, feat,
                                                   bound_total, pcanvas, pbound, skip);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const float* feat, uint32_t Np, const MlpDev& m, float* pred, hipStream_t s,
                            const int32_t* skip) {
    if (Np == 0) return hipSuccess;
    k_mlp_infer<<<(Np + 63) / 64, 256, 0, s>>>(feat, Np, m, pred, skip);
    return hipGetLastError();
}

hipError_t launch_caps(const float* pred, const int32_t* visible, uint32_t Np, float log_margin, int32_t slack,
                       int32_t* arena, int32_t* cap_n, int32_t* cap_r_total, hipStream_t s, const int32_t* pbound,
                       const int32_t* skip) {
    if (Np == 0) return hipSuccess;
    k_caps<<<blocks(Np), 256, 0, s>>>(pred, visible, Np, log_margin, slack, arena, cap_n, cap_r_total, pbound, skip);
    return hipGetLastError();
}

hipError_t launch_path_overflow(const int32_t* seg, const int32_t* cap_n, const int32_t* visible, uint32_t Np,
                                int32_t* n_over, hipStream_t s, const int32_t* skip) {
    if (Np == 0) return hipSuccess;
    k_path_overflow<<<blocks(Np), 256, 0, s>>>(seg, cap_n, visible, Np, n_over, skip);
    return hipGetLastError();
}

hipError_t launch_doc_winding_base(const int32_t* winding, const int32_t* sstart, const int32_t* doc_first_path,
                                   uint32_t Np, int32_t* wbase, hipStream_t s) {
    if (Np == 0) return hipSuccess;
    k_doc_winding_base<<<blocks(Np), 256, 0, s>>>(winding, sstart, doc_first_path, Np, wbase);
    return hipGetLastError();
}

hipError_t launch_doc_winding_base_bands(const int32_t* band_start, const int32_t* bpre, const int32_t* n_bands,
                                         const int32_t* sstart, const int32_t* doc_first_path, uint32_t Np,
                                         int32_t* wbase, hipStream_t s) {
    if (Np == 0) return hipSuccess;
    k_doc_winding_base_bands<<<blocks(Np), 256, 0, s>>>(band_start, bpre, n_bands, sstart, doc_first_path, Np, wbase);
    return hipGetLastError();
}

hipError_t launch_sorted_starts(const int32_t* boff, const int32_t* bucket_base, uint32_t Np, int32_t* sstart,
                                hipStream_t s) {
    k_sorted_starts<<<blocks((int64_t)Np + 1), 256, 0, s>>>(boff, bucket_base, Np, sstart);
    return hipGetLastError();
}

hipError_t launch_path_targets(const int32_t* n_or_seg, int seg_form, const int2* brec, const int32_t* bucket_base,
                               uint32_t Np, float* targets, int32_t* exact, hipStream_t s) {
    if (Np == 0) return hipSuccess;
    k_path_targets<<<blocks(Np), 256, 0, s>>>(n_or_seg, seg_form, brec, bucket_base, Np, targets, exact);
    return hipGetLastError();
}

hipError_t launch_dense_fwd(const float* X, int B, int K, const float* W, const float* b, int N, int relu, float* Y,
                            hipStream_t s) {
    k_dense_fwd<<<blocks((int64_t)B * N), 256, 0, s>>>(X, B, K, W, b, N, relu, Y);
    return hipGetLastError();
}
hipError_t launch_loss_grad(const float* Y, const float* T, int B, int N, float* dY, float* sq, hipStream_t s) {
    k_loss_grad<<<blocks((int64_t)B * N), 256, 0, s>>>(Y, T, B, N, dY, sq);
    return hipGetLastError();
}
hipError_t launch_dense_bwd_data(const float* dY, int B, int N, const float* W, int K, const float* A, float* dA,
                                 hipStream_t s) {
    k_dense_bwd_data<<<blocks((int64_t)B * K), 256, 0, s>>>(dY, B, N, W, K, A, dA);
    return hipGetLastError();
}
hipError_t launch_dense_bwd_weight(const float* X, const float* dY, int B, int K, int N, float* dW, float* db,
                                   hipStream_t s) {
    k_dense_bwd_weight<<<(unsigned)(((int64_t)(K + 1) * N + 3) / 4), 256, 0, s>>>(X, dY, B, K, N, dW, db);
    return hipGetLastError();
}
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, hipStream_t s) {
    k_adam<<<blocks(n), 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}
hipError_t launch_gather_rows(const float* src, const int32_t* idx, int B, int K, float* dst, hipStream_t s) {
    k_gather_rows<<<blocks((int64_t)B * K), 256, 0, s>>>(src, idx, B, K, dst);
    return hipGetLastError();
}

}  // namespace spr
