// k_records.hip -- the draw of a given record list (gfx950): the reference's MOCK_DATA replay,
// which loads (yx, width, fill, frag_index) records from a file and draws them with the line-list
// pipeline (src/core/scanline/scanline_rasterizer.cpp:261-281 load, :1057-1109 draw;
// shaders/scanline/surface/scanlinepr.vert:18-50).  A record fills rows y and y + 1 over columns
// [x, x + width) with its RGBA8 fill (x = yx & 0xFFFF, y = yx >> 16); records are drawn in list
// order (overwrite: the later one wins; src-over: folded in order), on a white canvas, the image
// stored top row first (canvas row h - 1 - y).
//
// One workgroup per image row and 1024-column chunk: the records are scanned 256 at a time, those
// touching the row and chunk are listed in LDS in list order (a ballot prefix per wave), and each
// thread paints its 4 register-resident pixels from the list.  The whole list is scanned per row,
// which suits debug-size lists (the reference's are a few thousand records).
#include "kernels.hpp"
#include "scan_common.hpp"

namespace spr {

namespace {

constexpr int kDrawThreads = 256, kDrawPx = 4, kDrawCols = kDrawThreads * kDrawPx;

__global__ __launch_bounds__(kDrawThreads) void k_draw_records(const int4* __restrict__ rec, int32_t R, int w,
                                                               int h, int compose, uint8_t* __restrict__ image) {
    __shared__ int4 s_list[kDrawThreads];
    __shared__ int32_t s_wave[kDrawThreads / 64];
    const int row = blockIdx.y;            // image row, top first
    const int yy = h - 1 - row;            // canvas row
    const int c0 = blockIdx.x * kDrawCols;  // first column of the chunk
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float4 acc[kDrawPx];
    uint32_t last[kDrawPx];
#pragma unroll
    for (int k = 0; k < kDrawPx; ++k) { acc[k] = make_float4(1.0f, 1.0f, 1.0f, 1.0f); last[k] = 0xFFFFFFFFu; }
    for (int32_t b = 0; b < R; b += kDrawThreads) {
        const int32_t i = b + threadIdx.x;
        bool hit = false;
        int4 r = make_int4(0, 0, 0, 0);
        if (i < R) {
            r = rec[i];
            const int x = r.x & 0xFFFF, y = r.x >> 16;
            const int x0 = max(x, 0), x1 = min(x + r.y, w);
            hit = yy >= y && yy < y + 2 && x0 < x1 && x1 > c0 && x0 < c0 + kDrawCols;
        }
        const unsigned long long m = __ballot(hit);
        if (lane == 0) s_wave[wave] = __popcll(m);
        __syncthreads();
        int32_t off = 0, n = 0;
#pragma unroll
        for (int k = 0; k < kDrawThreads / 64; ++k) {
            off += k < wave ? s_wave[k] : 0;
            n += s_wave[k];
        }
        if (hit) s_list[off + __popcll(m & ((1ull << lane) - 1ull))] = r;
        __syncthreads();
        for (int32_t j = 0; j < n; ++j) {  // in list order
            const int4 q = s_list[j];
            const int x = q.x & 0xFFFF;
            const int x0 = max(x, 0), x1 = min(x + q.y, w);
            const uint32_t c = (uint32_t)q.z;
#pragma unroll
            for (int k = 0; k < kDrawPx; ++k) {
                const int px = c0 + k * kDrawThreads + (int)threadIdx.x;
                if (px < x0 || px >= x1) continue;
                if (compose == 0) {
                    last[k] = c;
                } else {
                    const float a = (float)(c >> 24) / 255.0f, ia = 1.0f - a;
                    acc[k].x = (float)(c & 0xFFu) / 255.0f * a + acc[k].x * ia;
                    acc[k].y = (float)((c >> 8) & 0xFFu) / 255.0f * a + acc[k].y * ia;
                    acc[k].z = (float)((c >> 16) & 0xFFu) / 255.0f * a + acc[k].z * ia;
                    acc[k].w = a + acc[k].w * ia;
                }
            }
        }
        __syncthreads();
    }
    uint32_t* out = reinterpret_cast<uint32_t*>(image) + (size_t)row * w;
#pragma unroll
    for (int k = 0; k < kDrawPx; ++k) {
        const int px = c0 + k * kDrawThreads + (int)threadIdx.x;
        if (px >= w) continue;
        uint32_t v;
        if (compose == 0) {
            v = last[k];  // white where no record covers the pixel
        } else {
            auto u8 = [](float f) {
                float t = f * 255.0f + 0.5f;
                t = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
                return (uint32_t)t;
            };
            v = u8(acc[k].x) | (u8(acc[k].y) << 8) | (u8(acc[k].z) << 16) | (u8(acc[k].w) << 24);
        }
        out[px] = v;
    }
}

}  // namespace

hipError_t launch_draw_records(const int4* records, int32_t R, int w, int h, int compose, uint8_t* image,
                               hipStream_t s) {
    if (w <= 0 || h <= 0) return hipSuccess;
    const dim3 grid((unsigned)((w + kDrawCols - 1) / kDrawCols), (unsigned)h);
    k_draw_records<<<grid, kDrawThreads, 0, s>>>(records, R, w, h, compose, image);
    return hipGetLastError();
}

}  // namespace spr
