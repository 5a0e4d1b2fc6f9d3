// k_composite.hip -- records to RGBA8 pixels (gfx950).
//
// Replaces the reference's line-list draw (scanlinepr.vert:18-50, scanlinepr.frag:7-14, pipeline
// state scanline_rasterizer.cpp:881-944): every record (yx, width, fill) covers canvas rows y and
// y+1 over columns [x, x+width), records are applied in order (blending disabled: the later one
// wins), and the canvas row y lands on image row h-1-y.  All records sit on even rows, so the
// image is a stack of independent two-pixel "cell rows".
//
// Records are grouped into runs (one path, one cell row: contiguous and x-sorted, never
// overlapping).  Runs are bucketed by cell row; one workgroup per cell row sorts its run ids
// (= path order), and each thread owns a fixed group of columns in registers, finds the records
// of each run that touch its columns by binary search, and applies them in path order.  The
// image is written exactly once.
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxLdsRuns = 8192;

__device__ inline int rec_row(const int4& r) { return (r.x >> 16) >> 1; }

__global__ void k_run_flags(const int4* __restrict__ records, const uint32_t* __restrict__ rec_path,
                            const int32_t* __restrict__ totals, int64_t max_records,
                            int32_t* __restrict__ run_flag) {
    const int64_t R = totals[2];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < max_records; k += (int64_t)gridDim.x * blockDim.x) {
        int32_t f = 0;
        if (k < R) {
            f = k == 0 || rec_row(records[k]) != rec_row(records[k - 1]) || rec_path[k] != rec_path[k - 1];
        }
        run_flag[k] = f;
    }
}

__global__ void k_run_emit(const int4* __restrict__ records, const int32_t* __restrict__ run_scan,
                           const int32_t* __restrict__ totals, int64_t max_records, int2* __restrict__ runs) {
    const int64_t R = totals[2];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < R; k += (int64_t)gridDim.x * blockDim.x) {
        if (run_scan[k + 1] != run_scan[k]) runs[run_scan[k]] = make_int2(rec_row(records[k]), (int32_t)k);
    }
    (void)max_records;
}

__global__ void k_row_count(const int2* __restrict__ runs, const int32_t* __restrict__ run_total,
                            int32_t* __restrict__ row_count) {
    const int64_t NR = *run_total;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < NR; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&row_count[runs[i].x], 1);
}

__global__ void k_row_scatter(const int2* __restrict__ runs, const int32_t* __restrict__ run_total,
                              const int32_t* __restrict__ row_start, int32_t* __restrict__ row_fill,
                              int32_t* __restrict__ row_runs) {
    const int64_t NR = *run_total;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < NR; i += (int64_t)gridDim.x * blockDim.x) {
        const int row = runs[i].x;
        const int slot = atomicAdd(&row_fill[row], 1);
        row_runs[row_start[row] + slot] = (int32_t)i;
    }
}

// Ascending bitonic sort of n ints at `a` (LDS or global) by one workgroup.  Indices >= n are
// treated as +infinity; with every comparator moving the minimum down, such virtual elements
// only ever move upward, so nothing at or beyond n is read or written.
template <class P>
__device__ inline void block_sort_ints(P a, int n) {
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int k = 2; k <= n2; k <<= 1) {
        const int half = k >> 1;
        for (int q = threadIdx.x; q < n2 / 2; q += blockDim.x) {
            const int i = (q / half) * k + (q % half);
            const int l = (q / half) * k + k - 1 - (q % half);
            if (l < n) { int x = a[i], y = a[l]; if (y < x) { a[i] = y; a[l] = x; } }
        }
        __syncthreads();
        for (int j = k >> 2; j > 0; j >>= 1) {
            for (int q = threadIdx.x; q < n2 / 2; q += blockDim.x) {
                const int i = (q / j) * 2 * j + (q % j);
                const int l = i + j;
                if (l < n) { int x = a[i], y = a[l]; if (y < x) { a[i] = y; a[l] = x; } }
            }
            __syncthreads();
        }
    }
}

template <int CW, int COMPOSE>
__global__ __launch_bounds__(kBlock) void k_composite(const int4* __restrict__ records,
                                                      const int2* __restrict__ runs,
                                                      const int32_t* __restrict__ run_total,
                                                      const int32_t* __restrict__ totals,
                                                      const int32_t* __restrict__ row_start,
                                                      int32_t* __restrict__ row_runs, int w, int h,
                                                      uint8_t* __restrict__ image) {
    __shared__ int32_t ids[kMaxLdsRuns];
    const int r = blockIdx.x;
    const int y = 2 * r;
    const int s = row_start[r];
    const int cnt = row_start[r + 1] - s;
    const int NR = *run_total;
    const int R = totals[2];
    const bool lds = cnt <= kMaxLdsRuns;
    if (lds) {
        for (int i = threadIdx.x; i < cnt; i += kBlock) ids[i] = row_runs[s + i];
        __syncthreads();
        block_sort_ints(ids, cnt);
    } else {
        block_sort_ints(row_runs + s, cnt);
        __threadfence_block();
    }
    __syncthreads();
    const int c0 = threadIdx.x * CW;
    uint32_t pix[CW];
    float acc[COMPOSE ? CW : 1][4];
#pragma unroll
    for (int j = 0; j < CW; ++j) {
        pix[j] = 0xFFFFFFFFu;
        if (COMPOSE) { acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 1.0f; }
    }
    if (c0 < w) {
        for (int q = 0; q < cnt; ++q) {
            const int run = lds ? ids[q] : row_runs[s + q];
            const int rs = runs[run].y;
            const int re = run + 1 < NR ? runs[run + 1].y : R;
            // Inside a run, element f emits its cell [x_f, x_f+2) and then its span
            // [x_{f-1}+2, x_f): both have end + (span ? 2 : 0) == x_f + 2, which therefore never
            // decreases along the run.  Find the first record that can reach column c0.
            int lo = rs, hi = re;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                const int4 m = records[mid];
                if ((m.x & 0xFFFF) + m.y + (m.w == 0 ? 2 : 0) <= c0) lo = mid + 1; else hi = mid;
            }
            for (int k = lo; k < re; ++k) {
                const int4 rec = records[k];
                const int x = rec.x & 0xFFFF;
                if (x >= c0 + CW) {
                    // Later elements start right of this one; only the span of this same
                    // element (the next record, if it is a span) can still reach back.
                    if (rec.w == 0 || k + 1 >= re || records[k + 1].w != 0) break;
                    continue;
                }
                const int xe = x + rec.y;
                const uint32_t col = (uint32_t)rec.z;
                if (COMPOSE == 0) {
#pragma unroll
                    for (int j = 0; j < CW; ++j)
                        if (c0 + j >= x && c0 + j < xe) pix[j] = col;
                } else {
                    const float s0 = (float)(col & 0xFFu) / 255.0f, s1 = (float)((col >> 8) & 0xFFu) / 255.0f;
                    const float s2 = (float)((col >> 16) & 0xFFu) / 255.0f, a = (float)((col >> 24) & 0xFFu) / 255.0f;
                    const float ia = 1.0f - a;
#pragma unroll
                    for (int j = 0; j < CW; ++j) {
                        if (c0 + j >= x && c0 + j < xe) {
                            acc[j][0] = s0 * a + acc[j][0] * ia;
                            acc[j][1] = s1 * a + acc[j][1] * ia;
                            acc[j][2] = s2 * a + acc[j][2] * ia;
                            acc[j][3] = a + acc[j][3] * ia;
                        }
                    }
                }
            }
        }
        if (COMPOSE) {
#pragma unroll
            for (int j = 0; j < CW; ++j) {
                uint32_t o = 0;
#pragma unroll
                for (int ch = 0; ch < 4; ++ch) {
                    float v = acc[j][ch] * 255.0f + 0.5f;
                    v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
                    o |= (uint32_t)v << (8 * ch);
                }
                pix[j] = o;
            }
        }
        for (int yy = y; yy < y + 2 && yy < h; ++yy) {
            uint32_t* row = reinterpret_cast<uint32_t*>(image) + (int64_t)(h - 1 - yy) * w;
#pragma unroll
            for (int j = 0; j < CW; ++j)
                if (c0 + j < w) row[c0 + j] = pix[j];
        }
    }
}

inline unsigned stride_grid(int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

hipError_t launch_run_flags(const int4* records, const uint32_t* rec_path, const int32_t* totals,
                            int64_t max_records, int32_t* run_flag, hipStream_t s) {
    k_run_flags<<<stride_grid(max_records), kBlock, 0, s>>>(records, rec_path, totals, max_records, run_flag);
    return hipGetLastError();
}

hipError_t launch_run_emit(const int4* records, const int32_t* run_scan, const int32_t* totals,
                           int64_t max_records, int2* runs, hipStream_t s) {
    k_run_emit<<<stride_grid(max_records), kBlock, 0, s>>>(records, run_scan, totals, max_records, runs);
    return hipGetLastError();
}

hipError_t launch_row_count(const int2* runs, const int32_t* run_total, int64_t max_runs,
                            int32_t* row_count, hipStream_t s) {
    k_row_count<<<stride_grid(max_runs), kBlock, 0, s>>>(runs, run_total, row_count);
    return hipGetLastError();
}

hipError_t launch_row_scatter(const int2* runs, const int32_t* run_total, int64_t max_runs,
                              const int32_t* row_start, int32_t* row_fill, int32_t* row_runs,
                              hipStream_t s) {
    k_row_scatter<<<stride_grid(max_runs), kBlock, 0, s>>>(runs, run_total, row_start, row_fill, row_runs);
    return hipGetLastError();
}

hipError_t launch_composite(const int4* records, const int2* runs, const int32_t* run_total,
                            const int32_t* totals, const int32_t* row_start,
                            const int32_t* row_runs, int w, int h, int compose, uint8_t* image,
                            hipStream_t s) {
    const int rows = (h + 1) / 2;
    if (rows == 0 || w == 0) return hipSuccess;
    int32_t* rr = const_cast<int32_t*>(row_runs);
    if (w <= 256 * 4) {
        if (compose) k_composite<4, 1><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
        else k_composite<4, 0><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
    } else if (w <= 256 * 8) {
        if (compose) k_composite<8, 1><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
        else k_composite<8, 0><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
    } else if (w <= 256 * 16) {
        if (compose) k_composite<16, 1><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
        else k_composite<16, 0><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
    } else if (w <= 256 * 32) {
        if (compose) k_composite<32, 1><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
        else k_composite<32, 0><<<rows, kBlock, 0, s>>>(records, runs, run_total, totals, row_start, rr, w, h, image);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace spr
