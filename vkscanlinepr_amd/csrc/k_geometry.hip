// k_geometry.hip -- transform + cull, monotonisation + count, and fragment-key generation
// (gfx950).  The intersection write between count and keys lives in k_chain.hip.
#include "device_common.hpp"
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Frame summary: dst[i] = *src.p[i] (0 for a null source), one lane per value, so that one
// device-to-host copy returns every counter of the frame.
__global__ void k_gather_summary(SummarySrc src, int32_t* __restrict__ dst) {
    const int i = threadIdx.x;
    if (i < src.n) dst[i] = src.p[i] ? *src.p[i] : 0;
}

// Fill of min(n_max, mult * *n_dev + add) words (a count known only on the device).
__global__ void k_fill_u32_dev(uint32_t* __restrict__ p, int64_t n_max, const int32_t* __restrict__ n_dev,
                               int64_t mult, int64_t add, uint32_t v) {
    const int64_t n = min(n_max, mult * (int64_t)*n_dev + add);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// transform_pos.comp:33-85.  The per-path outcode OR is reduced inside the wave first (points of
// one path are contiguous), then one atomicOr per (wave, path).
// It also bounds each path's cell rows from its transformed control points (the curves stay
// inside their hull; the margin of 1 px plus 2^-20 of the magnitude covers the rounding of the
// curve evaluation), which sizes the path's (path, row) buckets of the sort.
__global__ __launch_bounds__(kBlock) void k_transform(const float2* __restrict__ pos,
                                                      const uint32_t* __restrict__ pos_path,
                                                      uint32_t P, Mvp m, float fw, float fh, int rmax,
                                                      float2* __restrict__ tpos,
                                                      int32_t* __restrict__ visible,
                                                      int32_t* __restrict__ row_lo,
                                                      int32_t* __restrict__ row_hi,
                                                      const int2* __restrict__ pcanvas,
                                                      const float* __restrict__ mvp_dev) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < P;
    if (mvp_dev) {  // the matrix in device memory (a captured frame that follows setMVP)
#pragma unroll
        for (int k = 0; k < 16; ++k) m.m[k] = mvp_dev[k];
    }
    int32_t bits = 0, lo = 0, hi = 0;
    uint32_t path = 0xFFFFFFFFu;
    if (valid) {
        if (pcanvas) {  // documents of a batch: each path has its own canvas
            const int2 cv = pcanvas[pos_path[i]];
            fw = (float)cv.x; fh = (float)cv.y; rmax = (cv.y - 1) / 2;
        }
        float2 p = pos[i];
        float o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            o[r] = ((p.x * m.m[4 * r] + p.y * m.m[4 * r + 1]) + 0.0f * m.m[4 * r + 2]) + 1.0f * m.m[4 * r + 3];
        float ox = o[0] / o[3], oy = o[1] / o[3];
        tpos[i] = make_float2(ox, oy);
        bits = outcode_bits(ox, oy, fw, fh);
        path = pos_path[i];
        const float mg = 1.0f + fabsf(oy) * 0x1p-20f;
        lo = sat_f2i(floorf((oy - mg) * 0.5f));
        hi = sat_f2i(floorf((oy + mg) * 0.5f));
        lo = lo < -1 ? -1 : (lo > rmax + 1 ? rmax + 1 : lo);
        hi = hi < -1 ? -1 : (hi > rmax + 1 ? rmax + 1 : hi);
        if (!(oy == oy)) { lo = -1; hi = rmax + 1; }
    }
    unsigned long long todo = __ballot(valid && bits != 0);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const uint32_t lp = __shfl(path, leader);
        const bool match = valid && bits != 0 && path == lp;
        todo &= ~__ballot(match);
        int32_t v = match ? bits : 0;
        int32_t vlo = match ? lo : 0x7FFFFFFF, vhi = match ? hi : (int32_t)0x80000000;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            v |= __shfl_xor(v, o);
            vlo = min(vlo, __shfl_xor(vlo, o));
            vhi = max(vhi, __shfl_xor(vhi, o));
        }
        if ((int)(threadIdx.x & 63) == leader) {
            atomicOr(&visible[lp], v);
            atomicMin(&row_lo[lp], vlo);
            atomicMax(&row_hi[lp], vhi);
        }
    }
}

// make_intersection_0.comp:226-411: monotone cuts, cut cache, per-curve intersection count.
__global__ __launch_bounds__(kBlock) void k_count(const float2* __restrict__ tpos,
                                                  const uint32_t* __restrict__ cpos,
                                                  const uint32_t* __restrict__ ctype,
                                                  const float* __restrict__ cweight,
                                                  const uint32_t* __restrict__ cpath,
                                                  const int32_t* __restrict__ visible, uint32_t C,
                                                  int w, int h, float4* __restrict__ cuts,
                                                  uint32_t* __restrict__ ncuts,
                                                  int32_t* __restrict__ counts,
                                                  const int2* __restrict__ pcanvas) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const uint32_t path = cpath[c];
    if (pcanvas) { w = pcanvas[path].x; h = pcanvas[path].y; }  // a batch: the path's document canvas
    const CurvePts cv = load_curve(tpos, cpos[c], ctype[c], cweight, c);
    const bool vis = visible_mask(visible[path]);
    float q[4];
    const uint32_t n = monotone_cuts(cv, vis, q);
    cuts[c] = make_float4(q[0], q[1], q[2], q[3]);
    ncuts[c] = n;
    int32_t cnt = 0;
    if (vis) {
        float p0x = cv.x[0], p0y = cv.y[0];
#pragma unroll
        for (uint32_t k = 0; k < 5; ++k) {
            if (k > n) break;
            const float t = k < n ? q[k] : 1.0f;
            const float p1x = cv.ex(t), p1y = cv.ey(t);
            cnt += piece_count(p0x, p0y, p1x, p1y, w, h);
            p0x = p1x; p0y = p1y;
        }
    }
    counts[c] = cnt;
}

// Per-path segment starts: seg[p] = entries of all curves of paths < p (the sort segments of
// gen_fragment.comp:227-244, derived from the curve offsets instead of entry boundaries).
__global__ void k_path_segments(const uint32_t* __restrict__ cpath, uint32_t C,
                                const int32_t* __restrict__ offsets, uint32_t Np,
                                int32_t* __restrict__ seg) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const uint32_t p = cpath[c];
    const uint32_t lo = c == 0 ? 0u : cpath[c - 1] + 1u;
    const int32_t off = offsets[c];
    for (uint32_t j = lo; j <= p; ++j) seg[j] = off;
    if (c == C - 1) {
        const int32_t total = offsets[C];
        for (uint32_t j = p + 1; j <= Np; ++j) seg[j] = total;
    }
}

__global__ void k_segment_stats(const int32_t* __restrict__ seg, uint32_t Np, int32_t* __restrict__ max_len) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    int32_t len = p < Np ? seg[p + 1] - seg[p] : 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) len = max(len, __shfl_xor(len, o));
    if ((threadIdx.x & 63) == 0 && len > 0) atomicMax(max_len, len);
}

// gen_fragment.comp:90-224: cell key of the segment between consecutive intersections of one
// curve, and its winding change across the cell's centre row.  Returns the entry's bucket.
__device__ inline uint32_t gen_one(const int2* __restrict__ inter, int64_t N, int64_t i,
                                   const uint32_t* __restrict__ cpos, const uint32_t* __restrict__ ctype,
                                   const float* __restrict__ cweight, const uint32_t* __restrict__ cpath,
                                   const float2* __restrict__ tpos, int w,
                                   int h, int32_t* __restrict__ key, int8_t* __restrict__ wn,
                                   uint32_t* __restrict__ epath, const int32_t* __restrict__ bucket_base,
                                   const int2* __restrict__ prow, int32_t* __restrict__ err,
                                   float4* __restrict__ chord) {
    const int2 a = inter[i];
    const int2 b = (i + 1 != N) ? inter[i + 1] : make_int2(-1, 0x3f800000);
    float t0 = glmax(bitsf((uint32_t)a.y & 0xFFFFFFFCu), 0.0f);
    float t1 = glmax(bitsf((uint32_t)b.y & 0xFFFFFFFCu), 0.0f);
    const uint32_t c = (uint32_t)a.x;
    const uint32_t path = cpath[c];
    if (a.x != b.x) t1 = 1.0f;
    uint32_t yx;
    int swn;
    if (t0 < t1) {
        const CurvePts cv = load_curve(tpos, cpos[c], ctype[c], cweight, c);
        const float pfx = cv.ex(t0), pfy = cv.ey(t0), plx = cv.ex(t1), ply = cv.ey(t1);
        segment_cell_pts(true, pfx, pfy, plx, ply, w, h, yx, swn);
        if (chord) chord[i] = make_float4(pfx, pfy, plx, ply);
    } else {
        yx = kSentinelKey;
        swn = 0;
    }
    key[i] = (int32_t)yx;
    wn[i] = (int8_t)swn;
    epath[i] = path;
    return (uint32_t)bucket_base[path] + bucket_local((int32_t)yx, prow[path], err);
}


// Keys of all entries, counted per (path, row) bucket while they are produced.
__global__ __launch_bounds__(kBlock) void k_gen_fragment(const int2* __restrict__ inter, int64_t N,
                                                         const uint32_t* __restrict__ cpos,
                                                         const uint32_t* __restrict__ ctype,
                                                         const float* __restrict__ cweight,
                                                         const uint32_t* __restrict__ cpath,
                                                         const float2* __restrict__ tpos, int w,
                                                         int h, int32_t* __restrict__ key,
                                                         int8_t* __restrict__ wn,
                                                         uint32_t* __restrict__ epath,
                                                         const int32_t* __restrict__ bucket_base,
                                                         const int2* __restrict__ prow,
                                                         int32_t* __restrict__ hist,
                                                         int32_t* __restrict__ err,
                                                         const int32_t* __restrict__ n_dev,
                                                         const int2* __restrict__ pcanvas,
                                                         float4* __restrict__ chord) {
    const int64_t n = n_dev ? min((int64_t)*n_dev, N) : N;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool valid = i < n;
    uint32_t bucket = 0;
    if (valid && inter[i].x < 0) {
        // a gap of a predicted-allocation layout: no fragment, no bucket
        epath[i] = kNoPath;
        valid = false;
    }
    if (valid) {
        int cw = w, ch = h;
        if (pcanvas) { const int2 cv = pcanvas[cpath[inter[i].x]]; cw = cv.x; ch = cv.y; }
        bucket = gen_one(inter, n, i, cpos, ctype, cweight, cpath, tpos, cw, ch, key, wn, epath, bucket_base, prow, err,
                         chord);
    }
    wave_count(hist, bucket, valid);
}

}  // namespace

hipError_t launch_fill_u32_dev(uint32_t* p, int64_t n_max, const int32_t* n_dev, int64_t mult, int64_t add,
                               uint32_t v, hipStream_t s) {
    if (n_max <= 0) return hipSuccess;
    int64_t g = (n_max + kBlock - 1) / kBlock;
    if (g > 4096) g = 4096;
    k_fill_u32_dev<<<(unsigned)g, kBlock, 0, s>>>(p, n_max, n_dev, mult, add, v);
    return hipGetLastError();
}

hipError_t launch_gather_summary(const SummarySrc& src, int32_t* dst, hipStream_t s) {
    if (src.n <= 0 || src.n > kSummaryMax) return hipErrorInvalidValue;
    k_gather_summary<<<1, 64, 0, s>>>(src, dst);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g > 4096) g = 4096;
    k_fill_u32<<<(unsigned)g, kBlock, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_transform(const float2* pos, const uint32_t* pos_path, uint32_t P, const Mvp& m,
                            int w, int h, float2* tpos, int32_t* visible, int32_t* row_lo, int32_t* row_hi,
                            hipStream_t s, const int2* pcanvas, const float* mvp_dev) {
    if (P == 0) return hipSuccess;
    k_transform<<<grid_for(P), kBlock, 0, s>>>(pos, pos_path, P, m, (float)w, (float)h, (h - 1) / 2, tpos,
                                               visible, row_lo, row_hi, pcanvas, mvp_dev);
    return hipGetLastError();
}

hipError_t launch_count(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                        const float* cweight, const uint32_t* cpath, const int32_t* visible, uint32_t C, int w, int h,
                        float4* cuts, uint32_t* ncuts, int32_t* counts, hipStream_t s, const int2* pcanvas) {
    if (C == 0) return hipSuccess;
    k_count<<<grid_for(C), kBlock, 0, s>>>(tpos, cpos, ctype, cweight, cpath, visible, C, w, h, cuts, ncuts, counts,
                                           pcanvas);
    return hipGetLastError();
}

hipError_t launch_path_segments(const uint32_t* cpath, uint32_t C, const int32_t* offsets,
                                uint32_t Np, int32_t* seg, hipStream_t s) {
    if (C == 0) return launch_fill_u32((uint32_t*)seg, (int64_t)Np + 1, 0u, s);
    k_path_segments<<<grid_for(C), kBlock, 0, s>>>(cpath, C, offsets, Np, seg);
    return hipGetLastError();
}

hipError_t launch_segment_stats(const int32_t* seg, uint32_t Np, int32_t* max_len, hipStream_t s) {
    if (Np == 0) return hipSuccess;
    k_segment_stats<<<grid_for(Np), kBlock, 0, s>>>(seg, Np, max_len);
    return hipGetLastError();
}

hipError_t launch_gen_fragment(const int2* inter, int64_t N, const uint32_t* cpos,
                               const uint32_t* ctype, const float* cweight, const uint32_t* cpath, const float2* tpos,
                               int w, int h, int32_t* key, int8_t* wn, uint32_t* epath,
                               const int32_t* bucket_base, const int2* prow, int32_t* hist, int32_t* err,
                               hipStream_t s, const int32_t* n_dev, const int2* pcanvas, float4* chord) {
    if (N == 0) return hipSuccess;
    k_gen_fragment<<<grid_for(N), kBlock, 0, s>>>(inter, N, cpos, ctype, cweight, cpath, tpos, w, h, key, wn, epath,
                                                  bucket_base, prow, hist, err, n_dev, pcanvas, chord);
    return hipGetLastError();
}

}  // namespace spr
