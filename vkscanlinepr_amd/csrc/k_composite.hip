// k_composite.hip -- records to RGBA8 pixels (gfx950).
//
// Replaces the reference's line-list draw (scanlinepr.vert:18-50, scanlinepr.frag:7-14, pipeline
// state scanline_rasterizer.cpp:881-944): every record (yx, width, fill) covers canvas rows y and
// y+1 over columns [x, x+width), records are applied in order (blending disabled: the later one
// wins), and canvas row y lands on image row h-1-y.  All records sit on even rows, so the image
// is a stack of independent two-pixel "cell rows".
//
// One workgroup owns one cell row, in column chunks of 4096; each thread keeps PIX consecutive
// pixels of the chunk in registers (16; 8 in the anti-aliased variants, which keep both rows of a
// cell row, in workgroups of 512 threads, so that they run as many waves per SIMD).  The per-column event words
// live in LDS column-interleaved (column c at (c % PIX) * (4096 / PIX) + c / PIX), so the 64 lanes
// of a wave touching "their" j-th column hit consecutive words instead of one bank.  The (path, row) bucket table gives each path's records on
// this row as one contiguous range of non-overlapping compact records (x | width << 16) with one
// colour.  Paths are taken in order, 64 at a time: a record toggles its path's bit at its first
// column and at its end column in a 64-bit event word per column (LDS), and a prefix XOR over
// the columns turns the events into each pixel's coverage mask (the records of one path never
// overlap, so the XOR of the toggles left of a pixel is exactly "inside one of them").  Each
// thread then folds the set bits of its pixels in path order (overwrite: the last one wins;
// src-over: d = s*a + d*(1-a)).  A record costs two LDS operations whatever its width, and each
// image row is written once.
//
// Gradient paints (no reference counterpart: the reference draws solid colours only): a path
// with a paint takes, at each covered pixel, the paint's colour at that pixel's centre, packed
// like a solid fill, so the two rows of a cell row get their own colours; the variant that
// evaluates them keeps a second row of pixels per thread.
//
// Anti-aliased coverage (src-over): a boundary cell record whose 64-sample mask (k_aa.hip) is not
// full appends, at each of its two columns, an entry (path of the batch, covered samples of row y,
// of row y + 1) to that column's list in LDS (at most kAaSlots per column); the fold scales such
// a path's premultiplied colour by count/16 in each row.  A batch whose lists would overflow is
// redone with half as many paths (kAaSlots paths can never overflow).
#include "device_common.hpp"
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kChunk = 4096;               // columns per pass (a workgroup of kChunk / PIX threads)
constexpr int kPixOne = 16;                // pixels per thread of the variants that keep one row
constexpr int kBatch = 64;                 // paths per coverage mask
constexpr int kAaSlots = 4;                // partial-coverage entries per column and batch
constexpr int kList = 512;                 // paths listed before their batches are folded
constexpr int kListStep = 256;             // candidate paths examined per listing step
constexpr int kUnroll = 4;                 // records loaded per thread before applying
typedef unsigned long long u64;

// LDS slot of chunk column c (0 <= c < kChunk) with PIX consecutive pixels per thread: thread t's
// column t * PIX + j sits at j * (kChunk / PIX) + t.
template <int PIX>
__device__ inline int ev_slot(int c) { return (c % PIX) * (kChunk / PIX) + c / PIX; }

__device__ inline u64 shfl_up_u64(u64 v, int d) {
    const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    return (u64)(uint32_t)__shfl_up((int)lo, d) | ((u64)(uint32_t)__shfl_up((int)hi, d) << 32);
}

// Exclusive prefix XOR of one value per thread over a workgroup of BLOCK threads.
template <int BLOCK>
__device__ inline u64 block_exclusive_xor(u64 v, u64* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u64 inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u64 u = shfl_up_u64(inc, d);
        if (lane >= d) inc ^= u;
    }
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    u64 off = 0;
#pragma unroll
    for (int k = 0; k < BLOCK / 64; ++k)
        if (k < wave) off ^= lds[k];
    __syncthreads();
    return off ^ inc ^ v;
}

// A gradient's colour at the centre of canvas pixel (x, y), packed like a solid fill (channel
// times 255 truncated, alpha times opacity, 0 when the alpha byte is 0).  Every float operation
// is written out in order (the library is built without contraction).
__device__ __forceinline__ uint32_t paint_color(const PaintArgs& pa, uint32_t pi, int x, int y) {
    const float* h = pa.inv;
    const float px = (float)x + 0.5f, py = (float)y + 0.5f;
    const float u = (h[0] * px + h[1] * py) + h[2];
    const float v = (h[3] * px + h[4] * py) + h[5];
    const float q = (h[6] * px + h[7] * py) + h[8];
    const float sx = u / q, sy = v / q;
    const Paint& g = pa.paints[pi];
    const float gx = (g.xform[0] * sx + g.xform[1] * sy) + g.xform[2];
    const float gy = (g.xform[3] * sx + g.xform[4] * sy) + g.xform[5];
    float t;
    bool degenerate = false;
    if (g.kind == kPaintLinear) {
        const float dx = g.geom[2] - g.geom[0], dy = g.geom[3] - g.geom[1];
        const float den = dx * dx + dy * dy;
        if (den > 0.0f) {
            t = ((gx - g.geom[0]) * dx + (gy - g.geom[1]) * dy) / den;
        } else {
            t = 1.0f;  // both points equal: the last stop's colour
            degenerate = true;
        }
    } else {
        // |g - f - t (c - f)| = t r: (d.d - r^2) t^2 - 2 (q.d) t + q.q = 0, a < 0, largest root
        const float qx = gx - g.geom[2], qy = gy - g.geom[3];
        const float dx = g.geom[0] - g.geom[2], dy = g.geom[1] - g.geom[3];
        const float r = g.geom[4];
        const float a = (dx * dx + dy * dy) - r * r;
        const float b = qx * dx + qy * dy;
        const float c = qx * qx + qy * qy;
        t = (b - sqrtf(b * b - a * c)) / a;
    }
    if (!(t == t)) t = 0.0f;
    if (!degenerate) {
        if (g.spread == kSpreadPad) {
            t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        } else if (g.spread == kSpreadRepeat) {
            t = t - floorf(t);
        } else if (g.spread == kSpreadReflect) {
            const float m = t - 2.0f * floorf(t * 0.5f);
            t = m > 1.0f ? 2.0f - m : m;
        } else if (t < 0.0f || t > 1.0f) {
            return 0u;  // transparent outside [0, 1]
        }
        if (!(t == t)) t = 0.0f;
    }
    const float* st = pa.stops + 5 * (size_t)g.first_stop;
    const uint32_t n = g.n_stops;
    float c[4];
    if (t <= st[0]) {
        c[0] = st[1]; c[1] = st[2]; c[2] = st[3]; c[3] = st[4];
    } else if (t >= st[5 * (n - 1)]) {
        const float* e = st + 5 * (n - 1);
        c[0] = e[1]; c[1] = e[2]; c[2] = e[3]; c[3] = e[4];
    } else {
        uint32_t i = 0;
        while (i + 1 < n && st[5 * (i + 1)] <= t) ++i;  // off[i] <= t < off[i + 1]
        const float* e0 = st + 5 * i;
        const float* e1 = e0 + 5;
        const float f = (t - e0[0]) / (e1[0] - e0[0]);
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = lerp(e0[1 + k], e1[1 + k], f);
    }
    c[3] = c[3] * g.opacity;
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float vv = c[k] * 255.0f;
        const int iv = (vv != vv) ? 0 : (vv <= 0.0f ? 0 : (vv >= 255.0f ? 255 : (int)vv));
        out |= (uint32_t)iv << (8 * k);
    }
    return (out & 0xFF000000u) ? out : 0u;
}

// The colours of pixels (x, y) and (x, y + 1), the two rows of a cell row.
__device__ __noinline__ uint2 paint_pair(const PaintArgs pa, uint32_t pi, int x, int y) {
    return make_uint2(paint_color(pa, pi, x, y), paint_color(pa, pi, x, y + 1));
}

// (r*a, g*a, b*a, a) of a packed colour, as the blend uses it.
__device__ inline float4 premul(uint32_t col) {
    const float a = (float)((col >> 24) & 0xFFu) / 255.0f;
    return make_float4(((float)(col & 0xFFu) / 255.0f) * a, ((float)((col >> 8) & 0xFFu) / 255.0f) * a,
                       ((float)((col >> 16) & 0xFFu) / 255.0f) * a, a);
}

// A premultiplied colour times a coverage share.
__device__ inline float4 scale(const float4 s, float c) { return make_float4(s.x * c, s.y * c, s.z * c, s.w * c); }

__device__ inline void blend(float* d, const float4 sa) {
    const float ia = 1.0f - sa.w;
    d[0] = sa.x + d[0] * ia;
    d[1] = sa.y + d[1] * ia;
    d[2] = sa.z + d[2] * ia;
    d[3] = sa.w + d[3] * ia;
}

template <int COMPOSE, int GRAD, int AA, int PIX>
__global__ __launch_bounds__(kChunk / PIX) __attribute__((amdgpu_waves_per_eu((AA && !GRAD) ? 4 : 1))) void k_composite(const uint32_t* __restrict__ records,
                                                      const int2* __restrict__ brec,
                                                      const int32_t* __restrict__ bucket_base,
                                                      const int2* __restrict__ prow,
                                                      const uint32_t* __restrict__ fill, uint32_t Np,
                                                      int w, int h, int64_t rec_cap,
                                                      uint8_t* __restrict__ image,
                                                      const DocDev* __restrict__ docs, int n_docs,
                                                      const PaintArgs pa,
                                                      const unsigned long long* __restrict__ cmask) {
    constexpr int TWO = (GRAD || AA) ? 1 : 0;            // the two rows of a cell row differ
    constexpr int BLOCK = kChunk / PIX;                  // threads; each keeps PIX consecutive pixels
    static_assert(PIX % 4 == 0 && BLOCK % 64 == 0 && BLOCK >= kListStep, "composite shape");
    // partial coverage: entry k of column c at lslot[k * kChunk + ev_slot<PIX>(c)], q | n0 << 6 | n1 << 11;
    // byte c of lcnt counts the column's entries (a thread's 16 columns in one 16-byte word)
    __shared__ uint16_t lslot[AA ? kAaSlots * kChunk : 1];
    __shared__ uint32_t lcnt[AA ? kChunk / 4 : 1];
    __shared__ int32_t lovf;
    __shared__ u64 ev[kChunk];
    __shared__ int2 list[kList];
    __shared__ uint32_t lcol[kList];
    __shared__ uint32_t lpaint[GRAD ? kList : 1];  // each listed path's paint (kSolidPaint: its colour)
    __shared__ u64 lgrad;                          // the batch's paths that have a paint
    __shared__ float4 lsa[kBatch];   // (r*a, g*a, b*a, a) of each path of the batch, as the blend uses them
    __shared__ int32_t wcount[BLOCK / 64];
    __shared__ int32_t bpre[kBatch + 1];
    __shared__ u64 xlds[BLOCK / 64];
    int r = blockIdx.x;
    uint32_t pbeg = 0, pend = Np;
    if (docs) {
        // batch: find this cell row's document (documents are numbered by first row)
        int lo = 0, hi = n_docs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (docs[mid].row0 <= r) lo = mid; else hi = mid;
        }
        const DocDev d = docs[lo];
        r -= d.row0;
        w = d.w; h = d.h;
        pbeg = (uint32_t)d.p0; pend = (uint32_t)d.p1;
        image += d.img_off;
    }
    const int y = 2 * r;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int xb = tid * PIX;  // first column of this thread inside the chunk
    for (int c0 = 0; c0 < w; c0 += kChunk) {
        const int cw = w - c0 < kChunk ? w - c0 : kChunk;
        uint32_t pix[PIX], pix1[TWO ? PIX : 1];  // rows y and, with paints or coverage, y + 1
        float acc[COMPOSE ? PIX : 1][4], acc1[(COMPOSE && TWO) ? PIX : 1][4];
#pragma unroll
        for (int j = 0; j < PIX; ++j) {
            pix[j] = 0xFFFFFFFFu;
            if constexpr (TWO != 0) pix1[j] = 0xFFFFFFFFu;
            if constexpr (COMPOSE != 0) { acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 1.0f; }
            if constexpr (COMPOSE != 0 && TWO != 0) { acc1[j][0] = acc1[j][1] = acc1[j][2] = acc1[j][3] = 1.0f; }
        }
        uint32_t p0 = pbeg;
        while (p0 < pend) {
            // ordered list of the paths with records on this cell row, 256 candidates at a time
            // until the list could overflow: batches are then full whatever the path numbering
            int n = 0;
            while (p0 < pend && n + kListStep <= kList) {
                const uint32_t p = p0 + tid;
                int2 rng = make_int2(0, 0);
                if (tid < kListStep && p < pend) {
                    const int2 pr = prow[p];
                    const int loc = r == 0 ? pr.y + 1 : ((r >= pr.x && r < pr.x + pr.y) ? r - pr.x : -1);
                    if (loc >= 0) rng = brec[bucket_base[p] + loc];
                    // ranges never reach past the records actually stored (a frame whose record
                    // budget overflowed must still read in bounds)
                    rng.y = (int64_t)rng.y < rec_cap ? rng.y : (int32_t)rec_cap;
                    rng.x = rng.x < rng.y ? rng.x : rng.y;
                }
                const bool keep = rng.y > rng.x;
                const u64 m = __ballot(keep);
                if (lane == 0) wcount[wave] = (int32_t)__popcll(m);
                __syncthreads();
                int off = n, add = 0;
#pragma unroll
                for (int k = 0; k < BLOCK / 64; ++k) {
                    if (k < wave) off += wcount[k];
                    add += wcount[k];
                }
                if (keep) {
                    const int slot = off + (int)__popcll(m & ((1ull << lane) - 1ull));
                    list[slot] = rng;
                    lcol[slot] = fill[p];
                    if constexpr (GRAD != 0) lpaint[slot] = pa.path_paint[p];
                }
                __syncthreads();
                n += add;
                p0 += kListStep;
            }
            int qn = 0;
            for (int q0 = 0; q0 < n; q0 += qn) {
                qn = n - q0 < kBatch ? n - q0 : kBatch;
            batch:
                if constexpr (AA != 0) {
                    for (int k = tid; k < kChunk / 4; k += BLOCK) lcnt[k] = 0u;
                    if (tid == 0) lovf = 0;
                }
                // clear the event words; the batch's colours; its records as one flat index space
#pragma unroll
                for (int j = 0; j < PIX; ++j) ev[j * BLOCK + tid] = 0ull;
                if (wave == 0) {
                    if constexpr (COMPOSE != 0) {
                        if (lane < qn) lsa[lane] = premul(lcol[q0 + lane]);
                    }
                    if constexpr (GRAD != 0) {
                        const u64 gm = __ballot(lane < qn && lpaint[q0 + lane] != kSolidPaint);
                        if (lane == 0) lgrad = gm;
                    }
                    const int2 rg = lane < qn ? list[q0 + lane] : make_int2(0, 0);
                    int inc = rg.y - rg.x;
#pragma unroll
                    for (int d = 1; d < 64; d <<= 1) {
                        const int u = __shfl_up(inc, d);
                        if (lane >= d) inc += u;
                    }
                    bpre[lane + 1] = inc;
                    if (lane == 0) bpre[0] = 0;
                }
                __syncthreads();
                const int total = bpre[qn];
                for (int f0 = tid; f0 < total; f0 += BLOCK * kUnroll) {
                    uint32_t rec[kUnroll];
                    int qi[kUnroll];
                    int64_t ri[kUnroll];
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        const int f = f0 + u * BLOCK;
                        qi[u] = -1;
                        if (f < total) {
                            int lo = 0, hi = qn;
                            while (hi - lo > 1) {
                                const int mid = (lo + hi) >> 1;
                                if (bpre[mid] <= f) lo = mid; else hi = mid;
                            }
                            qi[u] = lo;
                            ri[u] = (int64_t)list[q0 + lo].x + (f - bpre[lo]);
                            rec[u] = records[ri[u]];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        if (qi[u] < 0) continue;
                        const u64 bit = 1ull << qi[u];
                        int xs = (int)(rec[u] & 0xFFFFu) - c0;
                        int xe = xs + (int)(rec[u] >> 16);
                        xs = xs < 0 ? 0 : xs;
                        xe = xe > cw ? cw : xe;
                        if (xs < xe) {
                            atomicXor(&ev[ev_slot<PIX>(xs)], bit);
                            if (xe < cw) atomicXor(&ev[ev_slot<PIX>(xe)], bit);  // ends past the chunk change nothing
                        }
                        if constexpr (AA != 0) {
                            const unsigned long long cm = cmask[ri[u]];
                            if (cm != ~0ull) {  // a boundary cell (2 wide): its pixels' sample counts
                                const int x0 = (int)(rec[u] & 0xFFFFu) - c0;
#pragma unroll
                                for (int px = 0; px < 2; ++px) {
                                    const int col = x0 + px;
                                    if (col < 0 || col >= cw) continue;
                                    const uint32_t n0 = __popcll((cm >> (16 * px)) & 0xFFFFull);
                                    const uint32_t n1 = __popcll((cm >> (16 * (2 + px))) & 0xFFFFull);
                                    const int sl = ev_slot<PIX>(col);
                                    const uint32_t k = (atomicAdd(&lcnt[col >> 2], 1u << (8 * (col & 3))) >>
                                                        (8 * (col & 3))) & 0xFFu;
                                    if (k < (uint32_t)kAaSlots)
                                        lslot[k * kChunk + sl] = (uint16_t)((uint32_t)qi[u] | (n0 << 6) | (n1 << 11));
                                    else
                                        lovf = 1;
                                }
                            }
                        }
                    }
                }
                __syncthreads();
                if constexpr (AA != 0) {
                    if (lovf) {  // some column has more partial cells than slots: fewer paths
                        qn = qn / 2;
                        __syncthreads();
                        goto batch;
                    }
                }
                // coverage = prefix XOR of the events, then the fold in path order
                u64 run = 0;
#pragma unroll
                for (int j = 0; j < PIX; ++j) run ^= ev[j * BLOCK + tid];
                u64 cov = block_exclusive_xor<BLOCK>(run, xlds);
                const u64 gmask = GRAD ? lgrad : 0ull;
                uint32_t cnt4[PIX / 4];  // AA: the partial-entry counts of the thread's columns, bytes
#pragma unroll
                for (int k = 0; k < PIX / 4; ++k) cnt4[k] = AA ? lcnt[tid * (PIX / 4) + k] : 0u;
#pragma unroll
                for (int j = 0; j < PIX; ++j) {
                    cov ^= ev[j * BLOCK + tid];
                    u64 mm = cov;
                    if (mm == 0ull) continue;
                    const int x = c0 + xb + j;  // the pixel's canvas column
                    // this pixel's partial-coverage entries (AA), and the paths they belong to
                    uint32_t pe[AA ? kAaSlots : 1];
                    uint32_t np = 0;
                    u64 pm = 0ull;
                    if constexpr (AA != 0) {
                        np = (cnt4[j >> 2] >> (8 * (j & 3))) & 0xFFu;
#pragma unroll
                        for (int k = 0; k < kAaSlots; ++k) pe[k] = 0u;
                        if (np) {
                            const int sl = j * BLOCK + tid;
#pragma unroll
                            for (int k = 0; k < kAaSlots; ++k) {
                                if ((uint32_t)k < np) {
                                    pe[k] = (uint32_t)lslot[k * kChunk + sl];
                                    pm |= 1ull << (pe[k] & 63u);
                                }
                            }
                        }
                    }
                    if constexpr (COMPOSE == 0) {
                        const int q = 63 - __clzll((long long)mm);
                        if (GRAD && ((gmask >> q) & 1ull)) {
                            const uint2 pc = paint_pair(pa, lpaint[q0 + q], x, y);
                            pix[j] = pc.x;
                            pix1[j] = pc.y;
                        } else {
                            pix[j] = lcol[q0 + q];
                            if constexpr (GRAD != 0) pix1[j] = pix[j];
                        }
                    } else {
                        // src-over, d = s*a + d*(1-a) per channel, the paths in order
                        // the mask's two words in turn: 32-bit bit scans and clears
#pragma unroll
                        for (int hw = 0; hw < 2; ++hw) {
                            uint32_t m32 = (uint32_t)(mm >> (32 * hw));
                            const float4* L = lsa + 32 * hw;
                            while (m32) {
                                const int q = __builtin_ctz(m32);
                                m32 &= m32 - 1u;
                                if constexpr (TWO != 0) {
                                    float4 s0 = L[q], s1 = L[q];
                                    if (GRAD && ((gmask >> (32 * hw + q)) & 1ull)) {
                                        const uint2 pc = paint_pair(pa, lpaint[q0 + 32 * hw + q], x, y);
                                        s0 = premul(pc.x);
                                        s1 = premul(pc.y);
                                    }
                                    if constexpr (AA != 0) {
                                        const uint32_t qq = (uint32_t)(32 * hw + q);
                                        if ((pm >> qq) & 1ull) {
                                            uint32_t e = 0;
#pragma unroll
                                            for (int k = 0; k < kAaSlots; ++k)
                                                if ((uint32_t)k < np && (pe[k] & 63u) == qq) e = pe[k];
                                            s0 = scale(s0, (float)((e >> 6) & 31u) / 16.0f);
                                            s1 = scale(s1, (float)((e >> 11) & 31u) / 16.0f);
                                        }
                                    }
                                    blend(acc[j], s0);
                                    blend(acc1[j], s1);
                                } else {
                                    blend(acc[j], L[q]);
                                }
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        auto round8 = [](const float* a) -> uint32_t {
            uint32_t o = 0;
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
                float q = a[ch] * 255.0f + 0.5f;
                q = q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q);
                o |= (uint32_t)q << (8 * ch);
            }
            return o;
        };
        if constexpr (COMPOSE != 0) {
#pragma unroll
            for (int j = 0; j < PIX; ++j) {
                pix[j] = round8(acc[j]);
                if constexpr (TWO != 0) pix1[j] = round8(acc1[j]);
            }
        }
        for (int yy = y; yy < y + 2 && yy < h; ++yy) {
            uint32_t* row = reinterpret_cast<uint32_t*>(image) + (int64_t)(h - 1 - yy) * w + c0;
            uint32_t out[PIX];
#pragma unroll
            for (int j = 0; j < PIX; ++j) {
                if constexpr (TWO != 0) out[j] = yy == y ? pix[j] : pix1[j];
                else out[j] = pix[j];
            }
            if (xb + PIX <= cw && (((uintptr_t)(row + xb)) & 15u) == 0) {
                uint4* q = reinterpret_cast<uint4*>(row + xb);
#pragma unroll
                for (int k = 0; k < PIX / 4; ++k)
                    q[k] = make_uint4(out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3]);
            } else {
#pragma unroll
                for (int j = 0; j < PIX; ++j)
                    if (xb + j < cw) row[xb + j] = out[j];
            }
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_composite(const uint32_t* records, const int2* brec, const int32_t* bucket_base,
                            const int2* prow, const uint32_t* fill, uint32_t Np, int w, int h, int compose,
                            uint8_t* image, hipStream_t s, int64_t rec_cap, const DocDev* docs, int n_docs,
                            int total_rows, const PaintArgs* paint, const unsigned long long* cmask) {
    const int rows = docs ? total_rows : (h + 1) / 2;
    if (rows == 0 || (!docs && w == 0)) return hipSuccess;
    if (cmask && !compose) return hipErrorInvalidValue;  // coverage needs the src-over blend
    const PaintArgs pa = paint ? *paint : PaintArgs{nullptr, nullptr, nullptr, nullptr};
// pixels per thread: the coverage variants keep two rows of accumulators, so they run 8 per
// thread in workgroups of 512 threads (without paints: at most 128 VGPRs, 4 waves per SIMD)
#define SPR_COMPOSITE(C, G, A)                                                                              \
    k_composite<C, G, A, (A) ? 8 : kPixOne><<<rows, kChunk / ((A) ? 8 : kPixOne), 0, s>>>(                    \
        records, brec, bucket_base, prow, fill, Np, w, h, rec_cap, image, docs, n_docs, pa, cmask)
    if (cmask) {
        if (paint) SPR_COMPOSITE(1, 1, 1); else SPR_COMPOSITE(1, 0, 1);
    } else if (compose) {
        if (paint) SPR_COMPOSITE(1, 1, 0); else SPR_COMPOSITE(1, 0, 0);
    } else {
        if (paint) SPR_COMPOSITE(0, 1, 0); else SPR_COMPOSITE(0, 0, 0);
    }
#undef SPR_COMPOSITE
    return hipGetLastError();
}

}  // namespace spr
