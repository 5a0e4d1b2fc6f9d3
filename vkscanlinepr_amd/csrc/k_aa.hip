// k_aa.hip -- anti-aliased coverage of the boundary cells (gfx950).
//
// No reference counterpart: the reference leaves the boundary fragments' "comb-like
// multisampling" stencil mask TODO (gen_fragment.comp:202); its draw would fetch one mask per
// cell record (scanlinepr.vert:48-50, draw.w - 1) and set gl_SampleMask from it
// (scanlinepr.frag:13-14), and today sets all samples.  Here every boundary cell record gets a
// 64-bit mask: 16 samples in each of its 2x2 pixels, sample j of a pixel at
// ((7j mod 16) + 1/2)/16 across and (j + 1/2)/16 up (each sample row and column used once, an
// N-rooks "comb"), bit (2 py + px) * 16 + j.  Spans keep full coverage: a pixel that no edge of
// the path crosses lies in no boundary cell of that path, so the winding number is constant over
// it.
//
// The path inside a cell is its chord polygon (the segments pf -> pl between consecutive
// intersections, as gen_fragment forms them).  The winding number at a sample is the cell's
// centre-row winding number at its left edge (the exclusive scan at the cell's first sorted
// entry), plus the changes met on the way to the sample: along the centre row to the sample's
// column, the winding change of each chord crossing the row (gen_fragment's half-open rule),
// then along that column to the sample's height, +1 for a chord going right and -1 for a chord
// going left when moving up (the reverse when moving down).  Both legs stay inside the cell, so
// only the cell's own chords (one sorted run of entries) are met.
//
// One thread per record; the 64 winding numbers live in registers while the thread walks the
// cell's chords once.  Every float expression has a fixed operation order (no contraction).
#include "device_common.hpp"
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kBlock = 256;
constexpr int kCellSamples = 4 * kAaSamples;

__global__ __launch_bounds__(kBlock) void k_cell_cover(const int32_t* __restrict__ cell_first,
                                                       const int32_t* __restrict__ totals, int64_t rec_cap,
                                                       const int32_t* __restrict__ skey,
                                                       const int32_t* __restrict__ sidx,
                                                       const int32_t* __restrict__ winding,
                                                       const uint32_t* __restrict__ epath,
                                                       const int2* __restrict__ inter, int64_t N,
                                                       const float2* __restrict__ tpos,
                                                       const uint32_t* __restrict__ cpos,
                                                       const uint32_t* __restrict__ ctype,
                                                       const float* __restrict__ cweight,
                                                       const uint32_t* __restrict__ rule,
                                                       unsigned long long* __restrict__ cmask) {
    const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t R = min((int64_t)totals[0] + (int64_t)totals[1], rec_cap);
    if (o >= R) return;
    const int32_t f = cell_first[o];
    if (f < 0) {
        cmask[o] = ~0ull;  // a span: every sample
        return;
    }
    const int32_t key = skey[f];
    const uint32_t path = epath[f] & 0x3FFFFFFFu;
    const int rx = (key & 0xFFFF) - 0x7FFF, ry = ((key >> 16) & 0xFFFF) - 0x7FFF;
    const float yc = (float)(ry + 1);
    int32_t W[kCellSamples];
    const int32_t w0 = winding[f];
#pragma unroll
    for (int k = 0; k < kCellSamples; ++k) W[k] = w0;
    for (int64_t g = f; g < N && skey[g] == key && (epath[g] & 0x3FFFFFFFu) == path; ++g) {
        // the chord of entry i, as gen_fragment forms its segment
        const int32_t i = sidx[g];
        const int2 a = inter[i];
        const int2 b = (int64_t)i + 1 < N ? inter[i + 1] : make_int2(-1, 0x3f800000);
        const float t0 = glmax(bitsf((uint32_t)a.y & 0xFFFFFFFCu), 0.0f);
        float t1 = glmax(bitsf((uint32_t)b.y & 0xFFFFFFFCu), 0.0f);
        if (a.x != b.x) t1 = 1.0f;
        const uint32_t c = (uint32_t)a.x;
        const CurvePts cv = load_curve(tpos, cpos[c], ctype[c], cweight, c);
        const float ax = cv.ex(t0), ay = cv.ey(t0), bx = cv.ex(t1), by = cv.ey(t1);
        // the centre-row crossing (sample independent)
        int sr = 0;
        if (ay < yc && yc <= by) sr = -1;
        else if (by < yc && yc <= ay) sr = 1;
        const float xc = sr != 0 ? ax + (yc - ay) * ((bx - ax) / (by - ay)) : 0.0f;
#pragma unroll
        for (int py = 0; py < 2; ++py) {
#pragma unroll
            for (int px = 0; px < 2; ++px) {
#pragma unroll
                for (int j = 0; j < kAaSamples; ++j) {
                    const float xs = (float)(rx + px) + ((float)((7 * j) & 15) + 0.5f) / 16.0f;
                    const float ys = (float)(ry + py) + ((float)j + 0.5f) / 16.0f;
                    int32_t d = 0;
                    if (sr != 0 && xc < xs) d += sr;
                    int dc = 0;
                    if (ax < xs && xs <= bx) dc = 1;
                    else if (bx < xs && xs <= ax) dc = -1;
                    if (dc != 0) {
                        const float yx = ay + (xs - ax) * ((by - ay) / (bx - ax));
                        if (ys > yc) {
                            if (yc <= yx && yx < ys) d += dc;
                        } else if (ys < yx && yx <= yc) {
                            d -= dc;
                        }
                    }
                    W[(py * 2 + px) * kAaSamples + j] += d;
                }
            }
        }
    }
    const bool even_odd = rule[path] == 1u;
    unsigned long long m = 0;
#pragma unroll
    for (int k = 0; k < kCellSamples; ++k) {
        const bool in = even_odd ? (W[k] & 1) != 0 : W[k] != 0;
        m |= (unsigned long long)in << k;
    }
    cmask[o] = m;
}

}  // namespace

hipError_t launch_cell_cover(const int32_t* cell_first, const int32_t* totals, int64_t rec_cap,
                             const int32_t* skey, const int32_t* sidx, const int32_t* winding,
                             const uint32_t* epath, const int2* inter, int64_t N, const float2* tpos,
                             const uint32_t* cpos, const uint32_t* ctype, const float* cweight,
                             const uint32_t* rule, unsigned long long* cmask, hipStream_t s) {
    if (rec_cap <= 0 || N <= 0) return hipSuccess;
    const int64_t g = (rec_cap + kBlock - 1) / kBlock;
    k_cell_cover<<<(unsigned)g, kBlock, 0, s>>>(cell_first, totals, rec_cap, skey, sidx, winding, epath, inter, N,
                                                tpos, cpos, ctype, cweight, rule, cmask);
    return hipGetLastError();
}

}  // namespace spr
