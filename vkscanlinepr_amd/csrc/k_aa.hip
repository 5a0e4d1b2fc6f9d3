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
// intersections, as gen_fragment forms them; the fragment stage stores them for these frames).  The winding number at a sample is the cell's
// centre-row winding number at its left edge (the exclusive scan at the cell's first sorted
// entry), plus the changes met on the way to the sample: along the centre row to the sample's
// column, the winding change of each chord crossing the row (gen_fragment's half-open rule),
// then along that column to the sample's height, +1 for a chord going right and -1 for a chord
// going left when moving up (the reverse when moving down).  Both legs stay inside the cell, so
// only the cell's own chords (one sorted run of entries) are met.
//
// Four lanes per record, one per pixel: each walks the cell's chords (the lanes of a record load
// the same lines) with its 16 winding numbers in registers and stores its 16 mask bits.  Every
// float expression has a fixed operation order (no contraction).
#include "device_common.hpp"
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kBlock = 256;

// Lane t covers pixel t % 4 of record t / 4.
__device__ inline void cover_one(int64_t t, const int32_t* __restrict__ cell_first, const int32_t* __restrict__ skey,
                                 const int32_t* __restrict__ sidx, const int32_t* __restrict__ winding,
                                 const uint32_t* __restrict__ epath, const float4* __restrict__ chord, int64_t N,
                                 const uint32_t* __restrict__ rule, uint16_t* __restrict__ cmask16) {
    const int64_t o = t >> 2;            // the record
    const int pix = (int)(t & 3);        // its pixel: px = pix & 1, py = pix >> 1
    const int32_t f = cell_first[o];
    if (f < 0) {
        cmask16[t] = 0xFFFFu;  // a span: every sample
        return;
    }
    const int32_t key = skey[f];
    const uint32_t path = epath[f] & 0x3FFFFFFFu;
    const int rx = (key & 0xFFFF) - 0x7FFF, ry = ((key >> 16) & 0xFFFF) - 0x7FFF;
    const int px = pix & 1, py = pix >> 1;
    const float yc = (float)(ry + 1);
    const float x0 = (float)(rx + px), y0 = (float)(ry + py);
    int32_t W[kAaSamples];
    const int32_t w0 = winding[f];
#pragma unroll
    for (int k = 0; k < kAaSamples; ++k) W[k] = w0;
    // one chord's winding changes at the 16 samples
    auto add_chord = [&](const float4 ch) {
        const float ax = ch.x, ay = ch.y, bx = ch.z, by = ch.w;
        // the centre-row crossing (sample independent)
        int sr = 0;
        if (ay < yc && yc <= by) sr = -1;
        else if (by < yc && yc <= ay) sr = 1;
        const float xc = sr != 0 ? ax + (yc - ay) * ((bx - ax) / (by - ay)) : 0.0f;
        // the column crossings' slope, once per chord (read only when the chord spans a column)
        const float sl = (by - ay) / (bx - ax);
#pragma unroll
        for (int j = 0; j < kAaSamples; ++j) {
            const float xs = x0 + ((float)((7 * j) & 15) + 0.5f) / 16.0f;
            const float ys = y0 + ((float)j + 0.5f) / 16.0f;
            int32_t d = 0;
            if (sr != 0 && xc < xs) d += sr;
            int dc = 0;
            if (ax < xs && xs <= bx) dc = 1;
            else if (bx < xs && xs <= ax) dc = -1;
            if (dc != 0) {
                const float yx = ay + (xs - ax) * sl;
                if (ys > yc) {
                    if (yc <= yx && yx < ys) d += dc;
                } else if (ys < yx && yx <= yc) {
                    d -= dc;
                }
            }
            W[j] += d;
        }
    };
    // the cell's first kPre entries fetched together (keys, paths and indices, then chords): a
    // cell rarely has more, which the loop after them walks one at a time
    constexpr int kPre = 4;
    int32_t ii[kPre];
    int nv = 0;
    bool more = true;
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const int64_t g = f + u;
        const bool inb = g < N;
        const int32_t kg = inb ? skey[g] : ~key;
        const uint32_t pg = inb ? (epath[g] & 0x3FFFFFFFu) : ~path;
        ii[u] = inb ? sidx[g] : 0;
        more = more && kg == key && pg == path;
        nv += more ? 1 : 0;
    }
    float4 chs[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) chs[u] = u < nv ? chord[ii[u]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kPre; ++u)
        if (u < nv) add_chord(chs[u]);
    if (nv == kPre)
        for (int64_t g = f + kPre; g < N && skey[g] == key && (epath[g] & 0x3FFFFFFFu) == path; ++g)
            add_chord(chord[sidx[g]]);
    const bool even_odd = rule[path] == 1u;
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kAaSamples; ++k) {
        const bool in = even_odd ? (W[k] & 1) != 0 : W[k] != 0;
        m |= (uint32_t)in << k;
    }
    cmask16[t] = (uint16_t)m;  // bits (2 py + px) * 16 .. + 15 of the record's 64-bit mask
}

__global__ __launch_bounds__(kBlock) void k_cell_cover(const int32_t* __restrict__ cell_first,
                                                       const int32_t* __restrict__ totals, int64_t rec_cap,
                                                       const int32_t* __restrict__ skey,
                                                       const int32_t* __restrict__ sidx,
                                                       const int32_t* __restrict__ winding,
                                                       const uint32_t* __restrict__ epath,
                                                       const float4* __restrict__ chord, int64_t N,
                                                       const uint32_t* __restrict__ rule,
                                                       uint16_t* __restrict__ cmask16) {
    // grid-stride over the records actually emitted (the launch cannot know their number)
    const int64_t R = min((int64_t)totals[0] + (int64_t)totals[1], rec_cap);
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < 4 * R; t += (int64_t)gridDim.x * kBlock)
        cover_one(t, cell_first, skey, sidx, winding, epath, chord, N, rule, cmask16);
}

}  // namespace

hipError_t launch_cell_cover(const int32_t* cell_first, const int32_t* totals, int64_t rec_cap,
                             const int32_t* skey, const int32_t* sidx, const int32_t* winding,
                             const uint32_t* epath, const float4* chord, int64_t N, const uint32_t* rule,
                             unsigned long long* cmask, hipStream_t s) {
    if (rec_cap <= 0 || N <= 0) return hipSuccess;
    int64_t g = (4 * rec_cap + kBlock - 1) / kBlock;
    g = g < 8192 ? g : 8192;  // 32 waves per CU; the blocks stride over the live records
    k_cell_cover<<<(unsigned)g, kBlock, 0, s>>>(cell_first, totals, rec_cap, skey, sidx, winding, epath, chord, N,
                                                rule, reinterpret_cast<uint16_t*>(cmask));
    return hipGetLastError();
}

}  // namespace spr
