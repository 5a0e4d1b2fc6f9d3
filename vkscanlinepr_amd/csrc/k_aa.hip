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

// The first sample column c (0..15) of a pixel whose x, x0 + (c + 1/2)/16, lies above v; 16 when
// none does (NaN included).  x0 is an integer and every sample x is exact, so for v within a few
// columns of x0, u = 16 (v - x0) - 1/2 is exact and the columns above v are those with c > u; the
// two boundary compares keep the answer exact wherever u is not.
__device__ inline int first_above(float v, float x0) {
    if (!(v == v)) return 16;
    const float u = fminf(fmaxf((v - x0) * 16.0f - 0.5f, -2.0f), 17.0f);
    int k = (int)floorf(u) + 1;
    k = k < 0 ? 0 : (k > 16 ? 16 : k);
    if (k < 16 && !(x0 + ((float)k + 0.5f) / 16.0f > v)) ++k;
    if (k > 0 && x0 + ((float)(k - 1) + 0.5f) / 16.0f > v) --k;
    return k;
}

// Lane t covers pixel t % 4 of record t / 4.  Its 16 winding numbers are kept by sample column c
// (sample j = 7c mod 16 sits in column c, (7 * 7) mod 16 = 1), so every chord is one
// straight-line pass over the columns with fixed register indices:
//  - the centre-row crossing adds its change to the columns above its x (a suffix c >= kr);
//  - a chord spanning column c (ax < xs <= bx or bx < xs <= ax: c in [ka, kb)) crosses it at
//    height yx = ay + (xs - ax) * slope, which counts in the upper pixel when yc <= yx < ys and
//    in the lower one (with the opposite sign) when ys < yx <= yc; both tests are one form on
//    values multiplied by +1 or -1 (exact).
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
    const float sy = py ? 1.0f : -1.0f;  // the upper pixel tests [yc, ys), the lower one (ys, yc]
    const float syc = yc * sy;
    int32_t W[kAaSamples];  // by sample column
    const int32_t w0 = winding[f];
#pragma unroll
    for (int c = 0; c < kAaSamples; ++c) W[c] = w0;
    auto add_chord = [&](const float4 ch) {
        const float ax = ch.x, ay = ch.y, bx = ch.z, by = ch.w;
        // the centre-row crossing: its change at the columns right of its x
        const int sr = (ay < yc && yc <= by) ? -1 : ((by < yc && yc <= ay) ? 1 : 0);
        const float xc = ax + (yc - ay) * ((bx - ax) / (by - ay));
        const int kr = sr != 0 ? first_above(xc, x0) : 16;
        // the spanned columns and the sign of their crossings in this pixel
        const bool rightward = ax < bx;
        const float lo = rightward ? ax : bx, hi = rightward ? bx : ax;
        const int ka = first_above(lo, x0), kb = hi == hi ? first_above(hi, x0) : 0;
        const int sc = (rightward ? 1 : -1) * (py ? 1 : -1);
        const float sl = (by - ay) / (bx - ax);
#pragma unroll
        for (int c = 0; c < kAaSamples; ++c) {
            const float xs = x0 + ((float)c + 0.5f) / 16.0f;
            const float ys = y0 + ((float)((7 * c) & 15) + 0.5f) / 16.0f;
            const float yx = ay + (xs - ax) * sl;
            const float v = yx * sy;
            const bool hit = (c >= ka) & (c < kb) & (syc <= v) & (v < ys * sy);  // no short circuit: no branches
            W[c] += (c >= kr ? sr : 0) + (hit ? sc : 0);
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
        const int64_t g = f + u < N ? f + u : N - 1;  // loads in range, the run test below stops at N
        const int32_t kg = skey[g];
        const uint32_t pg = epath[g] & 0x3FFFFFFFu;
        ii[u] = sidx[g];
        more = more && f + u < N && kg == key && pg == path;
        nv += more ? 1 : 0;
    }
    float4 chs[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) chs[u] = chord[u < nv ? ii[u] : ii[0]];
#pragma unroll
    for (int u = 0; u < kPre; ++u)
        if (u < nv) add_chord(chs[u]);
    if (nv == kPre)
        for (int64_t g = f + kPre; g < N && skey[g] == key && (epath[g] & 0x3FFFFFFFu) == path; ++g)
            add_chord(chord[sidx[g]]);
    const bool even_odd = rule[path] == 1u;
    uint32_t m = 0;
#pragma unroll
    for (int c = 0; c < kAaSamples; ++c) {
        const bool in = even_odd ? (W[c] & 1) != 0 : W[c] != 0;
        m |= (uint32_t)in << ((7 * c) & 15);  // column c holds sample 7c mod 16
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
