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
// k_cell_cover gives spans their full masks, lists each workgroup's cells by number of chords,
// longest first, so that the cells of one wave have (nearly) the same number, and then runs four
// lanes per listed cell, one per pixel: each walks the cell's chords (the lanes of a cell load the
// same lines) with its 16 winding numbers in registers and stores its 16 mask bits.  Every float
// expression has a fixed operation order (no contraction).
#include "device_common.hpp"
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kBlock = 256;
// Cell records by number of chords (the run of sorted entries with the cell's key and path):
// class k < kCoverClasses - 1 holds the cells of k chords, the last class those of more.
constexpr int kCoverClasses = 32;
constexpr int kChunk = 4 * kBlock;  // records per workgroup pass of the cover kernel

// A lane covers one pixel of a cell.  Its 16 winding numbers are kept by sample column c (sample
// j = 7c mod 16 sits in column c, (7 * 7) mod 16 = 1), so every chord is one straight-line pass
// over the columns with fixed register indices, each column tested on its sample x, xs:
//  - the centre-row crossing adds its change to the columns whose xs lies above its x;
//  - a chord spanning column c (ax < xs <= bx or bx < xs <= ax) crosses it at height
//    yx = ay + (xs - ax) * slope, which counts in the upper pixel when yc <= yx < ys and in the
//    lower one (with the opposite sign) when ys < yx <= yc; both tests are one form on values
//    multiplied by +1 or -1 (exact).
__device__ inline void cover_one(int64_t o, int32_t f, int cls, int pix, const int32_t* __restrict__ skey,
                                 const int32_t* __restrict__ sidx, const int32_t* __restrict__ winding,
                                 const uint32_t* __restrict__ epath, const float4* __restrict__ chord, int64_t N,
                                 const uint32_t* __restrict__ rule, uint16_t* __restrict__ cmask16) {
    // o: a cell record, f: its first sorted entry, cls: its chord class (run_class), pix: its
    // pixel, px = pix & 1, py = pix >> 1
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
        // the centre-row crossing: its change sr at the columns right of its x, xc (sr = 0 for a
        // chord that does not cross the row, whose xc is never used)
        const int sr = (ay < yc && yc <= by) ? -1 : ((by < yc && yc <= ay) ? 1 : 0);
        const float xc = ax + (yc - ay) * ((bx - ax) / (by - ay));
        // the spanned columns (lo < xs <= hi) and the sign of their crossings in this pixel
        const bool rightward = ax < bx;
        const float lo = rightward ? ax : bx, hi = rightward ? bx : ax;
        const int sc = (rightward ? 1 : -1) * (py ? 1 : -1);
        const float sl = (by - ay) / (bx - ax);
        // yx * sy = ay * sy + (xs - ax) * (sl * sy) up to the sign of a zero (negation is exact)
        const float ays = ay * sy, sls = sl * sy;
#pragma unroll
        for (int c = 0; c < kAaSamples; ++c) {
            const float xs = x0 + ((float)c + 0.5f) / 16.0f;
            const float ys = y0 + ((float)((7 * c) & 15) + 0.5f) / 16.0f;
            const float v = ays + (xs - ax) * sls;
            // no short circuit: no branches
            const bool hit = (xs > lo) & (xs <= hi) & (syc <= v) & (v < ys * sy);
            W[c] += ((xs > xc) ? sr : 0) + (hit ? sc : 0);
        }
    };
    // the cell's first kPre chords fetched together (indices, then chords), the rest one at a time
    constexpr int kPre = 4;
    const int nv = cls;
    int32_t ii[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) ii[u] = sidx[f + (u < nv ? u : nv - 1)];  // loads within the run
    float4 chs[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) chs[u] = chord[ii[u]];
#pragma unroll
    for (int u = 0; u < kPre; ++u)
        if (u < nv) add_chord(chs[u]);
    for (int64_t g = f + kPre; g < f + nv; ++g) add_chord(chord[sidx[g]]);
    if (nv == kCoverClasses - 1)  // the open class: the rest of the run
        for (int64_t g = f + nv; g < N && skey[g] == key && (epath[g] & 0x3FFFFFFFu) == path; ++g)
            add_chord(chord[sidx[g]]);
    const bool even_odd = rule[path] == 1u;
    uint32_t m = 0;
#pragma unroll
    for (int c = 0; c < kAaSamples; ++c) {
        const bool in = even_odd ? (W[c] & 1) != 0 : W[c] != 0;
        m |= (uint32_t)in << ((7 * c) & 15);  // column c holds sample 7c mod 16
    }
    cmask16[4 * o + pix] = (uint16_t)m;  // bits (2 py + px) * 16 .. + 15 of the record's 64-bit mask
}

// The chord class of the cell whose run starts at entry f: the run's length, or
// kCoverClasses - 1 for a run at least that long.  Four entries are loaded at a time.
__device__ inline int run_class(int32_t f, const int32_t* __restrict__ skey, const uint32_t* __restrict__ epath,
                                int64_t N) {
    const int32_t key = skey[f];
    const uint32_t path = epath[f] & 0x3FFFFFFFu;
    int k = 1;
    bool more = true;
    while (more) {
        int32_t kg[4];
        uint32_t pg[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t g = f + k + u < N ? f + k + u : N - 1;  // loads in range, the test stops at N
            kg[u] = skey[g];
            pg[u] = epath[g] & 0x3FFFFFFFu;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            more = more && k < kCoverClasses - 1 && f + k < N && kg[u] == key && pg[u] == path;
            k += more ? 1 : 0;
        }
    }
    return k;
}

// A workgroup takes kChunk records at a time: spans get every sample at once; the cells are
// counted by chord class in LDS and listed there longest first, and the four waves then cover
// the listed cells four lanes per cell, so that the 16 cells of a wave have (nearly) the same
// number of chords: a wave runs as many chord passes as its longest cell needs.
// Five waves per SIMD (at most 96 VGPRs, the sample x and y values rematerialised): faster than
// four waves holding them all in registers.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
void k_cell_cover(const int32_t* __restrict__ cell_first, const int32_t* __restrict__ totals, int64_t rec_cap,
                  const int32_t* __restrict__ skey, const int32_t* __restrict__ sidx,
                  const int32_t* __restrict__ winding, const uint32_t* __restrict__ epath,
                  const float4* __restrict__ chord, int64_t N, const uint32_t* __restrict__ rule,
                  uint16_t* __restrict__ cmask16) {
    __shared__ int32_t cnt[kCoverClasses];  // per class: the count, then the next list place
    __shared__ int32_t list_f[kChunk];      // listed cells: first entry
    __shared__ uint16_t list_oc[kChunk];    // listed cells: record within the chunk << 5 | class
    const int64_t R = min((int64_t)totals[0] + (int64_t)totals[1], rec_cap);
    for (int64_t c0 = (int64_t)blockIdx.x * kChunk; c0 < R; c0 += (int64_t)gridDim.x * kChunk) {
        if (threadIdx.x < kCoverClasses) cnt[threadIdx.x] = 0;
        __syncthreads();
        int32_t fr[4];
        int kr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t o = c0 + u * kBlock + threadIdx.x;
            fr[u] = o < R ? cell_first[o] : -1;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t o = c0 + u * kBlock + threadIdx.x;
            kr[u] = 0;
            if (o < R) {
                if (fr[u] < 0) {
                    reinterpret_cast<unsigned long long*>(cmask16)[o] = ~0ull;  // a span: every sample
                } else {
                    kr[u] = run_class(fr[u], skey, epath, N);
                    atomicAdd(&cnt[kr[u]], 1);
                }
            }
        }
        __syncthreads();
        // list places, longest class first
        int32_t off = 0;
        if (threadIdx.x < kCoverClasses)
            for (int d = threadIdx.x + 1; d < kCoverClasses; ++d) off += cnt[d];
        __syncthreads();
        if (threadIdx.x < kCoverClasses) cnt[threadIdx.x] = off;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (kr[u]) {
                const int at = atomicAdd(&cnt[kr[u]], 1);
                list_f[at] = fr[u];
                list_oc[at] = (uint16_t)(((u * kBlock + threadIdx.x) << 5) | kr[u]);
            }
        __syncthreads();
        const int cells = cnt[1];  // class 1 is listed last: its end is the list's length
        for (int t = threadIdx.x; t < 4 * cells; t += kBlock) {
            const int oc = list_oc[t >> 2];
            cover_one(c0 + (oc >> 5), list_f[t >> 2], oc & 31, t & 3, skey, sidx, winding, epath, chord, N, rule,
                      cmask16);
        }
        __syncthreads();  // the next chunk reuses the list
    }
}

}  // namespace

hipError_t launch_cell_cover(const int32_t* cell_first, const int32_t* totals, int64_t rec_cap,
                             const int32_t* skey, const int32_t* sidx, const int32_t* winding,
                             const uint32_t* epath, const float4* chord, int64_t N, const uint32_t* rule,
                             unsigned long long* cmask, hipStream_t s) {
    if (rec_cap <= 0 || N <= 0) return hipSuccess;
    int64_t g = (rec_cap + kChunk - 1) / kChunk;
    g = g < 4096 ? g : 4096;  // the workgroups stride over the live records
    k_cell_cover<<<(unsigned)g, kBlock, 0, s>>>(cell_first, totals, rec_cap, skey, sidx, winding, epath, chord, N,
                                                rule, reinterpret_cast<uint16_t*>(cmask));
    return hipGetLastError();
}

}  // namespace spr
