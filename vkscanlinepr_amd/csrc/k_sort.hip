// k_sort.hip -- sort of the cell keys inside each (path, row) band, fused with the winding
// scan and, in frames that export no stage, with mark + emit (gfx950).
//
// Replaces naive_seg_sort_pairs.comp:26-97 (one workgroup per path running an O(n^2) odd-even
// transposition network), shuffle_fragment.comp:17-26 and the winding-number scan of
// scanline_rasterizer.cpp:478-506 (naive_scan.comp over the sorted winding changes).  The
// reference compares (key, index) as signed int32 inside each path segment; the sort word here is
//     ((key ^ 0x80000000) << 32) | (index << 2) | (winding change + 1)
// whose unsigned order equals the reference's (key, index) order, is unique, and carries the
// winding change, so the reference's separate gather pass disappears.
//
// After the band scatter every band (<= 2048 entries, a run of whole buckets of one path) only
// needs sorting in place: one workgroup loads it once, sorts it in LDS (8-element register
// network per thread, then merge-path rounds) and stores it once.  A band's sum of winding
// changes needs no sorting: the scatter adds the sums up per band and one scan turns them into
// each band's exclusive prefix before any band is sorted, so the bands are independent.  Bands
// of up to 512 entries are sorted by one-wave workgroups (so that the many short bands of a page
// of small paths do not leave most of a large workgroup idle), the longer ones, listed by a
// classifying pass, by 256-thread workgroups that take them in chunks, the first chunk by
// workgroup id and the rest from an atomic ticket (their cost varies with the band length, so
// the work is handed out as it frees up, and a frame without long bands takes no ticket at all).
// A band above 2048 entries is a single oversized (path, row) bucket; such bands are queued, and
// each is then sorted by one workgroup with LDS chunk sorts followed by merge-path passes
// through global memory, the last pass writing keys and winding numbers.
//
// Frames that export no stage sort the narrow 32-bit words of the scatter instead (bucket inside
// the path, x, winding change: see k_bucket.hip); the keys are rebuilt from the band's path rows.
// Each band of up to 2048 narrow words is sorted by one wave in registers (k_sort_bands_wave: a
// bitonic network generated at compile time, in-lane for short strides, DPP and ds_bpermute
// across lanes), in three size classes of 8, 16 and 32 words per lane; k_band_classify2 lists the
// bands of the two larger classes.  These frames also run mark + emit (mark_merged_fragment_and_span
// .comp, the 2N flag scan and gen_merged_fragment_and_span.comp) inside the band sort, on the
// sorted words and winding numbers still in registers: every flag of a band depends on its own entries
// only (a band starts at a bucket, i.e. a new cell row), so each band writes its records at twice
// its start, the record range of each of its buckets and its (cells, spans); the oversized bands
// do the same from their sorted keys (k_emit_big_bands).  Debug frames write keys, indices and
// winding numbers, and k_mark_emit (k_scan.hip) emits the records in the reference's order.
#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"
#include "scan_common.hpp"

namespace spr {

namespace {

constexpr int kItems = 8;  // sort words per thread
typedef unsigned long long u64;

// Workgroup shape of a band class: THREADS threads sort up to THREADS * kItems words, in two LDS
// buffers with one pad word per 8 words (so the 8-word runs of consecutive lanes start on
// different banks).
template <int THREADS>
struct Tile {
    static constexpr int kWaves = THREADS / 64;
    static constexpr int kSize = THREADS * kItems;
    static constexpr int kPadded = kSize + kSize / 8;
};
constexpr int kBigThreads = 256;  // oversized bands: chunks of one full tile
static_assert(Tile<kBigThreads>::kSize == kSortTile, "tile shape");

template <class T>
__device__ inline void cswap(T& a, T& b) {
    T lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo; b = hi;
}

// Ascending bitonic network over kItems registers (indices are compile-time constants).
template <class T>
__device__ inline void sort16(T (&v)[kItems]) {
#pragma unroll
    for (int k = 2; k <= kItems; k <<= 1) {
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            const int l = i ^ (k - 1);
            if (l > i && (i & (k >> 1)) == 0) cswap(v[i], v[l]);
        }
#pragma unroll
        for (int j = k >> 2; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < kItems; ++i) {
                const int l = i ^ j;
                if (l > i) cswap(v[i], v[l]);
            }
        }
    }
}

// Merge-path split: number of elements taken from A among the first `diag` outputs of the
// stable merge of A (length na) and B (length nb).
template <class PA, class PB>
__device__ inline int64_t merge_split(PA A, int64_t na, PB B, int64_t nb, int64_t diag) {
    int64_t lo = diag > nb ? diag - nb : 0;
    int64_t hi = diag < na ? diag : na;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[diag - 1 - mid]) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Sorted output, unpacked from the sort words: key, (optionally) index, and the winding number
// (exclusive prefix of the winding changes in sorted order).
struct Unpacked {
    int32_t* skey;
    int32_t* sidx;
    int32_t* winding;
};
// Wide words: key ^ 0x80000000, entry index, winding change + 1.
struct Wide {
    typedef u64 T;
    static constexpr T kPad = ~0ull;  // no word has 3 in its low two bits
    int2 pr;
    __device__ void band(const int2*, int64_t) {}
    __device__ int32_t key(T w) const { return (int32_t)((uint32_t)(w >> 32) ^ 0x80000000u); }
    __device__ void put(const Unpacked& o, int64_t i, T w) const {
        o.skey[i] = key(w);
        if (o.sidx) o.sidx[i] = (int32_t)((w >> 2) & 0x3FFFFFFFull);
    }
};
// Narrow words: bucket inside the band's path << 17 | (x + 2) << 2 | winding change + 1; the key
// is rebuilt from the path's rows pr = (first row, number of rows): buckets 0..pr.y-1 are rows
// pr.x.., bucket pr.y the off-canvas key, bucket pr.y+1 row 0.
struct Narrow {
    typedef uint32_t T;
    static constexpr T kPad = ~0u;
    int2 pr;
    __device__ void band(const int2* band_pr, int64_t b) { pr = band_pr[b]; }
    __device__ int32_t key(T w) const {
        const int loc = (int)(w >> 17);
        const uint32_t x = ((w >> 2) & 0x7FFFu) + kNarrowX0;
        const int y = loc < pr.y ? 2 * (pr.x + loc) : 0;
        return loc == pr.y ? (int32_t)kSentinelKey : (int32_t)(((uint32_t)(y + 0x7FFF) << 16) | x);
    }
    __device__ void put(const Unpacked& o, int64_t i, T w) const { o.skey[i] = key(w); }
};
template <class T>
__device__ inline int32_t wn_of(T w) { return (int32_t)(w & 3u) - 1; }
__device__ inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// LDS slot of element i (one pad word per 8 elements).
__device__ inline int pad(int i) { return i + (i >> 3); }

__device__ inline int key_x(int32_t k) { return (k & 0xFFFF) - 0x7FFF; }
__device__ inline int key_y(int32_t k) { return ((k >> 16) & 0xFFFF) - 0x7FFF; }

// What mark + emit needs of a band's path.
struct BandPath {
    uint32_t rule;
    int cw, ch;       // canvas
    int32_t wb, bb;   // winding base, first bucket
    int2 pr;          // rows: first, count
    int64_t r0;       // first record slot: 2 * band start
};

__device__ inline BandPath band_path_of(const EmitArgs& E, int64_t band, int2 pr, int64_t start) {
    BandPath b;
    const int32_t path = E.band_path[band];
    b.rule = E.rule[path];
    b.cw = E.w; b.ch = E.h;
    if (E.pcanvas) { const int2 c = E.pcanvas[path]; b.cw = c.x; b.ch = c.y; }
    b.wb = E.wbase ? E.wbase[path] : 0;
    b.bb = E.bucket_base[path];
    b.pr = pr;
    b.r0 = 2 * start;
    return b;
}

// Mark + emit of band entries [c0, c0 + m) of a band of n entries (mark_merged_fragment_and_
// span.comp:48-120, gen_merged_fragment_and_span.comp:152-222), THREADS threads with kItems
// consecutive entries each.  key(j) and wind(j) give entry j's key and winding number (j in
// [0, n)).  A cell is an in-canvas entry whose key differs from the previous entry's (the
// band's first entry follows another bucket, so its key differs); a span runs from the previous
// entry of the same cell row to this one when the winding number fills.  Records go to
// crec[r0 + carry + ...]; carry (records of earlier chunks) is advanced and the chunk's
// (cells, spans) are added to cnt (thread 0).
template <int THREADS, class KEY, class WIND>
__device__ inline void emit_chunk(const EmitArgs& E, const BandPath& b, KEY key, WIND wind, int64_t c0, int m,
                                  int64_t n, int32_t& carry, int2& cnt, int32_t* s_w) {
    const int64_t j0 = c0 + (int64_t)threadIdx.x * kItems;
    int32_t keys[kItems];
    uint32_t cbits = 0, sbits = 0;
    int32_t kprev = j0 > 0 && j0 - 1 < n ? key(j0 - 1) : 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t j = j0 + k;
        keys[k] = 0;
        if (j >= c0 + m) continue;
        const int32_t k1 = key(j);
        keys[k] = k1;
        const int x1 = key_x(k1), y1 = key_y(k1);
        const bool in_canvas = !(x1 < 0 || y1 < 0 || x1 >= b.cw || y1 >= b.ch);
        const bool cell = in_canvas && (j == 0 || kprev != k1);
        bool span = false;
        if (j > 0) {
            const int32_t wn = (int32_t)((uint32_t)wind(j) - (uint32_t)b.wb);
            const bool wf = (b.rule == 0u && wn != 0) || (b.rule == 1u && (wn & 1) != 0);
            span = key_y(kprev) == y1 && key_x(kprev) + 2 < x1 && wf;
        }
        cbits |= (uint32_t)cell << k;
        sbits |= (uint32_t)span << k;
        kprev = k1;
    }
    // chunk scan of (cells | spans << 16): a chunk holds at most 2048 of each
    int32_t tot;
    const int32_t ex = block_exclusive_add<int32_t, THREADS / 64>(__popc(cbits) | (__popc(sbits) << 16), s_w, tot);
    int64_t pos = b.r0 + carry + (ex & 0xFFFF) + (ex >> 16);
    kprev = j0 > 0 && j0 - 1 < n ? key(j0 - 1) : 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t j = j0 + k;
        if (j >= c0 + m) continue;
        const int32_t k1 = keys[k];
        const int cell = (cbits >> k) & 1, span = (sbits >> k) & 1;
        const int x1 = key_x(k1), y1 = key_y(k1);
        if (cell) E.crec[pos] = (uint32_t)x1 | (2u << 16);
        if (span) {
            const int x0 = max(0, key_x(kprev) + 2);
            E.crec[pos + cell] = (uint32_t)x0 | ((uint32_t)(x1 - x0) << 16);
        }
        // record range of this entry's (path, cell row) bucket, for the composite
        if ((uint32_t)k1 != kSentinelKey && y1 >= 0 && y1 < b.ch) {
            const int32_t kn = j + 1 < n ? (k + 1 < kItems ? keys[k + 1 < kItems ? k + 1 : k] : key(j + 1)) : 0;
            const bool first = j == 0 || (kprev >> 16) != (k1 >> 16);
            const bool last = j == n - 1 || (kn >> 16) != (k1 >> 16);
            if (first || last) {
                const int r = y1 >> 1;
                const int loc = r == 0 ? b.pr.y + 1 : r - b.pr.x;
                if (loc >= 0 && loc < b.pr.y + 2) {  // out of range only after a flagged bound error
                    int2* br = &E.brec[b.bb + loc];
                    if (first) br->x = (int32_t)pos;
                    if (last) br->y = (int32_t)(pos + cell + span);
                }
            }
        }
        pos += cell + span;
        kprev = k1;
    }
    carry += (tot & 0xFFFF) + (tot >> 16);
    if (threadIdx.x == 0) { cnt.x += tot & 0xFFFF; cnt.y += tot >> 16; }
}

// Sorts n <= THREADS * kItems words from src with one workgroup of THREADS threads (buf: 2 x
// kPadded words of LDS) into dst (sort words) or, when dst is null, unpacked into out at offset
// o0 with winding numbers from the band's exclusive prefix `pre`.
template <class W, int THREADS>
__device__ __forceinline__ void tile_sort(const W& wd, const typename W::T* __restrict__ src,
                                          typename W::T* __restrict__ dst, int n,
                                          typename W::T (*buf)[Tile<THREADS>::kPadded], const Unpacked& out,
                                          int64_t o0, int32_t pre, int32_t* s_w, const EmitArgs& E,
                                          bool emit, int64_t band) {
    typedef typename W::T T;
    constexpr T kPad = W::kPad;
    int P2 = kItems;
    while (P2 < n) P2 <<= 1;
    const int active = P2 / kItems;
    const int tid = threadIdx.x;
    T v[kItems];
    if (tid < active) {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int i = tid * kItems + j;
            v[j] = i < n ? src[i] : kPad;
        }
    }
    if (tid < active) {
        sort16(v);
#pragma unroll
        for (int j = 0; j < kItems; ++j) buf[0][pad(tid * kItems + j)] = v[j];
    }
    __syncthreads();
    int cur = 0;
    for (int run = kItems; run < P2; run <<= 1) {
        if (tid < active) {
            const int o = tid * kItems;
            const int pb = o & ~(2 * run - 1);  // run is a power of two
            const T* S = buf[cur];
            auto A = [&](int i) { return S[pad(pb + i)]; };
            auto B = [&](int j) { return S[pad(pb + run + j)]; };
            const int diag = o - pb;
            int lo = diag > run ? diag - run : 0, hi = diag < run ? diag : run;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (A(mid) <= B(diag - 1 - mid)) lo = mid + 1; else hi = mid;
            }
            int i = lo, j = diag - lo;
            T* D = buf[cur ^ 1];
            // a run's end reads as kPad, which no sort word equals (its low two bits never hold
            // 3), so the merge takes the smaller head without bounds tests; each step refills the
            // consumed side with one load at a clamped index (equal narrow words are alike, so
            // the side taken on a tie does not matter)
            T a = A(i < run ? i : run - 1), b = B(j < run ? j : run - 1);
            a = i < run ? a : kPad;
            b = j < run ? b : kPad;
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                const bool takeA = a < b;
                D[pad(o + k)] = takeA ? a : b;
                i += takeA ? 1 : 0;
                j += takeA ? 0 : 1;
                const int nx = takeA ? i : j;
                const T v = S[pad(pb + (takeA ? 0 : run) + (nx < run ? nx : run - 1))];
                const T w = nx < run ? v : kPad;
                a = takeA ? w : a;
                b = takeA ? b : w;
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    if (dst) {
        for (int i = tid; i < n; i += THREADS) dst[i] = buf[cur][pad(i)];
        __syncthreads();
        return;
    }
    // winding numbers: each thread scans its 8 consecutive sorted entries, the block scans the
    // thread sums, and the results go through the free LDS buffer to coalesced stores
    int32_t* wl = reinterpret_cast<int32_t*>(buf[cur ^ 1]);
    int32_t w8[kItems], sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = tid * kItems + k;
        w8[k] = sum;
        if (i < n) sum += wn_of(buf[cur][pad(i)]);
    }
    int32_t tot;
    const int32_t ex = wadd(block_exclusive_add<int32_t, Tile<THREADS>::kWaves>(sum, s_w, tot), pre);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = tid * kItems + k;
        if (i < n) wl[i] = wadd(ex, w8[k]);
    }
    __syncthreads();
    if (emit) {
        // mark + emit from LDS: sorted words and winding numbers stay on chip
        const BandPath b = band_path_of(E, band, wd.pr, o0);
        const T* S = buf[cur];
        int32_t carry = 0;
        int2 cnt = make_int2(0, 0);
        emit_chunk<THREADS>(E, b, [&](int64_t j) { return wd.key(S[pad((int)j)]); },
                            [&](int64_t j) { return wl[j]; }, 0, n, n, carry, cnt, s_w);
        if (tid == 0) E.band_cnt[band] = cnt;
        __syncthreads();
        return;
    }
    for (int i = tid; i < n; i += THREADS) {
        wd.put(out, o0 + i, buf[cur][pad(i)]);
        out.winding[o0 + i] = wl[i];
    }
    __syncthreads();
}

// Long bands (above one 64-thread tile) are listed in `list`, their number in *count, with one
// atomic per 1024-thread block (the waves' counts are added up in LDS); the order of the list
// does not matter.
constexpr int kShortMax = 64 * kItems;
constexpr int kClassifyThreads = 1024;
__global__ __launch_bounds__(kClassifyThreads) void k_band_classify(const int32_t* __restrict__ band_start,
                                                                    const int32_t* __restrict__ n_bands,
                                                                    int32_t* __restrict__ list,
                                                                    uint32_t* __restrict__ count) {
    __shared__ uint32_t s_n[kClassifyThreads / 64], s_base;
    const int64_t t = (int64_t)blockIdx.x * kClassifyThreads + threadIdx.x;
    const int32_t nb = *n_bands;
    const bool is_long = t < nb && band_start[t + 1] - band_start[t] > kShortMax;
    const unsigned long long m = __ballot(is_long);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s_n[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int k = 0; k < kClassifyThreads / 64; ++k) tot += s_n[k];
        s_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    if (is_long) {
        uint32_t off = s_base;
        for (int k = 0; k < wave; ++k) off += s_n[k];
        list[off + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)t;
    }
}

// ---- production bands: one wave per band, sorted in registers
//
// Frames that export no stage sort their narrow words band by band with one wave each.  A band of
// n <= 2048 words is padded to 64 R words (R = 1, 2, 4, ..., 32 per lane: the smallest that holds
// it) and sorted by a bitonic network over the element index i = l R + r (lane l, register r):
// stages whose stride is below R compare two registers of one lane, the others exchange register
// r with lane l ^ (stride / R) (ds_bpermute).  The words are loaded with coalesced accesses (word
// l + 64 q) and transposed into registers through a wave-private LDS region padded by one word
// every R words (no bank conflicts either way).  The sorted words then give the winding numbers
// (the lane's prefix, a DPP wave scan and the band's prefix) and mark + emit runs from registers.
// Bands of up to 512 words are sorted by one kernel (8 registers per lane: high occupancy), which
// lists the longer ones for a second kernel (up to 32 per lane).
constexpr int kWaveBandMax = 64 * 32;
constexpr int kMediumBandMax = 64 * 16;
constexpr int kSmallBandMax = 64 * 8;

// The value of lane l ^ D (D a power of two below 64; low: bit D of the lane is clear).  Inside a
// quad a DPP quad permutation, inside a 16-lane row a DPP row shift (the lower lane of a pair reads
// l + D, the upper one l - D): plain VALU operations.  Across rows ds_bpermute.
template <int D>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, bool low, int lane) {
    if constexpr (D == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (D == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else if constexpr (D == 4 || D == 8) {
        const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x100 + D, 0xF, 0xF, false);  // row_shl
        const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 + D, 0xF, 0xF, false);  // row_shr
        return low ? up : dn;
    } else {
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ D) << 2, (int)v);
    }
}

// Stage J (stride) of block K (ascending where bit K of the element index is 0; descending lanes
// have complemented their words for the blocks from R up).  The network is generated at compile
// time, one stage at a time.
template <int R, int K, int J>
__device__ __forceinline__ void bitonic_stage(uint32_t (&v)[R], int lane) {
    if constexpr (J > 0) {
        if constexpr (J < R) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int r2 = r ^ J;
                if (r2 <= r) continue;
                const bool asc = K < R ? (r & K) == 0 : true;
                const uint32_t a = v[r], b = v[r2];
                const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
                v[r] = asc ? lo : hi;
                v[r2] = asc ? hi : lo;
            }
        } else {
            constexpr int D = J / R;
            const bool keep_min = (lane & D) == 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t o = lane_xor<D>(v[r], keep_min, lane);
                const uint32_t lo = v[r] < o ? v[r] : o, hi = v[r] < o ? o : v[r];
                v[r] = keep_min ? lo : hi;
            }
        }
        bitonic_stage<R, K, J / 2>(v, lane);
    }
}

template <int R, int K>
__device__ __forceinline__ void bitonic_block(uint32_t (&v)[R], int lane) {
    if constexpr (K <= 64 * R) {
        // from R up the direction is a bit of the lane, the same for the whole block: a lane whose
        // block descends complements its words (reversing their order) and sorts ascending
        const uint32_t flip = K >= R && ((lane * R) & K) != 0 ? ~0u : 0u;
        if constexpr (K >= R) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] ^= flip;
        }
        bitonic_stage<R, K, K / 2>(v, lane);
        if constexpr (K >= R) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] ^= flip;
        }
        bitonic_block<R, 2 * K>(v, lane);
    }
}

template <int R>
__device__ __forceinline__ void wave_bitonic(uint32_t (&v)[R]) {
    bitonic_block<R, 2>(v, threadIdx.x & 63);
}

// One band of n words at src (R per lane; lds: the wave's 64 R + 64 words): sort, winding numbers
// from the band's prefix `pre`, mark + emit (mark_merged_fragment_and_span.comp:48-120,
// gen_merged_fragment_and_span.comp:152-222: a cell is an in-canvas entry whose key differs from
// the previous entry's, a span runs from the previous entry of the same cell row to this one when
// the winding number fills), the record range of each of the band's buckets and its (cells,
// spans).
template <int R>
__device__ __forceinline__ void wave_band(const uint32_t* __restrict__ src, int n, int64_t band, int64_t start,
                                          int32_t pre, int2 pr, const EmitArgs& EA, uint32_t* lds) {
    const int lane = threadIdx.x & 63;
    uint32_t v[R];
    if (R == 1) {
        v[0] = lane < n ? src[lane] : Narrow::kPad;
    } else {
        // coalesced loads, then the transpose: word i sits at LDS slot i + i / R
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = 64 * q + lane;
            lds[i + i / R] = i < n ? src[i] : Narrow::kPad;  // pads sort to the end
        }
        // one wave's LDS accesses run in order; the barrier keeps the compiler's order too
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = lds[lane * (R + 1) + r];
        __builtin_amdgcn_wave_barrier();
    }
    // the band's path tables (rule, canvas, winding base, first bucket) load while the words sort
    const BandPath b = band_path_of(EA, band, pr, start);
    wave_bitonic<R>(v);
    Narrow wd;
    wd.pr = pr;
    // keys and winding numbers are recomputed in each pass from the sorted words (registers
    // for occupancy): the lane's sum of winding changes first
    int32_t sum = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (lane * R + r < n) sum = wadd(sum, wn_of(v[r]));
    const int32_t lane_pre = wadd(pre, (int32_t)((uint32_t)wave_inclusive_add(sum) - (uint32_t)sum));
    const int32_t prev_last = __shfl_up(wd.key(v[R - 1]), 1);
    uint32_t cbits = 0, sbits = 0;
    {
        int32_t kprev = prev_last, wl = lane_pre;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane * R + r;
            if (i >= n) continue;
            const int32_t k1 = wd.key(v[r]);
            const int x1 = key_x(k1), y1 = key_y(k1);
            const bool in_canvas = !(x1 < 0 || y1 < 0 || x1 >= b.cw || y1 >= b.ch);
            const bool cell = in_canvas && (i == 0 || kprev != k1);
            bool span = false;
            if (i > 0) {
                const int32_t wn = (int32_t)((uint32_t)wl - (uint32_t)b.wb);
                const bool wf = (b.rule == 0u && wn != 0) || (b.rule == 1u && (wn & 1) != 0);
                span = key_y(kprev) == y1 && key_x(kprev) + 2 < x1 && wf;
            }
            cbits |= (uint32_t)cell << r;
            sbits |= (uint32_t)span << r;
            kprev = k1;
            wl = wadd(wl, wn_of(v[r]));
        }
    }
    // record slots: cells in the low half, spans in the high half (at most 2048 of each per band);
    // the records are staged in the wave's LDS region (2 per entry at most: 128 R words) at their
    // slot inside the band and stored with coalesced writes
    const int32_t mine = __popc(cbits) | (__popc(sbits) << 16);
    const int32_t inc = wave_inclusive_add(mine);
    const int32_t ex = inc - mine;
    int32_t rel = (ex & 0xFFFF) + (ex >> 16);  // slot inside the band's records (the global slot is r0 + rel)
    const int32_t next_first = __shfl_down(wd.key(v[0]), 1);
    int32_t kprev = prev_last, k1 = wd.key(v[0]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane * R + r;
        if (i >= n) continue;
        const int32_t kn = i + 1 < n ? (r + 1 < R ? wd.key(v[r + 1 < R ? r + 1 : r]) : next_first) : 0;
        const int cell = (cbits >> r) & 1, span = (sbits >> r) & 1;
        const int x1 = key_x(k1), y1 = key_y(k1);
        if (cell) lds[rel] = (uint32_t)x1 | (2u << 16);
        if (span) {
            const int x0 = max(0, key_x(kprev) + 2);
            lds[rel + cell] = (uint32_t)x0 | ((uint32_t)(x1 - x0) << 16);
        }
        // record range of this entry's (path, cell row) bucket, for the composite
        if ((uint32_t)k1 != kSentinelKey && y1 >= 0 && y1 < b.ch) {
            const bool first = i == 0 || (kprev >> 16) != (k1 >> 16);
            const bool last = i == n - 1 || (kn >> 16) != (k1 >> 16);
            if (first || last) {
                const int row = y1 >> 1;
                const int loc = row == 0 ? b.pr.y + 1 : row - b.pr.x;
                if (loc >= 0 && loc < b.pr.y + 2) {  // out of range only after a flagged bound error
                    int2* br = &EA.brec[b.bb + loc];
                    if (first) br->x = (int32_t)(b.r0 + rel);
                    if (last) br->y = (int32_t)(b.r0 + rel + cell + span);
                }
            }
        }
        rel += cell + span;
        kprev = k1;
        k1 = kn;
    }
    __builtin_amdgcn_wave_barrier();
    const int32_t total = __shfl(inc, 63);
    const int32_t n_rec = (total & 0xFFFF) + (total >> 16);
    for (int32_t q = lane; q < n_rec; q += 64) EA.crec[b.r0 + q] = lds[q];
    if (lane == 63) EA.band_cnt[band] = make_int2(inc & 0xFFFF, inc >> 16);
}

// Bands above 512 words in two lists: up to 1024 words from the front of `list`, longer ones (single
// buckets) from the back (list holds `cap` slots); counts in count[0] and count[1].  One atomic per
// list per 1024-thread block.
__global__ __launch_bounds__(kClassifyThreads) void k_band_classify2(const int32_t* __restrict__ band_start,
                                                                     const int32_t* __restrict__ n_bands,
                                                                     int32_t* __restrict__ list, int64_t cap,
                                                                     uint32_t* __restrict__ count,
                                                                     const int32_t* __restrict__ bpre,
                                                                     int32_t* __restrict__ big_list,
                                                                     int32_t* __restrict__ big_excl,
                                                                     int32_t* __restrict__ big_count) {
    __shared__ uint32_t s_n[2][kClassifyThreads / 64], s_base[2];
    const int64_t t = (int64_t)blockIdx.x * kClassifyThreads + threadIdx.x;
    const int32_t nb = *n_bands;
    const int32_t n = t < nb ? band_start[t + 1] - band_start[t] : 0;
    const bool med = n > kSmallBandMax && n <= kMediumBandMax;
    const bool large = n > kMediumBandMax && n <= kWaveBandMax;
    if (n > kWaveBandMax) {
        // a single oversized (path, row) bucket: queued with its exclusive prefix for the
        // global-merge kernels, which may run beside the wave sorts
        const int32_t q = atomicAdd(big_count, 1);
        big_list[q] = (int32_t)t;
        big_excl[q] = bpre[t];
    }
    const unsigned long long mm = __ballot(med), ml = __ballot(large);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s_n[0][wave] = (uint32_t)__popcll(mm); s_n[1][wave] = (uint32_t)__popcll(ml); }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t tot = 0;
        for (int k = 0; k < kClassifyThreads / 64; ++k) tot += s_n[threadIdx.x][k];
        s_base[threadIdx.x] = tot ? atomicAdd(&count[threadIdx.x], tot) : 0u;
    }
    __syncthreads();
    if (med || large) {
        const int c = large ? 1 : 0;
        uint32_t off = s_base[c];
        for (int k = 0; k < wave; ++k) off += s_n[c][k];
        off += (uint32_t)__popcll((large ? ml : mm) & ((1ull << lane) - 1ull));
        if (large) list[cap - 1 - off] = (int32_t)t;
        else list[off] = (int32_t)t;
    }
}

// CLS 0: every band of up to 512 words, one wave each (four per workgroup; the others are
// skipped).  CLS 1: the listed bands of 513..1024 words (16 per lane).  CLS 2: the listed bands of
// 1025..2048 words (32 per lane); k_band_classify2 queues the longer ones (single oversized
// buckets) for the global-merge kernels.
template <int CLS>
__global__ __launch_bounds__(256) void k_sort_bands_wave(const int32_t* __restrict__ band_start,
                                                         const int32_t* __restrict__ n_bands,
                                                         const uint32_t* __restrict__ src,
                                                         const int32_t* __restrict__ bpre,
                                                         const int2* __restrict__ band_pr,
                                                         int32_t* __restrict__ big_list,
                                                         int32_t* __restrict__ big_excl,
                                                         int32_t* __restrict__ big_count,
                                                         const int32_t* __restrict__ list, int64_t cap,
                                                         const uint32_t* __restrict__ count, EmitArgs EA) {
    constexpr int RMAX = CLS == 0 ? 8 : (CLS == 1 ? 16 : 32);
    __shared__ uint32_t s_lds[4][128 * RMAX];  // the transposed load (64 R + 64), then the records (<= 128 R)
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= (CLS == 0 ? (int64_t)*n_bands : (int64_t)count[CLS - 1])) return;  // wave-uniform
    const int64_t t = CLS == 0 ? w : (CLS == 1 ? list[w] : list[cap - 1 - w]);
    const int64_t s = band_start[t], n = band_start[t + 1] - s;
    if (CLS == 0 && n > kSmallBandMax) return;
    const uint32_t* p = src + s;
    const int2 pr = band_pr[t];
    const int32_t pre = bpre[t];
    const int m = (int)n;
    uint32_t* lds = s_lds[threadIdx.x >> 6];
    if constexpr (CLS == 2) {
        wave_band<32>(p, m, t, s, pre, pr, EA, lds);
    } else if constexpr (CLS == 1) {
        wave_band<16>(p, m, t, s, pre, pr, EA, lds);
    } else {
        if (m <= 64) wave_band<1>(p, m, t, s, pre, pr, EA, lds);
        else if (m <= 128) wave_band<2>(p, m, t, s, pre, pr, EA, lds);
        else if (m <= 256) wave_band<4>(p, m, t, s, pre, pr, EA, lds);
        else wave_band<8>(p, m, t, s, pre, pr, EA, lds);
    }
}

// BIG = false: the bands of up to THREADS * kItems entries, the workgroups striding over all
// bands.  BIG = true: the listed long bands, kChunk at a time (the first chunk by workgroup id,
// the rest from ticket[0]; ticket[1] holds the list length); this kernel also queues the
// oversized bands with their exclusive prefixes and writes the scan's total after the last entry.
constexpr int kChunk = 4;

template <class W, int THREADS, bool BIG>
__global__ __launch_bounds__(THREADS) void k_sort_bands(const int32_t* __restrict__ band_start,
                                                        const int32_t* __restrict__ n_bands,
                                                        const typename W::T* __restrict__ src, Unpacked out,
                                                        const int32_t* __restrict__ bpre,
                                                        const int32_t* __restrict__ list,
                                                        uint32_t* __restrict__ ticket,
                                                        int32_t* __restrict__ big_list,
                                                        int32_t* __restrict__ big_excl,
                                                        int32_t* __restrict__ big_count,
                                                        const int2* __restrict__ band_pr, EmitArgs E, bool emit) {
    constexpr int kSize = Tile<THREADS>::kSize;
    __shared__ typename W::T buf[2][Tile<THREADS>::kPadded];
    __shared__ int32_t s_w[Tile<THREADS>::kWaves], s_t;
    const int32_t nb = *n_bands;
    if (BIG && !emit && blockIdx.x == 0 && threadIdx.x == 0) {
        if (nb > 0) out.winding[band_start[nb]] = bpre[nb];  // the scan's total
        else out.winding[0] = 0;
    }
    const int64_t n_items = BIG ? (int64_t)ticket[1] : (int64_t)nb;
    int64_t c = blockIdx.x;
    while (true) {
        const int64_t q0 = BIG ? c * kChunk : c;
        if (q0 >= n_items) return;
        const int64_t q1 = BIG ? (q0 + kChunk < n_items ? q0 + kChunk : n_items) : q0 + 1;
        for (int64_t q = q0; q < q1; ++q) {
            const int64_t t = BIG ? list[q] : q;
            const int64_t s = band_start[t], n = band_start[t + 1] - s;
            if (!BIG && n > kSize) continue;  // a long band
            if (n > kSize) {
                if (BIG && threadIdx.x == 0) {
                    const int32_t q = atomicAdd(big_count, 1);
                    big_list[q] = (int32_t)t;
                    big_excl[q] = bpre[t];
                }
                continue;
            }
            W wd;
            wd.band(band_pr, t);
            tile_sort<W, THREADS>(wd, src + s, (typename W::T*)nullptr, (int)n, buf, out, s, bpre[t], s_w, E,
                                  emit, t);
        }
        // any order: only the hand-out of the chunks matters
        c = BIG ? (int64_t)gridDim.x + ordered_tile_id(ticket, &s_t) : c + gridDim.x;
    }
}

// One workgroup per oversized band: LDS chunk sorts into `a`, then merge-path passes between
// `a` and `b` (the scatter buffer, already consumed), the last pass writing keys and winding
// numbers (the band's exclusive prefix carried across its 2048-entry output blocks).
template <class W>
__global__ __launch_bounds__(kBigThreads) void k_sort_big_bands(const int32_t* __restrict__ band_start,
                                                             const int32_t* __restrict__ big_list,
                                                             const int32_t* __restrict__ big_excl,
                                                             const int32_t* __restrict__ big_count,
                                                             typename W::T* __restrict__ scatter,
                                                             typename W::T* __restrict__ a, Unpacked out,
                                                             const int2* __restrict__ band_pr) {
    typedef typename W::T T;
    constexpr int kThreads = kBigThreads;
    __shared__ T buf[2][Tile<kThreads>::kPadded];
    __shared__ int32_t s_w[Tile<kThreads>::kWaves];
    const int32_t nbig = *big_count;
    for (int32_t q = blockIdx.x; q < nbig; q += gridDim.x) {
        const int32_t band = big_list[q];
        const int64_t s = band_start[band], n = band_start[band + 1] - s;
        W wd;
        wd.band(band_pr, band);
        for (int64_t c = 0; c < n; c += kSortTile) {
            const int m = (int)(n - c < kSortTile ? n - c : kSortTile);
            tile_sort<W, kThreads>(wd, scatter + s + c, a + s + c, m, buf, out, 0, 0, s_w, EmitArgs{}, false, 0);
        }
        __threadfence_block();
        __syncthreads();
        T* cur = a + s;
        T* nxt = scatter + s;
        int32_t carry = big_excl[q];
        for (int64_t r = kSortTile; r < n; r <<= 1) {
            const bool last_pass = 2 * r >= n;
            for (int64_t base = 0; base < n; base += (int64_t)kThreads * kItems) {
                const int64_t o = base + (int64_t)threadIdx.x * kItems;
                T vals[kItems];
                int cnt = 0;
                if (o < n) {
                    const int64_t pb = (o / (2 * r)) * (2 * r);
                    const int64_t na = pb + r < n ? r : n - pb;
                    const int64_t nb = pb + r < n ? ((pb + 2 * r < n ? pb + 2 * r : n) - (pb + r)) : 0;
                    const T* A = cur + pb;
                    const T* B = A + na;
                    cnt = (int)(n - o < kItems ? n - o : kItems);
                    const int64_t diag = o - pb;
                    int64_t i = nb ? merge_split(A, na, B, nb, diag) : diag;
                    int64_t j = diag - i;
#pragma unroll
                    for (int k = 0; k < kItems; ++k) {
                        if (k >= cnt) break;
                        const bool takeA = j >= nb || (i < na && A[i] <= B[j]);
                        vals[k] = takeA ? A[i] : B[j];
                        i += takeA; j += !takeA;
                    }
                }
                if (!last_pass) {
                    for (int k = 0; k < cnt; ++k) nxt[o + k] = vals[k];
                    continue;
                }
                int32_t local = 0;
                for (int k = 0; k < cnt; ++k) local += wn_of(vals[k]);
                int32_t tot;
                int32_t run = wadd(carry, block_exclusive_add<int32_t, Tile<kThreads>::kWaves>(local, s_w, tot));
                for (int k = 0; k < cnt; ++k) {
                    wd.put(out, s + o + k, vals[k]);
                    out.winding[s + o + k] = run;
                    run = wadd(run, wn_of(vals[k]));
                }
                carry = wadd(carry, tot);
            }
            __threadfence_block();
            __syncthreads();
            if (last_pass) break;
            T* t = cur; cur = nxt; nxt = t;
        }
    }
}

}  // namespace

namespace {

// Mark + emit of each oversized band (one workgroup per band) from its sorted keys and winding
// numbers, in chunks of one tile with the records carried across.
__global__ __launch_bounds__(kBigThreads) void k_emit_big_bands(const int32_t* __restrict__ band_start,
                                                                 const int32_t* __restrict__ big_list,
                                                                 const int32_t* __restrict__ big_count,
                                                                 const int32_t* __restrict__ skey,
                                                                 const int32_t* __restrict__ winding,
                                                                 const int2* __restrict__ band_pr, EmitArgs E) {
    __shared__ int32_t s_w[Tile<kBigThreads>::kWaves];
    const int32_t nbig = *big_count;
    for (int32_t q = blockIdx.x; q < nbig; q += gridDim.x) {
        const int32_t band = big_list[q];
        const int64_t s = band_start[band], n = band_start[band + 1] - s;
        const BandPath b = band_path_of(E, band, band_pr[band], s);
        int32_t carry = 0;
        int2 cnt = make_int2(0, 0);
        for (int64_t c = 0; c < n; c += kSortTile) {
            const int m = (int)(n - c < kSortTile ? n - c : kSortTile);
            emit_chunk<kBigThreads>(E, b, [&](int64_t j) { return skey[s + j]; },
                                    [&](int64_t j) { return winding[s + j]; }, c, m, n, carry, cnt, s_w);
        }
        if (threadIdx.x == 0) E.band_cnt[band] = cnt;
    }
}

// totals (zeroed) += (cells, spans, records) summed over the bands' counts: a grid-stride
// loop, one sum per block, three atomics per block.
constexpr int kTotBlocks = 64, kTotThreads = 256;
__global__ __launch_bounds__(kTotThreads) void k_band_totals(const int2* __restrict__ band_cnt,
                                                             const int32_t* __restrict__ n_bands,
                                                             int32_t* __restrict__ totals) {
    __shared__ int32_t s_c[kTotThreads / 64], s_s[kTotThreads / 64];
    const int32_t nb = *n_bands;
    int32_t c = 0, sp = 0;
    for (int32_t t = blockIdx.x * kTotThreads + threadIdx.x; t < nb; t += kTotBlocks * kTotThreads) {
        const int2 v = band_cnt[t];
        c += v.x;
        sp += v.y;
    }
    c = wave_sum(c);
    sp = wave_sum(sp);
    if ((threadIdx.x & 63) == 0) { s_c[threadIdx.x >> 6] = c; s_s[threadIdx.x >> 6] = sp; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tc = 0, ts = 0;
        for (int k = 0; k < kTotThreads / 64; ++k) { tc += s_c[k]; ts += s_s[k]; }
        if (tc) atomicAdd(&totals[0], tc);
        if (ts) atomicAdd(&totals[1], ts);
        if (tc + ts) atomicAdd(&totals[2], tc + ts);
    }
}

// Workgroups per CU for a class: as many as the CU's 32 wave slots and 160 KiB of LDS allow.
template <class T, int THREADS>
unsigned class_grid(int64_t max_bands, int n_cu) {
    constexpr int kLds = (int)(2 * Tile<THREADS>::kPadded * sizeof(T)) + 256;
    constexpr int kByWaves = 32 / Tile<THREADS>::kWaves, kByLds = 160 * 1024 / kLds;
    constexpr int kPerCu = kByWaves < kByLds ? kByWaves : kByLds;
    const int64_t g = (int64_t)n_cu * kPerCu;
    return (unsigned)(max_bands < g ? max_bands : g);
}

template <class W>
hipError_t launch_classes(const int32_t* band_start, const int32_t* n_bands, int64_t max_bands,
                          const typename W::T* src, const Unpacked& out, const int32_t* bpre, uint32_t* ticket,
                          int32_t* big_list, int32_t* big_excl, int32_t* big_count, int n_cu, hipStream_t s,
                          const int2* band_pr, const EmitArgs* emit, SideStream side) {
    typedef typename W::T T;
    const EmitArgs E = emit ? *emit : EmitArgs{};
    int32_t* long_list = big_list + 2 * max_bands;
    // the short bands (no dependence on the long ones) on the second stream when there is one
    hipStream_t ss = s;
    if (side.s2) {
        hipError_t e = hipEventRecord(side.fork, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(side.s2, side.fork, 0);
        if (e != hipSuccess) return e;
        ss = side.s2;
    }
    k_sort_bands<W, 64, false><<<class_grid<T, 64>(max_bands, n_cu), 64, 0, ss>>>(
        band_start, n_bands, src, out, bpre, nullptr, nullptr, nullptr, nullptr, nullptr, band_pr, E, emit != nullptr);
    k_band_classify<<<(unsigned)((max_bands + kClassifyThreads - 1) / kClassifyThreads), kClassifyThreads, 0, s>>>(
        band_start, n_bands, long_list, ticket + 1);
    k_sort_bands<W, 256, true><<<class_grid<T, 256>((max_bands + kChunk - 1) / kChunk, n_cu), 256, 0, s>>>(
        band_start, n_bands, src, out, bpre, long_list, ticket, big_list, big_excl, big_count, band_pr, E,
        emit != nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && side.s2) {
        e = hipEventRecord(side.join, side.s2);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, side.join, 0);
    }
    return e;
}

}  // namespace

hipError_t launch_sort_bands(const int32_t* band_start, const int32_t* n_bands, int64_t max_bands,
                             const unsigned long long* src, int32_t* skey, int32_t* sidx, int32_t* winding,
                             int32_t* big_list, int32_t* big_excl, int32_t* big_count, const int32_t* bpre,
                             uint32_t* ticket, int n_cu, hipStream_t s, const int2* band_pr,
                             const EmitArgs* emit, SideStream side, int64_t n_entries, hipStream_t s_small) {
    if (max_bands <= 0) return hipSuccess;
    const Unpacked out{skey, sidx, winding};
    if (band_pr) {
        if (sidx) return hipErrorInvalidValue;  // narrow words hold no index
        if (emit) {
            // production frames (after launch_band_classify): one wave per band, four bands per
            // workgroup; the bands above 512 words (at most N / 513 of them) in later launches
            const uint32_t* w = reinterpret_cast<const uint32_t*>(src);
            int32_t* lists = big_list + 2 * max_bands;  // 2 * max_bands slots: both lists of long bands
            const int64_t cap = 2 * max_bands;
            k_sort_bands_wave<0><<<(unsigned)((max_bands + 3) / 4), 256, 0, s_small ? s_small : s>>>(
                band_start, n_bands, w, bpre, band_pr, big_list, big_excl, big_count, lists, cap, ticket + 1, *emit);
            const int64_t max_med = std::min<int64_t>(max_bands, n_entries / (kSmallBandMax + 1) + 1);
            k_sort_bands_wave<1><<<(unsigned)((max_med + 3) / 4), 256, 0, s>>>(
                band_start, n_bands, w, bpre, band_pr, big_list, big_excl, big_count, lists, cap, ticket + 1, *emit);
            const int64_t max_large = std::min<int64_t>(max_bands, n_entries / (kMediumBandMax + 1) + 1);
            k_sort_bands_wave<2><<<(unsigned)((max_large + 3) / 4), 256, 0, s>>>(
                band_start, n_bands, w, bpre, band_pr, big_list, big_excl, big_count, lists, cap, ticket + 1, *emit);
            return hipGetLastError();
        }
        return launch_classes<Narrow>(band_start, n_bands, max_bands, reinterpret_cast<const uint32_t*>(src), out,
                                      bpre, ticket, big_list, big_excl, big_count, n_cu, s, band_pr, emit, side);
    }
    if (emit) return hipErrorInvalidValue;  // wide words: the debug path keeps the separate mark + emit
    return launch_classes<Wide>(band_start, n_bands, max_bands, src, out, bpre, ticket, big_list, big_excl,
                                big_count, n_cu, s, nullptr, nullptr, side);
}

hipError_t launch_band_classify(const int32_t* band_start, const int32_t* n_bands, int64_t max_bands,
                                const int32_t* bpre, int32_t* big_list, int32_t* big_excl, int32_t* big_count,
                                uint32_t* ticket, hipStream_t s) {
    if (max_bands <= 0) return hipSuccess;
    int32_t* lists = big_list + 2 * max_bands;
    k_band_classify2<<<(unsigned)((max_bands + kClassifyThreads - 1) / kClassifyThreads), kClassifyThreads, 0, s>>>(
        band_start, n_bands, lists, 2 * max_bands, ticket + 1, bpre, big_list, big_excl, big_count);
    return hipGetLastError();
}

hipError_t launch_emit_big_bands(const int32_t* band_start, const int32_t* big_list, const int32_t* big_count,
                                 const int32_t* skey, const int32_t* winding, const int2* band_pr,
                                 const EmitArgs& emit, hipStream_t s) {
    k_emit_big_bands<<<64, kBigThreads, 0, s>>>(band_start, big_list, big_count, skey, winding, band_pr, emit);
    return hipGetLastError();
}

hipError_t launch_band_totals(const int2* band_cnt, const int32_t* n_bands, int32_t* totals, hipStream_t s) {
    k_band_totals<<<kTotBlocks, kTotThreads, 0, s>>>(band_cnt, n_bands, totals);
    return hipGetLastError();
}

hipError_t launch_sort_big_bands(const int32_t* band_start, const int32_t* big_list, const int32_t* big_excl,
                                 const int32_t* big_count, unsigned long long* scatter, unsigned long long* tmp,
                                 int32_t* skey, int32_t* sidx, int32_t* winding, hipStream_t s,
                                 const int2* band_pr) {
    const Unpacked out{skey, sidx, winding};
    if (band_pr) {
        if (sidx) return hipErrorInvalidValue;
        k_sort_big_bands<Narrow><<<64, kBigThreads, 0, s>>>(band_start, big_list, big_excl, big_count,
                                                          reinterpret_cast<uint32_t*>(scatter),
                                                          reinterpret_cast<uint32_t*>(tmp), out, band_pr);
        return hipGetLastError();
    }
    k_sort_big_bands<Wide><<<64, kBigThreads, 0, s>>>(band_start, big_list, big_excl, big_count, scatter, tmp, out,
                                                    nullptr);
    return hipGetLastError();
}

}  // namespace spr
