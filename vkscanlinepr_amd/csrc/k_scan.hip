// k_scan.hip -- single-pass device-wide scans (decoupled look-back) for gfx950.
//
// k_scan_i32 replaces the reference's single-workgroup naive_scan (naive_scan.comp:27-73): an
// exclusive prefix sum with the total at out[n].  Each 256-thread block owns a 4096-element
// tile read and written once (16 contiguous int32 per thread, 4 x 16-byte accesses), so the
// kernel moves 8 bytes per element and is bounded by HBM bandwidth.
//
// k_mark_emit fuses mark_merged_fragment_and_span.comp:48-120, the 2N-flag scan of
// scanline_rasterizer.cpp:544-573 and gen_merged_fragment_and_span.comp:152-222 into one pass:
// the cell/span flags of a tile are counted, the (cells, spans) pair is scanned with one
// look-back, and the records are written at their final positions.  R is left on the device,
// which removes the reference's third host synchronisation.  It runs in debug frames (and on
// canvases too wide for the narrow sort words); the other frames do mark + emit inside the band
// sort (k_sort.hip).
#include <mutex>

#include "device_common.hpp"
#include "kernels.hpp"
#include "scan_common.hpp"

namespace spr {

namespace {

// n_dev (optional) holds the element count on the device (the launch covers n_max); total_out
// (optional) also receives the total.  A tile is V x THREADS int4: load k of wave w brings in the 64
// consecutive int4 of segment (k, w) with one coalesced 1 KiB access; each wave scans its segment in
// registers (lane sums, DPP wave scan), wave 0 scans the V x W segment totals (up to four per lane)
// and does the tile's look-back, and the results go back with the same coalesced stores.  No LDS
// staging: many tiles stay in flight, which hides the look-back's latency across XCDs.
template <int V, int THREADS>
__global__ __launch_bounds__(THREADS) void k_scan_i32(const int32_t* in, int32_t* out, int64_t n_max,
                                                      const int32_t* __restrict__ n_dev,
                                                      int32_t* __restrict__ total_out,
                                                      unsigned long long* __restrict__ status,
                                                      uint32_t* __restrict__ counter) {
    constexpr int W = THREADS / 64, SEG = V * W, PER = (SEG + 63) / 64;
    static_assert(SEG <= 256, "wave 0 scans the segment totals, four per lane");
    constexpr int TILE = 4 * V * THREADS;
    __shared__ int32_t s_seg[64 * PER];
    __shared__ int32_t s_tile;
    const int64_t n = n_dev ? min((int64_t)*n_dev, n_max) : n_max;
    // a launch sized for a budget: the blocks past the live tiles leave without a ticket, so the
    // live tiles' tickets go to exactly the first blocks (no atomics spent on the surplus)
    if ((int64_t)blockIdx.x * TILE >= (n > 0 ? n : 1)) return;
    const int64_t tile = ordered_tile_id(counter, &s_tile);
    const int64_t base = tile * (int64_t)TILE;
    if (base >= n && tile != 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool vec = base + TILE <= n && ((((uintptr_t)(in + base)) & 15u) == 0) &&
                     ((((uintptr_t)(out + base)) & 15u) == 0);
    int4 q[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int idx = (k * W + wave) * 64 + lane;
        if (vec) {
            q[k] = reinterpret_cast<const int4*>(in + base)[idx];
        } else {
            const int64_t e = base + 4 * (int64_t)idx;
            q[k].x = e < n ? in[e] : 0;
            q[k].y = e + 1 < n ? in[e + 1] : 0;
            q[k].z = e + 2 < n ? in[e + 2] : 0;
            q[k].w = e + 3 < n ? in[e + 3] : 0;
        }
    }
    int32_t lex[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int32_t sm = wadd(wadd(q[k].x, q[k].y), wadd(q[k].z, q[k].w));
        const int32_t inc = wave_inclusive_add(sm);
        lex[k] = (int32_t)((uint32_t)inc - (uint32_t)sm);
        if (lane == 63) s_seg[k * W + wave] = inc;
    }
    __syncthreads();
    if (wave == 0) {
        int32_t t[PER], ls = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            t[j] = PER * lane + j < SEG ? s_seg[PER * lane + j] : 0;
            ls = wadd(ls, t[j]);
        }
        const int32_t inc = wave_inclusive_add(ls);
        const int32_t total = __shfl(inc, 63);
        const int32_t pre = lookback<OpI32>(status, tile, total);
        int32_t run = wadd(pre, (int32_t)((uint32_t)inc - (uint32_t)ls));
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (PER * lane + j < SEG) s_seg[PER * lane + j] = run;
            run = wadd(run, t[j]);
        }
        const bool last_tile = (n == 0) ? (tile == 0) : (base <= n - 1 && n - 1 < base + TILE);
        if (last_tile && lane == 0) {
            const int32_t tot = wadd(pre, total);
            out[n] = tot;
            if (total_out) *total_out = tot;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int idx = (k * W + wave) * 64 + lane;
        const int32_t a = wadd(s_seg[k * W + wave], lex[k]);
        const int32_t b = wadd(a, q[k].x), c = wadd(b, q[k].y), d = wadd(c, q[k].z);
        if (vec) {
            reinterpret_cast<int4*>(out + base)[idx] = make_int4(a, b, c, d);
        } else {
            const int64_t e = base + 4 * (int64_t)idx;
            if (e < n) out[e] = a;
            if (e + 1 < n) out[e + 1] = b;
            if (e + 2 < n) out[e + 2] = c;
            if (e + 3 < n) out[e + 3] = d;
        }
    }
}

// (cells, spans) packed as two 31-bit counts under a 2-bit state.
struct OpPair {
    using T = unsigned long long;  // cells in bits 0..31, spans in bits 32..63
    __device__ static T zero() { return 0ull; }
    __device__ static T add(T a, T b) { return a + b; }
    __device__ static uint64_t encode(uint64_t st, T v) {
        return (st << 62) | ((v >> 32) << 31) | (v & 0x7FFFFFFFull);
    }
    __device__ static uint64_t state(uint64_t w) { return w >> 62; }
    __device__ static T value(uint64_t w) { return (w & 0x7FFFFFFFull) | (((w >> 31) & 0x7FFFFFFFull) << 32); }
};

// Path lookup of sorted elements.  With sstart == nullptr the path is epath[f]; otherwise f's
// path is the p with sstart[p] <= f < sstart[p+1] (binary search once, then a forward walk).
struct PathCursor {
    const uint32_t* epath;
    const int32_t* sstart;
    uint32_t Np, p;
    __device__ PathCursor(const uint32_t* e, const int32_t* s, uint32_t n, int64_t f0)
        : epath(e), sstart(s), Np(n), p(0) {
        if (sstart) {
            uint32_t lo = 0, hi = Np;  // largest p with sstart[p] <= f0
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((int64_t)sstart[mid] <= f0) lo = mid; else hi = mid;
            }
            p = lo;
        }
    }
    __device__ uint32_t at(int64_t f) {
        if (!sstart) return epath[f] & 0x3FFFFFFFu;
        while (p + 1 < Np && (int64_t)sstart[p + 1] <= f) ++p;
        return p;
    }
};

__device__ inline int key_x(int32_t k) { return (k & 0xFFFF) - 0x7FFF; }
__device__ inline int key_y(int32_t k) { return ((k >> 16) & 0xFFFF) - 0x7FFF; }

// mark + emit tiles: 512 threads x 16 elements.  The tiles' look-back is one chain, so larger
// tiles (half as many as 4096-element ones) finish sooner; the 64 KiB record staging keeps two
// blocks (16 waves) per CU.
constexpr int kMarkThreads = 512, kMarkTile = kMarkThreads * kScanItems;

__global__ __launch_bounds__(kMarkThreads) void k_mark_emit(
    const int32_t* __restrict__ skey, const int32_t* __restrict__ winding,
    const uint32_t* __restrict__ epath, const uint32_t* __restrict__ rule,
    const uint32_t* __restrict__ fill, int64_t N_max, const int32_t* __restrict__ n_dev,
    const int32_t* __restrict__ sstart, uint32_t Np, int64_t rec_cap, int w, int h,
    unsigned long long* __restrict__ status, uint32_t* __restrict__ counter,
    int4* __restrict__ records, uint32_t* __restrict__ crec, int32_t* __restrict__ flags,
    int32_t* __restrict__ totals,
    const int32_t* __restrict__ bucket_base, const int2* __restrict__ prow, int2* __restrict__ brec,
    int32_t* __restrict__ err, const int2* __restrict__ pcanvas, const int32_t* __restrict__ wbase,
    int32_t* __restrict__ cell_first) {
    __shared__ int32_t s_tile;
    __shared__ unsigned long long s_w[kMarkThreads / 64];
    __shared__ unsigned long long s_prefix;
    __shared__ uint32_t s_rec[2 * kMarkTile];  // the tile's compact records, stored coalesced
    const int64_t N = n_dev ? min((int64_t)*n_dev, N_max) : N_max;
    if ((int64_t)blockIdx.x * kMarkTile >= N) return;  // past the live tiles: no ticket taken
    const int64_t tile = ordered_tile_id(counter, &s_tile);
    const int64_t base = tile * (int64_t)kMarkTile;
    if (base >= N) return;
    const int64_t i0 = base + (int64_t)threadIdx.x * kScanItems;
    // Path of sorted element f: epath[f] when sorted segments coincide with the intersection
    // segments (plain layout); otherwise from the sorted-space path starts sstart[0..Np].
    PathCursor cur(epath, sstart, Np, i0 > 0 ? i0 - 1 : 0);
    int32_t kprev = 0;
    uint32_t pprev = 0;
    if (i0 > 0 && i0 <= N) { kprev = skey[i0 - 1]; pprev = cur.at(i0 - 1); }
    int32_t keys[kScanItems], wind[kScanItems];
    uint32_t paths[kScanItems];
    // the thread's 16 consecutive keys, winding numbers and paths: 16-byte loads when whole
    if (i0 + kScanItems <= N) {
        const int4* kq = reinterpret_cast<const int4*>(skey + i0);
        const int4* wq = reinterpret_cast<const int4*>(winding + i0);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const int4 a = kq[q], b = wq[q];
            keys[4 * q] = a.x; keys[4 * q + 1] = a.y; keys[4 * q + 2] = a.z; keys[4 * q + 3] = a.w;
            wind[4 * q] = b.x; wind[4 * q + 1] = b.y; wind[4 * q + 2] = b.z; wind[4 * q + 3] = b.w;
        }
        if (!sstart) {
            const uint4* pq = reinterpret_cast<const uint4*>(epath + i0);
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q) {
                const uint4 c = pq[q];
                paths[4 * q] = c.x & 0x3FFFFFFFu; paths[4 * q + 1] = c.y & 0x3FFFFFFFu;
                paths[4 * q + 2] = c.z & 0x3FFFFFFFu; paths[4 * q + 3] = c.w & 0x3FFFFFFFu;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kScanItems; ++j) paths[j] = cur.at(i0 + j);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int64_t f = i0 + j;
            keys[j] = 0; wind[j] = 0; paths[j] = 0;
            if (f < N) { keys[j] = skey[f]; wind[j] = winding[f]; paths[j] = cur.at(f); }
        }
    }
    // the neighbours across the thread's range, and the per-path tables of its first and last
    // paths (sorted, so the 16 elements almost always have one or two paths), all loaded here
    const int32_t k_before = kprev;
    const uint32_t p_before = pprev;
    const int64_t inext = i0 + kScanItems;
    const int32_t knext_out = inext < N ? skey[inext] : 0;
    const uint32_t pnext_out = inext < N ? cur.at(inext) : 0xFFFFFFFFu;
    const uint32_t pa = paths[0], pz = paths[kScanItems - 1];
    const uint32_t rule_a = rule[pa], rule_z = rule[pz];
    const int2 pr_a = prow[pa], pr_z = prow[pz];
    const int32_t bb_a = bucket_base[pa], bb_z = bucket_base[pz];
    uint32_t cellbits = 0, spanbits = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t f = i0 + j;
        if (f >= N) continue;
        const int32_t k1 = keys[j];
        const uint32_t p1 = paths[j];
        const int x1 = key_x(k1), y1 = key_y(k1);
        int cw = w, ch = h;
        if (pcanvas) { const int2 cv = pcanvas[p1]; cw = cv.x; ch = cv.y; }
        const bool in_canvas = !(x1 < 0 || y1 < 0 || x1 >= cw || y1 >= ch);
        bool cell, span = false;
        if (f == 0) {
            cell = in_canvas;
        } else {
            const uint32_t r = p1 == pa ? rule_a : (p1 == pz ? rule_z : rule[p1]);
            const int x0 = key_x(kprev), y0 = key_y(kprev);
            cell = in_canvas && (pprev != p1 || kprev != k1);
            // documents of a batch each scan their own winding numbers from zero
            const int32_t wn = (int32_t)((uint32_t)wind[j] - (uint32_t)(wbase ? wbase[p1] : 0));
            const bool wf = (r == 0u && wn != 0) || (r == 1u && (wn & 1) != 0);
            span = (y0 == y1) && (x0 + 2 < x1) && (pprev == p1) && wf;
        }
        cellbits |= (uint32_t)cell << j;
        spanbits |= (uint32_t)span << j;
        kprev = k1; pprev = p1;
    }
    const unsigned long long mine = (unsigned long long)__popc(cellbits) | ((unsigned long long)__popc(spanbits) << 32);
    unsigned long long total;
    const unsigned long long texcl = block_exclusive_add<unsigned long long, kMarkThreads / 64>(mine, s_w, total);
    if (threadIdx.x < 64) {
        const unsigned long long pre = lookback<OpPair>(status, tile, total);
        if (threadIdx.x == 0) s_prefix = pre;
    }
    __syncthreads();
    const unsigned long long off = s_prefix + texcl;
    int32_t cells_before = (int32_t)(off & 0xFFFFFFFFull);
    int32_t spans_before = (int32_t)(off >> 32);
    const int64_t tile_rec0 = (int64_t)(s_prefix & 0xFFFFFFFFull) + (int64_t)(s_prefix >> 32);
    const int64_t tile_recs = (int64_t)(total & 0xFFFFFFFFull) + (int64_t)(total >> 32);
    if (i0 < N) {
        int32_t kp = (i0 > 0) ? k_before : 0;
        uint32_t pp = (i0 > 0) ? p_before : 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int64_t f = i0 + j;
            if (f >= N) break;
            const int cell = (cellbits >> j) & 1, span = (spanbits >> j) & 1;
            const int32_t k1 = keys[j];
            const uint32_t p1 = paths[j];
            if (flags) { flags[f] = cell; flags[N + f] = span; }
            if ((cell | span) && (int64_t)(cells_before + spans_before + cell + span) > rec_cap) {
                atomicOr(err, 2);  // record capacity exceeded
            } else if (cell | span) {
                const int32_t o = cells_before + spans_before;
                const int loc = (int)(o - tile_rec0);
                // compact record for the composite: x | width << 16 (row and colour are implied by
                // the (path, row) record range); the full int4 record only for debug export
                // anti-aliased frames: each record's first sorted entry (-1 for a span)
                if (cell_first) {
                    if (cell) cell_first[o] = (int32_t)f;
                    if (span) cell_first[o + cell] = -1;
                }
                if (cell) {
                    const int x1 = key_x(k1);
                    s_rec[loc] = (uint32_t)x1 | (2u << 16);
                    if (records)
                        records[o] = make_int4((int32_t)((uint32_t)key_y(k1) << 16) | x1, 2, (int32_t)fill[p1],
                                               cells_before + 1);
                }
                if (span) {
                    int x0 = key_x(kp) + 2;
                    const int x1 = key_x(k1);
                    x0 = max(0, x0);
                    s_rec[loc + cell] = (uint32_t)x0 | ((uint32_t)(x1 - x0) << 16);
                    if (records)
                        records[o + cell] = make_int4((int32_t)((uint32_t)key_y(kp) << 16) | x0, x1 - x0,
                                                      (int32_t)fill[p1], 0);
                }
            }
            // record range of this element's (path, cell row) bucket, for the composite
            const int y1 = key_y(k1);
            const int chh = pcanvas ? pcanvas[p1].y : h;
            if ((uint32_t)k1 != 0xFFFEFFFEu && y1 >= 0 && y1 < chh) {
                const int32_t kn = j + 1 < kScanItems ? keys[j + 1 < kScanItems ? j + 1 : j] : knext_out;
                const uint32_t pn = j + 1 < kScanItems ? paths[j + 1 < kScanItems ? j + 1 : j] : pnext_out;
                const bool first = (f == 0) || pp != p1 || (kp >> 16) != (k1 >> 16);
                const bool last = (f == N - 1) || pn != p1 || (kn >> 16) != (k1 >> 16);
                if (first || last) {
                    const int2 pr = p1 == pa ? pr_a : (p1 == pz ? pr_z : prow[p1]);
                    const int r = y1 >> 1;
                    const int loc = r == 0 ? pr.y + 1 : r - pr.x;
                    if (loc >= 0 && loc < pr.y + 2) {  // out of range only after a flagged bound error
                        const int32_t bb = p1 == pa ? bb_a : (p1 == pz ? bb_z : bucket_base[p1]);
                        int2* br = &brec[bb + loc];
                        const int64_t e = min((int64_t)(cells_before + spans_before + cell + span), rec_cap);
                        if (first) br->x = (int32_t)min((int64_t)(cells_before + spans_before), e);
                        if (last) br->y = (int32_t)e;
                    }
                }
            }
            cells_before += cell;
            spans_before += span;
            kp = k1;
            pp = p1;
        }
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < tile_recs; i += kMarkThreads)
        if (tile_rec0 + i < rec_cap) crec[tile_rec0 + i] = s_rec[i];
    if (threadIdx.x == 0 && base <= N - 1 && N - 1 < base + kMarkTile) {
        const unsigned long long t = s_prefix + total;
        totals[0] = (int32_t)(t & 0xFFFFFFFFull);
        totals[1] = (int32_t)(t >> 32);
        totals[2] = (int32_t)((t & 0xFFFFFFFFull) + (t >> 32));
    }
}

}  // namespace

// ---- the scan without allocation: its tile statuses and tile counter live in a static device
// array (one copy per GPU, held until the process ends).  Host code reaches that array only through
// the address hipGetSymbolAddress returns on each device; the array's name is not a device address.
namespace {
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned long long scratch[kNoAllocScanTiles + 1];
unsigned long long* g_scratch_ptr[kNoAllocMaxDevices] = {};
std::mutex g_scratch_mu;

// The calling thread's current device's copy of `scratch` (looked up once per device).
hipError_t scratch_for_current_device(unsigned long long** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

hipError_t launch_scan_i32_noalloc(const int32_t* in, int32_t* out, int64_t n, hipStream_t s) {
    if (n < 0 || n > kNoAllocScanMaxN) return hipErrorInvalidValue;
    unsigned long long* ws = nullptr;
    hipError_t e = scratch_for_current_device(&ws);
    if (e != hipSuccess) return e;
    const int64_t tiles = scan_tiles(n > 0 ? n : 1);
    // the statuses and the counter must start at zero: a fill kernel on the same stream (the
    // array keeps the previous call's contents)
    e = launch_fill_u32(reinterpret_cast<uint32_t*>(ws), 2 * (tiles + 1), 0u, s);
    if (e != hipSuccess) return e;
    ScanCtl c;
    c.status = ws;
    c.counter = reinterpret_cast<uint32_t*>(ws + tiles);
    c.epoch = 0;
    return launch_scan_i32(in, out, n, c, s);
}

hipError_t launch_scan_i32(const int32_t* in, int32_t* out, int64_t n, ScanCtl ctl, hipStream_t s,
                           const int32_t* n_dev, int32_t* total_out) {
    // the status array holds scan_tiles(n) (4096-element) entries, enough for either tile shape:
    // 32768-element tiles (8 int4 x 1024 threads) for large scans, 4096-element tiles otherwise
    if (n >= ((int64_t)1 << 20)) {
        auto go = [&](auto kern, int v, int threads) {
            const int64_t tile = 4 * (int64_t)v * threads;
            kern<<<(unsigned)((n + tile - 1) / tile), threads, 0, s>>>(in, out, n, n_dev, total_out, ctl.status,
                                                                       ctl.counter);
        };
        switch (ctl.shape) {
            case 1: go(k_scan_i32<8, 512>, 8, 512); break;
            case 2: go(k_scan_i32<4, 1024>, 4, 1024); break;
            case 3: go(k_scan_i32<16, 512>, 16, 512); break;
            default: go(k_scan_i32<8, 1024>, 8, 1024); break;
        }
    } else {
        const int64_t tiles = n > 0 ? scan_tiles(n) : 1;
        k_scan_i32<4, 256><<<(unsigned)tiles, 256, 0, s>>>(in, out, n, n_dev, total_out, ctl.status, ctl.counter);
    }
    return hipGetLastError();
}

hipError_t launch_mark_emit(const int32_t* skey, const int32_t* winding, const uint32_t* epath,
                            const uint32_t* rule, const uint32_t* fill, int64_t N, int w, int h,
                            ScanCtl ctl, int4* records, uint32_t* crec, int32_t* flags, int32_t* totals,
                            const int32_t* bucket_base, const int2* prow, int2* brec, hipStream_t s,
                            const int32_t* n_dev, const int32_t* sstart, uint32_t Np, int64_t rec_cap,
                            int32_t* err, const int2* pcanvas, const int32_t* wbase, int32_t* cell_first) {
    if (N == 0) return hipSuccess;
    k_mark_emit<<<(unsigned)((N + kMarkTile - 1) / kMarkTile), kMarkThreads, 0, s>>>(skey, winding, epath, rule, fill, N, n_dev,
                                                                 sstart, Np, rec_cap, w, h, ctl.status,
                                                                 ctl.counter, records, crec, flags, totals,
                                                                 bucket_base, prow, brec, err, pcanvas, wbase,
                                                                 cell_first);
    return hipGetLastError();
}

}  // namespace spr
