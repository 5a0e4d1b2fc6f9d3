// k_bucket.hip -- (path, cell row) buckets: the MSD step of the segmented sort (gfx950).
//
// A path's sort keys are (row, x) cell keys, so its entries split into one bucket per cell row,
// followed by an off-canvas bucket and the row-0 bucket (the reference compares keys as signed
// int32: rows 2..h-2, then the off-canvas sentinel, then row 0; gen_fragment.comp:180-187,
// naive_seg_sort_pairs.comp:69).  The transform pass bounds each path's rows from its control
// points (a Bezier curve stays inside their hull), so each path owns a small, exactly sized
// bucket range.  Entries are counted per bucket while their keys are generated, and the buckets
// are packed into bands of at most 2048 entries that one workgroup sorts in LDS.  The scatter
// only has to bring each entry into its band (the band sort orders by bucket, then x), so the
// entries of a run that fall into one band take their slots with one atomic and are stored
// contiguously.  After the sort the per-bucket ranges give the composite each path's records of
// each cell row without any further sorting.
#include "device_common.hpp"
#include "kernels.hpp"
#include "scan_common.hpp"

namespace spr {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    return (unsigned)(g < 1 ? 1 : g);
}

// Per-path bucket table from the path's cell-row bounds [lo, hi] (clamped to the canvas):
// in-canvas rows r >= 1 get buckets first..first+nrows-1, then the off-canvas bucket, then row 0.
__global__ void k_bucket_counts(const int32_t* __restrict__ row_lo, const int32_t* __restrict__ row_hi,
                                uint32_t Np, int rmax, int32_t* __restrict__ nb, int2* __restrict__ prow,
                                const int2* __restrict__ pcanvas) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= Np) return;
    if (pcanvas) rmax = (pcanvas[p].y - 1) / 2;
    int lo = row_lo[p], hi = row_hi[p];
    int first = lo < 1 ? 1 : lo;
    int last = hi > rmax ? rmax : hi;
    int nrows = last >= first ? last - first + 1 : 0;
    if (lo > hi) { first = 1; nrows = 0; }  // no points
    prow[p] = make_int2(first, nrows);
    nb[p] = nrows + 2;
}

// Scatter into band order.  The wide sort word orders exactly like the reference's signed
// (key, index) pair and carries the winding change in its two low bits.  The narrow word (used
// when no stage is exported) drops the index: (bucket inside its path) << 17 | (x + 2) << 2 |
// (winding change + 1); inside a band (whole buckets of one path) it orders like the key, and the
// order among equal keys changes no record, no span and no pixel (no span starts between equal
// keys, and the winding number after a run of equal keys is the sum over the whole run).  The
// band of bucket b is flag_scan[b + 1] - 1 (flag_scan: exclusive scan of the band-start flags).
// Each wave takes kRounds chunks of 64 consecutive entries and issues every load, then every
// slot atomic (one per run of equal bands, on the band's fill counter), of all chunks before it
// waits on any of them.  The head of a run also adds the run's winding changes to the band's sum
// (bsum), from which one scan gives every band its exclusive prefix before the band sort.
// PACKED (narrow only): the piece kernels already stored each entry's global bucket (in key) and
// its narrow word (in epath), so the path-table gathers are skipped.
constexpr int kRounds = 4;

template <bool NARROW, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_bucket_scatter(const int32_t* __restrict__ key,
                                                           const int8_t* __restrict__ wn,
                                                           const uint32_t* __restrict__ epath, int64_t N,
                                                           const int32_t* __restrict__ bucket_base,
                                                           const int2* __restrict__ prow,
                                                           const int32_t* __restrict__ flag_scan,
                                                           const int32_t* __restrict__ band_start,
                                                           int32_t* __restrict__ fill,
                                                           int32_t* __restrict__ bsum,
                                                           unsigned long long* __restrict__ out,
                                                           uint32_t* __restrict__ out32,
                                                           int32_t* __restrict__ err,
                                                           const int32_t* __restrict__ n_dev) {
    const int64_t n = n_dev ? min((int64_t)*n_dev, N) : N;
    const int lane = threadIdx.x & 63;
    const int64_t g0 = ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64 * kRounds);
    uint32_t pth[kRounds];
    int32_t k[kRounds], wv[kRounds];
    bool valid[kRounds];
    uint32_t b[kRounds], loc[kRounds];
    bool sent[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int64_t i = g0 + r * 64 + lane;
        valid[r] = i < n;
        pth[r] = kNoPath; k[r] = 0; wv[r] = 0;
        b[r] = 0; loc[r] = 0; sent[r] = false;
        if constexpr (PACKED) {
            // k = the narrow word, b = the global bucket
            if (valid[r]) {
                b[r] = (uint32_t)key[i];
                k[r] = (int32_t)epath[i];
                wv[r] = (int32_t)((uint32_t)k[r] & 3u) - 1;
            }
        } else {
            if (valid[r]) { pth[r] = epath[i]; k[r] = key[i]; wv[r] = wn[i]; }
            if (pth[r] == kNoPath) { valid[r] = false; wv[r] = 0; }  // an entry without a fragment
        }
    }
    if constexpr (!PACKED) {
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            if (valid[r]) {
                const int2 pr = prow[pth[r]];
                loc[r] = bucket_local(k[r], pr, err);
                sent[r] = loc[r] == (uint32_t)pr.y;
                b[r] = (uint32_t)bucket_base[pth[r]] + loc[r];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (valid[r]) b[r] = (uint32_t)(flag_scan[b[r] + 1] - 1);  // bucket -> band
    // runs of equal bands: heads take their run's slots with one atomic each, and add the run's
    // winding changes (an inclusive wave scan read at the run's last lane) to the band's sum
    unsigned long long vm[kRounds], hd[kRounds];
    int32_t base[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        vm[r] = __ballot(valid[r]);
        hd[r] = run_heads(b[r], valid[r], vm[r]);
        const int32_t inc = wave_inclusive_add(wv[r]);
        const bool head = (hd[r] >> lane) & 1ull;
        const int len = head ? run_length(lane, hd[r], vm[r]) : 1;
        const int32_t at_end = __shfl(inc, lane + len - 1);
        base[r] = 0;
        if (head) {
            base[r] = band_start[b[r]] + atomicAdd(&fill[b[r]], len);
            const int32_t rs = at_end - inc + wv[r];
            if (rs != 0) atomicAdd(&bsum[b[r]], rs);
        }
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const unsigned long long mine = hd[r] & ((2ull << lane) - 1ull);
        const int h = (valid[r] && mine) ? 63 - __clzll((long long)mine) : lane;
        const int32_t slot = __shfl(base[r], h) + (lane - h);
        if (valid[r]) {
            const int64_t i = g0 + r * 64 + lane;
            if constexpr (PACKED) {
                out32[slot] = (uint32_t)k[r];
            } else if constexpr (NARROW) {
                // the off-canvas bucket holds one key: its x is left at 0
                const uint32_t xl = sent[r] ? 0u : ((uint32_t)k[r] & 0xFFFFu) - kNarrowX0;
                out32[slot] = (loc[r] << 17) | (xl << 2) | (uint32_t)(wv[r] + 1);
            } else {
                const unsigned long long w = ((unsigned long long)((uint32_t)k[r] ^ 0x80000000u) << 32) |
                                             ((unsigned long long)i << 2) | (unsigned long long)(uint32_t)(wv[r] + 1);
                out[slot] = w;
            }
        }
    }
}

// Band starts: a band is a run of consecutive buckets of one path whose entries start in the
// same 512-entry window (kBandWindowLog); a bucket of more than 512 entries is a band of its own.
// Every band therefore holds at most 1024 entries (sorted by one wave, 16 per lane) unless it is a
// single larger bucket.
// One wave per path: the most recent non-empty bucket before each bucket is the nearest lane
// below it in the ballot of non-empty buckets (or, for the chunk's first ones, the last
// non-empty bucket of the earlier chunks), read with one shuffle.
__global__ __launch_bounds__(kBlock) void k_band_flags(const int32_t* __restrict__ bucket_base, uint32_t Np,
                                                       const int32_t* __restrict__ boff,
                                                       int32_t* __restrict__ flag, int window_log) {
    const uint32_t p = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= Np) return;
    const int32_t b0 = bucket_base[p], b1 = bucket_base[p + 1];
    bool have = false;         // a non-empty bucket in an earlier chunk
    int32_t cps = 0, cpz = 0;  // its start and size
    for (int32_t c = b0; c < b1; c += 64) {
        const int32_t b = c + lane;
        int32_t o = 0, sz = 0;
        if (b < b1) { o = boff[b]; sz = boff[b + 1] - o; }
        const unsigned long long ne = __ballot(sz > 0);
        const unsigned long long below = ne & ((1ull << lane) - 1ull);
        const int src = below ? 63 - __clzll((long long)below) : lane;
        const int32_t ps_l = __shfl(o, src), pz_l = __shfl(sz, src);
        if (b < b1) {
            int32_t f = 0;
            if (sz > 0) {
                const bool prev = below != 0ull || have;
                if (!prev) {
                    f = 1;
                } else {
                    const int32_t ps = below ? ps_l : cps, pz = below ? pz_l : cpz;
                    const int32_t win = 1 << window_log;
                    f = (o >> window_log) != (ps >> window_log) || pz > win || sz > win;
                }
            }
            flag[b] = f;
        }
        if (ne) {
            const int last = 63 - __clzll((long long)ne);
            cps = __shfl(o, last);
            cpz = __shfl(sz, last);
            have = true;
        }
    }
}

// Band starts, and each band's path rows (of the path whose bucket range holds the band's first
// bucket).
__global__ void k_band_emit(const int32_t* __restrict__ flag_scan, const int32_t* __restrict__ boff,
                            int32_t NB_max, int64_t N_max, int32_t* __restrict__ band_start,
                            const int32_t* __restrict__ nb_dev, const int32_t* __restrict__ n_dev,
                            const int32_t* __restrict__ bucket_base, uint32_t Np, const int2* __restrict__ prow,
                            int2* __restrict__ band_pr, int32_t* __restrict__ band_path) {
    const int32_t NB = nb_dev ? min(*nb_dev, NB_max) : NB_max;
    const int64_t N = n_dev ? min((int64_t)*n_dev, N_max) : N_max;
    // grid-stride over the live buckets (a launch for a bucket budget is capped)
    for (int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x; b < NB; b += (int64_t)gridDim.x * kBlock) {
        if (flag_scan[b + 1] != flag_scan[b]) {
            band_start[flag_scan[b]] = boff[b];
            if (band_pr) {
                uint32_t lo = 0, hi = Np;  // bucket_base[lo] <= b < bucket_base[hi]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (bucket_base[mid] <= b) lo = mid; else hi = mid;
                }
                band_pr[flag_scan[b]] = prow[lo];
                if (band_path) band_path[flag_scan[b]] = (int32_t)lo;
            }
        }
        if (b == NB - 1) band_start[flag_scan[NB]] = (int32_t)N;
    }
}

}  // namespace

hipError_t launch_bucket_counts(const int32_t* row_lo, const int32_t* row_hi, uint32_t Np, int rmax,
                                int32_t* nb, int2* prow, hipStream_t s, const int2* pcanvas) {
    if (Np == 0) return hipSuccess;
    k_bucket_counts<<<grid_for(Np), kBlock, 0, s>>>(row_lo, row_hi, Np, rmax, nb, prow, pcanvas);
    return hipGetLastError();
}

hipError_t launch_bucket_scatter(const int32_t* key, const int8_t* wn, const uint32_t* epath, int64_t N,
                                 const int32_t* bucket_base, const int2* prow, const int32_t* flag_scan,
                                 const int32_t* band_start, int32_t* fill, int32_t* bsum, unsigned long long* out,
                                 int32_t* err, hipStream_t s, const int32_t* n_dev, bool narrow, bool packed) {
    if (N == 0) return hipSuccess;
    if (packed && !narrow) return hipErrorInvalidValue;
    const unsigned g = grid_for((N + kRounds - 1) / kRounds);
    if (packed)
        k_bucket_scatter<true, true><<<g, kBlock, 0, s>>>(key, wn, epath, N, bucket_base, prow, flag_scan,
                                                          band_start, fill, bsum, nullptr,
                                                          reinterpret_cast<uint32_t*>(out), err, n_dev);
    else if (narrow)
        k_bucket_scatter<true, false><<<g, kBlock, 0, s>>>(key, wn, epath, N, bucket_base, prow, flag_scan,
                                                           band_start, fill, bsum, nullptr,
                                                           reinterpret_cast<uint32_t*>(out), err, n_dev);
    else
        k_bucket_scatter<false, false><<<g, kBlock, 0, s>>>(key, wn, epath, N, bucket_base, prow, flag_scan,
                                                            band_start, fill, bsum, out, nullptr, err, n_dev);
    return hipGetLastError();
}

hipError_t launch_band_flags(const int32_t* bucket_base, uint32_t Np, const int32_t* boff, int32_t* flag,
                             hipStream_t s, int window_log) {
    if (Np == 0) return hipSuccess;
    if (window_log < kBandWindowLog || window_log > 10) return hipErrorInvalidValue;  // bands <= max_bands_for
    k_band_flags<<<(Np + 3) / 4, kBlock, 0, s>>>(bucket_base, Np, boff, flag, window_log);
    return hipGetLastError();
}

hipError_t launch_band_emit(const int32_t* flag_scan, const int32_t* boff, int32_t NB, int64_t N,
                            int32_t* band_start, hipStream_t s, const int32_t* nb_dev, const int32_t* n_dev,
                            const int32_t* bucket_base, uint32_t Np, const int2* prow, int2* band_pr,
                            int32_t* band_path) {
    if (NB == 0) return hipSuccess;
    unsigned g = grid_for(NB);
    if (g > 2048u) g = 2048u;  // the kernel strides over the live buckets
    k_band_emit<<<g, kBlock, 0, s>>>(flag_scan, boff, NB, N, band_start, nb_dev, n_dev, bucket_base, Np, prow,
                                     band_pr, band_path);
    return hipGetLastError();
}

}  // namespace spr
