// k_sort.hip -- per-path segmented sort of the cell keys (gfx950).
//
// Replaces naive_seg_sort_pairs.comp:26-97 (one workgroup per path running an O(n^2) odd-even
// transposition network) and shuffle_fragment.comp:17-26.  The reference compares (key, index)
// as signed int32 inside each path segment; here the key is biased (key ^ 0x80000000) so the
// unsigned order equals the signed one, and the sort word is
//     (biased key << 32) | (index << 2) | (winding change + 1)
// which is unique, orders exactly like (key, index), and carries the winding change so the
// reference's separate gather pass disappears.
//
// Segments are cut into 4096-entry tiles.  A tile is loaded once from HBM, sorted in LDS (a
// 16-element register network per thread, then merge-path rounds in LDS) and stored once.
// Segments longer than one tile are finished by global merge-path passes.
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
static_assert(kThreads * kItems == kSortTile, "tile shape");
typedef unsigned long long u64;
constexpr u64 kPad = ~0ull;

__device__ inline u64 sort_word(int32_t key, int64_t idx, int8_t wn) {
    return ((u64)((uint32_t)key ^ 0x80000000u) << 32) | ((u64)idx << 2) | (u64)(uint32_t)(wn + 1);
}

__device__ inline void cswap(u64& a, u64& b) {
    u64 lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo; b = hi;
}

// Ascending bitonic network over 16 registers (indices are compile-time constants).
__device__ inline void sort16(u64 (&v)[kItems]) {
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
__device__ inline int merge_split(PA A, int na, PB B, int nb, int diag) {
    int lo = diag > nb ? diag - nb : 0;
    int hi = diag < na ? diag : na;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] <= B[diag - 1 - mid]) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Maps a block to (path, tile of that path) through the exclusive tile prefix.
__device__ inline bool block_tile(const int32_t* tile_prefix, uint32_t Np, int64_t b, uint32_t& p, int& t) {
    __shared__ int32_t s_p;
    if (threadIdx.x == 0) {
        const int32_t total = tile_prefix[Np];
        int32_t res = -1;
        if (b < total) {
            uint32_t lo = 0, hi = Np;  // largest p with tile_prefix[p] <= b
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (tile_prefix[mid] <= b) lo = mid; else hi = mid;
            }
            while (lo + 1 <= Np && tile_prefix[lo + 1] <= b) ++lo;
            res = (int32_t)lo;
        }
        s_p = res;
    }
    __syncthreads();
    const int32_t r = s_p;
    __syncthreads();
    if (r < 0) return false;
    p = (uint32_t)r;
    t = (int)(b - tile_prefix[p]);
    return true;
}

__global__ void k_tile_counts(const int32_t* __restrict__ seg, uint32_t Np, int32_t* __restrict__ tile_counts) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Np) return;
    const int32_t n = seg[p + 1] - seg[p];
    tile_counts[p] = (n + kSortTile - 1) / kSortTile;
}

__global__ __launch_bounds__(kThreads) void k_sort_tiles(const int32_t* __restrict__ key,
                                                         const int8_t* __restrict__ wn,
                                                         const int32_t* __restrict__ seg,
                                                         const int32_t* __restrict__ tile_prefix,
                                                         uint32_t Np, u64* __restrict__ out_final,
                                                         u64* __restrict__ out_big) {
    __shared__ u64 buf[2][kSortTile];
    uint32_t p;
    int t;
    if (!block_tile(tile_prefix, Np, blockIdx.x, p, t)) return;
    const int64_t s = seg[p], e = seg[p + 1];
    const int64_t len = e - s;
    const int64_t b0 = s + (int64_t)t * kSortTile;
    const int n = (int)(len - (int64_t)t * kSortTile < kSortTile ? len - (int64_t)t * kSortTile : kSortTile);
    u64* out = len > kSortTile ? out_big : out_final;
    int P2 = kItems;
    while (P2 < n) P2 <<= 1;
    const int active = P2 / kItems;
    const int tid = threadIdx.x;
    if (tid < active) {
        u64 v[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int i = tid * kItems + j;
            v[j] = i < n ? sort_word(key[b0 + i], b0 + i, wn[b0 + i]) : kPad;
        }
        sort16(v);
#pragma unroll
        for (int j = 0; j < kItems; ++j) buf[0][tid * kItems + j] = v[j];
    }
    __syncthreads();
    int src = 0;
    for (int run = kItems; run < P2; run <<= 1) {
        if (tid < active) {
            const int o = tid * kItems;
            const int pb = (o / (2 * run)) * (2 * run);
            const u64* A = &buf[src][pb];
            const u64* B = A + run;
            const int diag = o - pb;
            int i = merge_split(A, run, B, run, diag);
            int j = diag - i;
            u64* D = &buf[src ^ 1][o];
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                const bool takeA = j >= run || (i < run && A[i] <= B[j]);
                D[k] = takeA ? A[i] : B[j];
                i += takeA; j += !takeA;
            }
        }
        __syncthreads();
        src ^= 1;
    }
    for (int i = tid; i < n; i += kThreads) out[b0 + i] = buf[src][i];
}

// One merge pass over every segment longer than (kSortTile << pass): runs of length
// r = kSortTile << pass are merged pairwise from src into dst, or into out_final on the
// segment's last pass.  Blocks map to 4096-output slices of the segments.
__global__ __launch_bounds__(kThreads) void k_merge_pass(const int32_t* __restrict__ seg,
                                                         const int32_t* __restrict__ tile_prefix,
                                                         uint32_t Np, int pass,
                                                         const u64* __restrict__ src,
                                                         u64* __restrict__ dst,
                                                         u64* __restrict__ out_final) {
    uint32_t p;
    int t;
    if (!block_tile(tile_prefix, Np, blockIdx.x, p, t)) return;
    const int64_t s = seg[p], len = seg[p + 1] - s;
    const int64_t r = (int64_t)kSortTile << pass;
    if (len <= r) return;
    u64* out = (2 * r >= len) ? out_final : dst;
    const int64_t o = (int64_t)t * kSortTile + (int64_t)threadIdx.x * kItems;
    if (o >= len) return;
    const int64_t pb = (o / (2 * r)) * (2 * r);
    const int64_t na = (pb + r < len ? r : len - pb);
    const int64_t nb = pb + r < len ? ((pb + 2 * r < len ? pb + 2 * r : len) - (pb + r)) : 0;
    const u64* A = src + s + pb;
    const u64* B = A + na;
    const int cnt = (int)(len - o < kItems ? len - o : kItems);
    if (nb == 0) {
        for (int k = 0; k < cnt; ++k) out[s + o + k] = src[s + o + k];
        return;
    }
    const int diag = (int)(o - pb);
    int i = merge_split(A, (int)na, B, (int)nb, diag);
    int j = diag - i;
    for (int k = 0; k < cnt; ++k) {
        const bool takeA = j >= nb || (i < na && A[i] <= B[j]);
        out[s + o + k] = takeA ? A[i] : B[j];
        i += takeA; j += !takeA;
    }
}

__global__ void k_unpack_sorted(const u64* __restrict__ sorted, int64_t N, int32_t* __restrict__ skey,
                                int32_t* __restrict__ sidx, int32_t* __restrict__ swn) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const u64 w = sorted[i];
    skey[i] = (int32_t)((uint32_t)(w >> 32) ^ 0x80000000u);
    if (sidx) sidx[i] = (int32_t)((w >> 2) & 0x3FFFFFFFull);
    swn[i] = (int32_t)(w & 3ull) - 1;
}

}  // namespace

hipError_t launch_tile_counts(const int32_t* seg, uint32_t Np, int32_t* tile_counts, hipStream_t s) {
    if (Np == 0) return hipSuccess;
    k_tile_counts<<<(Np + 255) / 256, 256, 0, s>>>(seg, Np, tile_counts);
    return hipGetLastError();
}

hipError_t launch_sort_tiles(const int32_t* key, const int8_t* wn, const int32_t* seg,
                             const int32_t* tile_prefix, uint32_t Np, int64_t max_tiles,
                             unsigned long long* out_final, unsigned long long* out_big,
                             hipStream_t s) {
    if (max_tiles <= 0) return hipSuccess;
    k_sort_tiles<<<(unsigned)max_tiles, kThreads, 0, s>>>(key, wn, seg, tile_prefix, Np, out_final, out_big);
    return hipGetLastError();
}

hipError_t launch_merge_pass(const int32_t* seg, const int32_t* tile_prefix, uint32_t Np,
                             int64_t max_tiles, int pass, const unsigned long long* src,
                             unsigned long long* dst, unsigned long long* out_final,
                             hipStream_t s) {
    if (max_tiles <= 0) return hipSuccess;
    k_merge_pass<<<(unsigned)max_tiles, kThreads, 0, s>>>(seg, tile_prefix, Np, pass, src, dst, out_final);
    return hipGetLastError();
}

hipError_t launch_unpack_sorted(const unsigned long long* sorted, int64_t N, int32_t* skey,
                                int32_t* sidx, int32_t* swn, hipStream_t s) {
    if (N == 0) return hipSuccess;
    k_unpack_sorted<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(sorted, N, skey, sidx, swn);
    return hipGetLastError();
}

}  // namespace spr
