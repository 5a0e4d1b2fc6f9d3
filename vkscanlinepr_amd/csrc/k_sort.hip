// k_sort.hip -- sort of the cell keys inside each (path, row) band (gfx950).
//
// Replaces naive_seg_sort_pairs.comp:26-97 (one workgroup per path running an O(n^2) odd-even
// transposition network) and shuffle_fragment.comp:17-26.  The reference compares (key, index)
// as signed int32 inside each path segment; the sort word here is
//     ((key ^ 0x80000000) << 32) | (index << 2) | (winding change + 1)
// whose unsigned order equals the reference's (key, index) order, is unique, and carries the
// winding change, so the reference's separate gather pass disappears.
//
// After the bucket scatter every band (<= 2048 entries, a run of whole buckets of one path) only
// needs sorting in place: one 256-thread workgroup loads it once, sorts it in LDS (8-element
// register network per thread, then merge-path rounds) and stores it once.  A band above 2048
// entries is a single oversized (path, row) bucket; such bands are queued and each is sorted by
// one workgroup with LDS chunk sorts followed by merge-path passes through global memory.
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
static_assert(kThreads * kItems == kSortTile, "tile shape");
typedef unsigned long long u64;
constexpr u64 kPad = ~0ull;

__device__ inline void cswap(u64& a, u64& b) {
    u64 lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo; b = hi;
}

// Ascending bitonic network over kItems registers (indices are compile-time constants).
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
__device__ inline int64_t merge_split(PA A, int64_t na, PB B, int64_t nb, int64_t diag) {
    int64_t lo = diag > nb ? diag - nb : 0;
    int64_t hi = diag < na ? diag : na;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[diag - 1 - mid]) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Sorted output, unpacked from the sort words: key, winding change and (optionally) index.
struct Unpacked {
    int32_t* skey;
    int32_t* sidx;
    int32_t* swn;
};
__device__ inline void put(const Unpacked& o, int64_t i, u64 w) {
    o.skey[i] = (int32_t)((uint32_t)(w >> 32) ^ 0x80000000u);
    if (o.sidx) o.sidx[i] = (int32_t)((w >> 2) & 0x3FFFFFFFull);
    o.swn[i] = (int32_t)(w & 3ull) - 1;
}

// LDS slot of element i: one pad word per 8 elements, so the 8-element runs owned by
// consecutive lanes start on different banks.
constexpr int kPadded = kSortTile + kSortTile / 8;
__device__ inline int pad(int i) { return i + (i >> 3); }

// Sorts n <= kSortTile words from src with one workgroup (buf: 2 x kPadded words of LDS) into dst
// (sort words) or, when dst is null, unpacked into out at offset o0.
__device__ inline void tile_sort(const u64* __restrict__ src, u64* __restrict__ dst, int n, u64 (*buf)[kPadded],
                                 const Unpacked& out, int64_t o0) {
    int P2 = kItems;
    while (P2 < n) P2 <<= 1;
    const int active = P2 / kItems;
    const int tid = threadIdx.x;
    if (tid < active) {
        u64 v[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int i = tid * kItems + j;
            v[j] = i < n ? src[i] : kPad;
        }
        sort16(v);
#pragma unroll
        for (int j = 0; j < kItems; ++j) buf[0][pad(tid * kItems + j)] = v[j];
    }
    __syncthreads();
    int cur = 0;
    for (int run = kItems; run < P2; run <<= 1) {
        if (tid < active) {
            const int o = tid * kItems;
            const int pb = (o / (2 * run)) * (2 * run);
            const u64* S = buf[cur];
            auto A = [&](int i) { return S[pad(pb + i)]; };
            auto B = [&](int j) { return S[pad(pb + run + j)]; };
            const int diag = o - pb;
            int lo = diag > run ? diag - run : 0, hi = diag < run ? diag : run;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (A(mid) <= B(diag - 1 - mid)) lo = mid + 1; else hi = mid;
            }
            int i = lo, j = diag - lo;
            u64* D = buf[cur ^ 1];
            u64 a = i < run ? A(i) : kPad, b = j < run ? B(j) : kPad;
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                const bool takeA = j >= run || (i < run && a <= b);
                D[pad(o + k)] = takeA ? a : b;
                if (takeA) { ++i; a = i < run ? A(i) : kPad; } else { ++j; b = j < run ? B(j) : kPad; }
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    if (dst) {
        for (int i = tid; i < n; i += kThreads) dst[i] = buf[cur][pad(i)];
    } else {
        for (int i = tid; i < n; i += kThreads) put(out, o0 + i, buf[cur][pad(i)]);
    }
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_sort_bands(const int32_t* __restrict__ band_start,
                                                         const int32_t* __restrict__ n_bands,
                                                         const u64* __restrict__ src, Unpacked out,
                                                         int32_t* __restrict__ big_list,
                                                         int32_t* __restrict__ big_count) {
    __shared__ u64 buf[2][kPadded];
    const int32_t b = blockIdx.x;
    if (b >= *n_bands) return;
    const int64_t s = band_start[b], e = band_start[b + 1];
    const int64_t n = e - s;
    if (n <= 0) return;
    if (n > kSortTile) {
        if (threadIdx.x == 0) big_list[atomicAdd(big_count, 1)] = b;
        return;
    }
    if (n == 1) {
        if (threadIdx.x == 0) put(out, s, src[s]);
        return;
    }
    tile_sort(src + s, nullptr, (int)n, buf, out, s);
}

// One workgroup per oversized band: LDS chunk sorts into `a`, then merge-path passes between
// `a` and `b` (the scatter buffer, already consumed), the last pass writing `dst`.
__global__ __launch_bounds__(kThreads) void k_sort_big_bands(const int32_t* __restrict__ band_start,
                                                             const int32_t* __restrict__ big_list,
                                                             const int32_t* __restrict__ big_count,
                                                             u64* __restrict__ scatter, u64* __restrict__ a,
                                                             Unpacked out) {
    __shared__ u64 buf[2][kPadded];
    const int32_t nbig = *big_count;
    for (int32_t q = blockIdx.x; q < nbig; q += gridDim.x) {
        const int32_t band = big_list[q];
        const int64_t s = band_start[band], n = band_start[band + 1] - s;
        for (int64_t c = 0; c < n; c += kSortTile) {
            const int m = (int)(n - c < kSortTile ? n - c : kSortTile);
            tile_sort(scatter + s + c, a + s + c, m, buf, out, 0);
        }
        __threadfence_block();
        __syncthreads();
        u64* cur = a + s;
        u64* nxt = scatter + s;
        for (int64_t r = kSortTile; r < n; r <<= 1) {
            const bool last_pass = 2 * r >= n;
            u64* dstw = last_pass ? nullptr : nxt;
            for (int64_t o = (int64_t)threadIdx.x * kItems; o < n; o += (int64_t)kThreads * kItems) {
                const int64_t pb = (o / (2 * r)) * (2 * r);
                const int64_t na = pb + r < n ? r : n - pb;
                const int64_t nb = pb + r < n ? ((pb + 2 * r < n ? pb + 2 * r : n) - (pb + r)) : 0;
                const u64* A = cur + pb;
                const u64* B = A + na;
                const int cnt = (int)(n - o < kItems ? n - o : kItems);
                const int64_t diag = o - pb;
                int64_t i = nb ? merge_split(A, na, B, nb, diag) : diag;
                int64_t j = diag - i;
                for (int k = 0; k < cnt; ++k) {
                    const bool takeA = j >= nb || (i < na && A[i] <= B[j]);
                    const u64 v = takeA ? A[i] : B[j];
                    if (dstw) dstw[o + k] = v; else put(out, s + o + k, v);
                    i += takeA; j += !takeA;
                }
            }
            __threadfence_block();
            __syncthreads();
            if (last_pass) { cur = nullptr; break; }
            u64* t = cur; cur = nxt; nxt = t;
        }
        if (cur) {  // a single chunk: no merge pass ran
            for (int64_t i = threadIdx.x; i < n; i += kThreads) put(out, s + i, cur[i]);
            __syncthreads();
        }
    }
}

__global__ void k_unpack_sorted(const u64* __restrict__ sorted, int64_t N_max, int32_t* __restrict__ skey,
                                int32_t* __restrict__ sidx, int32_t* __restrict__ swn,
                                const int32_t* __restrict__ n_dev) {
    const int64_t N = n_dev ? min((int64_t)*n_dev, N_max) : N_max;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const u64 w = sorted[i];
    skey[i] = (int32_t)((uint32_t)(w >> 32) ^ 0x80000000u);
    if (sidx) sidx[i] = (int32_t)((w >> 2) & 0x3FFFFFFFull);
    swn[i] = (int32_t)(w & 3ull) - 1;
}

}  // namespace

hipError_t launch_sort_bands(const int32_t* band_start, const int32_t* n_bands, int64_t max_bands,
                             const unsigned long long* src, int32_t* skey, int32_t* sidx, int32_t* swn,
                             int32_t* big_list, int32_t* big_count, hipStream_t s) {
    if (max_bands <= 0) return hipSuccess;
    k_sort_bands<<<(unsigned)max_bands, kThreads, 0, s>>>(band_start, n_bands, src, Unpacked{skey, sidx, swn},
                                                           big_list, big_count);
    return hipGetLastError();
}

hipError_t launch_sort_big_bands(const int32_t* band_start, const int32_t* big_list, const int32_t* big_count,
                                 unsigned long long* scatter, unsigned long long* tmp, int32_t* skey, int32_t* sidx,
                                 int32_t* swn, hipStream_t s) {
    k_sort_big_bands<<<64, kThreads, 0, s>>>(band_start, big_list, big_count, scatter, tmp,
                                             Unpacked{skey, sidx, swn});
    return hipGetLastError();
}

hipError_t launch_unpack_sorted(const unsigned long long* sorted, int64_t N, int32_t* skey,
                                int32_t* sidx, int32_t* swn, hipStream_t s, const int32_t* n_dev) {
    if (N == 0) return hipSuccess;
    k_unpack_sorted<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(sorted, N, skey, sidx, swn, n_dev);
    return hipGetLastError();
}

}  // namespace spr
