// scan_common.hpp -- wave64 / block primitives and the decoupled look-back used by every
// single-pass scan (prefix sums over counts, winding numbers and the cell/span flags).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spr {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kWaves = kScanThreads / 64;

// Tile states in the high bits of a 64-bit status word (0 = not yet published).
constexpr uint64_t kStAggregate = 1, kStInclusive = 2;

__device__ inline unsigned long long load_status(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_status(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ inline T wave_inclusive_add(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    return v;
}

// int32: the same scan with DPP row shifts (1, 2, 4, 8 inside each 16-lane row) and the row
// broadcasts of lanes 15 and 31, all in vector registers.  A lane without a source adds 0.
__device__ inline int32_t wave_inclusive_add(int32_t v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2, 3
    return v;
}

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Exclusive block scan of one value per thread; returns the exclusive prefix and the block
// total.  `lds` holds NW values (NW: the block's waves).
template <class T, int NW = kWaves>
__device__ inline T block_exclusive_add(T v, T* lds, T& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T inc = wave_inclusive_add(v);
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    T wave_off = T(0), tot = T(0);
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        T x = lds[k];
        if (k < wave) wave_off += x;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return wave_off + inc - v;
}

// A tile id in dispatch order (blocks can only wait on tiles that were handed out earlier).
__device__ inline int64_t ordered_tile_id(uint32_t* counter, int32_t* lds_slot) {
    if (threadIdx.x == 0) *lds_slot = (int32_t)atomicAdd(counter, 1u);
    __syncthreads();
    int64_t t = *lds_slot;
    __syncthreads();
    return t;
}

// Decoupled look-back over a status array: called by the first wave only.  Op provides
//   using T;  static T zero();  static T add(T, T);
//   static uint64_t encode(uint64_t state, T);  static uint64_t state(uint64_t);
//   static T value(uint64_t);
// Publishes the tile's aggregate, walks back over its predecessors 64 at a time, publishes the
// inclusive prefix and returns the exclusive prefix (valid in every lane of wave 0).
template <class Op>
__device__ inline typename Op::T lookback(unsigned long long* status, int64_t tile, typename Op::T agg,
                                         bool published = false) {
    using T = typename Op::T;
    const int lane = threadIdx.x & 63;
    if (tile == 0) {
        if (lane == 0 && !published) store_status(&status[0], Op::encode(kStInclusive, agg));
        return Op::zero();
    }
    if (lane == 0 && !published) store_status(&status[tile], Op::encode(kStAggregate, agg));
    T excl = Op::zero();
    int64_t pred = tile - 1;
    while (true) {
        const int64_t idx = pred - lane;
        unsigned long long st = idx >= 0 ? load_status(&status[idx]) : Op::encode(kStInclusive, Op::zero());
        const uint64_t state = Op::state(st);
        const unsigned long long inc_mask = __ballot(state == kStInclusive);
        const unsigned long long bad_mask = __ballot(state == 0);
        const int first_inc = inc_mask ? __ffsll((long long)inc_mask) - 1 : 64;
        const unsigned long long upto = first_inc >= 63 ? ~0ull : ((2ull << first_inc) - 1ull);
        if (bad_mask & upto) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        T v = lane <= first_inc ? Op::value(st) : Op::zero();
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v = Op::add(v, __shfl_xor(v, o));
        excl = Op::add(excl, v);
        if (first_inc < 64) break;
        pred -= 64;
    }
    if (lane == 0) store_status(&status[tile], Op::encode(kStInclusive, Op::add(excl, agg)));
    return excl;
}

// The first half of lookback, for tiles that have other work before they need their prefix:
// publishes the tile's aggregate (tile 0: its inclusive prefix) right away.  The later
// lookback(..., published = true) walks back and publishes the inclusive prefix.
template <class Op>
__device__ inline void publish_aggregate(unsigned long long* status, int64_t tile, typename Op::T agg) {
    if ((threadIdx.x & 63) == 0)
        store_status(&status[tile], Op::encode(tile == 0 ? kStInclusive : kStAggregate, agg));
}

struct OpI32 {
    using T = int32_t;
    __device__ static T zero() { return 0; }
    __device__ static T add(T a, T b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
    __device__ static uint64_t encode(uint64_t st, T v) { return (st << 32) | (uint32_t)v; }
    __device__ static uint64_t state(uint64_t w) { return w >> 32; }
    __device__ static T value(uint64_t w) { return (int32_t)(uint32_t)w; }
};

}  // namespace spr
