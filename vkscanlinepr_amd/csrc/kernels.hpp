// kernels.hpp -- host-side launch wrappers of the gfx950 kernels.  Every wrapper enqueues on
// the given stream and returns hipGetLastError() of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spr {

struct Mvp { float m[16]; };

// Scan control: a tile-status array tagged with an epoch so it never needs clearing, and an
// atomic counter that hands out tile ids in dispatch order.
struct ScanCtl {
    unsigned long long* status;  // >= tiles entries
    uint32_t* counter;           // one u32, zero before the first use of the frame
    uint32_t epoch;              // distinct per scan invocation inside a frame sequence
};
constexpr int kScanTile = 4096;  // elements per scan tile (256 threads x 16)
inline int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

// ---- control
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);

// ---- geometry (a3, a4, a7, a8)
hipError_t launch_transform(const float2* pos, const uint32_t* pos_path, uint32_t P, const Mvp& m,
                            int w, int h, float2* tpos, int32_t* visible, int32_t* row_lo, int32_t* row_hi,
                            hipStream_t s);
hipError_t launch_count(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                        const uint32_t* cpath, const int32_t* visible, uint32_t C, int w, int h,
                        float4* cuts, uint32_t* ncuts, int32_t* counts, hipStream_t s);
hipError_t launch_path_segments(const uint32_t* cpath, uint32_t C, const int32_t* offsets,
                                uint32_t Np, int32_t* seg, hipStream_t s);
hipError_t launch_segment_stats(const int32_t* seg, uint32_t Np, int32_t* max_len, hipStream_t s);
// One (curve, piece, axis) chain of dependent crossing solves, and one monotone piece.
struct ChainDesc {
    int32_t curve, n, out;  // curve, number of crossings, first output slot
    int16_t side, cubic;
    float line0, step;      // first grid line and its step (+-2)
    float t0, t1;           // flagged piece start and end parameters
};
struct PieceDesc {
    int32_t base, nx, ny, curve;  // base = -1: no piece in this slot
    float t0;
};
constexpr int kChainClasses = 32;
hipError_t launch_chain_prep(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                             const uint32_t* cpath, const int32_t* visible, const float4* cuts,
                             const uint32_t* ncuts, const int32_t* offsets, uint32_t C, int w, int h,
                             ChainDesc* chains, PieceDesc* pieces, int32_t* class_count, hipStream_t s);
hipError_t launch_chain_order(const ChainDesc* chains, uint32_t C, const int32_t* class_count,
                              int32_t* class_off, int32_t* class_fill, int32_t* n_long, int32_t* order,
                              hipStream_t s);
hipError_t launch_chain_solve(const float2* tpos, const uint32_t* cpos, const ChainDesc* chains,
                              const int32_t* order, const int32_t* class_off, const int32_t* n_long,
                              int32_t* work, float* chain_t, uint8_t* mono, int n_cu, hipStream_t s);
hipError_t launch_piece_merge(const PieceDesc* pieces, uint32_t C, const ChainDesc* chains, const uint8_t* mono,
                              const float* chain_t, int32_t* raw, int8_t* start_mark, int2* inter,
                              int32_t* long_list, int32_t* long_count, uint8_t* pieces_seq, int n_cu,
                              hipStream_t s);
hipError_t launch_or_merge(const int32_t* raw, const int8_t* start_mark, int64_t N, int2* inter, hipStream_t s);
hipError_t launch_gen_fragment(const int2* inter, int64_t N, const uint32_t* cpos,
                               const uint32_t* ctype, const uint32_t* cpath, const float2* tpos,
                               int w, int h, int32_t* key, int8_t* wn, uint32_t* epath,
                               const int32_t* bucket_base, const int2* prow, int32_t* hist, int32_t* err,
                               hipStream_t s);

// ---- scans
hipError_t launch_scan_i32(const int32_t* in, int32_t* out, int64_t n, ScanCtl ctl, hipStream_t s);

// ---- (path, row) buckets and the band sort (a9 + a10): per path segment, ascending
// (key ^ 0x80000000, index); the winding change rides along in the low bits of the sort word.
constexpr int kSortTile = 2048;
hipError_t launch_bucket_counts(const int32_t* row_lo, const int32_t* row_hi, uint32_t Np, int rmax,
                                int32_t* nb, int2* prow, hipStream_t s);
hipError_t launch_bucket_scatter(const int32_t* key, const int8_t* wn, const uint32_t* epath, int64_t N,
                                 const int32_t* bucket_base, const int2* prow, const int32_t* boff,
                                 int32_t* fill, unsigned long long* out, int32_t* err, hipStream_t s);
hipError_t launch_band_flags(const int32_t* bucket_base, uint32_t Np, const int32_t* boff, int32_t* flag,
                             hipStream_t s);
hipError_t launch_band_emit(const int32_t* flag_scan, const int32_t* boff, int32_t NB, int64_t N,
                            int32_t* band_start, hipStream_t s);
hipError_t launch_sort_bands(const int32_t* band_start, const int32_t* n_bands, int64_t max_bands,
                             const unsigned long long* src, unsigned long long* dst, int32_t* big_list,
                             int32_t* big_count, hipStream_t s);
hipError_t launch_sort_big_bands(const int32_t* band_start, const int32_t* big_list, const int32_t* big_count,
                                 unsigned long long* scatter, unsigned long long* tmp, unsigned long long* dst,
                                 hipStream_t s);
hipError_t launch_unpack_sorted(const unsigned long long* sorted, int64_t N, int32_t* skey,
                                int32_t* sidx, int32_t* swn, hipStream_t s);

// ---- winding / mark / emit (a10 scan, a11); also writes each (path, row) bucket's record range
hipError_t launch_mark_emit(const int32_t* skey, const int32_t* winding, const uint32_t* epath,
                            const uint32_t* rule, const uint32_t* fill, int64_t N, int w, int h,
                            ScanCtl ctl, int4* records, int32_t* flags, int32_t* totals,
                            const int32_t* bucket_base, const int2* prow, int2* brec, hipStream_t s);

// ---- composite (a12)
hipError_t launch_composite(const int4* records, const int2* brec, const int32_t* bucket_base,
                            const int2* prow, const uint32_t* fill, uint32_t Np, int w, int h, int compose,
                            uint8_t* image, hipStream_t s);

}  // namespace spr
