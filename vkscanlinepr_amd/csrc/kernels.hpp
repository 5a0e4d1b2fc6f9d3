// kernels.hpp -- host-side launch wrappers of the gfx950 kernels.  Every wrapper enqueues on
// the given stream and returns hipGetLastError() of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "paint.hpp"

namespace spr {

struct Mvp { float m[16]; };

// Scan control: a tile-status array tagged with an epoch so it never needs clearing, and an
// atomic counter that hands out tile ids in dispatch order.
struct ScanCtl {
    unsigned long long* status;  // >= tiles entries
    uint32_t* counter;           // one u32, zero before the first use of the frame
    uint32_t epoch;              // distinct per scan invocation inside a frame sequence
    int shape = 0;               // tile of the scans of 2^20 elements and more (kScanShapes; tuned)
};
// Tile shapes (int4 per thread, threads) of the large scans: 8 x 1024 (the default), 8 x 512,
// 4 x 1024, 16 x 512.  The status array is sized by 4096-element tiles, enough for every shape.
constexpr int kScanShapes = 4;
constexpr int kScanTile = 4096;  // elements per scan tile (256 threads x 16)
inline int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

// ---- control
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
hipError_t launch_fill_u32_dev(uint32_t* p, int64_t n_max, const int32_t* n_dev, int64_t mult, int64_t add,
                               uint32_t v, hipStream_t s);
// One device-to-host copy per frame: the frame's counters gathered from wherever the kernels
// left them into one contiguous int32 array.
constexpr int kSummaryMax = 24;
struct SummarySrc {
    const int32_t* p[kSummaryMax];
    int n;
};
hipError_t launch_gather_summary(const SummarySrc& src, int32_t* dst, hipStream_t s);

// ---- geometry (a3, a4, a7, a8)
// mvp_dev (optional): the matrix read from device memory instead of m.
hipError_t launch_transform(const float2* pos, const uint32_t* pos_path, uint32_t P, const Mvp& m,
                            int w, int h, float2* tpos, int32_t* visible, int32_t* row_lo, int32_t* row_hi,
                            hipStream_t s, const int2* pcanvas = nullptr, const float* mvp_dev = nullptr);
hipError_t launch_count(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                        const float* cweight, const uint32_t* cpath, const int32_t* visible, uint32_t C, int w, int h,
                        float4* cuts, uint32_t* ncuts, int32_t* counts, hipStream_t s,
                        const int2* pcanvas = nullptr);
hipError_t launch_path_segments(const uint32_t* cpath, uint32_t C, const int32_t* offsets,
                                uint32_t Np, int32_t* seg, hipStream_t s);
hipError_t launch_segment_stats(const int32_t* seg, uint32_t Np, int32_t* max_len, hipStream_t s);
// One (curve, piece, axis) chain of dependent crossing solves, and one monotone piece.
struct ChainDesc {
    int32_t pos, n, out;    // the curve's first transformed point, number of crossings, first output slot
    int16_t flags, type;    // flags: bit 0 the axis (y), bit 1 the grid step is -2 (else +2);
                            // type: the curve's (kLine, kQuad, kCubic, kArc)
    float line0, w;         // first grid line; an arc's weight (0 for other curves)
    float t0, t1;           // flagged piece start and end parameters
};
__host__ __device__ inline int chain_side(const ChainDesc& d) { return d.flags & 1; }
__host__ __device__ inline float chain_step(const ChainDesc& d) { return (d.flags & 2) ? -2.0f : 2.0f; }
struct PieceDesc {
    int32_t base, nx, ny, curve;  // base = -1: no piece in this slot
    float t0;                     // flagged start parameter
    float tend;                   // 1 for the curve's last piece, -1 otherwise (the next piece's first entry)
};
// Inputs and outputs of the fragment keys produced from the merged entries of each piece.
struct FragArgs {
    const uint32_t* cpos;
    const uint32_t* ctype;
    const float* cweight;       // arc weights (read for arcs only; null when there is none)
    const uint32_t* cpath;
    const float2* tpos;
    int w, h;
    const int2* pcanvas;        // per-path canvas of a batch, or null
    int32_t* key;
    int8_t* wn;
    uint32_t* epath;
    const int32_t* bucket_base;
    const int2* prow;
    int32_t* hist;              // per-bucket entry counts
    int32_t* err;
    int2* inter;                // (curve, t bits) entries for export, or null
    float4* chord;              // each entry's chord (pf.x, pf.y, pl.x, pl.y) for coverage, or null
    uint32_t* nword;            // frames with narrow sort words: each entry's word, with key[i] = its
                                // global bucket (wn and epath are then not written), or null
};
// Chain length classes: cubic chains by quarter octave of their length (16 octaves x 4), then
// line chains by octave.  The chain control block holds, as int32: class counts, class fill
// counters, class offsets [kChainClasses + 1], the number of wave-solved chains, the work counter
// of the persistent solve, the count of queued heavy pieces, the first light class's place, the
// light solve's work counter, the count of queued wave pieces that wait on heavy chains and the
// number of wave-solved chains at the head of the light launch.
constexpr int kChainClasses = 80;
constexpr int kCcFill = kChainClasses, kCcOff = 2 * kChainClasses, kCcNLong = 3 * kChainClasses + 1,
              kCcWork = kCcNLong + 1, kCcHeavyPieces = kCcWork + 1, kCcLightBegin = kCcHeavyPieces + 1,
              kCcWorkLight = kCcLightBegin + 1, kCcWavePieces = kCcWorkLight + 1, kCcNLongLight = kCcWavePieces + 1,
              kCcWords = kCcNLongLight + 1;
// Chains of at least kHeavyCrossings crossings are "heavy": they are solved by their own launch
// (wave-solved and the longest lane-solved chains), and the pieces of at most kHeavyCrossings
// entries, which never wait on them, are finished by a concurrent stream meanwhile.
constexpr int kHeavyCrossings = 256;
// A frame whose buffers were sized before its counts were known (no host synchronisation): the
// chain prep compares the exact total offsets[C] with the budget; a frame over budget writes no
// chain and no piece (as if every path were culled), sets *over and leaves 0 in *live, the number
// of entry positions the later kernels read; a frame within budget leaves the total in *live.
struct Budget {
    int64_t entries;
    int32_t* live;
    int32_t* over;
};
hipError_t launch_chain_prep(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                             const float* cweight, const uint32_t* cpath, const int32_t* visible, const float4* cuts,
                             const uint32_t* ncuts, const int32_t* offsets, uint32_t C, int w, int h,
                             ChainDesc* chains, PieceDesc* pieces, uint8_t* chain_cls, int32_t* class_count,
                             int n_cu, hipStream_t s, const Budget* budget = nullptr, const int2* pcanvas = nullptr);
// n_long / n_long_light: the wave-solved chains at the head of the heavy / light launch's range.
hipError_t launch_chain_order(const uint8_t* chain_cls, uint32_t C, const int32_t* class_count,
                              int32_t* class_off, int32_t* class_fill, int32_t* n_long, int32_t* light_begin,
                              int32_t* n_long_light, int32_t* order, int n_cu, hipStream_t s);
// heavy = 1: the chains of kHeavyCrossings and more; heavy = 0: the rest.  Each launch
// wave-solves the first *n_long chains of its range and lane-solves the others.  The two
// launches are independent.
hipError_t launch_chain_solve(const float2* tpos, const uint32_t* cpos, const float* cweight,
                              const ChainDesc* chains,
                              const int32_t* order, const int32_t* class_off, const int32_t* n_long,
                              const int32_t* light_begin, int heavy, int32_t* work, float* chain_t, uint8_t* mono,
                              int n_cu, hipStream_t s);
// Merges each piece's x and y chains into its (curve, t bits) entries, OR-merging equal t (the
// reference's per-piece walk), and turns the entries into fragment keys, winding changes and
// bucket counts (gen_fragment) without storing the entries unless A.inter is set.
//  launch_piece_small: one lane per piece of at most 32 entries whose work depends on light
//    chains only; the other pieces are listed per group of 64 piece slots in lo lists (up to
//    1024 entries, light chains only) or queued as heavy pieces (the rest).
//  launch_piece_lo: one wave per piece of the lo lists.
//  launch_piece_heavy: one workgroup per piece of the heavy queue, after every chain is solved;
//    scratch holds 2 * n_entries words.
struct PieceLists {
    int32_t* lo;       // [groups * 64] slots: pieces whose chains are all light
    int32_t* lo_cnt;   // [groups]
    int32_t* lo_h;     // [slots] queue: pieces of up to kWaveStage entries that wait on a heavy chain
    int32_t* lo_h_cnt; // its length
    int32_t* heavy;    // [slots]
    int32_t* heavy_count;
};
hipError_t launch_piece_small(const PieceDesc* pieces, uint32_t C, const uint8_t* mono, const float* chain_t,
                              const PieceLists& L, const FragArgs& A, int n_cu, hipStream_t s);
hipError_t launch_piece_lo(const PieceDesc* pieces, uint32_t C, const uint8_t* mono, const float* chain_t,
                           const int32_t* list, const int32_t* cnt, const FragArgs& A, hipStream_t s);
// One wave per piece of a queue (the pieces that wait on heavy chains), the waves striding over it.
hipError_t launch_piece_waves(const PieceDesc* pieces, const uint8_t* mono, const float* chain_t,
                              const int32_t* queue, const int32_t* queue_count, const FragArgs& A, int n_cu,
                              hipStream_t s);
hipError_t launch_piece_heavy(const PieceDesc* pieces, const PieceLists& L, const uint8_t* mono,
                              const float* chain_t, uint32_t* scratch, int64_t n_entries, const FragArgs& A, int n_cu,
                              hipStream_t s);
hipError_t launch_gen_fragment(const int2* inter, int64_t N, const uint32_t* cpos,
                               const uint32_t* ctype, const float* cweight, const uint32_t* cpath, const float2* tpos,
                               int w, int h, int32_t* key, int8_t* wn, uint32_t* epath,
                               const int32_t* bucket_base, const int2* prow, int32_t* hist, int32_t* err,
                               hipStream_t s, const int32_t* n_dev = nullptr, const int2* pcanvas = nullptr,
                               float4* chord = nullptr);

// ---- scans
hipError_t launch_scan_i32(const int32_t* in, int32_t* out, int64_t n, ScanCtl ctl, hipStream_t s,
                           const int32_t* n_dev = nullptr, int32_t* total_out = nullptr);
// The same scan with its workspace in a static device array instead of caller or handle memory:
// no allocation at all.  n <= kNoAllocScanMaxN; calls that overlap on one GPU race on the array.
constexpr int64_t kNoAllocScanMaxN = (int64_t)1 << 28;
constexpr int64_t kNoAllocScanTiles = (kNoAllocScanMaxN + kScanTile - 1) / kScanTile;
hipError_t launch_scan_i32_noalloc(const int32_t* in, int32_t* out, int64_t n, hipStream_t s);

// ---- (path, row) buckets and the band sort (a9 + a10): per path segment, ascending
// (key ^ 0x80000000, index); the winding change rides along in the low bits of the sort word.
constexpr int kSortTile = 2048;
// Bands: runs of consecutive buckets of one path whose entries start in one 2^kBandWindowLog-entry
// window (a bucket above the window's size is a band of its own), so a band holds at most twice
// the window unless it is a single larger bucket.  Band starts come from path starts, window
// changes and the buckets on either side of a large bucket: at most Np + 3 N / window + 2.
constexpr int kBandWindowLog = 9;
inline int64_t max_bands_for(uint32_t Np, int64_t n_entries) {
    return (int64_t)Np + 3 * (n_entries >> kBandWindowLog) + 2;
}
hipError_t launch_bucket_counts(const int32_t* row_lo, const int32_t* row_hi, uint32_t Np, int rmax,
                                int32_t* nb, int2* prow, hipStream_t s, const int2* pcanvas = nullptr);
// narrow: 32-bit sort words (first half of `out`) without the entry index, for frames that
// export no stage and whose canvases are narrower than kNarrowMaxW.
constexpr uint32_t kNarrowX0 = 0x7FFDu;   // x field of the leftmost key (x = -2)
constexpr int kNarrowMaxW = (1 << 15) - 4;
// bsum (zeroed, one int32 per band) receives each band's sum of winding changes.
// packed (narrow only): the entries are (key = global bucket, epath = narrow word), as the piece
// kernels write them with FragArgs::nword; otherwise (key, wn, epath).
hipError_t launch_bucket_scatter(const int32_t* key, const int8_t* wn, const uint32_t* epath, int64_t N,
                                 const int32_t* bucket_base, const int2* prow, const int32_t* flag_scan,
                                 const int32_t* band_start, int32_t* fill, int32_t* bsum, unsigned long long* out,
                                 int32_t* err, hipStream_t s, const int32_t* n_dev = nullptr, bool narrow = false,
                                 bool packed = false);
// window_log: bands start in 2^window_log-entry windows (9 = kBandWindowLog or 10; tuned).
hipError_t launch_band_flags(const int32_t* bucket_base, uint32_t Np, const int32_t* boff, int32_t* flag,
                             hipStream_t s, int window_log = kBandWindowLog);
// band_pr (or null): each band's path rows (prow of its path), for the narrow sort.
// band_path (with band_pr): each band's path.
hipError_t launch_band_emit(const int32_t* flag_scan, const int32_t* boff, int32_t NB, int64_t N,
                            int32_t* band_start, hipStream_t s, const int32_t* nb_dev = nullptr,
                            const int32_t* n_dev = nullptr, const int32_t* bucket_base = nullptr, uint32_t Np = 0,
                            const int2* prow = nullptr, int2* band_pr = nullptr, int32_t* band_path = nullptr);
// Mark + emit fused into the narrow band sort (frames that export no stage): each band flags its
// cells and spans (every flag depends only on the band's own entries: bands split at (path, row)
// bucket boundaries), writes its compact records at [2 * band_start, ...) of crec, the record
// range of each of its buckets into brec and its (cells, spans) into band_cnt.
struct EmitArgs {
    uint32_t* crec;
    int2* brec;
    int2* band_cnt;
    const int32_t* band_path;
    const int32_t* bucket_base;
    const uint32_t* rule;
    const int2* pcanvas;  // per-path canvas of a batch, or null
    const int32_t* wbase; // per-path winding base of a batch, or null
    int w, h;
};
// The band sorts write the sorted keys, winding changes and (optionally) indices directly.
// Sorts every band and writes keys, (debug) indices and winding numbers (the exclusive scan of
// the sorted winding changes, total at winding[n]); bpre holds each band's exclusive prefix of
// the winding changes (total at bpre[n_bands]).  Oversized bands are queued (with their
// exclusive prefix) for launch_sort_big_bands.  big_list: 3 * max_bands int32.  ticket: two
// zeroed u32 counters.  narrow (band_pr non-null): 32-bit words from the narrow scatter; sidx
// must be null.
// emit (narrow only, or null): mark + emit fused (EmitArgs); the bands then write no keys or
// winding numbers, except the oversized ones, which launch_emit_big_bands reads back.  With
// emit, launch_band_classify must run first: it lists the long bands for the wave sorts and
// queues the oversized ones, so that launch_sort_big_bands and launch_emit_big_bands may run on
// another stream beside the wave sorts.
// side (optional): a second stream and two events; the short bands are then sorted there,
// beside the long ones, and s waits for them before it returns.  n_entries: an upper bound of the
// sorted entries (sizes the launch for the bands above 512 words of frames with emit).  s_small
// (emit only, optional): the stream of the sort of the bands of up to 512 words, which the caller
// forks from s after launch_band_classify and joins.
struct SideStream {
    hipStream_t s2;
    hipEvent_t fork, join;
};
hipError_t launch_sort_bands(const int32_t* band_start, const int32_t* n_bands, int64_t max_bands,
                             const unsigned long long* src, int32_t* skey, int32_t* sidx, int32_t* winding,
                             int32_t* big_list, int32_t* big_excl, int32_t* big_count, const int32_t* bpre,
                             uint32_t* ticket, int n_cu, hipStream_t s, const int2* band_pr = nullptr,
                             const EmitArgs* emit = nullptr, SideStream side = {nullptr, nullptr, nullptr},
                             int64_t n_entries = 0, hipStream_t s_small = nullptr);
// Lists the bands of 513..1024 and 1025..2048 words (ticket[1..2] count them) and queues the
// longer ones with their exclusive prefixes (big_count zeroed): the first launch of a sort with
// emit.
hipError_t launch_band_classify(const int32_t* band_start, const int32_t* n_bands, int64_t max_bands,
                                const int32_t* bpre, int32_t* big_list, int32_t* big_excl, int32_t* big_count,
                                uint32_t* ticket, hipStream_t s);
// Mark + emit of the oversized bands from their sorted keys and winding numbers.
hipError_t launch_emit_big_bands(const int32_t* band_start, const int32_t* big_list, const int32_t* big_count,
                                 const int32_t* skey, const int32_t* winding, const int2* band_pr,
                                 const EmitArgs& emit, hipStream_t s);
// totals[0..2] (zeroed) += cells, spans, records: the sums of band_cnt over the bands.
hipError_t launch_band_totals(const int2* band_cnt, const int32_t* n_bands, int32_t* totals, hipStream_t s);
hipError_t launch_sort_big_bands(const int32_t* band_start, const int32_t* big_list, const int32_t* big_excl,
                                 const int32_t* big_count, unsigned long long* scatter, unsigned long long* tmp,
                                 int32_t* skey, int32_t* sidx, int32_t* winding, hipStream_t s,
                                 const int2* band_pr = nullptr);

// ---- winding / mark / emit (a10 scan, a11); also writes each (path, row) bucket's record range
hipError_t launch_mark_emit(const int32_t* skey, const int32_t* winding, const uint32_t* epath,
                            const uint32_t* rule, const uint32_t* fill, int64_t N, int w, int h,
                            ScanCtl ctl, int4* records, uint32_t* crec, int32_t* flags, int32_t* totals,
                            const int32_t* bucket_base, const int2* prow, int2* brec, hipStream_t s,
                            const int32_t* n_dev, const int32_t* sstart, uint32_t Np, int64_t rec_cap,
                            int32_t* err, const int2* pcanvas = nullptr, const int32_t* wbase = nullptr,
                            int32_t* cell_first = nullptr);

// ---- anti-aliased coverage (k_aa.hip): the 64-sample mask of every boundary cell record (2x2
// pixels x 16 samples, bit (2 py + px) * 16 + j), ~0 for spans.  cell_first[o] = the record's
// first sorted entry (-1: a span); the record count is totals[0] + totals[1] on the device.
constexpr int kAaSamples = 16;
// chord[i]: entry i's chord as the fragment stage formed it (FragArgs::chord).
hipError_t launch_cell_cover(const int32_t* cell_first, const int32_t* totals, int64_t rec_cap,
                             const int32_t* skey, const int32_t* sidx, const int32_t* winding,
                             const uint32_t* epath, const float4* chord, int64_t N, const uint32_t* rule,
                             unsigned long long* cmask, hipStream_t s);

// ---- learned single-pass allocation (k_learned.hip)
constexpr int kMlpIn = 16, kMlpHidden = 64, kMlpOut = 2;
// flat FP32 parameter vector: W1[16][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden * kMlpOut + kMlpOut;
struct MlpDev {
    const uint16_t* w1;     // BF16 [16][64]
    const uint16_t* w2;     // BF16 [64][64]
    const uint16_t* w3;     // BF16 [64][16], columns >= 2 are zero
    const float* b1;        // [64]
    const float* b2;        // [64]
    const float* b3;        // [2]
    const float* mean;      // [16] feature normalisation
    const float* inv_std;   // [16]
};
// The prediction's kernels take an optional device flag `skip`: when *skip != 0 they return at
// once and their outputs keep the values of the last frame that ran them (a captured frame holds
// them and the flag decides at each replay).
hipError_t launch_path_features(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                                const uint32_t* pcurve, const int32_t* visible, const uint32_t* rule,
                                uint32_t Np, int w, int h, float* feat, hipStream_t s,
                                unsigned long long* bound_total = nullptr, const int2* pcanvas = nullptr,
                                int32_t* pbound = nullptr, const int32_t* skip = nullptr);
hipError_t launch_mlp_infer(const float* feat, uint32_t Np, const MlpDev& m, float* pred, hipStream_t s,
                            const int32_t* skip = nullptr);
hipError_t launch_caps(const float* pred, const int32_t* visible, uint32_t Np, float log_margin, int32_t slack,
                       int32_t* arena, int32_t* cap_n, int32_t* cap_r_total, hipStream_t s,
                       const int32_t* pbound = nullptr, const int32_t* skip = nullptr);
// *n_over (zeroed) += visible paths whose exact entry count (seg[p+1] - seg[p]) exceeds their
// predicted capacity cap_n[p].
hipError_t launch_path_overflow(const int32_t* seg, const int32_t* cap_n, const int32_t* visible, uint32_t Np,
                                int32_t* n_over, hipStream_t s, const int32_t* skip = nullptr);
hipError_t launch_sorted_starts(const int32_t* boff, const int32_t* bucket_base, uint32_t Np, int32_t* sstart,
                                hipStream_t s);
hipError_t launch_path_targets(const int32_t* n_or_seg, int seg_form, const int2* brec, const int32_t* bucket_base,
                               uint32_t Np, float* targets, int32_t* exact, hipStream_t s);
hipError_t launch_dense_fwd(const float* X, int B, int K, const float* W, const float* b, int N, int relu, float* Y,
                            hipStream_t s);
hipError_t launch_loss_grad(const float* Y, const float* T, int B, int N, float* dY, float* sq, hipStream_t s);
hipError_t launch_dense_bwd_data(const float* dY, int B, int N, const float* W, int K, const float* A, float* dA,
                                 hipStream_t s);
hipError_t launch_dense_bwd_weight(const float* X, const float* dY, int B, int K, int N, float* dW, float* db,
                                   hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, hipStream_t s);
hipError_t launch_gather_rows(const float* src, const int32_t* idx, int B, int K, float* dst, hipStream_t s);

// ---- batches of documents: per-document paths, canvas, image offset and first cell row
struct DocDev {
    int32_t p0, p1;    // paths [p0, p1)
    int32_t w, h;      // canvas
    int64_t img_off;   // byte offset of the document's RGBA8 image
    int32_t row0;      // first cell row of the document in the batch's row numbering
    int32_t pad;
};
// wbase[p] = winding[first sorted entry of p's document] (each document scans from zero)
// The same from the bands' exclusive prefixes (bpre, total at [n_bands]), before the band sort:
// every sorted path start is a band start or the end.
hipError_t launch_doc_winding_base_bands(const int32_t* band_start, const int32_t* bpre, const int32_t* n_bands,
                                         const int32_t* sstart, const int32_t* doc_first_path, uint32_t Np,
                                         int32_t* wbase, hipStream_t s);
hipError_t launch_doc_winding_base(const int32_t* winding, const int32_t* sstart, const int32_t* doc_first_path,
                                   uint32_t Np, int32_t* wbase, hipStream_t s);

// ---- record replay: draws a given list of (yx, width, fill, frag_index) records in list order
// (overwrite or src-over) into a white w x h RGBA8 image, top row first.
hipError_t launch_draw_records(const int4* records, int32_t R, int w, int h, int compose, uint8_t* image,
                               hipStream_t s);

// ---- gradient paints (paint.hpp); stops are read as 5 floats (offset, r, g, b, a).
struct PaintArgs {
    const uint32_t* path_paint;  // per path: kSolidPaint or an index into paints
    const Paint* paints;
    const float* stops;
    const float* inv;            // 9 floats: inverse of the MVP's 2D projective part, rows
};

// ---- composite (a12).  paint (optional): per-pixel gradient colours for the paths that have one.
// cmask (optional, src-over only): the records' sample coverage (launch_cell_cover).
hipError_t launch_composite(const uint32_t* records, const int2* brec, const int32_t* bucket_base,
                            const int2* prow, const uint32_t* fill, uint32_t Np, int w, int h, int compose,
                            uint8_t* image, hipStream_t s, int64_t rec_cap, const DocDev* docs = nullptr,
                            int n_docs = 0, int total_rows = 0, const PaintArgs* paint = nullptr,
                            const unsigned long long* cmask = nullptr);

}  // namespace spr
