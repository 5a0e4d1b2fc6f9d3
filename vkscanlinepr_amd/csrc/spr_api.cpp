// spr_api.cpp -- the C ABI of include/spr.h: handle lifecycle, device buffers, and the per-frame
// orchestration of the gfx950 kernels on one HIP stream.
//
// Frame (plain two-pass form, the drop-in for drawFrame, scanline_rasterizer.cpp:282-696):
//   transform+cull -> monotonise+count -> scan(counts) -> path segments
//   -> HOST SYNC: read the intersection total N (the reference's readback, :355-356) and grow
//      buffers -> chain solve + piece merge (intersections) -> fragment keys -> tile sort +
//      merge passes -> winding scan -> fused mark/scan/emit (R stays on the device; the
//      reference's second and third host syncs are gone) -> runs by cell row -> composite.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/spr.h"
#include "kernels.hpp"
#include "learned.hpp"
#include "scene.hpp"

static_assert(sizeof(spr_paint) == sizeof(spr::Paint) && sizeof(spr_stop) == sizeof(spr::Stop),
              "spr_paint and spr_stop share the layout of the internal paint records");

struct spr_scene {
    spr::Scene s;
};

namespace {

// Message of the last failure of a call without a handle.  A fixed buffer: setting it never
// allocates, so the entry points that allocate nothing can report through it too.
struct ErrorText {
    char buf[512] = {0};
    ErrorText& operator=(const char* m) {
        std::snprintf(buf, sizeof buf, "%s", m ? m : "");
        return *this;
    }
    ErrorText& operator=(const std::string& m) { return *this = m.c_str(); }
    const char* c_str() const { return buf; }
};
thread_local ErrorText g_error;

struct Fail {
    int code;
    std::string msg;
};

#define SPR_HIP(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw Fail{SPR_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)};      \
    } while (0)

// Grow-only device allocation.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    void ensure(size_t n) {
        if (n <= cap && p) return;
        size_t want = std::max<size_t>(n, 1);
        if (cap) want = std::max(want, cap + cap / 4);
        if (p) {
            // the old buffer may still be read by queued work
            SPR_HIP(hipDeviceSynchronize());
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        void* q = nullptr;
        SPR_HIP(hipMalloc(&q, want * sizeof(T)));
        p = (T*)q;
        cap = want;
    }
    void release() {
        if (p) { (void)hipFree(p); (void)hipGetLastError(); }
        p = nullptr;
        cap = 0;
    }
};

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// The MVP rows and their inverse, copied to the device as one block.
// the MVP rows (16), the inverse of its 2D projective part (9), and two words for the learned
// frame's prediction: skip it (the last one still holds), the element count of its arena scan
constexpr size_t kMvpBytes = 27 * sizeof(float);

// Inverse of the MVP's 2D projective part (a point (x, y, 0, 1) goes to rows 0, 1 and 3:
// [[m0, m1, m3], [m4, m5, m7], [m12, m13, m15]]), by cofactors in double, rounded to float;
// zeros when it is singular.  The composite maps pixel centres back to the scene with it.
void mvp_inverse(const float m[16], float out[9]) {
    const double a = m[0], b = m[1], c = m[3], d = m[4], e = m[5], f = m[7], g = m[12], h = m[13], k = m[15];
    const double A = e * k - f * h, B = -(d * k - f * g), Cc = d * h - e * g;
    const double det = a * A + b * B + c * Cc;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
        for (int i = 0; i < 9; ++i) out[i] = 0.0f;
        return;
    }
    const double inv[9] = {A, -(b * k - c * h), b * f - c * e, B, a * k - c * g, -(a * f - c * d), Cc,
                           -(a * h - b * g), a * e - b * d};
    for (int i = 0; i < 9; ++i) out[i] = (float)(inv[i] / det);
}

// Frame summary slots (int32), gathered on the device at the end of every frame and read back
// with one copy: record totals, the error flags, the kernel routes taken, and the learned frame's
// allocation counters.
enum {
    kSumCells = 0, kSumSpans, kSumErr, kSumWaveChains, kSumHeavyPieces, kSumLongBands, kSumBigBands,
    kSumTotal, kSumOver, kSumOverflow, kSumLive, kSumArena, kSumSorted, kSumCapR, kSumWaveChainsLight, kSumCount
};
static_assert(kSumCount <= spr::kSummaryMax, "summary slots");

}  // namespace

struct spr_handle {
    uint32_t w = 0, h = 0;
    spr_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // the light chain solve and the pieces that need only it
    hipEvent_t ev_fork = nullptr, ev_light = nullptr, ev_small = nullptr, ev_join = nullptr;
    hipEvent_t ev_sfork = nullptr, ev_sjoin = nullptr;  // the short-band sort on the second stream
    hipEvent_t ev_pfork = nullptr, ev_pjoin = nullptr;  // the learned frame's prediction on the second stream
    hipEvent_t ev_pseg = nullptr;  // its per-path counts, read by the overflow statistics there
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    std::vector<hipEvent_t> ev_stage;
    bool stage_timing = false;
    int n_marks = 0;  // stage events recorded in the last frame
    std::string err;

    // inputs
    bool loaded = false;
    uint32_t P = 0, C = 0, Np = 0;
    DBuf<float2> pos;
    DBuf<uint32_t> pos_path, cpos, ctype, cpath, rule, fill;
    std::vector<uint32_t> h_cpath;
    spr::Mvp mvp{};
    DBuf<float> cweight;            // arc weights, uploaded when the scene has an arc
    bool has_arcs = false;
    const float* arcw() const { return has_arcs ? cweight.p : nullptr; }
    DBuf<uint32_t> path_paint;      // gradient paints, uploaded when a path has one
    DBuf<spr::Paint> paints;
    DBuf<float> stops;              // 5 floats per stop
    bool has_paint = false;

    // frame buffers
    DBuf<float2> tpos;
    DBuf<int32_t> visible;
    DBuf<float4> cuts;
    DBuf<uint32_t> ncuts;
    DBuf<int32_t> counts, seg;
    DBuf<int2> inter;
    DBuf<float> chain_t;
    DBuf<spr::ChainDesc> chains;
    DBuf<spr::PieceDesc> pieces;
    DBuf<int32_t> chain_order, heavy_list, lo_list, lo_cnt, wave_list;
    DBuf<uint8_t> mono, chain_cls;
    int n_cu = 256;
    DBuf<int32_t> key;
    DBuf<int8_t> wn;
    DBuf<uint32_t> epath;
    DBuf<unsigned long long> sorted, sbig, stmp;
    DBuf<int32_t> skey, sidx, winding, flags;
    DBuf<int4> records;     // full records, kept only with the debug option
    DBuf<int32_t> cell_first;            // anti-aliased frames: each record's first sorted entry
    DBuf<unsigned long long> cmask;      // anti-aliased frames: each record's sample coverage
    DBuf<float4> chord;                  // anti-aliased frames: each entry's chord
    DBuf<uint32_t> crec;    // compact records (x | width << 16) read by the composite
    DBuf<int32_t> row_lo, row_hi, bucket_base, bflag, band_start, big_list;
    DBuf<int2> band_pr;
    DBuf<int2> band_cnt;      // (cells, spans) per band, with mark + emit fused into the sort
    DBuf<int32_t> band_path;  // path of each band
    DBuf<int2> prow;
    DBuf<uint8_t> image;
    DBuf<unsigned long long> zero;  // per-frame zeroed control region
    int32_t* h_pinned = nullptr;    // 128 words of pinned host memory: [0, 16) mid-frame readbacks,
                                    // [16, 16 + kSummaryMax) the last frame's summary, [48, 73) the
                                    // MVP rows and the inverse of its 2D projective part, [73, 75)
                                    // the prediction's skip flag and arena count
    DBuf<int32_t> summary;          // the frame summary on the device, copied to h_pinned + 16
    DBuf<unsigned long long> scan_ws;  // spr_scan_i32: tile statuses and the tile counter
    int32_t sum(int k) const { return h_pinned[16 + k]; }
    float* mvp_pinned() const { return reinterpret_cast<float*>(h_pinned + 48); }  // [48, 64): the MVP rows
    float* inv_pinned() const { return reinterpret_cast<float*>(h_pinned + 64); }  // [64, 73): its inverse
    DBuf<float> mvp_dev;            // [0, 16) the MVP, [16, 25) its inverse, [25, 27) the prediction's
                                    // flag and count: copied from mvp_pinned in frames without a host
                                    // synchronisation and for paints
    int32_t* pred_pinned() const { return h_pinned + 73; }  // [73, 75): skip flag, arena scan count
    bool pred_stale = true;         // the MVP changed since the learned frame last ran its prediction
    bool last_predicted = false;    // the last learned frame ran the prediction

    // injection
    std::vector<int32_t> inj_inter;
    bool inj = false;
    DBuf<int4> inj_rec;             // injected records (replay), on the device
    int64_t inj_rec_n = -1;         // their number; -1: no records injected
    int64_t inj_rec_cells = 0;

    // learned allocation
    spr::MlpModel model;
    spr::MlpDevice mlpdev;
    DBuf<uint32_t> pcurve;                          // first curve of each path, Np+1
    DBuf<float> feat, pred, train_feat, train_tgt;
    DBuf<int32_t> arena, capn, sstart, hist_l, fill_l, bsum_l, pbound;
    DBuf<int2> brec_l;
    int64_t lb_n = 0, lb_r = 0;                     // learned frame budgets (entries, records)
    bool need_prediction = true;                    // scene or model changed: read the prediction back
    bool lb_dirty = true;                           // scene, MVP, model or margin changed: the graph is stale
    int64_t n_live = 0;                             // entries of the last learned frame
    int64_t learned_fallbacks = 0;                  // learned frames over budget, rendered with the plain method
    int64_t predictions = 0;                        // learned frames that read the predicted allocation back
    int64_t captures = 0;                           // frames captured as a graph
    bool last_learned = false;
    bool last_fallback = false;
    int2* last_brec = nullptr;                      // per-bucket record ranges of the last frame
    bool last_fused = false;                        // last frame fused mark + emit into the sort
    bool last_packed = false;                       // last frame stored buckets and sort words, not keys

    // batch of documents (spr_load_batch): per-path canvas, per-document tables
    bool batch = false;
    uint32_t n_docs = 0;
    int total_rows = 0, max_h = 0;
    size_t img_bytes = 0;
    std::vector<spr::DocDev> h_docs;
    DBuf<int2> pcanvas;
    DBuf<spr::DocDev> docs;
    DBuf<int32_t> doc_first, wbase;

    // captured frames (no-host-sync frames only), one without and one with the prediction, and
    // launch knobs picked by spr_autotune
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int64_t graph_lb_n[2] = {-1, -1}, graph_lb_r[2] = {-1, -1};
    bool capturing = false;
    bool cap_predict = false;                       // the frame being captured runs the prediction
    int chain_k = 8, merge_k = 8;                   // persistent blocks per CU
    int scan_shape = 0;                             // tile shape of the large scans (spr::kScanShapes)
    int band_window = spr::kBandWindowLog;          // log2 of the band window (9 or 10)

    // last frame
    int64_t N = 0;
    spr_stats stats{};
    bool frame_done = false;
};

namespace {

// The composite's paint arguments (null without paints); the inverse MVP is read on the device.
const spr::PaintArgs* paint_args(const spr_handle* H, spr::PaintArgs& pa) {
    if (!H->has_paint) return nullptr;
    pa = spr::PaintArgs{H->path_paint.p, H->paints.p, H->stops.p, H->mvp_dev.p + 16};
    return &pa;
}

}  // namespace

namespace {

// Per-frame zeroed control region, in 64-bit words.  Phase A is cleared at frame start (before
// the host sync), phase B once N and the bucket count NB are known.
struct ZeroLayout {
    size_t a_words = 0, total = 0;
    size_t ctr_a = 0, st_count = 0, st_pbase = 0;
    size_t ctr = 0, totals = 0, misc = 0, chainctl = 0, st_hist = 0, st_band = 0, st_wind = 0, st_mark = 0;
    size_t hist = 0, fill = 0, bsum = 0, brec = 0;
};

ZeroLayout zero_layout(uint32_t C, uint32_t Np, int64_t N, int64_t NB) {
    ZeroLayout z;
    size_t o = 0;
    z.ctr_a = o; o += 1;               // two u32 counters of the phase-A scans
    z.st_count = o; o += spr::scan_tiles(std::max<int64_t>(C, 1));
    z.st_pbase = o; o += spr::scan_tiles(std::max<uint32_t>(Np, 1));
    z.a_words = o;
    z.ctr = o; o += 4;                 // eight u32 counters of the phase-B scans
    z.totals = o; o += 2;              // int32 cells, spans, R, spare
    z.misc = o; o += 1;                // int32 error flag, big-band count
    z.chainctl = o; o += (spr::kCcWords + 1) / 2;  // chain classes, offsets, counters (int32)
    z.st_hist = o; o += spr::scan_tiles(std::max<int64_t>(NB, 1));
    z.st_band = o; o += spr::scan_tiles(std::max<int64_t>(NB, 1));
    z.st_wind = o; o += (size_t)spr::max_bands_for(Np, N) + 1;  // statuses of the band-sum scan
    z.st_mark = o; o += spr::scan_tiles(std::max<int64_t>(N, 1));
    z.hist = o; o += (size_t)(NB + 2 + 1) / 2;  // NB+1 int32 (scanned in place)
    z.fill = o; o += (size_t)(NB + 1) / 2;      // NB int32
    z.bsum = o; o += ((size_t)spr::max_bands_for(Np, N) + 2) / 2 + 1;  // max_bands + 1 int32
    z.brec = o; o += (size_t)NB;                // NB int2
    z.total = o;
    return z;
}

// The short bands are sorted on the second stream, beside the long ones, when the paths hold
// 2048 or more entries on average (then the long bands are most of the sort and the short ones
// fit beside them); with small paths the short bands are the whole sort and the fork and join
// would only add their latency.
bool serial(const spr_handle* H) { return H->stage_timing && !H->capturing; }

spr::SideStream side_for(spr_handle* H, int64_t n_entries, uint32_t Np) {
    if (serial(H) || Np == 0 || n_entries / (int64_t)Np < 2048) return {nullptr, nullptr, nullptr};
    return {H->stream2, H->ev_sfork, H->ev_sjoin};
}

// The stream of the oversized bands' sort and emit in frames with emit: the second stream, beside
// the wave sorts, in frames large enough to have such bands (paths averaging 2048+ entries, 8 M
// entries in all); otherwise the main stream, as the fork and join would cost more than they hide.
hipStream_t big_band_stream(const spr_handle* H, int64_t n_entries, uint32_t Np) {
    if (serial(H) || Np == 0 || n_entries < ((int64_t)1 << 23) || n_entries / (int64_t)Np < 2048) return H->stream;
    return H->stream2;
}

spr::ScanCtl ctl_at(spr_handle* H, size_t st, size_t ctr_word, int counter) {
    spr::ScanCtl c;
    c.status = H->zero.p + st;
    c.counter = reinterpret_cast<uint32_t*>(H->zero.p + ctr_word) + counter;
    c.epoch = 0;
    c.shape = H->scan_shape;
    return c;
}

// Gathers the frame's counters and starts their copy to the host (read once the stream is idle).
void finish_summary(spr_handle* H, const spr::SummarySrc& S, hipStream_t s) {
    SPR_HIP(spr::launch_gather_summary(S, H->summary.p, s));
    SPR_HIP(hipMemcpyAsync(H->h_pinned + 16, H->summary.p, (size_t)S.n * 4, hipMemcpyDeviceToHost, s));
}

void stage_mark(spr_handle* H, int idx) {
    if (!H->stage_timing || H->capturing) return;
    if ((int)H->ev_stage.size() <= idx) return;
    SPR_HIP(hipEventRecord(H->ev_stage[idx], H->stream));
    H->n_marks = idx + 1;
}

// Chain solve and piece kernels of one frame, on two streams: the heavy solve runs on the main
// stream while the second stream solves the light chains and finishes the pieces that depend on
// them only; the main stream then finishes the pieces that wait on heavy chains.  Stage marks 4
// and 6 are the caller's; mark 5 is set when both solves are done.
void solve_and_pieces(spr_handle* H, uint32_t C, int32_t* d_cc, const spr::FragArgs& A, int64_t n_entries,
                      int n_cu_solve, int n_cu_merge) {
    // with per-stage timing every kernel runs on the main stream, so that the events between
    // stages bracket exactly the device time of each stage's own kernels
    hipStream_t s = H->stream, s2 = serial(H) ? H->stream : H->stream2;
    const int32_t* d_coff = d_cc + spr::kCcOff;
    const int32_t* d_nlong = d_cc + spr::kCcNLong;
    const int32_t* d_lb = d_cc + spr::kCcLightBegin;
    const spr::PieceLists L{H->lo_list.p, H->lo_cnt.p, H->wave_list.p, d_cc + spr::kCcWavePieces, H->heavy_list.p,
                            d_cc + spr::kCcHeavyPieces};
    SPR_HIP(hipEventRecord(H->ev_fork, s));
    SPR_HIP(spr::launch_chain_solve(H->tpos.p, H->cpos.p, H->arcw(), H->chains.p, H->chain_order.p, d_coff, d_nlong, d_lb, 1,
                                    d_cc + spr::kCcWork, H->chain_t.p, H->mono.p, n_cu_solve, s));
    SPR_HIP(hipStreamWaitEvent(s2, H->ev_fork, 0));
    SPR_HIP(spr::launch_chain_solve(H->tpos.p, H->cpos.p, H->arcw(), H->chains.p, H->chain_order.p, d_coff,
                                    d_cc + spr::kCcNLongLight, d_lb, 0,
                                    d_cc + spr::kCcWorkLight, H->chain_t.p, H->mono.p, n_cu_solve, s2));
    if (s2 == s) stage_mark(H, 5);  // serial: both solves, then every piece kernel
    SPR_HIP(hipEventRecord(H->ev_light, s2));
    SPR_HIP(spr::launch_piece_small(H->pieces.p, C, H->mono.p, H->chain_t.p, L, A, n_cu_merge, s2));
    SPR_HIP(hipEventRecord(H->ev_small, s2));
    SPR_HIP(spr::launch_piece_lo(H->pieces.p, C, H->mono.p, H->chain_t.p, L.lo, L.lo_cnt, A, s2));
    SPR_HIP(hipEventRecord(H->ev_join, s2));
    SPR_HIP(hipStreamWaitEvent(s, H->ev_light, 0));
    if (s2 != s) stage_mark(H, 5);
    SPR_HIP(hipStreamWaitEvent(s, H->ev_small, 0));
    // the pieces of up to kWaveStage entries that wait on heavy chains: one wave each
    SPR_HIP(spr::launch_piece_waves(H->pieces.p, H->mono.p, H->chain_t.p, L.lo_h, L.lo_h_cnt, A, n_cu_merge, s));
    // pieces longer than the LDS stage merge through the sort buffer, which is free until the scatter
    SPR_HIP(spr::launch_piece_heavy(H->pieces.p, L, H->mono.p, H->chain_t.p, (uint32_t*)H->sbig.p, n_entries, A,
                                    n_cu_merge, s));
    SPR_HIP(hipStreamWaitEvent(s, H->ev_join, 0));
}

void run_frame(spr_handle* H, bool record_begin = true) {
    if (!H->loaded) throw Fail{SPR_ERR_STATE, "render before load"};
    hipStream_t s = H->stream;
    const uint32_t P = H->P, C = H->C, Np = H->Np;
    const int w = (int)H->w, h = (int)H->h;
    const bool debug = H->opts.debug != 0;
    // anti-aliased coverage reads each cell's sorted entries and their intersections
    const bool aa = H->opts.aa != 0;
    const bool keep = debug || aa;
    // 32-bit sort words unless the sorted indices are exported or read
    const bool narrow = !keep && w <= spr::kNarrowMaxW;

    if (record_begin) SPR_HIP(hipEventRecord(H->ev_begin, s));
    H->n_marks = 0;
    stage_mark(H, 0);
    // ---- phase A: transform + cull + row bounds, count pass, bucket tables
    const ZeroLayout za = zero_layout(C, Np, 0, 0);
    H->zero.ensure(za.a_words);
    H->visible.ensure(std::max<uint32_t>(Np, 1));
    H->row_lo.ensure(std::max<uint32_t>(Np, 1));
    H->row_hi.ensure(std::max<uint32_t>(Np, 1));
    H->prow.ensure(std::max<uint32_t>(Np, 1));
    H->bucket_base.ensure((size_t)Np + 1);
    H->tpos.ensure(std::max<uint32_t>(P, 1));
    H->cuts.ensure(std::max<uint32_t>(C, 1));
    H->ncuts.ensure(std::max<uint32_t>(C, 1));
    H->counts.ensure((size_t)C + 1);
    H->seg.ensure((size_t)Np + 1);
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->zero.p, 2 * (int64_t)za.a_words, 0u, s));
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->visible.p, Np, 0u, s));
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->row_lo.p, Np, 0x7FFFFFFFu, s));
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->row_hi.p, Np, 0x80000000u, s));
    if (H->has_paint) SPR_HIP(hipMemcpyAsync(H->mvp_dev.p, H->mvp_pinned(), kMvpBytes, hipMemcpyHostToDevice, s));
    SPR_HIP(spr::launch_transform(H->pos.p, H->pos_path.p, P, H->mvp, w, h, H->tpos.p, H->visible.p,
                                  H->row_lo.p, H->row_hi.p, s));
    stage_mark(H, 1);
    SPR_HIP(spr::launch_count(H->tpos.p, H->cpos.p, H->ctype.p, H->arcw(), H->cpath.p, H->visible.p, C, w, h,
                              H->cuts.p, H->ncuts.p, H->counts.p, s));
    stage_mark(H, 2);
    SPR_HIP(spr::launch_scan_i32(H->counts.p, H->counts.p, C, ctl_at(H, za.st_count, za.ctr_a, 0), s));
    SPR_HIP(spr::launch_bucket_counts(H->row_lo.p, H->row_hi.p, Np, (h - 1) / 2, H->bucket_base.p, H->prow.p, s));
    SPR_HIP(spr::launch_scan_i32(H->bucket_base.p, H->bucket_base.p, Np, ctl_at(H, za.st_pbase, za.ctr_a, 1), s));
    int64_t N = 0;
    if (!H->inj) {
        SPR_HIP(spr::launch_path_segments(H->cpath.p, C, H->counts.p, Np, H->seg.p, s));
        SPR_HIP(hipMemcpyAsync(H->h_pinned, H->counts.p + C, 4, hipMemcpyDeviceToHost, s));
    } else {
        N = (int64_t)H->inj_inter.size() / 2;
        std::vector<int32_t> hseg((size_t)Np + 1, 0);
        for (int64_t i = 0; i < N; ++i) {
            int32_t c0 = H->inj_inter[2 * i];
            int32_t c1 = i + 1 != N ? H->inj_inter[2 * i + 2] : -1;
            uint32_t p0 = H->h_cpath[c0];
            if (c0 != c1) {
                uint32_t p1 = c1 >= 0 ? H->h_cpath[c1] : Np;
                for (uint32_t j = p0 + 1; j <= p1; ++j) hseg[j] = (int32_t)(i + 1);
            }
            if (i == 0) for (uint32_t j = 0; j <= p0; ++j) hseg[j] = 0;
        }
        SPR_HIP(hipMemcpyAsync(H->seg.p, hseg.data(), hseg.size() * 4, hipMemcpyHostToDevice, s));
    }
    SPR_HIP(hipMemcpyAsync(H->h_pinned + 1, H->bucket_base.p + Np, 4, hipMemcpyDeviceToHost, s));
    // ---- HOST SYNC: the count-then-write readback (scanline_rasterizer.cpp:355-356)
    SPR_HIP(hipStreamSynchronize(s));
    if (!H->inj) N = H->h_pinned[0];
    const int64_t NB = H->h_pinned[1];
    if (N < 0 || N > (int64_t)(1 << 30)) throw Fail{SPR_ERR_CAPACITY, "intersection count exceeds 2^30"};
    if (NB < 0 || NB > (int64_t)(1 << 30)) throw Fail{SPR_ERR_CAPACITY, "bucket count exceeds 2^30"};
    H->N = N;
    stage_mark(H, 3);

    // ---- phase B: size buffers for N and NB
    const ZeroLayout z = zero_layout(C, Np, N, NB);
    H->zero.ensure(z.total);
    const size_t n1 = (size_t)std::max<int64_t>(N, 1), nb1 = (size_t)std::max<int64_t>(NB, 1);
    const int64_t max_bands = spr::max_bands_for(Np, N);
    H->inter.ensure(n1);
    H->chain_t.ensure(n1);
    H->key.ensure(n1);
    H->wn.ensure(n1);
    H->epath.ensure(n1);
    H->sbig.ensure(n1);
    H->stmp.ensure(n1);
    H->skey.ensure(n1);
    H->winding.ensure(n1 + 1);
    H->bflag.ensure(nb1 + 1);
    H->band_start.ensure((size_t)max_bands + 1);
    H->band_pr.ensure((size_t)max_bands + 1);
    H->band_cnt.ensure((size_t)max_bands + 1);
    H->band_path.ensure((size_t)max_bands + 1);
    H->big_list.ensure(4 * (size_t)max_bands);  // oversized band ids, their winding prefixes, two lists of long bands
    if (debug) { H->flags.ensure(2 * n1); H->records.ensure(2 * n1); }
    if (keep) H->sidx.ensure(n1);
    if (aa) { H->cell_first.ensure(2 * n1); H->cmask.ensure(2 * n1); H->chord.ensure(n1); }
    H->crec.ensure(2 * n1);
    H->image.ensure((size_t)4 * H->w * H->h);
    SPR_HIP(spr::launch_fill_u32((uint32_t*)(H->zero.p + z.a_words), 2 * (int64_t)(z.total - z.a_words), 0u, s));
    int32_t* d_totals = reinterpret_cast<int32_t*>(H->zero.p + z.totals);
    int32_t* d_err = reinterpret_cast<int32_t*>(H->zero.p + z.misc);
    int32_t* d_bigcount = d_err + 1;
    int32_t* d_hist = reinterpret_cast<int32_t*>(H->zero.p + z.hist);
    int32_t* d_fill = reinterpret_cast<int32_t*>(H->zero.p + z.fill);
    int32_t* d_bsum = reinterpret_cast<int32_t*>(H->zero.p + z.bsum);
    int2* d_brec = reinterpret_cast<int2*>(H->zero.p + z.brec);

    const bool gen = !H->inj && N > 0;
    int32_t* d_cc = reinterpret_cast<int32_t*>(H->zero.p + z.chainctl);  // class counts
    int32_t* d_cfill = d_cc + spr::kCcFill;
    int32_t* d_coff = d_cc + spr::kCcOff;
    int32_t* d_nlong = d_cc + spr::kCcNLong;
    int32_t* d_work = d_cc + spr::kCcWork;
    if (H->inj) {
        SPR_HIP(hipMemcpyAsync(H->inter.p, H->inj_inter.data(), (size_t)N * 8, hipMemcpyHostToDevice, s));
    } else if (gen) {
        H->chains.ensure((size_t)C * 10);
        H->pieces.ensure((size_t)C * 5);
        H->chain_order.ensure((size_t)C * 10);
        H->heavy_list.ensure((size_t)C * 5);
        H->lo_list.ensure((size_t)C * 5 + 64);
        H->lo_cnt.ensure(((size_t)C * 5 + 63) / 64);
        H->wave_list.ensure((size_t)C * 5);
        H->chain_cls.ensure((size_t)C * 10);
        H->mono.ensure((size_t)C * 10);
        SPR_HIP(spr::launch_chain_prep(H->tpos.p, H->cpos.p, H->ctype.p, H->arcw(), H->cpath.p, H->visible.p, H->cuts.p,
                                       H->ncuts.p, H->counts.p, C, w, h, H->chains.p, H->pieces.p, H->chain_cls.p,
                                       d_cc, H->n_cu, s));
        SPR_HIP(spr::launch_chain_order(H->chain_cls.p, C, d_cc, d_coff, d_cfill, d_nlong, d_cc + spr::kCcLightBegin,
                                        d_cc + spr::kCcNLongLight,
                                        H->chain_order.p, H->n_cu, s));
    }
    stage_mark(H, 4);
    if (gen) {
        // chain solve, then piece merge and fragment keys in one pass
        // narrow frames: the piece kernels store each entry's bucket (key) and sort word (epath)
        const spr::FragArgs A{H->cpos.p, H->ctype.p, H->arcw(), H->cpath.p, H->tpos.p, w, h, nullptr,
                              H->key.p, H->wn.p, H->epath.p, H->bucket_base.p, H->prow.p, d_hist, d_err,
                              keep ? H->inter.p : nullptr, aa ? H->chord.p : nullptr,
                              narrow ? H->epath.p : nullptr};
        solve_and_pieces(H, C, d_cc, A, (int64_t)n1, H->n_cu * H->chain_k / 8, H->n_cu * H->merge_k / 8);
    } else {
        stage_mark(H, 5);
    }
    stage_mark(H, 6);
    if (!gen)  // injected entries (or none)
        SPR_HIP(spr::launch_gen_fragment(H->inter.p, N, H->cpos.p, H->ctype.p, H->arcw(), H->cpath.p, H->tpos.p, w, h,
                                         H->key.p, H->wn.p, H->epath.p, H->bucket_base.p, H->prow.p, d_hist, d_err, s,
                                         nullptr, nullptr, aa ? H->chord.p : nullptr));
    stage_mark(H, 7);
    // ---- sort: bucket offsets, bands, scatter into band order, band sort in LDS
    if (N > 0) {
        SPR_HIP(spr::launch_scan_i32(d_hist, d_hist, NB, ctl_at(H, z.st_hist, z.ctr, 0), s));
        SPR_HIP(spr::launch_band_flags(H->bucket_base.p, Np, d_hist, H->bflag.p, s, H->band_window));
        SPR_HIP(spr::launch_scan_i32(H->bflag.p, H->bflag.p, NB, ctl_at(H, z.st_band, z.ctr, 1), s));
        SPR_HIP(spr::launch_band_emit(H->bflag.p, d_hist, (int32_t)NB, N, H->band_start.p, s, nullptr, nullptr,
                                      H->bucket_base.p, Np, H->prow.p, narrow ? H->band_pr.p : nullptr,
                                      narrow ? H->band_path.p : nullptr));
        SPR_HIP(spr::launch_bucket_scatter(H->key.p, H->wn.p, H->epath.p, N, H->bucket_base.p, H->prow.p,
                                           H->bflag.p, H->band_start.p, d_fill, d_bsum, H->sbig.p, d_err, s, nullptr,
                                           narrow, narrow && gen));
        // each band's exclusive prefix of the winding changes (total at [n_bands])
        SPR_HIP(spr::launch_scan_i32(d_bsum, d_bsum, max_bands, ctl_at(H, z.st_wind, z.ctr, 2), s, H->bflag.p + NB));
    }
    stage_mark(H, 8);
    // narrow frames fuse mark + emit into the band sort: records of band t at [2 * start, ...)
    const spr::EmitArgs E{H->crec.p, d_brec, H->band_cnt.p, H->band_path.p, H->bucket_base.p, H->rule.p,
                          nullptr, nullptr, w, h};
    if (N > 0) {
        // band sort fused with the winding numbers (global exclusive scan of the sorted winding
        // changes, scanline_rasterizer.cpp:478-506)
        uint32_t* ticket = ctl_at(H, z.st_wind, z.ctr, 4).counter;
        if (narrow) {
            // the oversized bands (queued by the classify pass) are sorted and emitted on the second
            // stream, followed there by the bands of up to 512 words, beside the longer bands
            hipStream_t s2 = big_band_stream(H, N, Np);
            SPR_HIP(spr::launch_band_classify(H->band_start.p, H->bflag.p + NB, max_bands, d_bsum, H->big_list.p,
                                              H->big_list.p + max_bands, d_bigcount, ticket, s));
            if (s2 != s) {
                SPR_HIP(hipEventRecord(H->ev_sfork, s));
                SPR_HIP(hipStreamWaitEvent(s2, H->ev_sfork, 0));
            }
            SPR_HIP(spr::launch_sort_big_bands(H->band_start.p, H->big_list.p, H->big_list.p + max_bands, d_bigcount,
                                               H->sbig.p, H->stmp.p, H->skey.p, nullptr, H->winding.p, s2,
                                               H->band_pr.p));
            SPR_HIP(spr::launch_emit_big_bands(H->band_start.p, H->big_list.p, d_bigcount, H->skey.p, H->winding.p,
                                               H->band_pr.p, E, s2));
            SPR_HIP(spr::launch_sort_bands(H->band_start.p, H->bflag.p + NB, max_bands, H->sbig.p, H->skey.p,
                                           nullptr, H->winding.p, H->big_list.p, H->big_list.p + max_bands,
                                           d_bigcount, d_bsum, ticket, H->n_cu, s, H->band_pr.p, &E,
                                           side_for(H, N, Np), N, s2));
            if (s2 != s) {
                SPR_HIP(hipEventRecord(H->ev_sjoin, s2));  // after the short bands' sort on s2
                SPR_HIP(hipStreamWaitEvent(s, H->ev_sjoin, 0));
            }
            SPR_HIP(spr::launch_band_totals(H->band_cnt.p, H->bflag.p + NB, d_totals, s));
        } else {
            SPR_HIP(spr::launch_sort_bands(H->band_start.p, H->bflag.p + NB, max_bands, H->sbig.p, H->skey.p,
                                           keep ? H->sidx.p : nullptr, H->winding.p, H->big_list.p,
                                           H->big_list.p + max_bands, d_bigcount, d_bsum, ticket, H->n_cu, s,
                                           nullptr, nullptr, side_for(H, N, Np), N));
            SPR_HIP(spr::launch_sort_big_bands(H->band_start.p, H->big_list.p, H->big_list.p + max_bands, d_bigcount,
                                               H->sbig.p, H->stmp.p, H->skey.p, keep ? H->sidx.p : nullptr,
                                               H->winding.p, s, nullptr));
        }
    }
    stage_mark(H, 9);
    stage_mark(H, 10);
    if (!narrow)
        SPR_HIP(spr::launch_mark_emit(H->skey.p, H->winding.p, H->epath.p, H->rule.p, H->fill.p, N, w, h,
                                      ctl_at(H, z.st_mark, z.ctr, 3), debug ? H->records.p : nullptr, H->crec.p,
                                      debug ? H->flags.p : nullptr, d_totals, H->bucket_base.p, H->prow.p, d_brec,
                                      s, nullptr, nullptr, Np, (int64_t)(2 * n1), d_err, nullptr, nullptr,
                                      aa ? H->cell_first.p : nullptr));
    if (aa)
        SPR_HIP(spr::launch_cell_cover(H->cell_first.p, d_totals, (int64_t)(2 * n1), H->skey.p, H->sidx.p,
                                       H->winding.p, H->epath.p, H->chord.p, N, H->rule.p, H->cmask.p, s));
    H->last_brec = d_brec;
    H->last_learned = false;
    H->last_fused = narrow;
    H->last_packed = narrow && gen;
    stage_mark(H, 11);
    // ---- composite
    spr::PaintArgs pa;
    SPR_HIP(spr::launch_composite(H->crec.p, d_brec, H->bucket_base.p, H->prow.p, H->fill.p, Np, w, h,
                                  (int)H->opts.compose, H->image.p, s, (int64_t)(2 * n1), nullptr, 0, 0,
                                  paint_args(H, pa), aa ? H->cmask.p : nullptr));
    stage_mark(H, 12);
    spr::SummarySrc S{};
    S.n = kSumCount;
    S.p[kSumCells] = d_totals;
    S.p[kSumSpans] = d_totals + 1;
    S.p[kSumErr] = d_err;
    S.p[kSumWaveChains] = gen ? d_nlong : nullptr;
    S.p[kSumWaveChainsLight] = gen ? d_cc + spr::kCcNLongLight : nullptr;
    S.p[kSumHeavyPieces] = gen ? d_cc + spr::kCcHeavyPieces : nullptr;
    S.p[kSumLongBands] = reinterpret_cast<const int32_t*>(ctl_at(H, z.st_wind, z.ctr, 4).counter + 1);
    S.p[kSumBigBands] = d_bigcount;
    S.p[kSumTotal] = H->inj ? nullptr : H->counts.p + C;
    finish_summary(H, S, s);
    SPR_HIP(hipEventRecord(H->ev_end, s));
    H->frame_done = true;
}

// ---- learned single-pass frame
//
// The stages of run_frame with the same entry layout (exact per-curve counts, scanned on the
// device), but no host synchronisation: every buffer is sized by a budget the host sets before
// the frame, and every kernel after the count reads its element counts from device memory.
//  - The budget comes from the predictor.  The first frame after a load or a model change runs
//    the 16-64-64-2 MLP on its paths and reads the sum of the predicted per-path capacities back
//    (for a batch of documents: the proven upper bound, as a batch has no plain frame to fall back
//    on); every frame then grows the budget to cover its own exact need and its prediction, with
//    headroom, so that the next frames (after a new MVP too) need no readback.
//  - The features, the MLP and the capacities run on the second stream, beside the count.
//  - The chain prep checks the exact total against the budget on the device.  A frame over its
//    budget is culled there (no chain, no piece: nothing is written past any buffer), and the
//    host renders that frame with the plain method: the output is always the plain method's.
struct LearnedLayout {
    size_t total = 0, ctl = 0, misc = 0, capr = 0, counters = 0, totals = 0, slots = 0, bound = 0;
    size_t st_count = 0, st_arena = 0, st_pb = 0, st_hist = 0, st_band = 0, st_wind = 0, st_mark = 0, chainctl = 0;
};

// Scan tile counters of the learned frame (u32 slots of the counters block).
enum { kLcArena = 0, kLcBucketBase, kLcHist, kLcBandFlags, kLcBandSums, kLcMark, kLcSortTicket, kLcSortTicket2,
       kLcSortTicket3, kLcCount, kLcWords };

LearnedLayout learned_layout(uint32_t C, uint32_t Np, int64_t nb_budget, int64_t n_budget) {
    LearnedLayout z;
    size_t o = 0;
    z.counters = o; o += (kLcWords + 1) / 2;  // u32 scan tile counters
    z.ctl = o; o += 2;          // int32 live entries, over-budget flag, overflowing paths, spare
    z.misc = o; o += 1;         // int32 error flags, big-band count
    z.capr = o; o += 1;         // int32 predicted record total, spare
    z.totals = o; o += 2;       // int32 cells, spans, R, spare
    z.slots = o; o += 2;        // int32 sorted entries, bands, spare x2
    z.bound = o; o += 1;        // u64 upper bound of the intersections
    z.chainctl = o; o += (spr::kCcWords + 1) / 2;
    z.st_count = o; o += spr::scan_tiles(std::max<uint32_t>(C, 1));
    z.st_arena = o; o += spr::scan_tiles(std::max<uint32_t>(Np, 1));
    z.st_pb = o; o += spr::scan_tiles(std::max<uint32_t>(Np, 1));
    z.st_hist = o; o += spr::scan_tiles(std::max<int64_t>(nb_budget, 1));
    z.st_band = o; o += spr::scan_tiles(std::max<int64_t>(nb_budget, 1));
    z.st_wind = o; o += (size_t)spr::max_bands_for(Np, std::max<int64_t>(n_budget, 1)) + 1;  // band-sum scan
    z.st_mark = o; o += spr::scan_tiles(std::max<int64_t>(n_budget, 1));
    z.total = o;
    return z;
}

// The budget only grows, and only when a frame's exact need comes within 1/16 of it or the
// predicted total exceeds it: then to the need plus an eighth (and the prediction plus a
// sixteenth), so that frames whose counts drift a little (a new MVP) keep the captured frame.
// Records take at most two per entry.
void grow_budget(spr_handle* H, int64_t need, int64_t predicted) {
    if (need > H->lb_n - H->lb_n / 16 || predicted > H->lb_n)
        H->lb_n = std::max<int64_t>(std::max<int64_t>(need + need / 8, predicted + predicted / 16) + 1024, H->lb_n);
    H->lb_r = 2 * H->lb_n;
}

void learned_finish(spr_handle* H);

// The prediction (features, MLP, capacities) is due after a load, a model or margin change or a
// new MVP.  A batch without a model runs only the features, for the proven bound that sizes its
// first budget: later frames of the same batch and view need none of it.
bool prediction_due(const spr_handle* H) {
    return (!H->model.ready && !H->batch) || H->need_prediction || H->pred_stale;
}

// Whether the next learned frame runs the prediction, published in the pinned words that the
// frame copies to the device with the MVP.  Every earlier frame has ended (frames end with a stream
// synchronisation), so no copy is reading them.
bool publish_prediction(spr_handle* H) {
    const bool predict = H->capturing ? H->cap_predict : prediction_due(H);
    H->pred_pinned()[0] = predict ? 0 : 1;
    H->pred_pinned()[1] = predict ? (int32_t)H->Np : 0;
    H->last_predicted = predict;
    if (predict && !H->capturing) H->pred_stale = false;
    return predict;
}

// Enqueues one frame without any host synchronisation unless the predicted budget must first be
// read back (H->need_prediction).  With H->capturing the work is being recorded into a graph: no
// events, no waits, and every buffer already has its size.
void run_learned_frame(spr_handle* H) {
    if (!H->loaded) throw Fail{SPR_ERR_STATE, "render before load"};
    const bool use_model = H->model.ready;
    if (!use_model && !H->batch)
        throw Fail{SPR_ERR_STATE, "learned mode needs a trained or loaded model (spr_learned_train or spr_learned_load)"};
    if (H->capturing && H->need_prediction) throw Fail{SPR_ERR_STATE, "internal: capture of a frame without budgets"};
    hipStream_t s = H->stream;
    hipStream_t s2 = serial(H) ? s : H->stream2;
    if (use_model && !H->mlpdev.valid()) H->mlpdev.upload(H->model, s);
    const uint32_t P = H->P, C = H->C, Np = H->Np;
    const int w = (int)H->w, h = (int)H->h;
    const bool debug = H->opts.debug != 0;
    const int rmax = ((H->batch ? H->max_h : h) - 1) / 2;
    const int64_t nb_budget = (int64_t)Np * (rmax + 2);
    if (nb_budget > (int64_t)1 << 30) throw Fail{SPR_ERR_CAPACITY, "bucket bound exceeds 2^30"};
    const int2* pcv = H->batch ? H->pcanvas.p : nullptr;
    // 32-bit sort words unless the sorted indices are exported (canvases are at most 8192 wide)
    const bool narrow = !debug && w <= spr::kNarrowMaxW;

    if (!H->capturing) SPR_HIP(hipEventRecord(H->ev_begin, s));
    H->n_marks = 0;
    stage_mark(H, 0);
    H->visible.ensure(std::max<uint32_t>(Np, 1));
    H->row_lo.ensure(std::max<uint32_t>(Np, 1));
    H->row_hi.ensure(std::max<uint32_t>(Np, 1));
    H->prow.ensure(std::max<uint32_t>(Np, 1));
    H->bucket_base.ensure((size_t)Np + 1);
    H->tpos.ensure(std::max<uint32_t>(P, 1));
    H->cuts.ensure(std::max<uint32_t>(C, 1));
    H->ncuts.ensure(std::max<uint32_t>(C, 1));
    H->counts.ensure((size_t)C + 1);
    H->seg.ensure((size_t)Np + 1);
    H->feat.ensure((size_t)std::max<uint32_t>(Np, 1) * spr::kMlpIn);
    H->pred.ensure((size_t)std::max<uint32_t>(Np, 1) * spr::kMlpOut);
    H->arena.ensure((size_t)Np + 1);
    H->capn.ensure((size_t)Np + 1);
    H->sstart.ensure((size_t)Np + 1);
    H->pbound.ensure((size_t)Np + 1);

    // phase 1: transform; the prediction on the second stream, the exact count on the first
    LearnedLayout z = learned_layout(C, Np, nb_budget, std::max<int64_t>(H->lb_n, 1));
    H->zero.ensure(z.total);
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->zero.p, 2 * (int64_t)z.total, 0u, s));
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->visible.p, Np, 0u, s));
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->row_lo.p, Np, 0x7FFFFFFFu, s));
    SPR_HIP(spr::launch_fill_u32((uint32_t*)H->row_hi.p, Np, 0x80000000u, s));
    // the prediction is a function of the scene, the MVP and the model: it runs after a change of
    // one of them and is otherwise left out (its outputs still hold); a captured frame with it
    // reads the same decision from the pinned words at every replay
    const bool predict = publish_prediction(H);
    // the MVP goes through a copy from pinned host memory, which a captured frame repeats at every
    // replay: spr_set_mvp only rewrites that memory, and the graph follows it
    SPR_HIP(hipMemcpyAsync(H->mvp_dev.p, H->mvp_pinned(), kMvpBytes, hipMemcpyHostToDevice, s));
    const int32_t* d_skip = reinterpret_cast<const int32_t*>(H->mvp_dev.p + 25);
    const int32_t* d_arena_n = reinterpret_cast<const int32_t*>(H->mvp_dev.p + 26);
    SPR_HIP(spr::launch_transform(H->pos.p, H->pos_path.p, P, H->mvp, w, h, H->tpos.p, H->visible.p,
                                  H->row_lo.p, H->row_hi.p, s, pcv, H->mvp_dev.p));
    stage_mark(H, 1);
    int32_t* d_capr = reinterpret_cast<int32_t*>(H->zero.p + z.capr);
    if (!predict) s2 = s;  // nothing to run beside the count
    if (s2 != s) {
        SPR_HIP(hipEventRecord(H->ev_pfork, s));
        SPR_HIP(hipStreamWaitEvent(s2, H->ev_pfork, 0));
    }
    if (predict) {
        SPR_HIP(spr::launch_path_features(H->tpos.p, H->cpos.p, H->ctype.p, H->pcurve.p, H->visible.p, H->rule.p,
                                          Np, w, h, H->feat.p, s2, H->zero.p + z.bound, pcv, H->pbound.p, d_skip));
        if (use_model) {
            SPR_HIP(spr::launch_mlp_infer(H->feat.p, Np, H->mlpdev.dev(), H->pred.p, s2, d_skip));
            SPR_HIP(spr::launch_caps(H->pred.p, H->visible.p, Np, H->model.log_margin, H->model.slack, H->arena.p,
                                     H->capn.p, d_capr, s2, H->pbound.p, d_skip));
        } else {
            SPR_HIP(spr::launch_fill_u32((uint32_t*)H->arena.p, (int64_t)Np + 1, 0u, s2));
            SPR_HIP(spr::launch_fill_u32((uint32_t*)H->capn.p, (int64_t)Np + 1, 0u, s2));
        }
        SPR_HIP(spr::launch_scan_i32(H->arena.p, H->arena.p, Np, ctl_at(H, z.st_arena, z.counters, kLcArena), s2,
                                     d_arena_n));
    }
    SPR_HIP(spr::launch_count(H->tpos.p, H->cpos.p, H->ctype.p, H->arcw(), H->cpath.p, H->visible.p, C, w, h, H->cuts.p,
                              H->ncuts.p, H->counts.p, s, pcv));
    stage_mark(H, 2);
    SPR_HIP(spr::launch_scan_i32(H->counts.p, H->counts.p, C, ctl_at(H, z.st_count, z.counters, kLcCount), s));
    SPR_HIP(spr::launch_path_segments(H->cpath.p, C, H->counts.p, Np, H->seg.p, s));
    // the prediction's paths that overflowed their capacity (statistics only): on the second stream
    // once the per-path counts exist, unless this frame reads the prediction back (and re-lays the
    // control region), in which case they are counted at its end
    const bool ovf_side = use_model && predict && s2 != s && !H->need_prediction;
    if (ovf_side) {
        SPR_HIP(hipEventRecord(H->ev_pseg, s));
        SPR_HIP(hipStreamWaitEvent(s2, H->ev_pseg, 0));
        SPR_HIP(spr::launch_path_overflow(H->seg.p, H->capn.p, H->visible.p, Np,
                                          reinterpret_cast<int32_t*>(H->zero.p + z.ctl) + 2, s2, d_skip));
    }
    if (s2 != s) SPR_HIP(hipEventRecord(H->ev_pjoin, s2));
    SPR_HIP(spr::launch_bucket_counts(H->row_lo.p, H->row_hi.p, Np, rmax, H->bucket_base.p, H->prow.p, s, pcv));
    SPR_HIP(spr::launch_scan_i32(H->bucket_base.p, H->bucket_base.p, Np, ctl_at(H, z.st_pb, z.counters, kLcBucketBase),
                                 s));
    bool joined = s2 == s;
    if (H->need_prediction && !use_model) H->need_prediction = false;  // a batch without a model: no prediction
    if (H->need_prediction) {
        // the first frame after a load or a model change: read the predicted allocation back once
        if (!joined) { SPR_HIP(hipStreamWaitEvent(s, H->ev_pjoin, 0)); joined = true; }
        SPR_HIP(hipMemcpyAsync(H->h_pinned + 12, H->arena.p + Np, 4, hipMemcpyDeviceToHost, s));
        SPR_HIP(hipMemcpyAsync(H->h_pinned + 14, H->zero.p + z.bound, 8, hipMemcpyDeviceToHost, s));
        SPR_HIP(hipStreamSynchronize(s));
        const int64_t caps = std::max<int32_t>(H->h_pinned[12], 0);
        grow_budget(H, 0, caps);
        H->need_prediction = false;
        ++H->predictions;
        // the status arrays of phase 2 are sized by the budget: re-lay the control region (phase 1
        // is complete; of its results only the counts, segments and bucket bases, outside the
        // region, are needed later)
        z = learned_layout(C, Np, nb_budget, H->lb_n);
        H->zero.ensure(z.total);
        SPR_HIP(spr::launch_fill_u32((uint32_t*)H->zero.p, 2 * (int64_t)z.total, 0u, s));
        d_capr = reinterpret_cast<int32_t*>(H->zero.p + z.capr);
    }
    stage_mark(H, 3);
    int32_t* d_ctl = reinterpret_cast<int32_t*>(H->zero.p + z.ctl);
    int32_t* d_live = d_ctl, *d_over = d_ctl + 1, *d_novf = d_ctl + 2;
    int32_t* d_err = reinterpret_cast<int32_t*>(H->zero.p + z.misc);
    int32_t* d_bigcount = d_err + 1;
    int32_t* d_totals = reinterpret_cast<int32_t*>(H->zero.p + z.totals);
    int32_t* d_nsorted = reinterpret_cast<int32_t*>(H->zero.p + z.slots);
    int32_t* d_nbands = d_nsorted + 1;
    // entry offsets are int32 and so are record slots, two per entry
    if (H->lb_n > (int64_t)1 << 30 || H->lb_r > ((int64_t)1 << 31) - 2)
        throw Fail{SPR_ERR_CAPACITY, "intersection budget exceeds 2^30"};
    const int64_t nl = std::max<int64_t>(H->lb_n, 1), nr = std::max<int64_t>(H->lb_r, 1);
    const int64_t max_bands = spr::max_bands_for(Np, nl);
    H->inter.ensure(nl); H->chain_t.ensure(nl); H->key.ensure(nl); H->wn.ensure(nl); H->epath.ensure(nl);
    H->sbig.ensure(nl); H->stmp.ensure(nl); H->skey.ensure(nl);
    H->winding.ensure(nl + 1);
    H->bflag.ensure((size_t)nb_budget + 1);
    H->band_start.ensure((size_t)max_bands + 1);
    H->band_pr.ensure((size_t)max_bands + 1);
    H->band_cnt.ensure((size_t)max_bands + 1);
    H->band_path.ensure((size_t)max_bands + 1);
    H->big_list.ensure(4 * (size_t)max_bands);  // oversized band ids, their winding prefixes, two lists of long bands
    if (debug) { H->sidx.ensure(nl); H->flags.ensure(2 * nl); H->records.ensure(nr); }
    // narrow frames fuse mark + emit into the band sort: the records of band t sit at
    // [2 * start, ...), so the record buffer covers two per sorted entry
    const int64_t rec_cap = narrow ? 2 * nl : nr;
    H->crec.ensure(rec_cap);
    H->image.ensure(H->batch ? H->img_bytes : (size_t)4 * H->w * H->h);
    H->hist_l.ensure((size_t)nb_budget + 1);
    H->fill_l.ensure((size_t)nb_budget + 1);
    H->bsum_l.ensure((size_t)max_bands + 1);
    H->brec_l.ensure((size_t)nb_budget + 1);
    // per-bucket tables: cleared up to the live bucket count (bucket_base[Np])
    const int32_t* d_nb = H->bucket_base.p + Np;
    SPR_HIP(spr::launch_fill_u32_dev((uint32_t*)H->hist_l.p, nb_budget + 1, d_nb, 1, 1, 0u, s));
    SPR_HIP(spr::launch_fill_u32_dev((uint32_t*)H->fill_l.p, nb_budget, d_nb, 1, 0, 0u, s));
    SPR_HIP(spr::launch_fill_u32_dev((uint32_t*)H->bsum_l.p, max_bands, d_nb, 1, 0, 0u, s));  // bands <= buckets
    SPR_HIP(spr::launch_fill_u32_dev((uint32_t*)H->brec_l.p, 2 * nb_budget, d_nb, 2, 0, 0u, s));

    // phase 2: the plain stages on device counts, within the budget
    int32_t* d_cc = reinterpret_cast<int32_t*>(H->zero.p + z.chainctl);
    int32_t* d_cfill = d_cc + spr::kCcFill;
    int32_t* d_coff = d_cc + spr::kCcOff;
    int32_t* d_nlong = d_cc + spr::kCcNLong;
    if (C) {
        H->chains.ensure((size_t)C * 10);
        H->pieces.ensure((size_t)C * 5);
        H->chain_order.ensure((size_t)C * 10);
        H->heavy_list.ensure((size_t)C * 5);
        H->lo_list.ensure((size_t)C * 5 + 64);
        H->lo_cnt.ensure(((size_t)C * 5 + 63) / 64);
        H->wave_list.ensure((size_t)C * 5);
        H->chain_cls.ensure((size_t)C * 10);
        H->mono.ensure((size_t)C * 10);
        const spr::Budget B{nl, d_live, d_over};
        SPR_HIP(spr::launch_chain_prep(H->tpos.p, H->cpos.p, H->ctype.p, H->arcw(), H->cpath.p, H->visible.p, H->cuts.p,
                                       H->ncuts.p, H->counts.p, C, w, h, H->chains.p, H->pieces.p, H->chain_cls.p,
                                       d_cc, H->n_cu, s, &B, pcv));
        SPR_HIP(spr::launch_chain_order(H->chain_cls.p, C, d_cc, d_coff, d_cfill, d_nlong, d_cc + spr::kCcLightBegin,
                                        d_cc + spr::kCcNLongLight,
                                        H->chain_order.p, H->n_cu, s));
    }
    stage_mark(H, 4);
    if (C) {
        const spr::FragArgs A{H->cpos.p, H->ctype.p, H->arcw(), H->cpath.p, H->tpos.p, w, h, pcv,
                              H->key.p, H->wn.p, H->epath.p, H->bucket_base.p, H->prow.p, H->hist_l.p, d_err,
                              debug ? H->inter.p : nullptr, nullptr, narrow ? H->epath.p : nullptr};
        solve_and_pieces(H, C, d_cc, A, nl, H->n_cu * H->chain_k / 8, H->n_cu * H->merge_k / 8);
    } else {
        stage_mark(H, 5);
    }
    stage_mark(H, 6);
    stage_mark(H, 7);
    SPR_HIP(spr::launch_scan_i32(H->hist_l.p, H->hist_l.p, nb_budget, ctl_at(H, z.st_hist, z.counters, kLcHist), s,
                                 d_nb, d_nsorted));
    SPR_HIP(spr::launch_band_flags(H->bucket_base.p, Np, H->hist_l.p, H->bflag.p, s, H->band_window));
    SPR_HIP(spr::launch_scan_i32(H->bflag.p, H->bflag.p, nb_budget, ctl_at(H, z.st_band, z.counters, kLcBandFlags),
                                 s, d_nb, d_nbands));
    SPR_HIP(spr::launch_band_emit(H->bflag.p, H->hist_l.p, (int32_t)nb_budget, nl, H->band_start.p, s, d_nb,
                                  d_nsorted, H->bucket_base.p, Np, H->prow.p, narrow ? H->band_pr.p : nullptr,
                                  narrow ? H->band_path.p : nullptr));
    SPR_HIP(spr::launch_bucket_scatter(H->key.p, H->wn.p, H->epath.p, nl, H->bucket_base.p, H->prow.p,
                                       H->bflag.p, H->band_start.p, H->fill_l.p, H->bsum_l.p, H->sbig.p, d_err, s,
                                       d_live, narrow, narrow));
    SPR_HIP(spr::launch_scan_i32(H->bsum_l.p, H->bsum_l.p, max_bands, ctl_at(H, z.st_wind, z.counters, kLcBandSums),
                                 s, d_nbands));
    // the sorted start of each path: read by the documents' winding bases and the separate mark + emit
    if (H->batch || !narrow) SPR_HIP(spr::launch_sorted_starts(H->hist_l.p, H->bucket_base.p, Np, H->sstart.p, s));
    // documents of a batch scan their winding numbers from their own first entry: with the fused
    // mark + emit that base comes from the band prefixes, before the sort
    if (H->batch && narrow)
        SPR_HIP(spr::launch_doc_winding_base_bands(H->band_start.p, H->bsum_l.p, d_nbands, H->sstart.p,
                                                   H->doc_first.p, Np, H->wbase.p, s));
    stage_mark(H, 8);
    const spr::EmitArgs E{H->crec.p, H->brec_l.p, H->band_cnt.p, H->band_path.p, H->bucket_base.p, H->rule.p,
                          pcv, H->batch ? H->wbase.p : nullptr, w, h};
    uint32_t* ticket = ctl_at(H, z.st_wind, z.counters, kLcSortTicket).counter;
    if (narrow) {
        // the oversized bands (queued by the classify pass) are sorted and emitted on the second
        // stream, followed there by the bands of up to 512 words, beside the longer bands
        hipStream_t sb = big_band_stream(H, nl, Np);
        SPR_HIP(spr::launch_band_classify(H->band_start.p, d_nbands, max_bands, H->bsum_l.p, H->big_list.p,
                                          H->big_list.p + max_bands, d_bigcount, ticket, s));
        if (sb != s) {
            SPR_HIP(hipEventRecord(H->ev_sfork, s));
            SPR_HIP(hipStreamWaitEvent(sb, H->ev_sfork, 0));
        }
        SPR_HIP(spr::launch_sort_big_bands(H->band_start.p, H->big_list.p, H->big_list.p + max_bands, d_bigcount,
                                           H->sbig.p, H->stmp.p, H->skey.p, nullptr, H->winding.p, sb,
                                           H->band_pr.p));
        SPR_HIP(spr::launch_emit_big_bands(H->band_start.p, H->big_list.p, d_bigcount, H->skey.p, H->winding.p,
                                           H->band_pr.p, E, sb));
        SPR_HIP(spr::launch_sort_bands(H->band_start.p, d_nbands, max_bands, H->sbig.p, H->skey.p, nullptr,
                                       H->winding.p, H->big_list.p, H->big_list.p + max_bands, d_bigcount,
                                       H->bsum_l.p, ticket, H->n_cu, s, H->band_pr.p, &E, side_for(H, nl, Np), nl,
                                       sb));
        if (sb != s) {
            SPR_HIP(hipEventRecord(H->ev_sjoin, sb));  // after the short bands' sort on sb
            SPR_HIP(hipStreamWaitEvent(s, H->ev_sjoin, 0));
        }
        SPR_HIP(spr::launch_band_totals(H->band_cnt.p, d_nbands, d_totals, s));
    } else {
        SPR_HIP(spr::launch_sort_bands(H->band_start.p, d_nbands, max_bands, H->sbig.p, H->skey.p,
                                       debug ? H->sidx.p : nullptr, H->winding.p, H->big_list.p,
                                       H->big_list.p + max_bands, d_bigcount, H->bsum_l.p, ticket, H->n_cu, s,
                                       nullptr, nullptr, side_for(H, nl, Np), nl));
        SPR_HIP(spr::launch_sort_big_bands(H->band_start.p, H->big_list.p, H->big_list.p + max_bands, d_bigcount,
                                           H->sbig.p, H->stmp.p, H->skey.p, debug ? H->sidx.p : nullptr,
                                           H->winding.p, s, nullptr));
    }
    stage_mark(H, 9);
    if (H->batch && !narrow)
        SPR_HIP(spr::launch_doc_winding_base(H->winding.p, H->sstart.p, H->doc_first.p, Np, H->wbase.p, s));
    stage_mark(H, 10);
    if (!narrow)
        SPR_HIP(spr::launch_mark_emit(H->skey.p, H->winding.p, H->epath.p, H->rule.p, H->fill.p, nl, w, h,
                                      ctl_at(H, z.st_mark, z.counters, kLcMark), debug ? H->records.p : nullptr,
                                      H->crec.p, debug ? H->flags.p : nullptr, d_totals, H->bucket_base.p, H->prow.p,
                                      H->brec_l.p, s, d_nsorted, H->sstart.p, Np, nr, d_err, pcv,
                                      H->batch ? H->wbase.p : nullptr));
    stage_mark(H, 11);
    spr::PaintArgs pa;
    SPR_HIP(spr::launch_composite(H->crec.p, H->brec_l.p, H->bucket_base.p, H->prow.p, H->fill.p, Np, w, h,
                                  (int)H->opts.compose, H->image.p, s, rec_cap, H->batch ? H->docs.p : nullptr,
                                  (int)H->n_docs, H->total_rows, paint_args(H, pa)));
    stage_mark(H, 12);
    if (!joined) SPR_HIP(hipStreamWaitEvent(s, H->ev_pjoin, 0));
    if (use_model && predict && !ovf_side)
        SPR_HIP(spr::launch_path_overflow(H->seg.p, H->capn.p, H->visible.p, Np, d_novf, s, d_skip));
    spr::SummarySrc S{};
    S.n = kSumCount;
    S.p[kSumCells] = d_totals;
    S.p[kSumSpans] = d_totals + 1;
    S.p[kSumErr] = d_err;
    S.p[kSumWaveChains] = C ? d_nlong : nullptr;
    S.p[kSumWaveChainsLight] = C ? d_cc + spr::kCcNLongLight : nullptr;
    S.p[kSumHeavyPieces] = C ? d_cc + spr::kCcHeavyPieces : nullptr;
    S.p[kSumLongBands] = reinterpret_cast<const int32_t*>(ticket + 1);
    S.p[kSumBigBands] = d_bigcount;
    S.p[kSumTotal] = H->counts.p + C;
    S.p[kSumOver] = d_over;
    S.p[kSumOverflow] = d_novf;
    S.p[kSumLive] = d_live;
    S.p[kSumArena] = H->arena.p + Np;
    S.p[kSumSorted] = d_nsorted;
    S.p[kSumCapR] = d_capr;
    finish_summary(H, S, s);
    if (H->capturing) return;
    SPR_HIP(hipEventRecord(H->ev_end, s));
    learned_finish(H);
}

// The end of a no-host-sync frame (eager or replayed graph): wait, check the device flags, fall
// back to the plain method when the frame was over its budget, and grow the budget.
void learned_finish(spr_handle* H) {
    SPR_HIP(hipStreamSynchronize(H->stream));
    const int32_t err = H->sum(kSumErr);
    const int64_t caps = std::max<int32_t>(H->sum(kSumArena), 0);
    const int64_t total = H->sum(kSumTotal);
    const int64_t overflow_paths = H->sum(kSumOverflow);
    H->frame_done = true;
    H->last_learned = true;
    H->last_fallback = false;
    H->last_brec = H->brec_l.p;
    H->last_fused = !H->opts.debug && H->w <= spr::kNarrowMaxW;  // as the frame's narrow words
    H->last_packed = H->last_fused;
    if (err & 1) {
        H->h_pinned[16 + kSumErr] = 0;
        throw Fail{SPR_ERR_DEVICE, "internal: an entry fell outside its path's row bound"};
    }
    if (H->sum(kSumOver)) {
        // the exact total did not fit: the device culled the frame.  A single scene is rendered
        // with the plain method; a batch, which has no plain frame, is rendered again on a budget
        // grown to its exact total (read here, as the plain method reads its count).  Either way
        // the budget covers it from now on.
        ++H->learned_fallbacks;
        H->last_fallback = true;
        H->lb_dirty = true;
        if (H->batch) {
            if (total < 0 || total > ((int64_t)1 << 30) - ((int64_t)1 << 27))
                throw Fail{SPR_ERR_CAPACITY, "the batch needs more than 2^30 intersections: split it"};
            grow_budget(H, total, caps);
            run_learned_frame(H);
            return;
        }
        run_frame(H);
        grow_budget(H, H->N, caps);
        if (H->last_predicted) H->stats.n_overflow_paths = overflow_paths;
        H->stats.alloc_intersections = caps;
        return;
    }
    if (H->last_predicted) H->stats.n_overflow_paths = overflow_paths;
    H->stats.alloc_intersections = caps;
    H->N = H->sum(kSumSorted);
    H->n_live = H->sum(kSumLive);
    H->last_fallback = false;
    grow_budget(H, total, caps);
}

void drop_graph(spr_handle* H) {
    for (int k = 0; k < 2; ++k) {
        if (H->gexec[k]) { (void)hipGraphExecDestroy(H->gexec[k]); H->gexec[k] = nullptr; }
        if (H->graph[k]) { (void)hipGraphDestroy(H->graph[k]); H->graph[k] = nullptr; }
        H->graph_lb_n[k] = H->graph_lb_r[k] = -1;
    }
}

// Records the frame's device work (kernels and the final device-to-host copies) into graph k:
// 1 with the prediction, 0 without.
void capture_frame(spr_handle* H, int k) {
    hipStream_t s = H->stream;
    if (H->gexec[k]) { (void)hipGraphExecDestroy(H->gexec[k]); H->gexec[k] = nullptr; }
    if (H->graph[k]) { (void)hipGraphDestroy(H->graph[k]); H->graph[k] = nullptr; }
    H->graph_lb_n[k] = H->graph_lb_r[k] = -1;
    SPR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    H->capturing = true;
    H->cap_predict = k == 1;
    try {
        run_learned_frame(H);
    } catch (...) {
        H->capturing = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
    }
    H->capturing = false;
    SPR_HIP(hipStreamEndCapture(s, &H->graph[k]));
    SPR_HIP(hipGraphInstantiate(&H->gexec[k], H->graph[k], nullptr, nullptr, 0));
    H->graph_lb_n[k] = H->lb_n;
    H->graph_lb_r[k] = H->lb_r;
    ++H->captures;
}

// The captured frame that the next learned frame would replay: with the prediction when it is
// due, else without.
int graph_slot(const spr_handle* H) { return prediction_due(H) ? 1 : 0; }

// Frames without a host synchronisation inside (learned mode, batches).  While the scene, model
// and budget stay the same, the frame is captured once and later frames replay the graph: one
// launch per frame.  Frames that run the prediction (after a new MVP: the graph copies it from
// pinned memory) and frames that do not have a graph each.
void render_nosync(spr_handle* H) {
    hipStream_t s = H->stream;
    const int k = graph_slot(H);
    if (!H->lb_dirty && !H->need_prediction && H->gexec[k] && H->graph_lb_n[k] == H->lb_n && !H->stage_timing) {
        publish_prediction(H);  // read by the graph's copy of the MVP block
        SPR_HIP(hipEventRecord(H->ev_begin, s));
        SPR_HIP(hipGraphLaunch(H->gexec[k], s));
        SPR_HIP(hipEventRecord(H->ev_end, s));
        H->n_marks = 0;
        learned_finish(H);
        return;
    }
    drop_graph(H);
    const bool predicting = H->need_prediction;
    const int64_t n0 = H->lb_n;
    run_learned_frame(H);
    if (H->last_fallback) return;  // the plain frame rendered it; the next learned frame runs eagerly
    H->lb_dirty = false;
    // capture once a frame ran eagerly on an unchanged budget
    if (!predicting && H->lb_n == n0 && !H->stage_timing) capture_frame(H, k);
}

template <class F>
int guarded(spr_handle* H, F&& f) {
    try {
        f();
        if (H) H->err.clear();
        return SPR_OK;
    } catch (const Fail& e) {
        if (H) H->err = e.msg; else g_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (H) H->err = "host out of memory"; else g_error = "host out of memory";
        return SPR_ERR_CAPACITY;
    } catch (const std::exception& e) {
        if (H) H->err = e.what(); else g_error = e.what();
        return SPR_ERR_INVALID;
    }
}

void sync_stats(spr_handle* H) {
    if (!H->frame_done) return;
    SPR_HIP(hipStreamSynchronize(H->stream));
    H->stats.n_points = H->P;
    H->stats.n_curves = H->C;
    H->stats.n_paths = H->Np;
    H->stats.n_intersections = H->N;
    if (H->sum(kSumErr) != 0) {
        H->h_pinned[16 + kSumErr] = 0;
        throw Fail{SPR_ERR_DEVICE, "internal: an entry fell outside its path's row bound"};
    }
    H->stats.n_cells = H->sum(kSumCells);
    H->stats.n_spans = H->sum(kSumSpans);
    H->stats.n_wave_chains = H->sum(kSumWaveChains) + H->sum(kSumWaveChainsLight);
    H->stats.n_heavy_pieces = H->sum(kSumHeavyPieces);
    H->stats.n_long_bands = H->sum(kSumLongBands);
    H->stats.n_big_bands = H->sum(kSumBigBands);
    H->stats.graph_captures = H->captures;
    H->stats.learned_fallbacks = H->learned_fallbacks;
    float ms = 0.0f;
    SPR_HIP(hipEventElapsedTime(&ms, H->ev_begin, H->ev_end));
    H->stats.frame_ms = ms;
    for (int k = 0; k < 16; ++k) H->stats.stage_ms[k] = 0.0;
    if (H->stage_timing) {
        for (int k = 0; k + 1 < H->n_marks && k < 16; ++k) {
            float t = 0.0f;
            SPR_HIP(hipEventElapsedTime(&t, H->ev_stage[k], H->ev_stage[k + 1]));
            H->stats.stage_ms[k] = t;
        }
    }
}

void drop_graph(spr_handle* H);

// Arc weights and gradient paints of a scene (each optional).
struct Extras {
    const float* cweight = nullptr;
    const uint32_t* path_paint = nullptr;
    const spr::Paint* paints = nullptr;
    uint32_t n_paints = 0;
    const spr::Stop* stops = nullptr;
    uint32_t n_stops = 0;
};

Extras extras_of(const spr::Scene& s) {
    Extras e;
    e.cweight = s.curve_weight.empty() ? nullptr : s.curve_weight.data();
    e.path_paint = s.path_paint.empty() ? nullptr : s.path_paint.data();
    e.paints = s.paints.data(); e.n_paints = (uint32_t)s.paints.size();
    e.stops = s.stops.data(); e.n_stops = (uint32_t)s.stops.size();
    return e;
}

void load_arrays(spr_handle* H, const float* pos, const uint32_t* pos_path, uint32_t P, const uint32_t* cpos,
                 const uint32_t* ctype, const uint32_t* cpath, uint32_t C, const uint32_t* rule,
                 const uint32_t* fill, uint32_t Np, const Extras& ex = Extras{}) {
    std::string bad = spr::validate(pos, pos_path, P, cpos, ctype, cpath, C, rule, Np);
    if (!bad.empty()) throw Fail{SPR_ERR_INVALID, bad};
    bad = spr::validate_extras(ctype, C, ex.cweight, Np, ex.path_paint, ex.paints, ex.n_paints, ex.stops,
                               ex.n_stops);
    if (!bad.empty()) throw Fail{SPR_ERR_INVALID, bad};
    bool arcs = false, paint = false;
    for (uint32_t c = 0; c < C && !arcs; ++c) arcs = ctype[c] == 0x13u;  // an arc
    if (ex.path_paint)
        for (uint32_t p = 0; p < Np && !paint; ++p) paint = ex.path_paint[p] != spr::kSolidPaint;
    if (Np && !fill) throw Fail{SPR_ERR_INVALID, "null fill_info"};
    if (Np > (1u << 30) - 1) throw Fail{SPR_ERR_CAPACITY, "too many paths"};
    SPR_HIP(hipSetDevice(H->device));
    H->pos.ensure(std::max<uint32_t>(P, 1));
    H->pos_path.ensure(std::max<uint32_t>(P, 1));
    H->cpos.ensure(std::max<uint32_t>(C, 1));
    H->ctype.ensure(std::max<uint32_t>(C, 1));
    H->cpath.ensure(std::max<uint32_t>(C, 1));
    H->rule.ensure(std::max<uint32_t>(Np, 1));
    H->fill.ensure(std::max<uint32_t>(Np, 1));
    hipStream_t s = H->stream;
    if (P) {
        SPR_HIP(hipMemcpyAsync(H->pos.p, pos, (size_t)P * 8, hipMemcpyHostToDevice, s));
        SPR_HIP(hipMemcpyAsync(H->pos_path.p, pos_path, (size_t)P * 4, hipMemcpyHostToDevice, s));
    }
    if (C) {
        SPR_HIP(hipMemcpyAsync(H->cpos.p, cpos, (size_t)C * 4, hipMemcpyHostToDevice, s));
        SPR_HIP(hipMemcpyAsync(H->ctype.p, ctype, (size_t)C * 4, hipMemcpyHostToDevice, s));
        SPR_HIP(hipMemcpyAsync(H->cpath.p, cpath, (size_t)C * 4, hipMemcpyHostToDevice, s));
    }
    if (Np) {
        SPR_HIP(hipMemcpyAsync(H->rule.p, rule, (size_t)Np * 4, hipMemcpyHostToDevice, s));
        SPR_HIP(hipMemcpyAsync(H->fill.p, fill, (size_t)Np * 4, hipMemcpyHostToDevice, s));
    }
    std::vector<uint32_t> pc((size_t)Np + 1, 0);
    for (uint32_t c = 0, p = 0; p <= Np; ++p) {
        while (c < C && cpath[c] < p) ++c;
        pc[p] = c;
    }
    H->pcurve.ensure((size_t)Np + 1);
    SPR_HIP(hipMemcpyAsync(H->pcurve.p, pc.data(), pc.size() * 4, hipMemcpyHostToDevice, s));
    if (arcs) {
        H->cweight.ensure(C);
        SPR_HIP(hipMemcpyAsync(H->cweight.p, ex.cweight, (size_t)C * 4, hipMemcpyHostToDevice, s));
    }
    if (paint) {
        H->path_paint.ensure(Np);
        H->paints.ensure(std::max<uint32_t>(ex.n_paints, 1));
        H->stops.ensure(5 * (size_t)std::max<uint32_t>(ex.n_stops, 1));
        SPR_HIP(hipMemcpyAsync(H->path_paint.p, ex.path_paint, (size_t)Np * 4, hipMemcpyHostToDevice, s));
        SPR_HIP(hipMemcpyAsync(H->paints.p, ex.paints, (size_t)ex.n_paints * sizeof(spr::Paint),
                               hipMemcpyHostToDevice, s));
        SPR_HIP(hipMemcpyAsync(H->stops.p, ex.stops, (size_t)ex.n_stops * sizeof(spr::Stop), hipMemcpyHostToDevice,
                               s));
    }
    SPR_HIP(hipStreamSynchronize(s));
    H->has_arcs = arcs;
    H->has_paint = paint;
    H->h_cpath.assign(cpath, cpath + C);
    H->P = P; H->C = C; H->Np = Np;
    H->loaded = true;
    H->frame_done = false;
    H->inj = false;
    H->inj_inter.clear();
    H->lb_dirty = true;
    H->need_prediction = true;
    H->batch = false;
    drop_graph(H);
}

}  // namespace

extern "C" {

const char* spr_global_error(void) { return g_error.c_str(); }

const char* spr_last_error(const spr_handle* h) { return h ? h->err.c_str() : g_error.c_str(); }

int spr_create(uint32_t width, uint32_t height, const spr_opts* opts, spr_handle** out) {
    if (!out) { g_error = "null out"; return SPR_ERR_INVALID; }
    *out = nullptr;
    if (width == 0 || height == 0 || width > 8192 || height > 16384) {
        g_error = "canvas must be 1..8192 x 1..16384";
        return SPR_ERR_CAPACITY;
    }
    spr_handle* H = new (std::nothrow) spr_handle();
    if (!H) { g_error = "host out of memory"; return SPR_ERR_CAPACITY; }
    int rc = guarded(nullptr, [&] {
        H->w = width; H->h = height;
        if (opts) H->opts = *opts;
        if (H->opts.compose > 1 || H->opts.mode > 1) throw Fail{SPR_ERR_INVALID, "bad compose or mode"};
        if (H->opts.aa != 0 && H->opts.aa != (uint32_t)spr::kAaSamples)
            throw Fail{SPR_ERR_INVALID, "aa must be 0 or 16 (samples per pixel)"};
        if (H->opts.aa && (H->opts.compose != SPR_COMPOSE_SRC_OVER || H->opts.mode != SPR_MODE_PLAIN))
            throw Fail{SPR_ERR_INVALID, "anti-aliased coverage needs compose src-over and the plain mode"};
        int dev = 0;
        if (H->opts.device >= 0) dev = H->opts.device; else SPR_HIP(hipGetDevice(&dev));
        H->device = dev;
        SPR_HIP(hipSetDevice(dev));
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            H->n_cu = ncu;
        (void)hipGetLastError();
        SPR_HIP(hipStreamCreateWithFlags(&H->stream, hipStreamNonBlocking));
        SPR_HIP(hipStreamCreateWithFlags(&H->stream2, hipStreamNonBlocking));
        for (hipEvent_t* e : {&H->ev_fork, &H->ev_light, &H->ev_small, &H->ev_join, &H->ev_sfork, &H->ev_sjoin,
                              &H->ev_pfork, &H->ev_pjoin, &H->ev_pseg})
            SPR_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
        SPR_HIP(hipEventCreate(&H->ev_begin));
        SPR_HIP(hipEventCreate(&H->ev_end));
        H->ev_stage.resize(16);
        for (auto& e : H->ev_stage) SPR_HIP(hipEventCreate(&e));
        SPR_HIP(hipHostMalloc((void**)&H->h_pinned, 512, hipHostMallocDefault));
        std::memset(H->h_pinned, 0, 512);
        H->summary.ensure(spr::kSummaryMax);
        H->mvp_dev.ensure(32);  // before any capture: a frame never allocates it
        const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        std::memcpy(H->mvp.m, id, sizeof(id));
        std::memcpy(H->mvp_pinned(), id, sizeof(id));
        mvp_inverse(id, H->inv_pinned());
        H->image.ensure((size_t)4 * width * height);
    });
    if (rc != SPR_OK) { spr_destroy(H); return rc; }
    *out = H;
    return SPR_OK;
}

int spr_destroy(spr_handle* H) {
    if (!H) return SPR_OK;
    if (H->stream) (void)hipStreamSynchronize(H->stream);
    DBuf<int>* unused = nullptr; (void)unused;
    H->pos.release(); H->pos_path.release(); H->cpos.release(); H->ctype.release(); H->cpath.release();
    H->rule.release(); H->fill.release(); H->tpos.release(); H->visible.release(); H->cuts.release();
    H->ncuts.release(); H->counts.release(); H->seg.release(); H->inter.release();
    H->chain_t.release(); H->chains.release(); H->pieces.release(); H->chain_order.release();
    H->heavy_list.release(); H->lo_list.release(); H->lo_cnt.release(); H->wave_list.release(); H->chain_cls.release(); H->mono.release(); H->key.release(); H->wn.release(); H->epath.release(); H->sorted.release();
    H->sbig.release(); H->stmp.release(); H->skey.release(); H->sidx.release();
    H->winding.release(); H->flags.release(); H->records.release(); H->crec.release();
    H->cell_first.release(); H->cmask.release(); H->chord.release();
    H->row_lo.release(); H->row_hi.release(); H->bucket_base.release(); H->bflag.release();
    H->band_start.release(); H->band_pr.release(); H->big_list.release(); H->prow.release(); H->image.release(); H->zero.release();
    drop_graph(H);
    H->pcanvas.release(); H->docs.release(); H->doc_first.release(); H->wbase.release();
    H->mlpdev.release();
    H->pcurve.release(); H->feat.release(); H->pred.release(); H->train_feat.release(); H->train_tgt.release();
    H->arena.release(); H->capn.release(); H->sstart.release();
    H->hist_l.release(); H->fill_l.release(); H->bsum_l.release(); H->brec_l.release(); H->pbound.release();
    H->summary.release();
    H->scan_ws.release();
    H->inj_rec.release();
    H->mvp_dev.release();
    H->cweight.release(); H->path_paint.release(); H->paints.release(); H->stops.release();
    if (H->h_pinned) (void)hipHostFree(H->h_pinned);
    for (auto& e : H->ev_stage) if (e) (void)hipEventDestroy(e);
    if (H->ev_begin) (void)hipEventDestroy(H->ev_begin);
    if (H->ev_end) (void)hipEventDestroy(H->ev_end);
    for (hipEvent_t e : {H->ev_fork, H->ev_light, H->ev_small, H->ev_join, H->ev_sfork, H->ev_sjoin, H->ev_pfork,
                         H->ev_pjoin, H->ev_pseg})
        if (e) (void)hipEventDestroy(e);
    if (H->stream2) (void)hipStreamDestroy(H->stream2);
    if (H->stream) (void)hipStreamDestroy(H->stream);
    (void)hipGetLastError();  // nothing of a destroyed handle may leak into later calls
    delete H;
    return SPR_OK;
}

int spr_load_paths(spr_handle* H, const float* pos_xy, const uint32_t* pos_path_idx, uint32_t n_points,
                   const uint32_t* curve_pos_map, const uint32_t* curve_type, const uint32_t* curve_path_idx,
                   uint32_t n_curves, const uint32_t* fill_rule, const uint32_t* fill_info, uint32_t n_paths) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        load_arrays(H, pos_xy, pos_path_idx, n_points, curve_pos_map, curve_type, curve_path_idx, n_curves,
                    fill_rule, fill_info, n_paths);
    });
}

int spr_load_scene(spr_handle* H, const spr_scene* sc) {
    if (!H || !sc) { g_error = "null argument"; return SPR_ERR_INVALID; }
    const spr::Scene& s = sc->s;
    return guarded(H, [&] {
        load_arrays(H, s.pos.data(), s.pos_path.data(), s.n_points(), s.curve_pos.data(), s.curve_type.data(),
                    s.curve_path.data(), s.n_curves(), s.rule.data(), s.fill.data(), s.n_paths(), extras_of(s));
    });
}

int spr_load_rvg(spr_handle* H, const char* path) {
    if (!H || !path) { g_error = "null argument"; return SPR_ERR_INVALID; }
    spr::Scene s;
    int rc = guarded(H, [&] {
        try { s = spr::load_rvg(path); } catch (const std::exception& e) { throw Fail{SPR_ERR_IO, e.what()}; }
    });
    if (rc != SPR_OK) return rc;
    return guarded(H, [&] {
        load_arrays(H, s.pos.data(), s.pos_path.data(), s.n_points(), s.curve_pos.data(), s.curve_type.data(),
                    s.curve_path.data(), s.n_curves(), s.rule.data(), s.fill.data(), s.n_paths(), extras_of(s));
    });
}

int spr_set_mvp(spr_handle* H, const float m_rows[16]) {
    if (!H || !m_rows) { g_error = "null argument"; return SPR_ERR_INVALID; }
    // every frame has ended when this returns (frames end with a stream synchronisation), so no
    // copy is reading the pinned matrix; a captured frame picks the new one up at its next replay
    std::memcpy(H->mvp.m, m_rows, sizeof(H->mvp.m));
    std::memcpy(H->mvp_pinned(), m_rows, sizeof(H->mvp.m));
    H->pred_stale = true;  // the prediction depends on the transformed points
    mvp_inverse(m_rows, H->inv_pinned());
    return SPR_OK;
}

// Record replay: draws the injected records into the handle's image.
void run_records_frame(spr_handle* H) {
    hipStream_t s = H->stream;
    H->image.ensure((size_t)4 * H->w * H->h);
    SPR_HIP(hipEventRecord(H->ev_begin, s));
    SPR_HIP(spr::launch_draw_records(H->inj_rec.p, (int32_t)H->inj_rec_n, (int)H->w, (int)H->h,
                                     (int)H->opts.compose, H->image.p, s));
    SPR_HIP(hipEventRecord(H->ev_end, s));
    SPR_HIP(hipStreamSynchronize(s));
    H->n_marks = 0;
    H->N = 0;
    std::memset(H->h_pinned + 16, 0, sizeof(int32_t) * kSumCount);
    H->h_pinned[16 + kSumCells] = (int32_t)H->inj_rec_cells;
    H->h_pinned[16 + kSumSpans] = (int32_t)(H->inj_rec_n - H->inj_rec_cells);
    H->frame_done = true;
    H->last_learned = false;
    H->last_brec = nullptr;
}

int spr_render_frame(spr_handle* H, uint8_t* rgba_out) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        SPR_HIP(hipSetDevice(H->device));
        if (H->inj_rec_n >= 0 && !H->batch) {
            run_records_frame(H);
            if (rgba_out) {
                SPR_HIP(hipMemcpyAsync(rgba_out, H->image.p, (size_t)4 * H->w * H->h, hipMemcpyDeviceToHost,
                                       H->stream));
                SPR_HIP(hipStreamSynchronize(H->stream));
            }
            return;
        }
        if (H->batch) {
            if (rgba_out) throw Fail{SPR_ERR_INVALID, "a batch has one image per document: use spr_batch_image"};
            render_nosync(H);
            return;
        }
        if (H->opts.mode == SPR_MODE_LEARNED && !H->inj)
            render_nosync(H);
        else
            run_frame(H);
        if (rgba_out) {
            SPR_HIP(hipMemcpyAsync(rgba_out, H->image.p, (size_t)4 * H->w * H->h, hipMemcpyDeviceToHost, H->stream));
            SPR_HIP(hipStreamSynchronize(H->stream));
        }
    });
}

int spr_image_device_ptr(spr_handle* H, void** dptr) {
    if (!H || !dptr) { g_error = "null argument"; return SPR_ERR_INVALID; }
    *dptr = H->image.p;
    return SPR_OK;
}

int spr_synchronize(spr_handle* H) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    return guarded(H, [&] { SPR_HIP(hipStreamSynchronize(H->stream)); });
}

int spr_set_stage_timing(spr_handle* H, int enable) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    H->stage_timing = enable != 0;
    return SPR_OK;
}

int spr_get_stats(const spr_handle* Hc, spr_stats* out) {
    spr_handle* H = const_cast<spr_handle*>(Hc);
    if (!H || !out) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        sync_stats(H);
        *out = H->stats;
    });
}

int spr_scan_i32(spr_handle* H, const int32_t* d_in, int32_t* d_out, uint64_t n, void* stream) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (!d_out || (n && !d_in)) throw Fail{SPR_ERR_INVALID, "null device pointer"};
        if (n > (uint64_t)INT32_MAX - 1) throw Fail{SPR_ERR_CAPACITY, "scan length exceeds 2^31 - 2"};
        SPR_HIP(hipSetDevice(H->device));
        hipStream_t s = stream ? (hipStream_t)stream : H->stream;
        const int64_t tiles = spr::scan_tiles(std::max<int64_t>((int64_t)n, 1));
        if ((size_t)tiles + 1 > H->scan_ws.cap) {
            // the old workspace may still be in use by a scan queued on another stream
            SPR_HIP(hipDeviceSynchronize());
            H->scan_ws.ensure((size_t)tiles + 1);
        }
        // tile statuses and the tile counter start at zero (a fill kernel: capturable)
        SPR_HIP(spr::launch_fill_u32((uint32_t*)H->scan_ws.p, 2 * (tiles + 1), 0u, s));
        spr::ScanCtl c;
        c.status = H->scan_ws.p;
        c.counter = reinterpret_cast<uint32_t*>(H->scan_ws.p + tiles);
        c.epoch = 0;
        c.shape = H->scan_shape;
        SPR_HIP(spr::launch_scan_i32(d_in, d_out, (int64_t)n, c, s));
    });
}

int spr_scan_i32_noalloc(const int32_t* d_in, int32_t* d_out, uint64_t n, void* stream) {
    // no allocation on any path: errors are string literals copied into a fixed buffer
    if (!d_out || (n && !d_in)) { g_error = "spr_scan_i32_noalloc: null device pointer"; return SPR_ERR_INVALID; }
    if (n > (uint64_t)spr::kNoAllocScanMaxN) {
        g_error = "spr_scan_i32_noalloc: n exceeds 268435456, the size of the static workspace";
        return SPR_ERR_CAPACITY;
    }
    const hipError_t e = spr::launch_scan_i32_noalloc(d_in, d_out, (int64_t)n, (hipStream_t)stream);
    if (e != hipSuccess) {
        g_error = hipGetErrorString(e);
        return SPR_ERR_DEVICE;
    }
    return SPR_OK;
}

int spr_debug_export(spr_handle* H, int stage, void* dst, size_t cap, size_t* nbytes) {
    if (!H || !nbytes) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (!H->frame_done) throw Fail{SPR_ERR_STATE, "no frame rendered"};
        SPR_HIP(hipStreamSynchronize(H->stream));
        const bool debug = H->opts.debug != 0;
        const int64_t N = H->N;
        // learned frames keep intersections in per-path arenas: position-indexed stages span
        // the live positions (gaps have curve -1), sorted stages span the N entries
        const int64_t NP = H->last_learned ? H->n_live : N;
        const void* src = nullptr;
        size_t n = 0;
        std::vector<uint8_t> tmp;
        auto from_dev = [&](const void* p, size_t bytes) {
            tmp.resize(bytes);
            if (bytes) SPR_HIP(hipMemcpy(tmp.data(), p, bytes, hipMemcpyDeviceToHost));
            src = tmp.data(); n = bytes;
        };
        switch (stage) {
            case SPR_STAGE_TPOS: from_dev(H->tpos.p, (size_t)H->P * 8); break;
            case SPR_STAGE_VISIBLE: from_dev(H->visible.p, (size_t)H->Np * 4); break;
            case SPR_STAGE_CUTS: {
                std::vector<float4> q(H->C);
                std::vector<uint32_t> nc(H->C);
                if (H->C) {
                    SPR_HIP(hipMemcpy(q.data(), H->cuts.p, (size_t)H->C * 16, hipMemcpyDeviceToHost));
                    SPR_HIP(hipMemcpy(nc.data(), H->ncuts.p, (size_t)H->C * 4, hipMemcpyDeviceToHost));
                }
                tmp.resize((size_t)H->C * 20);
                float* o = (float*)tmp.data();
                for (uint32_t c = 0; c < H->C; ++c) {
                    o[5 * c] = q[c].x; o[5 * c + 1] = q[c].y; o[5 * c + 2] = q[c].z; o[5 * c + 3] = q[c].w;
                    std::memcpy(&o[5 * c + 4], &nc[c], 4);
                }
                src = tmp.data(); n = tmp.size();
                break;
            }
            case SPR_STAGE_COUNTS: from_dev(H->counts.p, ((size_t)H->C + 1) * 4); break;
            case SPR_STAGE_INTER:
                // the merged entries go straight into fragment keys unless they are kept
                if (!debug && !H->inj) throw Fail{SPR_ERR_STATE, "intersections are kept only with the debug option"};
                from_dev(H->inter.p, (size_t)NP * 8);
                break;
            case SPR_STAGE_KEY:
                if (H->last_packed) throw Fail{SPR_ERR_STATE, "fragment keys are kept only with the debug option"};
                from_dev(H->key.p, (size_t)NP * 4);
                break;
            case SPR_STAGE_WN: {
                if (H->last_packed) throw Fail{SPR_ERR_STATE, "winding changes are kept only with the debug option"};
                std::vector<int8_t> v((size_t)NP);
                if (NP) SPR_HIP(hipMemcpy(v.data(), H->wn.p, (size_t)NP, hipMemcpyDeviceToHost));
                tmp.resize((size_t)NP * 4);
                for (int64_t i = 0; i < NP; ++i) ((int32_t*)tmp.data())[i] = v[i];
                src = tmp.data(); n = tmp.size();
                break;
            }
            case SPR_STAGE_SEG: from_dev(H->seg.p, ((size_t)H->Np + 1) * 4); break;
            case SPR_STAGE_SKEY:
                if (H->last_fused) throw Fail{SPR_ERR_STATE, "sorted keys are kept only with the debug option"};
                from_dev(H->skey.p, (size_t)N * 4);
                break;
            case SPR_STAGE_SIDX:
                if (!debug) throw Fail{SPR_ERR_STATE, "sorted indices are kept only with the debug option"};
                from_dev(H->sidx.p, (size_t)N * 4);
                break;
            case SPR_STAGE_WINDING:
                if (H->last_fused) throw Fail{SPR_ERR_STATE, "winding numbers are kept only with the debug option"};
                from_dev(H->winding.p, ((size_t)N + 1) * 4);
                break;
            case SPR_STAGE_FLAGS:
                if (!debug) throw Fail{SPR_ERR_STATE, "flags are kept only with the debug option"};
                from_dev(H->flags.p, (size_t)N * 8);
                break;
            case SPR_STAGE_RECORDS: {
                if (!debug) throw Fail{SPR_ERR_STATE, "full records are kept only with the debug option"};
                int64_t R = (int64_t)H->sum(kSumCells) + H->sum(kSumSpans);
                from_dev(H->records.p, (size_t)R * 16);
                break;
            }
            case SPR_STAGE_IMAGE: from_dev(H->image.p, (size_t)4 * H->w * H->h); break;
            case SPR_STAGE_COVERAGE: {
                if (!H->opts.aa || H->last_learned || H->inj_rec_n >= 0)
                    throw Fail{SPR_ERR_STATE, "sample coverage is computed by anti-aliased frames"};
                int64_t R = (int64_t)H->sum(kSumCells) + H->sum(kSumSpans);
                from_dev(H->cmask.p, (size_t)R * 8);
                break;
            }
            case SPR_STAGE_FEATURES:
                if (!H->last_learned) throw Fail{SPR_ERR_STATE, "features are computed by learned frames"};
                from_dev(H->feat.p, (size_t)H->Np * spr::kMlpIn * 4);
                break;
            case SPR_STAGE_PRED:
                if (!H->last_learned) throw Fail{SPR_ERR_STATE, "predictions are computed by learned frames"};
                from_dev(H->pred.p, (size_t)H->Np * spr::kMlpOut * 4);
                break;
            case SPR_STAGE_PATH_COUNTS: {
                if (!H->last_brec) throw Fail{SPR_ERR_STATE, "no frame rendered"};
                if (!H->last_learned && H->inj) throw Fail{SPR_ERR_STATE, "not available with injected intersections"};
                H->train_tgt.ensure((size_t)std::max<uint32_t>(H->Np, 1) * 2);
                DBuf<int32_t> exact;
                exact.ensure((size_t)std::max<uint32_t>(H->Np, 1) * 2);
                SPR_HIP(spr::launch_path_targets(H->seg.p, 1,
                                                 H->last_brec, H->bucket_base.p, H->Np, H->train_tgt.p, exact.p,
                                                 H->stream));
                SPR_HIP(hipStreamSynchronize(H->stream));  // the copy below runs on another stream
                from_dev(exact.p, (size_t)H->Np * 8);
                exact.release();
                break;
            }
            default: throw Fail{SPR_ERR_INVALID, "unknown stage"};
        }
        *nbytes = n;
        if (dst) {
            if (cap < n) throw Fail{SPR_ERR_CAPACITY, "destination too small"};
            if (n) std::memcpy(dst, src, n);
        }
    });
}

int spr_debug_inject(spr_handle* H, int stage, const void* src, size_t nbytes) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (stage == SPR_STAGE_RECORDS) {
            if (nbytes == 0) { H->inj_rec_n = -1; H->inj_rec.release(); return; }
            if (!src || nbytes % 16) throw Fail{SPR_ERR_INVALID, "records are int4 (yx, width, fill, frag_index)"};
            const size_t n = nbytes / 16;
            if (n > ((size_t)1 << 22)) throw Fail{SPR_ERR_CAPACITY, "a replayed record list holds at most 2^22 records"};
            SPR_HIP(hipSetDevice(H->device));
            SPR_HIP(hipStreamSynchronize(H->stream));
            H->inj_rec.ensure(std::max<size_t>(n, 1));
            if (n) SPR_HIP(hipMemcpy(H->inj_rec.p, src, nbytes, hipMemcpyHostToDevice));
            const int32_t* v = (const int32_t*)src;
            int64_t cells = 0;
            for (size_t i = 0; i < n; ++i) cells += v[4 * i + 3] != 0;  // cells carry their index, spans 0
            H->inj_rec_n = (int64_t)n;
            H->inj_rec_cells = cells;
            return;
        }
        if (stage != SPR_STAGE_INTER) throw Fail{SPR_ERR_INVALID, "only SPR_STAGE_INTER and SPR_STAGE_RECORDS can be injected"};
        if (nbytes == 0) { H->inj = false; H->inj_inter.clear(); return; }
        if (!src || nbytes % 8) throw Fail{SPR_ERR_INVALID, "intersections are int2 records"};
        if (!H->loaded) throw Fail{SPR_ERR_STATE, "inject before load"};
        const int32_t* v = (const int32_t*)src;
        const size_t n = nbytes / 8;
        for (size_t i = 0; i < n; ++i) {
            if (v[2 * i] < 0 || (uint32_t)v[2 * i] >= H->C) throw Fail{SPR_ERR_INVALID, "curve index out of range"};
            if (i && v[2 * i] < v[2 * i - 2]) throw Fail{SPR_ERR_INVALID, "intersections must be in curve order"};
        }
        H->inj_inter.assign(v, v + 2 * n);
        H->inj = true;
    });
}

// ---- batches of documents
int spr_load_batch(spr_handle* H, const spr_scene* const* docs, const uint32_t* widths, const uint32_t* heights,
                   uint32_t n_docs) {
    if (!H || !docs || !widths || !heights) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (n_docs == 0) throw Fail{SPR_ERR_INVALID, "empty batch"};
        if (H->opts.aa) throw Fail{SPR_ERR_INVALID, "batches render without anti-aliased coverage"};
        std::vector<float> pos, cweight;
        std::vector<uint32_t> pos_path, cpos, ctype, cpath, rule, fill, path_paint;
        std::vector<spr::Paint> paints;
        std::vector<spr::Stop> stops;
        bool any_arcs = false, any_paint = false;
        for (uint32_t d = 0; d < n_docs; ++d) {
            if (!docs[d]) continue;
            any_arcs = any_arcs || !docs[d]->s.curve_weight.empty();
            any_paint = any_paint || !docs[d]->s.path_paint.empty();
        }
        std::vector<spr::DocDev> dd(n_docs);
        std::vector<int2> pcv;
        std::vector<int32_t> first;
        size_t img = 0;
        int rows = 0, max_h = 1;
        for (uint32_t d = 0; d < n_docs; ++d) {
            if (!docs[d]) throw Fail{SPR_ERR_INVALID, "null document"};
            const uint32_t w = widths[d], h = heights[d];
            if (w == 0 || h == 0 || w > 8192 || h > 16384) throw Fail{SPR_ERR_CAPACITY, "canvas must be 1..8192 x 1..16384"};
            const spr::Scene& S = docs[d]->s;
            const uint32_t p_off = (uint32_t)rule.size(), pt_off = (uint32_t)pos_path.size();
            pos.insert(pos.end(), S.pos.begin(), S.pos.end());
            for (uint32_t v : S.pos_path) pos_path.push_back(v + p_off);
            for (uint32_t v : S.curve_pos) cpos.push_back(v + pt_off);
            ctype.insert(ctype.end(), S.curve_type.begin(), S.curve_type.end());
            for (uint32_t v : S.curve_path) cpath.push_back(v + p_off);
            rule.insert(rule.end(), S.rule.begin(), S.rule.end());
            fill.insert(fill.end(), S.fill.begin(), S.fill.end());
            if (any_arcs) {
                if (S.curve_weight.empty()) cweight.insert(cweight.end(), S.n_curves(), 1.0f);
                else cweight.insert(cweight.end(), S.curve_weight.begin(), S.curve_weight.end());
            }
            if (any_paint) {
                // the document's paints and stops after those of the earlier documents
                const uint32_t g_off = (uint32_t)paints.size(), s_off = (uint32_t)stops.size();
                for (uint32_t p = 0; p < S.n_paths(); ++p) {
                    const uint32_t q = S.path_paint.empty() ? spr::kSolidPaint : S.path_paint[p];
                    path_paint.push_back(q == spr::kSolidPaint ? q : q + g_off);
                }
                for (spr::Paint g : S.paints) { g.first_stop += s_off; paints.push_back(g); }
                stops.insert(stops.end(), S.stops.begin(), S.stops.end());
            }
            dd[d].p0 = (int32_t)p_off;
            dd[d].p1 = (int32_t)rule.size();
            dd[d].w = (int32_t)w;
            dd[d].h = (int32_t)h;
            dd[d].img_off = (int64_t)img;
            dd[d].row0 = rows;
            dd[d].pad = 0;
            img += (size_t)4 * w * h;
            rows += (int)(h + 1) / 2;
            max_h = std::max<int>(max_h, (int)h);
            for (uint32_t p = p_off; p < rule.size(); ++p) { pcv.push_back(make_int2((int)w, (int)h)); first.push_back((int32_t)p_off); }
        }
        if (rule.size() > (1u << 30) || pos_path.size() > (1u << 31)) throw Fail{SPR_ERR_CAPACITY, "batch too large"};
        const uint32_t Np = (uint32_t)rule.size();
        load_arrays(H, pos.data(), pos_path.data(), (uint32_t)pos_path.size(), cpos.data(), ctype.data(), cpath.data(),
                    (uint32_t)ctype.size(), rule.data(), fill.data(), Np,
                    Extras{any_arcs ? cweight.data() : nullptr, any_paint ? path_paint.data() : nullptr, paints.data(),
                           (uint32_t)paints.size(), stops.data(), (uint32_t)stops.size()});
        first.push_back(0);
        pcv.push_back(make_int2(1, 1));
        H->pcanvas.ensure(pcv.size());
        H->doc_first.ensure(first.size());
        H->docs.ensure(n_docs);
        H->wbase.ensure((size_t)Np + 1);
        SPR_HIP(hipMemcpyAsync(H->pcanvas.p, pcv.data(), pcv.size() * sizeof(int2), hipMemcpyHostToDevice, H->stream));
        SPR_HIP(hipMemcpyAsync(H->doc_first.p, first.data(), first.size() * 4, hipMemcpyHostToDevice, H->stream));
        SPR_HIP(hipMemcpyAsync(H->docs.p, dd.data(), dd.size() * sizeof(spr::DocDev), hipMemcpyHostToDevice, H->stream));
        SPR_HIP(hipStreamSynchronize(H->stream));
        H->batch = true;
        H->n_docs = n_docs;
        H->h_docs = dd;
        H->total_rows = rows;
        H->max_h = max_h;
        H->img_bytes = img;
        H->lb_dirty = true;
        H->need_prediction = true;
    });
}

int spr_batch_image(spr_handle* H, uint32_t doc, uint8_t* rgba_out, size_t cap) {
    if (!H || !rgba_out) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (!H->batch) throw Fail{SPR_ERR_STATE, "no batch loaded"};
        if (doc >= H->n_docs) throw Fail{SPR_ERR_INVALID, "document index out of range"};
        if (!H->frame_done) throw Fail{SPR_ERR_STATE, "no frame rendered"};
        const spr::DocDev& d = H->h_docs[doc];
        const size_t bytes = (size_t)4 * d.w * d.h;
        if (cap < bytes) throw Fail{SPR_ERR_CAPACITY, "destination too small"};
        SPR_HIP(hipMemcpyAsync(rgba_out, H->image.p + d.img_off, bytes, hipMemcpyDeviceToHost, H->stream));
        SPR_HIP(hipStreamSynchronize(H->stream));
    });
}

int spr_autotune(spr_handle* H, uint32_t reps, spr_tune_report* out) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (!H->loaded) throw Fail{SPR_ERR_STATE, "autotune before load"};
        if (H->inj) throw Fail{SPR_ERR_STATE, "autotune with injected intersections"};
        if (reps == 0) reps = 3;
        const bool nosync = H->batch || (H->opts.mode == SPR_MODE_LEARNED && H->model.ready);
        const bool timing = H->stage_timing;
        H->stage_timing = false;
        // device time of the current knobs: captured-graph replays for frames without a host
        // synchronisation, eager frames otherwise
        auto measure = [&]() {
            if (nosync) {
                drop_graph(H);
                render_nosync(H);  // eager; sizes the budget when needed
                render_nosync(H);  // eager, then captured
                if (!H->gexec[graph_slot(H)]) render_nosync(H);
            } else {
                run_frame(H);
            }
            double acc = 0.0;
            for (uint32_t r = 0; r < reps; ++r) {
                if (nosync) render_nosync(H); else run_frame(H);
                SPR_HIP(hipStreamSynchronize(H->stream));
                float ms = 0.0f;
                SPR_HIP(hipEventElapsedTime(&ms, H->ev_begin, H->ev_end));
                acc += ms;
            }
            return acc / reps;
        };
        // coordinate descent from the defaults: each knob in turn over its values, the others at
        // their best so far
        int* knob[4] = {&H->chain_k, &H->merge_k, &H->scan_shape, &H->band_window};
        const std::vector<int> values[4] = {{8, 4, 16}, {8, 4, 16}, {0, 1, 2, 3}, {spr::kBandWindowLog, 10}};
        for (int k = 0; k < 4; ++k) *knob[k] = values[k][0];
        const double dflt = measure();
        double best = dflt;
        int tried = 1;
        for (int k = 0; k < 4; ++k) {
            const int keep = *knob[k];
            int best_v = keep;
            for (size_t i = 1; i < values[k].size(); ++i) {
                *knob[k] = values[k][i];
                const double ms = measure();
                ++tried;
                if (ms < best) { best = ms; best_v = values[k][i]; }
            }
            *knob[k] = best_v;
        }
        drop_graph(H);
        H->stage_timing = timing;
        if (out) {
            out->chain_blocks_per_cu = H->chain_k;
            out->merge_blocks_per_cu = H->merge_k;
            out->n_tried = tried;
            out->reserved = 0;
            out->default_ms = dflt;
            out->best_ms = best;
            out->scan_shape = H->scan_shape;
            out->band_window_log = H->band_window;
        }
    });
}

int spr_set_tuning(spr_handle* H, const spr_tune_report* k) {
    if (!H || !k) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        auto one_of = [](int v, std::initializer_list<int> ok) {
            for (int o : ok) if (v == o) return true;
            return false;
        };
        if (!one_of(k->chain_blocks_per_cu, {4, 8, 16}) || !one_of(k->merge_blocks_per_cu, {4, 8, 16}) ||
            k->scan_shape < 0 || k->scan_shape >= spr::kScanShapes || !one_of(k->band_window_log, {spr::kBandWindowLog, 10}))
            throw Fail{SPR_ERR_INVALID, "tuning knob out of range"};
        H->chain_k = k->chain_blocks_per_cu;
        H->merge_k = k->merge_blocks_per_cu;
        H->scan_shape = k->scan_shape;
        H->band_window = k->band_window_log;
        drop_graph(H);  // the captured frame holds the old launch shapes
    });
}

// ---- learned allocation
int spr_learned_train(spr_handle* H, const spr_scene* sc, const spr_train_opts* o, spr_train_report* out) {
    if (!H || !sc) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        SPR_HIP(hipSetDevice(H->device));
        const spr::Scene& S = sc->s;
        const uint32_t Np = S.n_paths();
        if (Np < 16) throw Fail{SPR_ERR_INVALID, "training scene needs at least 16 paths"};
        const uint32_t chunk = o && o->chunk_paths ? o->chunk_paths : 5000u;
        H->train_feat.ensure((size_t)Np * spr::kMlpIn);
        H->train_tgt.ensure((size_t)Np * spr::kMlpOut);
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t first = 0; first < Np; first += chunk) {
            const uint32_t cnt = std::min(chunk, Np - first);
            const spr::Scene sub = spr::subset(S, first, cnt);
            load_arrays(H, sub.pos.data(), sub.pos_path.data(), sub.n_points(), sub.curve_pos.data(),
                        sub.curve_type.data(), sub.curve_path.data(), sub.n_curves(), sub.rule.data(),
                        sub.fill.data(), sub.n_paths(), extras_of(sub));
            run_frame(H);
            SPR_HIP(spr::launch_path_features(H->tpos.p, H->cpos.p, H->ctype.p, H->pcurve.p, H->visible.p,
                                              H->rule.p, cnt, (int)H->w, (int)H->h,
                                              H->train_feat.p + (size_t)first * spr::kMlpIn, H->stream));
            SPR_HIP(spr::launch_path_targets(H->seg.p, 1, H->last_brec, H->bucket_base.p, cnt,
                                             H->train_tgt.p + (size_t)first * spr::kMlpOut, nullptr, H->stream));
            SPR_HIP(hipStreamSynchronize(H->stream));
        }
        const double collect_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        spr::TrainOptions to;
        if (o) {
            if (o->epochs) to.epochs = (int)o->epochs;
            if (o->batch) to.batch = (int)o->batch;
            if (o->lr > 0) to.lr = o->lr;
            to.seed = o->seed;
            if (o->overflow_quantile > 0) to.overflow_quantile = o->overflow_quantile;
        }
        spr::TrainReport rep;
        spr::mlp_train(H->model, H->train_feat.p, H->train_tgt.p, (int)Np, to, rep, H->stream);
        H->mlpdev.upload(H->model, H->stream);
        H->loaded = false;
        H->frame_done = false;
        H->last_brec = nullptr;
        H->lb_dirty = true;
        H->need_prediction = true;
        if (out) {
            out->samples = rep.samples;
            out->steps = rep.steps;
            out->first_loss = rep.first_loss;
            out->train_loss = rep.train_loss;
            out->val_loss = rep.val_loss;
            out->log_margin = rep.log_margin;
            out->collect_ms = collect_ms;
            out->train_ms = rep.ms;
        }
    });
}

int spr_learned_save(spr_handle* H, const char* path) {
    if (!H || !path) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        try { spr::mlp_save(H->model, path); } catch (const std::exception& e) { throw Fail{SPR_ERR_IO, e.what()}; }
    });
}

int spr_learned_load(spr_handle* H, const char* path) {
    if (!H || !path) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        try { spr::mlp_load(H->model, path); } catch (const std::exception& e) { throw Fail{SPR_ERR_IO, e.what()}; }
        H->mlpdev.release();
        H->lb_dirty = true;
        H->need_prediction = true;
    });
}

int spr_learned_set_margin(spr_handle* H, float log2_margin, int32_t slack) {
    if (!H) { g_error = "null handle"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (!(log2_margin > -30.0f && log2_margin < 30.0f) || slack < 0) throw Fail{SPR_ERR_INVALID, "bad margin"};
        // the capacities change: the budgets and any captured frame (whose kernels hold the old
        // margin as an argument) are stale
        if (H->model.log_margin != log2_margin || H->model.slack != slack) H->lb_dirty = H->need_prediction = true;
        H->model.log_margin = log2_margin;
        H->model.slack = slack;
    });
}

int spr_mlp_get_params(spr_handle* H, float* dst, size_t cap, size_t* n) {
    if (!H || !n) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        const size_t total = (size_t)spr::kMlpParams + 2 * spr::kMlpIn;
        *n = total;
        if (!dst) return;
        if (!H->model.ready) throw Fail{SPR_ERR_STATE, "no model"};
        if (cap < total) throw Fail{SPR_ERR_CAPACITY, "destination too small"};
        std::memcpy(dst, H->model.params.data(), spr::kMlpParams * 4);
        std::memcpy(dst + spr::kMlpParams, H->model.mean.data(), spr::kMlpIn * 4);
        std::memcpy(dst + spr::kMlpParams + spr::kMlpIn, H->model.inv_std.data(), spr::kMlpIn * 4);
    });
}

int spr_mlp_set_params(spr_handle* H, const float* src, size_t n) {
    if (!H || !src) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (n != (size_t)spr::kMlpParams + 2 * spr::kMlpIn) throw Fail{SPR_ERR_INVALID, "expected 5410 floats"};
        H->model.params.assign(src, src + spr::kMlpParams);
        H->model.mean.assign(src + spr::kMlpParams, src + spr::kMlpParams + spr::kMlpIn);
        H->model.inv_std.assign(src + spr::kMlpParams + spr::kMlpIn, src + n);
        H->model.ready = true;
        H->mlpdev.release();
        H->lb_dirty = true;
        H->need_prediction = true;
    });
}

int spr_mlp_predict(spr_handle* H, const float* features, uint32_t n, float* out) {
    if (!H || (!features && n) || (!out && n)) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        SPR_HIP(hipSetDevice(H->device));
        if (!H->mlpdev.valid()) H->mlpdev.upload(H->model, H->stream);
        if (!n) return;
        DBuf<float> f, p;
        f.ensure((size_t)n * spr::kMlpIn);
        p.ensure((size_t)n * spr::kMlpOut);
        SPR_HIP(hipMemcpyAsync(f.p, features, (size_t)n * spr::kMlpIn * 4, hipMemcpyHostToDevice, H->stream));
        SPR_HIP(spr::launch_mlp_infer(f.p, n, H->mlpdev.dev(), p.p, H->stream));
        SPR_HIP(hipMemcpyAsync(out, p.p, (size_t)n * spr::kMlpOut * 4, hipMemcpyDeviceToHost, H->stream));
        SPR_HIP(hipStreamSynchronize(H->stream));
        f.release();
        p.release();
    });
}

int spr_mlp_gradients(spr_handle* H, const float* features, const float* targets, uint32_t n, float* grads,
                      float* loss) {
    if (!H || !features || !targets || !grads || !loss) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        SPR_HIP(hipSetDevice(H->device));
        if (!H->model.ready) throw Fail{SPR_ERR_STATE, "no model"};
        spr::mlp_gradients(H->model, features, targets, (int)n, grads, loss, H->stream);
    });
}

// ---- scenes
int spr_scene_load_rvg(const char* path, spr_scene** out) {
    if (!path || !out) { g_error = "null argument"; return SPR_ERR_INVALID; }
    *out = nullptr;
    try {
        spr_scene* s = new spr_scene();
        try { s->s = spr::load_rvg(path); } catch (...) { delete s; throw; }
        *out = s;
        return SPR_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return SPR_ERR_IO;
    }
}

int spr_scene_generate(int kind, uint64_t seed, uint32_t n_paths, uint32_t canvas, spr_scene** out) {
    if (!out) { g_error = "null argument"; return SPR_ERR_INVALID; }
    *out = nullptr;
    try {
        spr_scene* s = new spr_scene();
        try { s->s = spr::generate(kind, seed, n_paths, canvas); } catch (...) { delete s; throw; }
        *out = s;
        return SPR_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return SPR_ERR_INVALID;
    }
}

int spr_scene_subset(const spr_scene* s, uint32_t first, uint32_t count, spr_scene** out) {
    if (!s || !out) { g_error = "null argument"; return SPR_ERR_INVALID; }
    *out = nullptr;
    try {
        spr_scene* o = new spr_scene();
        try { o->s = spr::subset(s->s, first, count); } catch (...) { delete o; throw; }
        *out = o;
        return SPR_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return SPR_ERR_INVALID;
    }
}

int spr_scene_view_get(const spr_scene* sc, spr_scene_view* v) {
    if (!sc || !v) { g_error = "null argument"; return SPR_ERR_INVALID; }
    const spr::Scene& s = sc->s;
    v->pos_xy = s.pos.data(); v->pos_path_idx = s.pos_path.data(); v->n_points = s.n_points();
    v->curve_pos_map = s.curve_pos.data(); v->curve_type = s.curve_type.data();
    v->curve_path_idx = s.curve_path.data(); v->n_curves = s.n_curves();
    v->fill_rule = s.rule.data(); v->fill_info = s.fill.data(); v->n_paths = s.n_paths();
    v->width = s.width; v->height = s.height;
    v->curve_weight = s.curve_weight.empty() ? nullptr : s.curve_weight.data();
    v->path_paint = s.path_paint.empty() ? nullptr : s.path_paint.data();
    v->paints = reinterpret_cast<const spr_paint*>(s.paints.data());
    v->n_paints = (uint32_t)s.paints.size();
    v->stops = reinterpret_cast<const spr_stop*>(s.stops.data());
    v->n_stops = (uint32_t)s.stops.size();
    return SPR_OK;
}

int spr_scene_create(const spr_scene_view* v, spr_scene** out) {
    if (!v || !out) { g_error = "null argument"; return SPR_ERR_INVALID; }
    *out = nullptr;
    return guarded(nullptr, [&] {
        const uint32_t P = v->n_points, C = v->n_curves, Np = v->n_paths;
        std::string bad = spr::validate(v->pos_xy, v->pos_path_idx, P, v->curve_pos_map, v->curve_type,
                                        v->curve_path_idx, C, v->fill_rule, Np);
        if (bad.empty() && Np && !v->fill_info) bad = "null fill_info";
        if (bad.empty())
            bad = spr::validate_extras(v->curve_type, C, v->curve_weight, Np, v->path_paint,
                                       reinterpret_cast<const spr::Paint*>(v->paints), v->n_paints,
                                       reinterpret_cast<const spr::Stop*>(v->stops), v->n_stops);
        if (!bad.empty()) throw Fail{SPR_ERR_INVALID, bad};
        spr_scene* sc = new spr_scene();
        spr::Scene& s = sc->s;
        s.pos.assign(v->pos_xy, v->pos_xy + 2 * (size_t)P);
        s.pos_path.assign(v->pos_path_idx, v->pos_path_idx + P);
        s.curve_pos.assign(v->curve_pos_map, v->curve_pos_map + C);
        s.curve_type.assign(v->curve_type, v->curve_type + C);
        s.curve_path.assign(v->curve_path_idx, v->curve_path_idx + C);
        s.rule.assign(v->fill_rule, v->fill_rule + Np);
        s.fill.assign(v->fill_info, v->fill_info + Np);
        bool arcs = false;
        for (uint32_t c = 0; c < C; ++c) arcs = arcs || v->curve_type[c] == 0x13u;
        if (arcs) s.curve_weight.assign(v->curve_weight, v->curve_weight + C);
        if (v->path_paint) {
            s.path_paint.assign(v->path_paint, v->path_paint + Np);
            const spr::Paint* g = reinterpret_cast<const spr::Paint*>(v->paints);
            const spr::Stop* t = reinterpret_cast<const spr::Stop*>(v->stops);
            s.paints.assign(g, g + v->n_paints);
            s.stops.assign(t, t + v->n_stops);
        }
        s.width = v->width; s.height = v->height;
        *out = sc;
    });
}

int spr_scene_write_rvg(const spr_scene* s, const char* path) {
    if (!s || !path) { g_error = "null argument"; return SPR_ERR_INVALID; }
    try { spr::write_rvg(s->s, path); return SPR_OK; }
    catch (const std::exception& e) { g_error = e.what(); return SPR_ERR_IO; }
}

void spr_scene_free(spr_scene* s) { delete s; }

int spr_write_png_rgba(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height) {
    if (!path || !rgba) { g_error = "null argument"; return SPR_ERR_INVALID; }
    try {
        spr::png_write_rgba(path, rgba, width, height);
        return SPR_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return SPR_ERR_IO;
    }
}

int spr_write_png(spr_handle* H, const char* path) {
    if (!H || !path) { g_error = "null argument"; return SPR_ERR_INVALID; }
    return guarded(H, [&] {
        if (!H->frame_done) throw Fail{SPR_ERR_STATE, "no frame rendered"};
        if (H->batch) throw Fail{SPR_ERR_STATE, "a batch has one image per document: use spr_batch_image"};
        std::vector<uint8_t> img((size_t)4 * H->w * H->h);
        SPR_HIP(hipMemcpyAsync(img.data(), H->image.p, img.size(), hipMemcpyDeviceToHost, H->stream));
        SPR_HIP(hipStreamSynchronize(H->stream));
        try {
            spr::png_write_rgba(path, img.data(), H->w, H->h);
        } catch (const std::exception& e) {
            throw Fail{SPR_ERR_IO, e.what()};
        }
    });
}

}  // extern "C"

