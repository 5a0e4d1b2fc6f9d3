/* spr.h -- C ABI of the MI355X-native scanline path rasteriser (libspr.so).
 *
 * This is the drop-in boundary for the hot path of VkScanlinePR.  Every entry point below names
 * the reference interface it replaces (paths relative to the reference root, src/ =
 * VkScanlinePR/src/).  All functions return 0 on success and a negative spr_status otherwise;
 * spr_last_error() gives the message.  No C++ exception crosses this boundary.
 *
 * Threading: one handle is used by one host thread at a time; different handles are
 * independent and each owns its own HIP stream and device memory.
 * Ownership: input arrays stay owned by the caller and are copied during the call; the handle
 * owns every device buffer.  Caller data is never modified (the reference's loadVG mutates the
 * container's colours, src/core/scanline/scanline_rasterizer.cpp:107-118; this ABI takes the
 * already-packed RGBA8 fill_info instead).
 */
#ifndef SPR_H_
#define SPR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spr_handle spr_handle;
typedef struct spr_scene spr_scene;

typedef enum spr_status {
    SPR_OK = 0,
    SPR_ERR_INVALID = -1,  /* bad argument or malformed input */
    SPR_ERR_DEVICE = -2,   /* HIP runtime error */
    SPR_ERR_STATE = -3,    /* call out of order (e.g. render before load) */
    SPR_ERR_IO = -4,       /* file could not be read or parsed */
    SPR_ERR_CAPACITY = -5  /* a size exceeds a fixed limit of the library */
} spr_status;

/* Curve types and fill rules, src/core/vg/vg_container.h:9-19. */
enum { SPR_CURVE_LINE = 2, SPR_CURVE_CUBIC = 4 };
enum { SPR_FILL_NONZERO = 0, SPR_FILL_EVEN_ODD = 1 };

/* Composite modes.  OVERWRITE is the reference's (blending disabled,
 * src/core/scanline/scanline_rasterizer.cpp:893-895); SRC_OVER is the back-to-front alpha
 * composite d = s*a + d*(1-a) in path order (new behaviour, no reference counterpart). */
enum { SPR_COMPOSE_OVERWRITE = 0, SPR_COMPOSE_SRC_OVER = 1 };

/* Allocation modes.  PLAIN is the reference's count-then-write form (count pass, host readback
 * of the intersection total, write pass; scanline_rasterizer.cpp:317-395).  LEARNED removes the
 * host readback: the frame's buffers are sized beforehand from the per-path counts predicted by
 * the built-in MLP (read back once per scene or model), the exact counts are scanned on the device,
 * and a frame whose exact total exceeds that allocation is culled on the device and rendered with
 * the plain method instead, so the output is identical. */
enum { SPR_MODE_PLAIN = 0, SPR_MODE_LEARNED = 1 };

typedef struct spr_opts {
    uint32_t compose;   /* SPR_COMPOSE_* */
    uint32_t mode;      /* SPR_MODE_* */
    int32_t device;     /* HIP device ordinal, -1 = current */
    uint32_t debug;     /* 1 = keep every intermediate for spr_debug_export (slower) */
    uint32_t aa;        /* 0 = the reference's binary coverage of the 2x2 cells; 16 = anti-aliased:
                         * every boundary cell gets a 16-samples-per-pixel coverage mask (the
                         * "comb-like multisampling" stencil mask the reference leaves TODO,
                         * gen_fragment.comp:202, scanlinepr.vert:48-50) and each path's colour is
                         * weighted by its pixel coverage.  Needs SPR_COMPOSE_SRC_OVER and
                         * SPR_MODE_PLAIN. */
    uint32_t reserved[3];
} spr_opts;

/* Per-frame counters and stage times, filled by spr_get_stats. */
typedef struct spr_stats {
    int64_t n_points, n_curves, n_paths;
    int64_t n_intersections;   /* N: the reference's n_fragments */
    int64_t n_cells;           /* merged cell records */
    int64_t n_spans;           /* span records */
    int64_t n_overflow_paths;  /* LEARNED: paths whose exact count exceeded their predicted capacity */
    int64_t alloc_intersections; /* LEARNED: predicted capacity summed over paths */
    double frame_ms;           /* device time of the last frame (HIP events) */
    double stage_ms[16];       /* per-stage device time when timing is enabled */
    int64_t learned_fallbacks; /* LEARNED: frames whose exact total exceeded the allocation and were
                                  rendered with the plain method (the allocation then grows) */
    /* Which kernel routes the last frame took (each route has its own kernel or mode): */
    int64_t n_wave_chains;     /* crossing chains solved by a whole wave each (896 or more crossings,
                                  and each launch's costliest classes while they hold at most 512) */
    int64_t n_heavy_pieces;    /* monotone pieces merged by a 256-thread workgroup each (more than
                                  1024 entries, or whose last segment waits on a chain of 256 or
                                  more crossings of the next piece) */
    int64_t n_long_bands;      /* sort bands of more than 512 words (one wave at 16 or 32 words per
                                  lane; 256-thread LDS sorts in frames that export a stage) */
    int64_t n_big_bands;       /* sort bands of more than 2048 words (global merge-path sort) */
    int64_t graph_captures;    /* LEARNED and batches: frames captured as a HIP graph so far (a new
                                  scene, model, margin, tuning or budget needs a new capture; a
                                  still frame and a frame after a new MVP have a graph each, and
                                  the latter reads the MVP at each replay) */
} spr_stats;

/* Debug stage identifiers for spr_debug_export / spr_debug_inject. */
enum {
    SPR_STAGE_TPOS = 0,        /* float2 x P: transformed positions */
    SPR_STAGE_VISIBLE = 1,     /* int32 x Np: per-path outcode mask */
    SPR_STAGE_CUTS = 2,        /* float x 5C: cut cache (4 cuts + n_cuts bits) */
    SPR_STAGE_COUNTS = 3,      /* int32 x (C+1): exclusive-scanned per-curve counts */
    SPR_STAGE_INTER = 4,       /* int2 x N: (curve, t bits) intersections (debug) */
    SPR_STAGE_KEY = 5,         /* int32 x N: cell keys in intersection order (debug, anti-aliased
                                  or injected frames, or canvases wider than 32,764) */
    SPR_STAGE_WN = 6,          /* int32 x N: winding change in intersection order (as KEY) */
    SPR_STAGE_SEG = 7,         /* int32 x (Np+1): per-path segment starts */
    SPR_STAGE_SKEY = 8,        /* int32 x N: keys sorted inside each path segment (debug; without
                                  it, canvases up to 32764 wide fuse mark + emit into the sort and
                                  keep no sorted keys) */
    SPR_STAGE_SIDX = 9,        /* int32 x N: sorted intersection indices (debug) */
    SPR_STAGE_WINDING = 10,    /* int32 x N: exclusive-scanned winding numbers, sorted order
                                  (debug; as SPR_STAGE_SKEY: a frame without debug on a canvas
                                  wider than 32764 keeps them, but entries of equal key are then
                                  in no set order, so only each run's first number is the
                                  reference's) */
    SPR_STAGE_FLAGS = 11,      /* int32 x 2N: cell flags, then span flags (debug) */
    SPR_STAGE_RECORDS = 12,    /* int4 x R: (yx, width, fill, frag_index) output records (debug).
                                  Injected, the following frames draw these records instead of
                                  rendering the scene (the reference's MOCK_DATA replay of a
                                  record file, scanline_rasterizer.cpp:261-281, :1057-1109) */
    SPR_STAGE_IMAGE = 13,      /* uint8 x 4WH: RGBA8 image, top row first */
    SPR_STAGE_FEATURES = 14,   /* float x 16Np: per-path predictor features (last learned frame) */
    SPR_STAGE_PRED = 15,       /* float x 2Np: predicted log2(1 + intersections), log2(1 + records) */
    SPR_STAGE_PATH_COUNTS = 16,/* int32 x 2Np: exact (intersections, records) per path, last frame */
    SPR_STAGE_COVERAGE = 17,   /* uint64 x R: per record, the 64-sample coverage mask (aa frames;
                                  bit (2 py + px) * 16 + j = sample j of pixel (px, py) of the 2x2 cell,
                                  sample j at ((7j mod 16) + 1/2)/16, (j + 1/2)/16; spans all ones) */
    SPR_STAGE_COUNT_ = 18
};

/* --- lifecycle: VGRasterizer::initialize(void* window, uint32_t w, uint32_t h),
 *     src/core/rasterizer.h:11, implemented at src/core/scanline/scanline_rasterizer.cpp:42-55.
 *     No window: the target is an RGBA8 image of w x h in device memory. */
int spr_create(uint32_t width, uint32_t height, const spr_opts* opts, spr_handle** out);
/* ~ScanlineVGRasterizer, src/core/scanline/scanline_rasterizer.h:45-47 */
int spr_destroy(spr_handle* h);
const char* spr_last_error(const spr_handle* h);
/* Message of the last failure of a call that had no handle (spr_create, spr_scene_*). */
const char* spr_global_error(void);

/* --- input: VGRasterizer::loadVG(shared_ptr<VGContainer>), src/core/rasterizer.h:15,
 *     implemented at src/core/scanline/scanline_rasterizer.cpp:67-189.  Takes exactly the
 *     flattened SoA that loadVG builds (src/core/scanline/vk_vg_data.h:14-35):
 *     per point pos_xy[2P] and pos_path_idx[P]; per curve curve_pos_map[C] (first point),
 *     curve_type[C] (2 = line, 4 = cubic) and curve_path_idx[C] (non-decreasing); per path
 *     fill_rule[Np] and fill_info[Np] (RGBA8, R in the low byte). */
int spr_load_paths(spr_handle* h, const float* pos_xy, const uint32_t* pos_path_idx,
                   uint32_t n_points, const uint32_t* curve_pos_map, const uint32_t* curve_type,
                   const uint32_t* curve_path_idx, uint32_t n_curves, const uint32_t* fill_rule,
                   const uint32_t* fill_info, uint32_t n_paths);
/* RVG::load + loadVG in one call: src/core/vg/rvg.cpp:9-26, src/app/vg_app.cpp:172-192. */
int spr_load_rvg(spr_handle* h, const char* path);
/* Same, from a scene object (parsed or generated by the spr_scene_* calls below). */
int spr_load_scene(spr_handle* h, const spr_scene* s);

/* --- VGRasterizer::setMVP(const glm::mat4&), src/core/rasterizer.h:17, stored as the
 *     TransPosIn rows m0..m3 (src/core/scanline/compute_ubo.h:9-13,
 *     scanline_rasterizer.cpp:191-197): m_rows[4*r + c], x' = dot((x,y,0,1), row 0). */
int spr_set_mvp(spr_handle* h, const float m_rows[16]);

/* --- VGRasterizer::render(), src/core/rasterizer.h:13 -> drawFrame,
 *     src/core/scanline/scanline_rasterizer.cpp:282-696.  Renders one frame into the handle's
 *     device image; copies it to rgba_out (4*w*h bytes, top row first) when not NULL. */
int spr_render_frame(spr_handle* h, uint8_t* rgba_out_or_null);
/* Device pointer of the handle's RGBA8 image (valid until the next resize or destroy). */
int spr_image_device_ptr(spr_handle* h, void** dptr);
/* Blocks until the handle's stream is idle. */
int spr_synchronize(spr_handle* h);
/* Enables per-stage HIP-event timing (adds event records between stages). */
int spr_set_stage_timing(spr_handle* h, int enable);
int spr_get_stats(const spr_handle* h, spr_stats* out);

/* --- the reference's standalone exclusive scan (shaders/common/naive_scan.comp:27-73, called at
 *     src/core/scanline/scanline_rasterizer.cpp:337-358, :478-506 and :544-573) on device memory:
 *     d_out[i] = d_in[0] + ... + d_in[i-1] for i <= n, so d_out holds n + 1 int32 and its last
 *     value is the total (d_out may equal d_in: the scan then runs in place, as the reference's
 *     does).  Sums wrap modulo 2^32, as the reference's int arithmetic does.  n <= 2^31 - 2.
 *     Runs on `stream` (a hipStream_t; NULL = the handle's stream) and returns without waiting;
 *     the handle's scan workspace grows on demand, so calls on one handle must not overlap. */
int spr_scan_i32(spr_handle* h, const int32_t* d_in, int32_t* d_out, uint64_t n, void* stream);
/* --- the same scan, allocating nothing.  No counterpart in the reference: it is for callers
 *     that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 *     forbid hipMalloc).  Its workspace (tile statuses and the tile counter) is a static device
 *     array of 512 KiB per GPU, held until the process ends; so n <= 2^28 (268,435,456), and
 *     larger n is rejected with SPR_ERR_CAPACITY.  Runs on the calling thread's current device,
 *     on `stream` (a hipStream_t; NULL = that device's null stream).  Every call on one GPU
 *     shares the array: calls that overlap on one GPU (on different streams) race; calls on
 *     one stream are ordered.  Errors are read with spr_global_error(). */
int spr_scan_i32_noalloc(const int32_t* d_in, int32_t* d_out, uint64_t n, void* stream);

/* --- debug access to intermediate buffers (the reference's drawDebug dumps,
 *     src/core/scanline/scanline_rasterizer.cpp:698-802).  export copies the last frame's stage
 *     into dst (cap bytes) and stores the full size in *nbytes; with dst == NULL only the size
 *     is returned.  inject(SPR_STAGE_INTER) makes the following frames use the given
 *     intersections (int2 x N, in curve order) instead of generating them, so the integer
 *     stages can be compared on identical intersection inputs; inject(SPR_STAGE_RECORDS) makes
 *     them draw the given records (int4 x R, at most 2^22) on the handle's canvas, no scene
 *     needed; inject with size 0 clears it. */
int spr_debug_export(spr_handle* h, int stage, void* dst, size_t cap, size_t* nbytes);
int spr_debug_inject(spr_handle* h, int stage, const void* src, size_t nbytes);

/* --- presentation: the reference presents its frames in a GLFW swapchain
 *     (src/core/scanline/scanline_rasterizer.cpp:611-666); headless, a frame goes to a PNG file
 *     (8-bit RGBA, top row first).  spr_write_png writes the handle's last frame (a batch's
 *     documents are read with spr_batch_image), spr_write_png_rgba any image in host memory. */
int spr_write_png(spr_handle* h, const char* path);
int spr_write_png_rgba(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height);

/* --- scenes: the VGContainer of src/core/vg/vg_container.h:21-87 plus the flattening of
 *     scanline_rasterizer.cpp:67-134, and the synthetic workload generators.
 *     Curve types are the reference's CurveType values (vg_container.h:9-15): 2 line (2 points),
 *     3 quadratic (3 points), 4 cubic (4 points), 0x13 arc (3 points: start, control point,
 *     end, a rational quadratic whose control point has weight curve_weight[c] > 0).  The
 *     reference declares QUADRIC and ARC but leaves them TODO in its shaders; here they are
 *     rasterised natively.  A path's colour is fill_info unless path_paint names a gradient. */
#define SPR_PAINT_LINEAR 1u
#define SPR_PAINT_RADIAL 2u
#define SPR_SPREAD_PAD 0u
#define SPR_SPREAD_REPEAT 1u
#define SPR_SPREAD_REFLECT 2u
#define SPR_SPREAD_TRANSPARENT 3u
#define SPR_SOLID_PAINT 0xFFFFFFFFu
/* A gradient.  At a pixel centre p (canvas coordinates), the scene point s = MVP^-1 p, the
 * gradient point g = xform * s; linear: t = ((g - p1) . d) / (d . d) with d = p2 - p1 (t = 1
 * when d = 0); radial: t = the largest root of |g - f - t (c - f)| = t r (the focus f strictly
 * inside the circle).  The spread maps t to [0, 1] (pad clamps, repeat wraps, reflect mirrors,
 * transparent gives a transparent colour outside), the stops are interpolated linearly in
 * non-premultiplied RGBA, and the colour is packed like fill_info with alpha times opacity. */
typedef struct spr_paint {
    uint32_t kind;                  /* SPR_PAINT_LINEAR or SPR_PAINT_RADIAL */
    uint32_t spread;                /* SPR_SPREAD_* */
    uint32_t first_stop, n_stops;   /* stops [first_stop, first_stop + n_stops), n_stops >= 1 */
    float opacity;
    float xform[6];                 /* scene -> gradient space: gx = x0 x + x1 y + x2, gy = x3 x + x4 y + x5 */
    float geom[5];                  /* linear: x1, y1, x2, y2, unused; radial: cx, cy, fx, fy, r */
} spr_paint;
typedef struct spr_stop {
    float offset, r, g, b, a;       /* offsets non-decreasing within a paint */
} spr_stop;

typedef struct spr_scene_view {
    const float* pos_xy; const uint32_t* pos_path_idx; uint32_t n_points;
    const uint32_t* curve_pos_map; const uint32_t* curve_type; const uint32_t* curve_path_idx;
    uint32_t n_curves;
    const uint32_t* fill_rule; const uint32_t* fill_info; uint32_t n_paths;
    uint32_t width, height;  /* suggested canvas (the RVG viewport, or the generator's) */
    const float* curve_weight;      /* per curve, read for arcs only; NULL when there is no arc */
    const uint32_t* path_paint;     /* per path: SPR_SOLID_PAINT or a paint index; NULL = all solid */
    const spr_paint* paints; uint32_t n_paints;
    const spr_stop* stops; uint32_t n_stops;
} spr_scene_view;

/* RVG::load, src/core/vg/rvg.cpp:9-255 (M/L/C/Z, solid rgb/rgba paint, nzfill/ofill), plus the
 * format's parts that parser skips: Q and A segments (an A segment of weight below 1/2 is split
 * into halves, so stored arcs have 1/2 <= w), dyn_affine path, paint and scene transforms, and
 * linear and radial gradients. */
int spr_scene_load_rvg(const char* path, spr_scene** out);
/* A scene from arrays (copied and validated), curve_weight and paints included. */
int spr_scene_create(const spr_scene_view* view, spr_scene** out);
/* Synthetic workloads.  kind: 0 = random cubic scribbles (64 paths, 1024^2 by default),
 * 1 = glyph page (quadratic glyphs elevated to cubics, even-odd), 2 = illustration (mixed
 * rules, alpha, per-path affine).  n_paths / canvas = 0 select the defaults. */
int spr_scene_generate(int kind, uint64_t seed, uint32_t n_paths, uint32_t canvas, spr_scene** out);
/* Paths [first, first+count) of a scene as a new scene (path indices renumbered). */
int spr_scene_subset(const spr_scene* s, uint32_t first, uint32_t count, spr_scene** out);
int spr_scene_view_get(const spr_scene* s, spr_scene_view* out);
int spr_scene_write_rvg(const spr_scene* s, const char* path);
void spr_scene_free(spr_scene* s);

/* --- learned single-pass allocation.  No reference counterpart: north of the reference's
 *     count pass (make_intersection_0.comp) this predicts each path's intersection and record
 *     counts with a 16-64-64-2 MLP (BF16 matrix-core inference, FP32 training kernels).  The
 *     network is fitted on the exact per-path counts of a training scene, which is rendered on
 *     this handle with the plain method chunk by chunk; afterwards no scene is loaded. */
typedef struct spr_train_opts {
    uint32_t epochs;          /* 0 = 60 */
    uint32_t batch;           /* 0 = 512; epochs are added until at least 4000 steps ran */
    float lr;                 /* Adam step, 0 = 3e-3 */
    uint32_t chunk_paths;     /* training paths rendered per frame, 0 = 5000 */
    uint64_t seed;            /* initialisation and shuffling */
    float overflow_quantile;  /* capacity margin keeps this fraction of paths in their arena, 0 = 0.95 */
    uint32_t reserved[3];
} spr_train_opts;
typedef struct spr_train_report {
    int64_t samples, steps;
    double first_loss, train_loss, val_loss; /* mean squared error of log2(1 + count) */
    double log_margin;                       /* chosen capacity margin (log2) */
    double collect_ms, train_ms;
} spr_train_report;
int spr_learned_train(spr_handle* h, const spr_scene* training_scene, const spr_train_opts* opts,
                      spr_train_report* out);
int spr_learned_save(spr_handle* h, const char* path);
int spr_learned_load(spr_handle* h, const char* path);
/* capacity = 2^(prediction + log2_margin) - 1 + slack per path */
int spr_learned_set_margin(spr_handle* h, float log2_margin, int32_t slack);
/* The model as floats: 5378 parameters (W1[16][64], b1, W2[64][64], b2, W3[64][2], b3), then
 * the feature mean[16] and 1/std[16] (5410 floats). */
int spr_mlp_get_params(spr_handle* h, float* dst, size_t cap, size_t* n);
int spr_mlp_set_params(spr_handle* h, const float* src, size_t n);
/* BF16 matrix-core inference on n raw feature rows (16 floats each) -> 2 floats per row. */
int spr_mlp_predict(spr_handle* h, const float* features, uint32_t n, float* out);
/* FP32 forward + backward of the training kernels on one batch: gradients of the mean squared
 * error in parameter order (5378 floats) and the loss. */
int spr_mlp_gradients(spr_handle* h, const float* features, const float* targets, uint32_t n, float* grads,
                      float* loss);

/* --- batches of documents.  No reference counterpart (the reference renders one scene).  Many
 *     independent documents of different sizes and canvases are loaded into one handle and
 *     rendered by one frame without host synchronisation: every stage runs once over all
 *     documents, each document keeps its own canvas and winding numbers, and once the scene is
 *     unchanged the frame is replayed as one captured HIP graph.  Uses the handle's learned
 *     model for the per-path arenas when one is loaded; otherwise every path takes an exact block
 *     of the overflow pool.  spr_render_frame(h, NULL) renders the batch; the images are read
 *     with spr_batch_image (4*w*h bytes each, top row first). */
int spr_load_batch(spr_handle* h, const spr_scene* const* docs, const uint32_t* widths, const uint32_t* heights,
                   uint32_t n_docs);
int spr_batch_image(spr_handle* h, uint32_t doc, uint8_t* rgba_out, size_t cap);
/* Project-written autotuner: times the loaded scene's frame (captured-graph replays for a batch or
 * a learned scene, eager frames for a plain one; `reps` frames per candidate) and keeps the
 * fastest launch knobs, tuned one at a time from the defaults (coordinate descent):
 *  - chain solve and long-piece merge: persistent blocks per CU (8, 4, 16 each);
 *  - the tile of the large scans (the reference's fixed BATCH_SIZE = 1024 of
 *    shaders/common/naive_scan.comp:2): 8 int4 x 1024 threads, 8 x 512, 4 x 1024, 16 x 512;
 *  - the sort's band window, which sets the band (sort tile) size (the reference's fixed
 *    BATCH_SIZE of shaders/common/naive_seg_sort_pairs.comp:3): 512 or 1024 entries.
 * The knobs stay with the handle for later frames. */
typedef struct spr_tune_report {
    int32_t chain_blocks_per_cu, merge_blocks_per_cu, n_tried, reserved;
    double default_ms, best_ms;
    int32_t scan_shape;        /* 0: 8 x 1024, 1: 8 x 512, 2: 4 x 1024, 3: 16 x 512 */
    int32_t band_window_log;   /* bands start in 2^band_window_log-entry windows (9 or 10) */
} spr_tune_report;
int spr_autotune(spr_handle* h, uint32_t reps, spr_tune_report* out);
/* Sets the launch knobs of spr_autotune explicitly (a tuned configuration restored; the timing
 * fields are ignored).  Every value is checked against the ones the autotuner tries. */
int spr_set_tuning(spr_handle* h, const spr_tune_report* knobs);

#ifdef __cplusplus
}
#endif

#endif /* SPR_H_ */
