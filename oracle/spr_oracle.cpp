// spr_oracle.cpp -- CPU restatement of the VkScanlinePR scanline path-rasterisation pipeline.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product path is compared
// against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
// product library never links, loads or calls it, and it contains no code of the product.
//
// It restates, stage by stage and in the reference's operation order, the compute shaders of
// VkScanlinePR (paths relative to the reference root; `shaders/` = workdir/shaders/,
// `src/` = VkScanlinePR/src/):
//   transform_pos        shaders/scanline/compute/transform_pos.comp:33-85
//   monotonise + count   shaders/scanline/compute/make_intersection_0.comp:226-411
//   exclusive scan       shaders/common/naive_scan.comp:27-73 (semantics: exclusive prefix sum)
//   write intersections  shaders/scanline/compute/make_intersection_1.comp:217-447
//   gen_fragment         shaders/scanline/compute/gen_fragment.comp:90-245
//   segmented sort       shaders/common/naive_seg_sort_pairs.comp:26-97 (semantics: sort by
//                        signed (key, value) inside each path segment)
//   shuffle_fragment     shaders/scanline/compute/shuffle_fragment.comp:17-26
//   mark                 shaders/scanline/compute/mark_merged_fragment_and_span.comp:48-120
//   gen_merged           shaders/scanline/compute/gen_merged_fragment_and_span.comp:33-103
//   line-list draw       shaders/scanline/surface/scanlinepr.vert:18-50, scanlinepr.frag:7-14,
//                        pipeline state src/core/scanline/scanline_rasterizer.cpp:881-944
//   host sequencing      src/core/scanline/scanline_rasterizer.cpp:282-608
//
// Floating-point rules (build with -ffp-contract=off, no fast-math): every LERP is a+t*(b-a)
// in that order, no fused multiply-add, IEEE division and square root.  GLSL min/max are taken
// literally from the GLSL spec (max(x,y) = x<y ? y : x), which keeps the sign of a zero first
// argument.  Float->int conversions saturate (NaN -> 0) and integer products wrap, which fixes
// the value of cases the GLSL spec leaves undefined.
//
// Deliberate deviations from the reference, each forced by undefined reference behaviour:
//   * path_visible is zeroed at the start of every frame (the reference never clears it).
//   * the per-path outcode OR is an exact OR (the reference's read-modify-write can lose bits).
//   * cut-cache slots at or beyond n_cuts are written as 0 (the reference leaves LDS garbage).
//   * scans define only indices 0..n (the reference's last batch writes up to n+1022).
//   * the draw uses the real canvas size instead of the hard-coded 1200x1024 viewport.
// The reference quirk of the 4-cut insertion sort (the 4th cut is compared with itself and so
// never moves, make_intersection_0.comp:338-341) is reproduced on purpose.
//
// Beyond the reference, whose shaders leave the QUADRIC and ARC cases TODO and whose draw has
// solid colours only: quadratic and arc (rational quadratic) curves, evaluated, monotonised and
// bisected as defined at eval1 / monotone_cuts / solve_crossing, and gradient paints in the
// draw (paint_rgba).  These follow the definitions documented in include/spr.h; the tests pin
// them to closed forms (points on circles, analytic extrema and crossings, gradient values).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int FRAG = 2;  // FRAG_SIZE, make_intersection_0.comp:3
constexpr int CUBIC_ITERS = 24;  // CUBIC_ITERATION_NUMBER, make_intersection_1.comp:7
constexpr uint32_t T_LINE = 0x02, T_CUBIC = 0x04;  // make_intersection_0.comp:22-25
// QUADRIC and ARC (make_intersection_0.comp:23-25; their shader cases are TODO in the reference)
constexpr uint32_t T_QUAD = 0x03, T_ARC = 0x13;
constexpr int32_t KEY_SENTINEL = (int32_t)0xFFFEFFFE;  // gen_fragment.comp:111

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline float gl_max(float x, float y) { return x < y ? y : x; }
inline float gl_min(float x, float y) { return y < x ? y : x; }
inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
// float2int_rd, make_intersection_1.comp:76-79 / gen_fragment.comp:54-57
inline int32_t float2int_rd(float x) { return x >= 0.0f ? f2i(x) : f2i(x - 1.0f); }
inline float lerp(float a, float b, float t) { return a + t * (b - a); }  // LERP, :10

// make_intersection_0.comp:14-20
inline bool path_visible(int32_t m) {
    bool inv = ((m & 0x11111000) == 0) || ((m & 0x01101011) == 0) || ((m & 0x00011111) == 0) ||
               ((m & 0x11010110) == 0);
    return !inv;
}

// Curve evaluation along one axis: interpolateGeneralCurve, make_intersection_1.comp:117-148
// (the count pass's variant, make_intersection_0.comp:131-165, agrees for LINE and CUBIC).
// QUADRIC: de Casteljau on c[0..2].  ARC, a rational quadratic held as c = (x0, w*x1, x2, w):
// de Casteljau on the homogeneous values and on the weights (1, w, 1), then their quotient.
inline float eval1(uint32_t type, float t, const float* c) {
    if (type == T_LINE) return lerp(c[0], c[1], t);
    if (type == T_QUAD || type == T_ARC) {
        float a = lerp(c[0], c[1], t), b = lerp(c[1], c[2], t);
        float num = lerp(a, b, t);
        if (type == T_QUAD) return num;
        float wa = lerp(1.0f, c[3], t), wb = lerp(c[3], 1.0f, t);
        float den = lerp(wa, wb, t);
        return num / den;
    }
    float q0 = lerp(c[0], c[1], t), q1 = lerp(c[1], c[2], t), q2 = lerp(c[2], c[3], t);
    float l0 = lerp(q0, q1, t), l1 = lerp(q1, q2, t);
    return lerp(l0, l1, t);
}

// solveQuadEquation, make_intersection_0.comp:76-129 (citardauq form)
inline void solve_quad(float a, float b, float c, float& r0, float& r1) {
    if (a == 0.0f) {
        float x = -c / b;
        r0 = x; r1 = x;
        return;
    }
    float A = a, B = b * 0.5f, C = c;
    float R = B * B - A * C;
    if (R > 0.0f) {
        float SR = std::sqrt(R);
        if (B > 0.0f) {
            float TB = B + SR;
            r0 = -C / TB; r1 = -TB / A;
        } else {
            float TB = -B + SR;
            r0 = TB / A; r1 = C / TB;
        }
    } else {
        r0 = 0.0f; r1 = 0.0f;
    }
}

// udpate_cut_range, make_intersection_0.comp:186-221 (identical in make_intersection_1.comp:174-213)
inline void cut_range(int w, int h, int& xb, int& xe, int& yb, int& ye, int& nx, int& ny) {
    const int xmin = 0, xmax = (w & ~1) + FRAG, ymin = 0, ymax = (h & ~1) + FRAG;
    if ((xb < xmin && xe < xmin) || (xb > xmax && xe > xmax) || (xb > xe)) {
        nx = 0;
    } else {
        xb = std::min(std::max(xb, xmin), xmax);
        xe = std::min(std::max(xe, xmin), xmax);
        nx = std::max((xe - xb) / FRAG + 1, 0);
    }
    if ((yb < ymin && ye < ymin) || (yb > ymax && ye > ymax) || (yb > ye)) {
        ny = 0;
    } else {
        yb = std::min(std::max(yb, ymin), ymax);
        ye = std::min(std::max(ye, ymin), ymax);
        ny = std::max((ye - yb) / FRAG + 1, 0);
    }
}

// get_xy_begin_end (count pass, floor based), make_intersection_0.comp:167-183
inline void begin_end_floor(float p0x, float p0y, float p1x, float p1y, int& xb, int& xe, int& yb,
                            int& ye) {
    auto g = [](float v) { return f2i(std::floor(v / (float)FRAG) * (float)FRAG); };
    if (p0x <= p1x) { xb = wadd(g(p0x), FRAG); xe = g(p1x); }
    else { xb = wadd(g(p1x), FRAG); xe = g(p0x); }
    if (p0y <= p1y) { yb = wadd(g(p0y), FRAG); ye = g(p1y); }
    else { yb = wadd(g(p1y), FRAG); ye = g(p0y); }
}

// get_xy_begin_end_delta (write pass, float2int_rd based), make_intersection_1.comp:150-170
inline void begin_end_rd(float p0x, float p0y, float p1x, float p1y, int& xb, int& xe, int& yb,
                         int& ye, float& dx, float& dy) {
    auto g = [](float v) { return wmul(float2int_rd(v / (float)FRAG), FRAG); };
    if (p0x <= p1x) { xb = wadd(g(p0x), FRAG); xe = g(p1x); dx = (float)FRAG; }
    else { xb = wadd(g(p1x), FRAG); xe = g(p0x); dx = -(float)FRAG; }
    if (p0y <= p1y) { yb = wadd(g(p0y), FRAG); ye = g(p1y); dy = (float)FRAG; }
    else { yb = wadd(g(p1y), FRAG); ye = g(p0y); dy = -(float)FRAG; }
}

struct Curve {
    uint32_t type;
    float x[4], y[4];
};

// An arc's control point is stored in Euclidean form with its weight beside; evaluation takes
// the homogeneous form (w*x1, w*y1) and the weight in slot 3.
inline void to_eval_form(Curve& cv, float w) {
    if (cv.type != T_ARC) return;
    cv.x[1] = w * cv.x[1]; cv.y[1] = w * cv.y[1];
    cv.x[3] = w; cv.y[3] = w;
}

inline Curve load_curve(const float* tpos, const uint32_t* cpos, const uint32_t* ctype, const float* cweight,
                        uint32_t c) {
    Curve cv;
    cv.type = ctype[c];
    uint32_t np = cv.type & 7;  // make_intersection_0.comp:251-257
    uint32_t base = cpos[c];
    for (uint32_t i = 0; i < 4; ++i) {
        if (i < np) { cv.x[i] = tpos[2 * (base + i)]; cv.y[i] = tpos[2 * (base + i) + 1]; }
        else { cv.x[i] = 0.0f; cv.y[i] = 0.0f; }
    }
    to_eval_form(cv, cv.type == T_ARC ? cweight[c] : 1.0f);
    return cv;
}

// Monotone cut points of one curve, before the t=1 append: make_intersection_0.comp:259-372.
// Returns n_cuts; q[0..3] receive the cache contents (slots >= n_cuts are 0).
inline uint32_t monotone_cuts(const Curve& cv, bool visible, float q[4]) {
    float t1q[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t n = 0;
    float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f, q3 = 0.0f;
    if (visible) {
        if (cv.type == T_CUBIC) {
            for (int axis = 0; axis < 2; ++axis) {
                const float* p = axis == 0 ? cv.x : cv.y;
                float x0 = p[0], x1 = p[1], x2 = p[2], x3 = p[3];
                float r0 = 0.0f, r1 = 0.0f;
                solve_quad(3.0f * (x1 - x2) + (x3 - x0), 2.0f * ((x0 - x1) + (x2 - x1)), x1 - x0, r0, r1);
                if (r0 > 0.0f && r0 < 1.0f) t1q[n++] = r0;
                if (r1 > 0.0f && r1 < 1.0f && r1 != r0) t1q[n++] = r1;
            }
        } else if (cv.type == T_QUAD || cv.type == T_ARC) {
            // stationary points per axis.  QUADRIC: B'(t)/2 = (x0 - 2 x1 + x2) t + (x1 - x0).
            // ARC: with homogeneous values (a0, a1, a2) and weights (1, w, 1), the numerator of
            // x'(t) is 2 [e0 (1-t)^2 + e1 t (1-t) + e2 t^2], e0 = a1 - a0 w, e1 = a2 - a0,
            // e2 = a2 w - a1; in powers of t: (e0 - e1 + e2) t^2 + (e1 - 2 e0) t + e0.
            for (int axis = 0; axis < 2; ++axis) {
                const float* p = axis == 0 ? cv.x : cv.y;
                float A, B, Cc;
                if (cv.type == T_QUAD) {
                    A = 0.0f;
                    B = (p[0] - p[1]) + (p[2] - p[1]);
                    Cc = p[1] - p[0];
                } else {
                    float w = p[3];
                    float e0 = p[1] - p[0] * w, e1 = p[2] - p[0], e2 = p[2] * w - p[1];
                    A = (e0 - e1) + e2;
                    B = e1 - 2.0f * e0;
                    Cc = e0;
                }
                float r0 = 0.0f, r1 = 0.0f;
                solve_quad(A, B, Cc, r0, r1);
                if (r0 > 0.0f && r0 < 1.0f) t1q[n++] = r0;
                if (r1 > 0.0f && r1 < 1.0f && r1 != r0) t1q[n++] = r1;
            }
            // conic cuts in ascending order (no reference ordering to follow)
            std::sort(t1q, t1q + n);
            q[0] = t1q[0]; q[1] = t1q[1]; q[2] = t1q[2]; q[3] = t1q[3];
            return n;
        }
        q0 = t1q[0]; q1 = t1q[1]; q2 = t1q[2]; q3 = t1q[3];
        if (n >= 2) {  // :315-322
            float t1 = q1, t0 = q0;
            if (t1 < t0) { q1 = t0; q0 = t1; }
        }
        if (n >= 3) {  // :323-337
            float t2 = q2, t1 = q1;
            if (t2 < t1) {
                q2 = t1;
                float t0 = q0;
                if (t2 < t0) { q1 = t0; q0 = t2; } else { q1 = t2; }
            }
        }
        if (n >= 4) {  // :338-359, reproduces the t2 = q3 typo: the comparison is always false
            float t3 = q3, t2 = q3;
            if (t3 < t2) {
                q3 = t2;
                float t1 = q1;
                if (t3 < t1) {
                    q2 = t1;
                    float t0 = q0;
                    if (t3 < t0) { q1 = t0; q0 = t3; } else { q1 = t3; }
                } else {
                    q2 = t3;
                }
            }
        }
    }
    q[0] = q0; q[1] = q1; q[2] = q2; q[3] = q3;
    return n;
}

// One crossing solve, make_intersection_1.comp:377-437.  Returns the unmasked t_solve.
inline float solve_crossing(const Curve& cv, int side, float c, float t_min, float t1_ms) {
    const float* p = side == 0 ? cv.x : cv.y;
    if (cv.type == T_LINE) {
        float x0 = p[0], x1 = p[1];
        float a = x1 - x0;
        a = (a != 0.0f ? (1.0f / a) : 0.0f);
        return gl_min(gl_max((c - x0) * a, t_min), t1_ms);
    }
    // quadratics and arcs: the cubic's bisection on their own evaluation
    float t0 = t_min, t1 = t1_ms;
    float vt0 = eval1(cv.type, t0, p);
    float t_solve = t0;
    if (vt0 != c) {
        float raw_t0 = t0;
        float last_vtm = 0.0f;
        for (int j = 0; j < CUBIC_ITERS; ++j) {
            float tm = (t0 + t1) * 0.5f;
            float vtm = eval1(cv.type, tm, p);
            t_solve = tm;
            last_vtm = vtm;
            if ((int32_t)(fbits(vtm - c) ^ fbits(vt0 - c)) >= 0) { t0 = tm; vt0 = vtm; }
            else { t1 = tm; }
        }
        if (std::fabs(last_vtm - c) > 1.0f) t_solve = raw_t0;
    }
    return t_solve;
}

// A gradient paint, 16 words: kind (1 linear, 2 radial), spread (0 pad, 1 repeat, 2 reflect,
// 3 transparent), first stop, stop count, then as floats the opacity, the scene-to-gradient
// affine map (gx = x0 x + x1 y + x2, gy = x3 x + x4 y + x5) and the geometry (linear x1, y1,
// x2, y2; radial cx, cy, fx, fy, r).  Stops are 5 floats: offset, r, g, b, a.
struct PaintRec {
    uint32_t kind, spread, first, count;
    float opacity, xf[6], geom[5];
};

struct Frame {
    // inputs (borrowed)
    const float* pos; const uint32_t* pos_path; uint32_t P;
    const uint32_t* cpos; const uint32_t* ctype; const uint32_t* cpath; uint32_t C;
    const uint32_t* rule; const uint32_t* fill; uint32_t Np;
    const float* cweight = nullptr;          // arc weights (read for arcs only)
    const uint32_t* path_paint = nullptr;    // per path: 0xFFFFFFFF or a paint index
    const PaintRec* paints = nullptr;
    const float* stops = nullptr;
    float m[16]; int w, h; int compose;
    int aa = 0;                               // 16: anti-aliased coverage of boundary cells
    float inv[9];                             // inverse of the MVP's 2D projective part
    // stages
    std::vector<float> tpos;
    std::vector<int32_t> visible;
    std::vector<float> cuts5;
    std::vector<int32_t> counts;  // C+1, exclusive-scanned, [C] = N
    std::vector<int32_t> inter;   // 2N
    std::vector<int32_t> key, idx, path, wn, seg;  // gen_fragment outputs (unsorted key)
    std::vector<int32_t> skey, sidx, winding;      // sorted key, sorted index, scanned winding
    std::vector<int32_t> flags;   // 2N: cell flags then span flags
    std::vector<int32_t> flag_scan;  // 2N+1 exclusive
    std::vector<int32_t> records;    // 4R
    std::vector<int32_t> rec_path;   // R: the path of each record
    std::vector<uint64_t> cmask;     // R: sample coverage of each record (anti-aliased mode)
    int32_t n_cells = 0, n_spans = 0;
    std::vector<uint8_t> image;
    bool injected = false;
};

int nthreads_setting = 1;

// transform_pos.comp:33-85 (op.x = dot(vec4(x,y,0,1), m0) evaluated left to right)
void stage_transform(Frame& F) {
    F.tpos.assign(2 * (size_t)F.P, 0.0f);
    F.visible.assign(F.Np, 0);
    const float* m = F.m;
    const float fw = (float)F.w, fh = (float)F.h;
    for (uint32_t p = 0; p < F.P; ++p) {
        float x = F.pos[2 * p], y = F.pos[2 * p + 1];
        float o[4];
        for (int r = 0; r < 4; ++r) {
            const float* mr = m + 4 * r;
            o[r] = ((x * mr[0] + y * mr[1]) + 0.0f * mr[2]) + 1.0f * mr[3];
        }
        o[0] /= o[3]; o[1] /= o[3]; o[2] /= o[3];
        int xf = o[0] < 0 ? 0 : (o[0] < fw ? 1 : 2);
        int yf = o[1] < 0 ? 0 : (o[1] < fh ? 1 : 2);
        int32_t bits = 0;
        switch ((yf << 4) | xf) {
            case 0x00: bits = 0x10000000; break;
            case 0x01: bits = 0x01000000; break;
            case 0x02: bits = 0x00100000; break;
            case 0x10: bits = 0x00010000; break;
            case 0x11: bits = 0x10000001; break;
            case 0x12: bits = 0x00001000; break;
            case 0x20: bits = 0x00000100; break;
            case 0x21: bits = 0x00000010; break;
            case 0x22: bits = 0x00000001; break;
            default: break;
        }
        F.visible[F.pos_path[p]] |= bits;
        F.tpos[2 * p] = o[0];
        F.tpos[2 * p + 1] = o[1];
    }
}

// make_intersection_0.comp:226-411
void stage_count(Frame& F) {
    F.cuts5.assign(5 * (size_t)F.C, 0.0f);
    F.counts.assign((size_t)F.C + 1, 0);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads_setting)
    for (int64_t ci = 0; ci < (int64_t)F.C; ++ci) {
        uint32_t c = (uint32_t)ci;
        Curve cv = load_curve(F.tpos.data(), F.cpos, F.ctype, F.cweight, c);
        bool vis = path_visible(F.visible[F.cpath[c]]);
        float q[5];
        uint32_t n = monotone_cuts(cv, vis, q);
        for (int i = 0; i < 4; ++i) F.cuts5[5 * (size_t)c + i] = q[i];
        F.cuts5[5 * (size_t)c + 4] = bitsf(n);
        if (vis) q[n++] = 1.0f;
        float p0x = cv.x[0], p0y = cv.y[0];
        int32_t pcnt = 0;
        for (uint32_t i = 0; i < n; ++i) {
            float t1 = q[i];
            float p1x = eval1(cv.type, t1, cv.x), p1y = eval1(cv.type, t1, cv.y);
            int xb, xe, yb, ye, nx = 0, ny = 0;
            begin_end_floor(p0x, p0y, p1x, p1y, xb, xe, yb, ye);
            cut_range(F.w, F.h, xb, xe, yb, ye, nx, ny);
            pcnt += 1 + nx + ny;
            p0x = p1x; p0y = p1y;
        }
        F.counts[c] = pcnt;
    }
}

// naive_scan.comp semantics: exclusive prefix sum, v[n] = total.
void exclusive_scan_inplace(std::vector<int32_t>& v, size_t n) {
    int32_t run = 0;
    for (size_t i = 0; i < n; ++i) { int32_t x = v[i]; v[i] = run; run = wadd(run, x); }
    v[n] = run;
}

// make_intersection_1.comp:217-447
void stage_write(Frame& F) {
    const int32_t N = F.counts[F.C];
    F.inter.assign(2 * (size_t)N, 0);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads_setting)
    for (int64_t ci = 0; ci < (int64_t)F.C; ++ci) {
        uint32_t c = (uint32_t)ci;
        Curve cv = load_curve(F.tpos.data(), F.cpos, F.ctype, F.cweight, c);
        bool vis = path_visible(F.visible[F.cpath[c]]);
        float q[5];
        for (int i = 0; i < 4; ++i) q[i] = F.cuts5[5 * (size_t)c + i];
        uint32_t n = fbits(F.cuts5[5 * (size_t)c + 4]);
        if (vis) q[n++] = 1.0f;
        float t0_ms = 0.0f;
        float p0x = cv.x[0], p0y = cv.y[0];
        int32_t pcnt = F.counts[c];
        int32_t* out = F.inter.data();
        for (uint32_t pi = 0; pi < n; ++pi) {
            float t1_ms = q[pi];
            float p1x = eval1(cv.type, t1_ms, cv.x), p1y = eval1(cv.type, t1_ms, cv.y);
            t1_ms = bitsf(fbits(t1_ms) & 0xFFFFFFFCu);
            if (std::floor(p1x) == p1x) t1_ms = bitsf((fbits(t1_ms) & 0xFFFFFFFCu) | 2u);
            else t1_ms = bitsf(fbits(t1_ms) | 3u);
            int xb, xe, yb, ye, nx = 0, ny = 0;
            float dx, dy;
            begin_end_rd(p0x, p0y, p1x, p1y, xb, xe, yb, ye, dx, dy);
            cut_range(F.w, F.h, xb, xe, yb, ye, nx, ny);
            int n_loop = nx + ny + 1;
            float x = (float)(dx < 0 ? xe : xb);
            float y = (float)(dy < 0 ? ye : yb);
            float tx = t0_ms, ty = t0_ms;  // point_coords[8], [9]
            int32_t last = (int32_t)fbits(-1.0f);
            for (int i = -1; i < n_loop; i++) {
                float t_solve = 0.0f, cc = 0.0f;
                int side = 0;
                float t_min;
                if (i == -1) {
                    t_min = tx;
                    if (nx == 0) { side = 0; t_solve = 2.0f; }
                    else if (ny == 0) { side = 1; t_solve = 2.0f; }
                    else { side = 0; --nx; cc = x; x += dx; }
                } else if (tx <= ty) {
                    t_min = tx; side = 0;
                    if (nx > 0) { --nx; cc = x; x += dx; } else { t_solve = 2.0f; }
                } else {
                    t_min = ty; side = 1;
                    if (ny > 0) { --ny; cc = y; y += dy; } else { t_solve = 2.0f; }
                }
                if (i >= 0) {
                    int32_t o = (int32_t)fbits(t_min);
                    int32_t tt = o & (int32_t)0xFFFFFFFC;
                    if (tt == (last & (int32_t)0xFFFFFFFC)) {
                        o = o | last;
                        out[2 * (size_t)(pcnt - 1)] = (int32_t)c;
                        out[2 * (size_t)(pcnt - 1) + 1] = o;
                    }
                    out[2 * (size_t)pcnt] = (int32_t)c;
                    out[2 * (size_t)pcnt + 1] = o;
                    last = o;
                    ++pcnt;
                }
                if (t_solve < 2.0f) t_solve = solve_crossing(cv, side, cc, t_min, t1_ms);
                float stored = bitsf((fbits(t_solve) & 0xFFFFFFFCu) | (uint32_t)side);
                if (side == 0) tx = stored; else ty = stored;
            }
            t0_ms = t1_ms;
            p0x = p1x; p0y = p1y;
        }
    }
}

// gen_fragment.comp:90-245
void stage_gen_fragment(Frame& F) {
    const int32_t N = (int32_t)(F.inter.size() / 2);
    F.key.assign(N, 0); F.idx.assign(N, 0); F.path.assign(N, 0); F.wn.assign(N, 0);
    F.seg.assign((size_t)F.Np + 1, 0);
    const int w = F.w, h = F.h;
#pragma omp parallel for schedule(static, 4096) num_threads(nthreads_setting)
    for (int64_t fi = 0; fi < N; ++fi) {
        int32_t c0 = F.inter[2 * fi], b0 = F.inter[2 * fi + 1];
        int32_t c1, b1;
        if (fi + 1 != N) { c1 = F.inter[2 * fi + 2]; b1 = F.inter[2 * fi + 3]; }
        else { c1 = -1; b1 = 0x3f800000; }
        float t0 = bitsf((uint32_t)b0 & 0xFFFFFFFCu), t1 = bitsf((uint32_t)b1 & 0xFFFFFFFCu);
        t0 = gl_max(t0, 0.0f);
        t1 = gl_max(t1, 0.0f);
        uint32_t path_idx = F.cpath[c0];
        int32_t yx = KEY_SENTINEL;
        if (c0 != c1) t1 = 1.0f;
        int32_t swn = 0;
        if (t0 < t1) {
            Curve cv = load_curve(F.tpos.data(), F.cpos, F.ctype, F.cweight, (uint32_t)c0);
            float pfx = eval1(cv.type, t0, cv.x), pfy = eval1(cv.type, t0, cv.y);
            float plx = eval1(cv.type, t1, cv.x), ply = eval1(cv.type, t1, cv.y);
            int raw_x = wmul(float2int_rd(((pfx + plx) * 0.5f) / (float)FRAG), FRAG);
            int raw_y = wmul(float2int_rd(((pfy + ply) * 0.5f) / (float)FRAG), FRAG);
            int pos_x = std::min(std::max(raw_x, -FRAG), (w & ~1) + FRAG);
            int pos_y = std::min(std::max(raw_y, -FRAG), (h & ~1) + FRAG);
            int32_t y_shift = (int32_t)((uint32_t)(pos_y + 0x7FFF) << 16);
            int32_t x_shift = (pos_x + 0x7FFF) & 0xFFFF;
            if ((uint32_t)raw_y < (uint32_t)h) yx = y_shift | x_shift;
            int wn_y = pos_y + 1;
            float fwy = (float)wn_y;
            if (pfy == ply) swn = 0;
            else if (pfy < fwy && fwy <= ply) swn = -1;
            else if (ply < fwy && fwy <= pfy) swn = 1;
        }
        F.key[fi] = yx;
        F.idx[fi] = (int32_t)fi;
        F.path[fi] = (int32_t)path_idx;
        F.wn[fi] = swn;
    }
    // sort segment offsets, gen_fragment.comp:227-244
    for (int64_t fi = 0; fi < N; ++fi) {
        int32_t c0 = F.inter[2 * fi];
        int32_t c1 = (fi + 1 != N) ? F.inter[2 * fi + 2] : -1;
        uint32_t path_idx = F.cpath[c0];
        if (c0 != c1) {
            uint32_t p1 = c1 >= 0 ? F.cpath[c1] : F.Np;
            for (uint32_t j = path_idx + 1; j <= p1; ++j) F.seg[j] = (int32_t)(fi + 1);
        }
        if (fi == 0) for (uint32_t j = 0; j <= path_idx; ++j) F.seg[j] = 0;
    }
}

// naive_seg_sort_pairs.comp:26-97: per segment, odd-even transposition on signed (key, value);
// the fixed point of that network is the ascending order of the (key, value) pairs.
void stage_sort(Frame& F) {
    const int32_t N = (int32_t)F.key.size();
    F.skey = F.key;
    F.sidx = F.idx;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_setting)
    for (int64_t s = 0; s < (int64_t)F.Np; ++s) {
        int32_t b = F.seg[s], e = F.seg[s + 1];
        if (N == 0 || e - b <= 1) continue;
        std::vector<std::pair<int32_t, int32_t>> kv((size_t)(e - b));
        for (int32_t i = b; i < e; ++i) kv[i - b] = {F.skey[i], F.sidx[i]};
        std::sort(kv.begin(), kv.end());
        for (int32_t i = b; i < e; ++i) { F.skey[i] = kv[i - b].first; F.sidx[i] = kv[i - b].second; }
    }
}

// shuffle_fragment.comp:17-26 then the winding scan, scanline_rasterizer.cpp:478-506 (global,
// not segmented)
void stage_winding(Frame& F) {
    const int32_t N = (int32_t)F.skey.size();
    F.winding.assign((size_t)N + 1, 0);
    for (int32_t i = 0; i < N; ++i) F.winding[i] = F.wn[F.sidx[i]];
    exclusive_scan_inplace(F.winding, N);
}

// mark_merged_fragment_and_span.comp:48-120
void stage_mark(Frame& F) {
    const int32_t N = (int32_t)F.skey.size();
    const int w = F.w, h = F.h;
    F.flags.assign(2 * (size_t)N, 0);
    for (int32_t f = 0; f < N; ++f) {
        int32_t cell = 0, span = 0;
        int32_t yx1 = F.skey[f];
        int x1 = (yx1 & 0xFFFF) - 0x7FFF;
        int y1 = ((yx1 >> 16) & 0xFFFF) - 0x7FFF;
        if (f == 0) {
            cell = (x1 < 0 || y1 < 0 || x1 >= w || y1 >= h) ? 0 : 1;
        } else {
            int32_t p0 = F.path[f - 1] & 0x3FFFFFFF, p1 = F.path[f] & 0x3FFFFFFF;
            uint32_t rule = F.rule[p1];
            int32_t yx0 = F.skey[f - 1];
            int x0 = (yx0 & 0xFFFF) - 0x7FFF;
            int y0 = ((yx0 >> 16) & 0xFFFF) - 0x7FFF;
            if (x1 < 0 || y1 < 0 || x1 >= w || y1 >= h) cell = 0;
            else if (p0 != p1 || yx0 != yx1) cell = 1;
            else cell = 0;
            int32_t wnv = F.winding[f];
            bool wf = (rule == 0 && wnv != 0) || (rule == 1 && (wnv & 1) != 0);
            span = (y0 == y1 && (x0 + FRAG) < x1 && p0 == p1 && wf) ? 1 : 0;
        }
        F.flags[f] = cell;
        F.flags[(size_t)N + f] = span;
    }
    F.flag_scan.assign(2 * (size_t)N + 1, 0);
    std::copy(F.flags.begin(), F.flags.end(), F.flag_scan.begin());
    exclusive_scan_inplace(F.flag_scan, 2 * (size_t)N);
    F.n_cells = F.flag_scan[N];
    F.n_spans = F.flag_scan[2 * (size_t)N] - F.n_cells;
}

// Anti-aliased coverage of a boundary cell (no reference counterpart: the reference leaves the
// cells' "comb-like multisampling" stencil mask TODO, gen_fragment.comp:202, scanlinepr.vert:48-50).
// 16 samples per pixel, sample j of a pixel at ((7j mod 16) + 1/2)/16 across and (j + 1/2)/16 up
// (one per sample row and column); bit (2 py + px) * 16 + j of the cell's 64-bit mask is sample j
// of its pixel (px, py).  The path is its chord polygon, the chords being the segments between
// consecutive intersections that gen_fragment forms (pf -> pl).  The winding number at a sample
// is the cell's centre-row winding number at its left edge (the exclusive scan at the cell's
// first sorted entry) plus the changes met on the way from there to the sample: along the
// centre row to the sample's column, the winding change of each chord crossing the row (the
// half-open rule of gen_fragment.comp:189-198), then along that column to the sample's height,
// +1 for each chord going right and -1 for each chord going left when moving up (the reverse
// when moving down).  Both legs stay inside the cell, so only the cell's own chords are met.
uint64_t cell_mask(const Frame& F, int32_t f) {
    const int32_t N = (int32_t)F.skey.size();
    const int32_t NI = (int32_t)(F.inter.size() / 2);
    const int32_t key = F.skey[f];
    const int32_t p = F.path[f] & 0x3FFFFFFF;
    const int rx = (key & 0xFFFF) - 0x7FFF, ry = ((key >> 16) & 0xFFFF) - 0x7FFF;
    const float yc = (float)(ry + 1);
    struct Chord { float ax, ay, bx, by; };
    std::vector<Chord> ch;
    for (int32_t g = f; g < N && F.skey[g] == key && (F.path[g] & 0x3FFFFFFF) == p; ++g) {
        const int32_t i = F.sidx[g];
        int32_t c0 = F.inter[2 * (size_t)i], b0 = F.inter[2 * (size_t)i + 1];
        int32_t c1 = -1, b1 = 0x3f800000;
        if (i + 1 < NI) { c1 = F.inter[2 * (size_t)i + 2]; b1 = F.inter[2 * (size_t)i + 3]; }
        float t0 = gl_max(bitsf((uint32_t)b0 & 0xFFFFFFFCu), 0.0f);
        float t1 = gl_max(bitsf((uint32_t)b1 & 0xFFFFFFFCu), 0.0f);
        if (c0 != c1) t1 = 1.0f;
        Curve cv = load_curve(F.tpos.data(), F.cpos, F.ctype, F.cweight, (uint32_t)c0);
        ch.push_back({eval1(cv.type, t0, cv.x), eval1(cv.type, t0, cv.y), eval1(cv.type, t1, cv.x),
                      eval1(cv.type, t1, cv.y)});
    }
    const int32_t W0 = F.winding[f];
    const uint32_t rule = F.rule[p];
    uint64_t m = 0;
    for (int py = 0; py < 2; ++py) {
        for (int px = 0; px < 2; ++px) {
            for (int j = 0; j < 16; ++j) {
                float xs = (float)(rx + px) + ((float)((7 * j) & 15) + 0.5f) / 16.0f;
                float ys = (float)(ry + py) + ((float)j + 0.5f) / 16.0f;
                int32_t W = W0;
                for (const Chord& c : ch) {
                    int sr = 0;
                    if (c.ay < yc && yc <= c.by) sr = -1;
                    else if (c.by < yc && yc <= c.ay) sr = 1;
                    if (sr != 0) {
                        float xc = c.ax + (yc - c.ay) * ((c.bx - c.ax) / (c.by - c.ay));
                        if (xc < xs) W += sr;
                    }
                    int dc = 0;
                    if (c.ax < xs && xs <= c.bx) dc = 1;
                    else if (c.bx < xs && xs <= c.ax) dc = -1;
                    if (dc != 0) {
                        float yx = c.ay + (xs - c.ax) * ((c.by - c.ay) / (c.bx - c.ax));
                        if (ys > yc) { if (yc <= yx && yx < ys) W += dc; }
                        else if (ys < yx && yx <= yc) W -= dc;
                    }
                }
                bool in = rule == 0 ? W != 0 : (W & 1) != 0;
                if (in) m |= 1ull << ((py * 2 + px) * 16 + j);
            }
        }
    }
    return m;
}

// gen_merged_fragment_and_span.comp:152-222 (the host reads n_output_fragments and n_spans back,
// scanline_rasterizer.cpp:572-580)
void stage_emit(Frame& F) {
    const int32_t N = (int32_t)F.skey.size();
    const int32_t R = F.n_cells + F.n_spans;
    F.records.assign(4 * (size_t)R, 0);
    F.rec_path.assign((size_t)R, 0);
    F.cmask.assign(F.aa ? (size_t)R : 0, ~0ull);
    const int32_t* S = F.flag_scan.data();
    for (int32_t f = 0; f < N; ++f) {
        int32_t ff = F.flags[f];
        int32_t fidx1 = S[f + 1];
        int32_t sf = F.flags[(size_t)N + f];
        int32_t sidx1 = S[(size_t)N + f + 1] - F.n_cells;
        int32_t nfb = fidx1 - ff, nsb = sidx1 - sf;
        if (ff != 0) {
            int32_t o = nfb + nsb;
            int32_t raw = F.skey[f];
            int rx = (raw & 0xFFFF) - 0x7FFF, ry = ((raw >> 16) & 0xFFFF) - 0x7FFF;
            int32_t pidx = F.path[f] & 0x3FFFFFFF;
            int32_t* r = &F.records[4 * (size_t)o];
            r[0] = (int32_t)((uint32_t)ry << 16) | rx;
            r[1] = 2;
            r[2] = (int32_t)F.fill[pidx];
            r[3] = fidx1;
            F.rec_path[o] = pidx;
            if (F.aa) F.cmask[o] = cell_mask(F, f);
        }
        if (sf != 0) {
            int32_t o = nfb + nsb + ff;
            int32_t q0 = F.skey[f - 1], q1 = F.skey[f];
            int32_t pidx = F.path[f] & 0x3FFFFFFF;
            int y0 = ((q0 >> 16) & 0xFFFF) - 0x7FFF;
            int x0 = (q0 & 0xFFFF) - 0x7FFF + FRAG;
            int x1 = (q1 & 0xFFFF) - 0x7FFF;
            x0 = std::max(0, x0);
            int32_t* r = &F.records[4 * (size_t)o];
            r[0] = (int32_t)((uint32_t)y0 << 16) | x0;
            r[1] = x1 - x0;
            r[2] = (int32_t)F.fill[pidx];
            r[3] = 0;
            F.rec_path[o] = pidx;
        }
    }
}

// Inverse of the MVP's 2D projective part: a point (x, y, 0, 1) meets rows 0, 1 and 3, so the
// map is [[m0, m1, m3], [m4, m5, m7], [m12, m13, m15]]; cofactors and determinant in double,
// rounded to float; all zeros when singular.
void mvp_inverse(const float* m, float* out) {
    double a = m[0], b = m[1], c = m[3], d = m[4], e = m[5], f = m[7], g = m[12], h = m[13], k = m[15];
    double c00 = e * k - f * h, c01 = -(d * k - f * g), c02 = d * h - e * g;
    double det = a * c00 + b * c01 + c * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
        for (int i = 0; i < 9; ++i) out[i] = 0.0f;
        return;
    }
    double adj[9] = {c00, -(b * k - c * h), b * f - c * e, c01, a * k - c * g, -(a * f - c * d), c02,
                     -(a * h - b * g), a * e - b * d};
    for (int i = 0; i < 9; ++i) out[i] = (float)(adj[i] / det);
}

// loadVG's colour packing (scanline_rasterizer.cpp:107-118): alpha times opacity, channels
// times 255 truncated into bytes, 0 when the alpha byte is 0.
uint32_t pack_rgba(float r, float g, float b, float a, float opacity) {
    float ch[4] = {r, g, b, a * opacity};
    uint32_t out = 0;
    for (int i = 0; i < 4; ++i) {
        float v = ch[i] * 255.0f;
        int iv = (v != v) ? 0 : (v <= 0.0f ? 0 : (v >= 255.0f ? 255 : (int)v));
        out |= (uint32_t)iv << (8 * i);
    }
    return (out & 0xFF000000u) ? out : 0u;
}

// Colour of gradient paint pi at the centre of canvas pixel (x, y).  The centre goes back to
// the scene through MVP^-1, then to gradient space through the paint's affine map.  Linear:
// t = ((g - p1) . d) / (d . d), d = p2 - p1 (equal points: the last stop).  Radial: the t >= 0
// for which g lies on the circle of centre f + t (c - f) and radius t r, the larger root of
// (d.d - r^2) t^2 - 2 (q.d) t + q.q = 0 with q = g - f, d = c - f (focus inside: d.d < r^2).
// NaN t counts as 0.  Spread: pad clamps to [0, 1], repeat keeps t - floor(t), reflect folds
// t mod 2 back into [0, 1], transparent gives colour 0 outside [0, 1].  Between stops the colour
// is interpolated per channel (non-premultiplied), then packed.
uint32_t paint_rgba(const Frame& F, uint32_t pi, int x, int y) {
    const float* H = F.inv;
    float px = (float)x + 0.5f, py = (float)y + 0.5f;
    float u = (H[0] * px + H[1] * py) + H[2];
    float v = (H[3] * px + H[4] * py) + H[5];
    float q = (H[6] * px + H[7] * py) + H[8];
    float sx = u / q, sy = v / q;
    const PaintRec& P = F.paints[pi];
    float gx = (P.xf[0] * sx + P.xf[1] * sy) + P.xf[2];
    float gy = (P.xf[3] * sx + P.xf[4] * sy) + P.xf[5];
    float t = 0.0f;
    bool last_stop = false;
    if (P.kind == 1) {
        float dx = P.geom[2] - P.geom[0], dy = P.geom[3] - P.geom[1];
        float dd = dx * dx + dy * dy;
        if (dd > 0.0f) t = ((gx - P.geom[0]) * dx + (gy - P.geom[1]) * dy) / dd;
        else { t = 1.0f; last_stop = true; }
    } else {
        float qx = gx - P.geom[2], qy = gy - P.geom[3];
        float dx = P.geom[0] - P.geom[2], dy = P.geom[1] - P.geom[3];
        float r = P.geom[4];
        float A = (dx * dx + dy * dy) - r * r;  // < 0
        float B = qx * dx + qy * dy;             // the root formula's -b/2
        float Cq = qx * qx + qy * qy;
        t = (B - std::sqrt(B * B - A * Cq)) / A;
    }
    if (t != t) t = 0.0f;
    if (!last_stop) {
        switch (P.spread) {
            case 0: t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t); break;
            case 1: t = t - std::floor(t); break;
            case 2: {
                float m2 = t - 2.0f * std::floor(t * 0.5f);
                t = m2 > 1.0f ? 2.0f - m2 : m2;
                break;
            }
            default:
                if (t < 0.0f || t > 1.0f) return 0u;
        }
        if (t != t) t = 0.0f;
    }
    const float* S = F.stops + 5 * (size_t)P.first;
    uint32_t n = P.count;
    float col[4];
    if (t <= S[0]) {
        for (int k = 0; k < 4; ++k) col[k] = S[1 + k];
    } else if (t >= S[5 * (n - 1)]) {
        for (int k = 0; k < 4; ++k) col[k] = S[5 * (n - 1) + 1 + k];
    } else {
        uint32_t i = 0;
        while (i + 1 < n && !(t < S[5 * (i + 1)])) ++i;
        float f = (t - S[5 * i]) / (S[5 * (i + 1)] - S[5 * i]);
        for (int k = 0; k < 4; ++k) col[k] = lerp(S[5 * i + 1 + k], S[5 * (i + 1) + 1 + k], f);
    }
    return pack_rgba(col[0], col[1], col[2], col[3], P.opacity);
}

// Line-list draw: scanlinepr.vert:18-50 + scanlinepr.frag:7-14 with the pipeline state of
// scanline_rasterizer.cpp:881-944 (line width 2, blending off, one sample, clear to white).
// Record (yx, width) becomes a 2-px-wide horizontal line centred on canvas row y+1 from x to
// x+width; its pixel centres cover canvas columns [x, x+width) and rows y, y+1.  The viewport
// maps canvas row y to image row h-1-y (the vertex shader flips y).  Image bytes are R,G,B,A.
// compose 0 = overwrite (the reference); compose 1 = src-over d = s*a + d*(1-a) per colour
// channel, alpha a + d_a*(1-a), accumulated in fp32 in record order and rounded once.
// A record of a path with a gradient paint (no reference counterpart) takes each pixel's
// colour from the paint instead of its fill.
void stage_draw(Frame& F) {
    const int w = F.w, h = F.h;
    const int32_t R = F.n_cells + F.n_spans;
    F.image.assign(4 * (size_t)w * h, 255);
    auto paint_of = [&](int32_t k) -> uint32_t {
        if (!F.path_paint || F.rec_path.empty()) return 0xFFFFFFFFu;
        return F.path_paint[F.rec_path[k]];
    };
    if (F.compose == 0) {
        for (int32_t k = 0; k < R; ++k) {
            const int32_t* r = &F.records[4 * (size_t)k];
            int x = r[0] & 0xFFFF, y = r[0] >> 16;
            int x0 = std::max(x, 0), x1 = std::min(x + r[1], w);
            uint32_t c = (uint32_t)r[2];
            const uint32_t pi = paint_of(k);
            for (int yy = y; yy < y + 2; ++yy) {
                if (yy < 0 || yy >= h) continue;
                uint8_t* row = &F.image[4 * (size_t)(h - 1 - yy) * w];
                for (int xx = x0; xx < x1; ++xx) {
                    if (pi != 0xFFFFFFFFu) c = paint_rgba(F, pi, xx, yy);
                    std::memcpy(row + 4 * xx, &c, 4);
                }
            }
        }
        return;
    }
    std::vector<float> acc(4 * (size_t)w * h, 1.0f);
    for (int32_t k = 0; k < R; ++k) {
        const int32_t* r = &F.records[4 * (size_t)k];
        int x = r[0] & 0xFFFF, y = r[0] >> 16;
        int x0 = std::max(x, 0), x1 = std::min(x + r[1], w);
        const uint32_t pi = paint_of(k);
        const uint64_t cm = F.aa ? F.cmask[k] : ~0ull;
        for (int yy = y; yy < y + 2; ++yy) {
            if (yy < 0 || yy >= h) continue;
            float* row = &acc[4 * (size_t)(h - 1 - yy) * w];
            for (int xx = x0; xx < x1; ++xx) {
                uint32_t c = pi != 0xFFFFFFFFu ? paint_rgba(F, pi, xx, yy) : (uint32_t)r[2];
                float s[4] = {(float)(c & 0xFF) / 255.0f, (float)((c >> 8) & 0xFF) / 255.0f,
                              (float)((c >> 16) & 0xFF) / 255.0f, (float)((c >> 24) & 0xFF) / 255.0f};
                float a = s[3], ia = 1.0f - a;
                float* d = row + 4 * xx;
                if (cm != ~0ull) {
                    // a boundary cell of the anti-aliased mode: the premultiplied colour times the
                    // pixel's share of covered samples
                    int bit = ((yy - y) * 2 + (xx - x)) * 16;
                    float cov = (float)__builtin_popcountll((cm >> bit) & 0xFFFFull) / 16.0f;
                    float s0 = (s[0] * a) * cov, s1 = (s[1] * a) * cov, s2 = (s[2] * a) * cov, sw = a * cov;
                    float iw = 1.0f - sw;
                    d[0] = s0 + d[0] * iw;
                    d[1] = s1 + d[1] * iw;
                    d[2] = s2 + d[2] * iw;
                    d[3] = sw + d[3] * iw;
                    continue;
                }
                d[0] = s[0] * a + d[0] * ia;
                d[1] = s[1] * a + d[1] * ia;
                d[2] = s[2] * a + d[2] * ia;
                d[3] = a + d[3] * ia;
            }
        }
    }
    for (size_t i = 0; i < acc.size(); ++i) {
        float v = acc[i] * 255.0f + 0.5f;
        v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
        F.image[i] = (uint8_t)v;
    }
}

}  // namespace

extern "C" {

// Stage identifiers; the product's spr_debug_export uses the same numbering.
enum {
    ORC_TPOS = 0, ORC_VISIBLE = 1, ORC_CUTS = 2, ORC_COUNTS = 3, ORC_INTER = 4, ORC_KEY = 5,
    ORC_WN = 6, ORC_SEG = 7, ORC_SKEY = 8, ORC_SIDX = 9, ORC_WINDING = 10, ORC_FLAGS = 11,
    ORC_RECORDS = 12, ORC_IMAGE = 13, ORC_PATH = 14, ORC_CMASK = 15
};

void oracle_set_threads(int n) { nthreads_setting = n < 1 ? 1 : n; }

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Runs the whole frame.  inter_override (optional, 2*n_inter ints) replaces the two-pass
// intersection generation, so the integer stages can be checked on identical intersections.
// cweight (arc weights), path_paint, paints (16 words each) and stops (5 floats each) may be
// null for scenes without arcs or gradients.
void* oracle_run(const float* pos, const uint32_t* pos_path, uint32_t P, const uint32_t* cpos,
                 const uint32_t* ctype, const uint32_t* cpath, uint32_t C, const uint32_t* rule,
                 const uint32_t* fill, uint32_t Np, const float* mvp_rows, int w, int h,
                 int compose, const int32_t* inter_override, int64_t n_inter, const float* cweight,
                 const uint32_t* path_paint, const void* paints, const float* stops, int aa) {
    Frame* F = new Frame();
    F->pos = pos; F->pos_path = pos_path; F->P = P;
    F->cpos = cpos; F->ctype = ctype; F->cpath = cpath; F->C = C;
    F->rule = rule; F->fill = fill; F->Np = Np;
    F->cweight = cweight;
    F->path_paint = path_paint;
    F->paints = (const PaintRec*)paints;
    F->stops = stops;
    std::memcpy(F->m, mvp_rows, sizeof(F->m));
    mvp_inverse(F->m, F->inv);
    F->w = w; F->h = h; F->compose = compose;
    F->aa = aa;
    stage_transform(*F);
    stage_count(*F);
    exclusive_scan_inplace(F->counts, C);
    if (inter_override) {
        F->inter.assign(inter_override, inter_override + 2 * n_inter);
        F->injected = true;
    } else {
        stage_write(*F);
    }
    stage_gen_fragment(*F);
    stage_sort(*F);
    stage_winding(*F);
    stage_mark(*F);
    stage_emit(*F);
    stage_draw(*F);
    return F;
}

// Pointer and byte size of one stage's result (valid until oracle_free).
int oracle_get(void* ctx, int stage, const void** ptr, int64_t* nbytes) {
    Frame* F = (Frame*)ctx;
    auto put = [&](const void* p, size_t n) { *ptr = p; *nbytes = (int64_t)n; return 0; };
    switch (stage) {
        case ORC_TPOS: return put(F->tpos.data(), F->tpos.size() * 4);
        case ORC_VISIBLE: return put(F->visible.data(), F->visible.size() * 4);
        case ORC_CUTS: return put(F->cuts5.data(), F->cuts5.size() * 4);
        case ORC_COUNTS: return put(F->counts.data(), F->counts.size() * 4);
        case ORC_INTER: return put(F->inter.data(), F->inter.size() * 4);
        case ORC_KEY: return put(F->key.data(), F->key.size() * 4);
        case ORC_WN: return put(F->wn.data(), F->wn.size() * 4);
        case ORC_SEG: return put(F->seg.data(), F->seg.size() * 4);
        case ORC_SKEY: return put(F->skey.data(), F->skey.size() * 4);
        case ORC_SIDX: return put(F->sidx.data(), F->sidx.size() * 4);
        case ORC_WINDING: return put(F->winding.data(), F->winding.size() * 4);
        case ORC_FLAGS: return put(F->flags.data(), F->flags.size() * 4);
        case ORC_RECORDS: return put(F->records.data(), F->records.size() * 4);
        case ORC_IMAGE: return put(F->image.data(), F->image.size());
        case ORC_PATH: return put(F->path.data(), F->path.size() * 4);
        case ORC_CMASK: return put(F->cmask.data(), F->cmask.size() * 8);
        default: return -1;
    }
}

void oracle_free(void* ctx) { delete (Frame*)ctx; }

// Single-stage entry points used by the known-answer tests.
void oracle_solve_quad(float a, float b, float c, float* r0, float* r1) { solve_quad(a, b, c, *r0, *r1); }

// The curve entry points take the control values as stored: for an arc (x0, x1, x2, w) with
// x1 the Euclidean control point and the weight in slot 3 (ys[3] is ignored).
Curve stored_curve(uint32_t type, const float* xs, const float* ys) {
    Curve cv; cv.type = type;
    for (int i = 0; i < 4; ++i) { cv.x[i] = xs[i]; cv.y[i] = ys[i]; }
    to_eval_form(cv, xs[3]);
    return cv;
}

int oracle_curve_cuts(uint32_t type, const float* xs, const float* ys, int visible, float* q4) {
    Curve cv = stored_curve(type, xs, ys);
    return (int)monotone_cuts(cv, visible != 0, q4);
}

// One axis at t; for an arc c4 = (x0, x1, x2, w) as stored.
float oracle_eval(uint32_t type, float t, const float* c4) {
    float c[4] = {c4[0], c4[1], c4[2], c4[3]};
    if (type == T_ARC) c[1] = c4[3] * c4[1];
    return eval1(type, t, c);
}

float oracle_solve_crossing(uint32_t type, const float* xs, const float* ys, int side, float c,
                            float t_min, float t1) {
    Curve cv = stored_curve(type, xs, ys);
    return solve_crossing(cv, side, c, t_min, t1);
}

// One gradient colour at canvas pixel (x, y) under the given MVP: the composite's paint rule.
uint32_t oracle_paint(const void* paint, const float* stops, const float* mvp_rows, int x, int y) {
    Frame F;
    F.paints = (const PaintRec*)paint;
    F.stops = stops;
    std::memcpy(F.m, mvp_rows, sizeof(F.m));
    mvp_inverse(F.m, F.inv);
    return paint_rgba(F, 0, x, y);
}

int oracle_path_visible(int32_t mask) { return path_visible(mask) ? 1 : 0; }

// Draw stage alone (scanlinepr.vert/.frag + pipeline state): records (4 ints each) to RGBA8.
void oracle_draw(const int32_t* records, int32_t R, int w, int h, int compose, uint8_t* out) {
    Frame F;
    F.w = w; F.h = h; F.compose = compose;
    F.records.assign(records, records + 4 * (size_t)R);
    F.n_cells = R; F.n_spans = 0;
    stage_draw(F);
    std::memcpy(out, F.image.data(), F.image.size());
}

int32_t oracle_float2int_rd(float x) { return float2int_rd(x); }

}  // extern "C"
