// device_common.hpp -- arithmetic shared by the gfx950 kernels.
//
// Every floating-point expression here keeps the operation order of the reference shaders
// (the library is built with -ffp-contract=off, IEEE division and square root), so that the
// integer results derived from it (counts, cells, winding) match the reference algorithm bit
// for bit.  GLSL min/max follow the GLSL definition (max(x,y) = x<y ? y : x), conversions to
// int saturate, and integer products wrap.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spr {

constexpr int kFrag = 2;           // FRAG_SIZE (make_intersection_0.comp:3)
constexpr int kCubicIters = 24;    // CUBIC_ITERATION_NUMBER (make_intersection_1.comp:7)
// CurveType (src/core/vg/vg_container.h:9-15).  A curve owns type & 7 consecutive points
// (make_intersection_0.comp:251-257): line 2, quadratic 3, cubic 4, arc 3.
constexpr uint32_t kLine = 2u, kQuad = 3u, kCubic = 4u, kArc = 0x13u;
constexpr uint32_t kSentinelKey = 0xFFFEFFFEu;  // off-canvas cell key (gen_fragment.comp:111)
constexpr uint32_t kNoPath = 0xFFFFFFFFu;      // entry path of a gap in a predicted-allocation layout
constexpr uint32_t kTwoBits = 0x40000000u;      // bit pattern of 2.0f: an exhausted chain

__host__ __device__ inline uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
__host__ __device__ inline float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ inline float glmax(float x, float y) { return x < y ? y : x; }
__host__ __device__ inline float glmin(float x, float y) { return y < x ? y : x; }
__host__ __device__ inline int32_t sat_f2i(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 0x7FFFFFFF;
    if (x <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)x;
}
__host__ __device__ inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
__host__ __device__ inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
// float2int_rd (make_intersection_1.comp:76-79)
__host__ __device__ inline int32_t f2i_rd(float x) { return x >= 0.0f ? sat_f2i(x) : sat_f2i(x - 1.0f); }
__host__ __device__ inline float lerp(float a, float b, float t) { return a + t * (b - a); }

// PATH_VISIBLE (make_intersection_0.comp:14-20): a path is culled when all its points lie
// beyond one canvas edge.
__host__ __device__ inline bool visible_mask(int32_t m) {
    return !(((m & 0x11111000) == 0) || ((m & 0x01101011) == 0) || ((m & 0x00011111) == 0) ||
             ((m & 0x11010110) == 0));
}

// Outcode bit of one transformed point (transform_pos.comp:57-82).
__host__ __device__ inline int32_t outcode_bits(float x, float y, float fw, float fh) {
    int xf = x < 0 ? 0 : (x < fw ? 1 : 2);
    int yf = y < 0 ? 0 : (y < fh ? 1 : 2);
    switch ((yf << 4) | xf) {
        case 0x00: return 0x10000000;
        case 0x01: return 0x01000000;
        case 0x02: return 0x00100000;
        case 0x10: return 0x00010000;
        case 0x11: return 0x10000001;
        case 0x12: return 0x00001000;
        case 0x20: return 0x00000100;
        case 0x21: return 0x00000010;
        case 0x22: return 0x00000001;
        default: return 0;
    }
}

// One axis of a curve at t.  Line and cubic: interpolateGeneralCurve
// (make_intersection_1.comp:117-148).  The reference declares QUADRIC and ARC (a rational
// quadratic) but leaves both cases TODO there; here a quadratic is de Casteljau on its three
// control values, and an arc de Casteljau on its homogeneous values (c0, w*x1, c2; c1 holds the
// product) and on its weights (1, w, 1; c3 holds w), then one division.
__host__ __device__ inline float eval_axis(uint32_t type, float t, float c0, float c1, float c2, float c3) {
    if (type == kCubic) {
        float q0 = lerp(c0, c1, t), q1 = lerp(c1, c2, t), q2 = lerp(c2, c3, t);
        float l0 = lerp(q0, q1, t), l1 = lerp(q1, q2, t);
        return lerp(l0, l1, t);
    }
    if (type == kLine) return lerp(c0, c1, t);
    const float n0 = lerp(c0, c1, t), n1 = lerp(c1, c2, t);
    const float n = lerp(n0, n1, t);
    if (type == kQuad) return n;
    const float d0 = lerp(1.0f, c3, t), d1 = lerp(c3, 1.0f, t);
    return n / lerp(d0, d1, t);
}

// Control values of a curve in evaluation form: an arc's x[1], y[1] hold w times its control
// point and x[3] = y[3] = w; unused slots are 0.
struct CurvePts {
    float x[4], y[4];
    uint32_t type;
    __host__ __device__ inline float ex(float t) const { return eval_axis(type, t, x[0], x[1], x[2], x[3]); }
    __host__ __device__ inline float ey(float t) const { return eval_axis(type, t, y[0], y[1], y[2], y[3]); }
};

// Curve c from the transformed points; an arc's weight (cweight[c], w > 0) is read only for arcs.
__device__ inline CurvePts load_curve(const float2* __restrict__ tpos, uint32_t base, uint32_t type,
                                      const float* __restrict__ cweight, uint32_t c) {
    CurvePts cv;
    cv.type = type;
    const float2 p0 = tpos[base], p1 = tpos[base + 1];
    cv.x[0] = p0.x; cv.y[0] = p0.y; cv.x[1] = p1.x; cv.y[1] = p1.y;
    cv.x[2] = cv.y[2] = cv.x[3] = cv.y[3] = 0.0f;
    if (type == kCubic) {
        const float2 p2 = tpos[base + 2], p3 = tpos[base + 3];
        cv.x[2] = p2.x; cv.y[2] = p2.y; cv.x[3] = p3.x; cv.y[3] = p3.y;
    } else if (type != kLine) {
        const float2 p2 = tpos[base + 2];
        cv.x[2] = p2.x; cv.y[2] = p2.y;
        if (type == kArc) {
            const float w = cweight[c];
            cv.x[1] = w * p1.x; cv.y[1] = w * p1.y;
            cv.x[3] = w; cv.y[3] = w;
        }
    }
    return cv;
}

// Roots of a*t^2 + b*t + c in citardauq form (solveQuadEquation, make_intersection_0.comp:76-129).
__host__ __device__ inline void solve_quad(float a, float b, float c, float& r0, float& r1) {
    if (a == 0.0f) {
        float x = -c / b;
        r0 = x; r1 = x;
        return;
    }
    float B = b * 0.5f;
    float R = B * B - a * c;
    if (R > 0.0f) {
        float SR = sqrtf(R);
        if (B > 0.0f) {
            float TB = B + SR;
            r0 = -c / TB; r1 = -TB / a;
        } else {
            float TB = -B + SR;
            r0 = TB / a; r1 = c / TB;
        }
    } else {
        r0 = 0.0f; r1 = 0.0f;
    }
}

// Interior monotone cuts of a curve: the roots in (0, 1) of each axis's derivative, x roots
// first.  A cubic's are ordered as the reference orders them (including its 4-cut insertion
// that never moves the fourth cut, make_intersection_0.comp:338-341).  A quadratic's derivative
// is linear; an arc's derivative numerator N'D - ND' is, per axis with homogeneous values
// (a0, a1, a2) and weights (1, w, 1), 2[e0 (1-t)^2 + e1 t(1-t) + e2 t^2] with e0 = a1 - a0 w,
// e1 = a2 - a0, e2 = a2 w - a1, solved in power form; conic cuts are sorted ascending.  Slots
// at or beyond the returned count are 0.
__host__ __device__ inline uint32_t monotone_cuts(const CurvePts& cv, bool visible, float q[4]) {
    q[0] = q[1] = q[2] = q[3] = 0.0f;
    if (!visible || cv.type == kLine) return 0;
    uint32_t n = 0;
    float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int axis = 0; axis < 2; ++axis) {
        const float* p = axis == 0 ? cv.x : cv.y;
        float a, b, c;
        if (cv.type == kCubic) {
            a = 3.0f * (p[1] - p[2]) + (p[3] - p[0]);
            b = 2.0f * ((p[0] - p[1]) + (p[2] - p[1]));
            c = p[1] - p[0];
        } else if (cv.type == kQuad) {
            a = 0.0f;
            b = (p[0] - p[1]) + (p[2] - p[1]);
            c = p[1] - p[0];
        } else {
            const float w = p[3];
            const float e0 = p[1] - p[0] * w, e1 = p[2] - p[0], e2 = p[2] * w - p[1];
            a = (e0 - e1) + e2;
            b = e1 - 2.0f * e0;
            c = e0;
        }
        float r0, r1;
        solve_quad(a, b, c, r0, r1);
        if (r0 > 0.0f && r0 < 1.0f) t[n++] = r0;
        if (r1 > 0.0f && r1 < 1.0f && r1 != r0) t[n++] = r1;
    }
    if (cv.type != kCubic) {
#pragma unroll
        for (int i = 1; i < 4; ++i)
#pragma unroll
            for (int j = i; j > 0; --j)
                if (i < (int)n && t[j] < t[j - 1]) { const float s = t[j]; t[j] = t[j - 1]; t[j - 1] = s; }
        q[0] = t[0]; q[1] = t[1]; q[2] = t[2]; q[3] = t[3];
        return n;
    }
    float q0 = t[0], q1 = t[1], q2 = t[2], q3 = t[3];
    if (n >= 2 && q1 < q0) { float s = q0; q0 = q1; q1 = s; }
    if (n >= 3) {
        float t2 = q2, t1 = q1;
        if (t2 < t1) {
            q2 = t1;
            float t0 = q0;
            if (t2 < t0) { q1 = t0; q0 = t2; } else { q1 = t2; }
        }
    }
    q[0] = q0; q[1] = q1; q[2] = q2; q[3] = q3;
    return n;
}

// udpate_cut_range (make_intersection_0.comp:186-221): clamps the grid range of one axis to the
// canvas (+1 cell) and returns how many grid lines it crosses.
__host__ __device__ inline int clamp_axis(int lim, int& b, int& e) {
    const int mx = (lim & ~1) + kFrag;
    if ((b < 0 && e < 0) || (b > mx && e > mx) || (b > e)) return 0;
    b = b < 0 ? 0 : (b > mx ? mx : b);
    e = e < 0 ? 0 : (e > mx ? mx : e);
    int n = (e - b) / kFrag + 1;
    return n > 0 ? n : 0;
}

// Grid-line count of one monotone piece (count pass, floor based: make_intersection_0.comp:167-183).
__host__ __device__ inline int piece_count(float p0x, float p0y, float p1x, float p1y, int w, int h) {
    auto g = [](float v) { return sat_f2i(floorf(v / (float)kFrag) * (float)kFrag); };
    int xb, xe, yb, ye;
    if (p0x <= p1x) { xb = wadd(g(p0x), kFrag); xe = g(p1x); } else { xb = wadd(g(p1x), kFrag); xe = g(p0x); }
    if (p0y <= p1y) { yb = wadd(g(p0y), kFrag); ye = g(p1y); } else { yb = wadd(g(p1y), kFrag); ye = g(p0y); }
    return 1 + clamp_axis(w, xb, xe) + clamp_axis(h, yb, ye);
}

// Geometry of one monotone piece for the write pass (get_xy_begin_end_delta + udpate_cut_range,
// make_intersection_1.comp:150-213, 278-302).
struct PieceGrid {
    int nx, ny;
    float x0, y0;  // first grid line crossed on each axis
    float dx, dy;  // +-2: direction of travel
};

__host__ __device__ inline PieceGrid piece_grid(float p0x, float p0y, float p1x, float p1y, int w, int h) {
    auto g = [](float v) { return wmul(f2i_rd(v / (float)kFrag), kFrag); };
    PieceGrid G;
    int xb, xe, yb, ye;
    if (p0x <= p1x) { xb = wadd(g(p0x), kFrag); xe = g(p1x); G.dx = (float)kFrag; }
    else { xb = wadd(g(p1x), kFrag); xe = g(p0x); G.dx = -(float)kFrag; }
    if (p0y <= p1y) { yb = wadd(g(p0y), kFrag); ye = g(p1y); G.dy = (float)kFrag; }
    else { yb = wadd(g(p1y), kFrag); ye = g(p0y); G.dy = -(float)kFrag; }
    G.nx = clamp_axis(w, xb, xe);
    G.ny = clamp_axis(h, yb, ye);
    G.x0 = (float)(G.dx < 0 ? xe : xb);
    G.y0 = (float)(G.dy < 0 ? ye : yb);
    return G;
}

// End-of-piece parameter with its side flag (make_intersection_1.comp:271-276): low bits 2 when
// the end point's x is integral, 3 otherwise.
__host__ __device__ inline float flag_piece_end(float t_raw, float p1x) {
    uint32_t b = fbits(t_raw) & 0xFFFFFFFCu;
    return (floorf(p1x) == p1x) ? bitsf(b | 2u) : bitsf(b | 3u);
}

// Control values of one axis of a curve, in the evaluation form of CurvePts.
struct CurveAxis {
    float c[4];
};

// One grid-crossing solve (make_intersection_1.comp:377-437): a line is solved by its
// reciprocal slope, a cubic by 24 bisection steps on [t_min, t1] keyed on the sign of B(t)-c.
// Quadratics and arcs (TODO in the reference) take the cubic's bisection on their own
// evaluation.
template <uint32_t T>
__host__ __device__ inline float bisect_axis(const CurveAxis& ax, float c, float t_min, float t1) {
    float t0 = t_min;
    const float c0 = ax.c[0], c1 = ax.c[1], c2 = ax.c[2], c3 = ax.c[3];
    float vt0 = eval_axis(T, t0, c0, c1, c2, c3);
    if (vt0 == c) return t0;
    // The sign of vt0-c never changes (vt0 is only replaced by a value of equal sign), so the
    // comparison is against the initial sign bit.
    const uint32_t s0 = fbits(vt0 - c) & 0x80000000u;
    float tm = t0, vtm = vt0;
#pragma unroll 4
    for (int j = 0; j < kCubicIters; ++j) {
        tm = (t0 + t1) * 0.5f;
        vtm = eval_axis(T, tm, c0, c1, c2, c3);
        if ((fbits(vtm - c) & 0x80000000u) == s0) t0 = tm; else t1 = tm;
    }
    return fabsf(vtm - c) > 1.0f ? t_min : tm;
}

// The chain solves take a shorter, bit-identical form of bisect_axis for cubics and quadratics.
//   (1) Every control value is finite with |c_i| < 2^17 (fast_axis).  Values are then finite,
//       and the sign bit of fl(v - c) is set exactly when v < c (fl(v - v) = +0).  So with the
//       start below the line a node turns right where v < c, and with the start above it where
//       v >= c (a value on the line counts as above).
//   (2) Evaluation commutes exactly with negation (lerp(-a, -b, t) = -lerp(a, b, t): every
//       rounding is symmetric), so a lane-solved chain stepping down runs on the negated axis,
//       where "v >= c" is "v' <= c'", that is v' < next_up(c'): the same single compare as a
//       chain stepping up (bisect_below).
//   (3) Within a chain, the start side is known after a crossing that did not fall back.  The
//       value at that crossing is within 1 of its grid line c - step.  The next bracket starts
//       at that parameter with its two low bits replaced (at most 3 * 2^-23 away; fast_bracket
//       keeps brackets in [0, 1 + 2^-20]).  The float evaluation differs from the exact
//       polynomial by at most 15.01 * 2^-24 * M (M = max |c_i|: three levels of lerps, 5 u M
//       each), and |B'| <= 6M.  So the start value lies within 1 + 66.01 * 2^-24 * M < 1.52 of
//       c - step, |step| = 2: on the side the chain comes from, never on the line.
constexpr float kFastMaxCoord = 131072.0f;               // 2^17
constexpr float kFastMaxT = 1.00000095367431640625f;     // 1 + 2^-20

__host__ __device__ inline bool fast_axis(uint32_t type, const CurveAxis& ax) {
    return (type == kCubic || type == kQuad) && fabsf(ax.c[0]) < kFastMaxCoord &&
           fabsf(ax.c[1]) < kFastMaxCoord && fabsf(ax.c[2]) < kFastMaxCoord && fabsf(ax.c[3]) < kFastMaxCoord;
}

// t in [+0, 1 + 2^-20] (negative values, -0 and NaN fail): one unsigned compare.
__host__ __device__ inline bool fast_t(float t) { return fbits(t) <= fbits(kFastMaxT); }

__host__ __device__ inline bool fast_bracket(float t_min, float t1) { return fast_t(t_min) && fast_t(t1); }

__host__ __device__ inline CurveAxis negated(const CurveAxis& a) {
    CurveAxis n;
#pragma unroll
    for (int i = 0; i < 4; ++i) n.c[i] = -a.c[i];
    return n;
}

// The least float above a finite x (above -0 and +0: the least positive denormal; the kernels
// keep fp32 denormals).
__host__ __device__ inline float next_up(float x) {
    const uint32_t b = fbits(x);
    if (b == 0x80000000u) return bitsf(1u);
    return bitsf((int32_t)b >= 0 ? b + 1u : b - 1u);
}

// bisect_axis under (1), with every decision "v < cut" (cut = c for a start below the line,
// next_up(c) on the negated axis); c itself decides the fallback to t_min (|B(tm) - c| > 1),
// reported in *ok.
// FULL unrolls the 24 steps (the straight-line form of the common path).
template <uint32_t T, bool FULL = false>
__host__ __device__ inline float bisect_below(const CurveAxis& ax, float cut, float c, float t_min, float t1, bool* ok) {
    const float c0 = ax.c[0], c1 = ax.c[1], c2 = ax.c[2], c3 = ax.c[3];
    float t0 = t_min, tm = t_min, vtm = 0.0f;
    auto step = [&]() {
        tm = (t0 + t1) * 0.5f;
        vtm = eval_axis(T, tm, c0, c1, c2, c3);
        const bool right = vtm < cut;
        t0 = right ? tm : t0;
        t1 = right ? t1 : tm;
    };
    if constexpr (FULL) {
#pragma unroll
        for (int j = 0; j < kCubicIters; ++j) step();
    } else {
#pragma unroll 4
        for (int j = 0; j < kCubicIters; ++j) step();
    }
    *ok = !(fabsf(vtm - c) > 1.0f);
    return *ok ? tm : t_min;
}

__host__ __device__ inline float solve_line(const CurveAxis& ax, float c, float t_min, float t1) {
    float a = ax.c[1] - ax.c[0];
    a = (a != 0.0f ? (1.0f / a) : 0.0f);
    return glmin(glmax((c - ax.c[0]) * a, t_min), t1);
}

__device__ inline float solve_axis(uint32_t type, const CurveAxis& ax, float c, float t_min, float t1) {
    if (type == kLine) return solve_line(ax, c, t_min, t1);
    if (type == kCubic) return bisect_axis<kCubic>(ax, c, t_min, t1);
    if (type == kQuad) return bisect_axis<kQuad>(ax, c, t_min, t1);
    return bisect_axis<kArc>(ax, c, t_min, t1);
}

// One lane's chain of crossings on one axis (make_intersection_1.comp:377-437 per crossing), in
// order: next() returns the raw parameter of the crossing with the given line, bracketed by the
// previous crossing (flagged) and the piece end.  Cubics and quadratics that satisfy (1) take
// bisect_below, on the negated axis when the chain steps down, and skip the start evaluation
// when (3) gives the start side; everything else takes solve_line or bisect_axis.  common()
// names the usual case (such a chain, start side known, t_min in range), for which
// next_common() is the same computation in straight-line form: a wave whose lanes are all in
// that case runs it without divergent branches.
template <uint32_t T>
struct ChainSolver {
    CurveAxis a, an;
    bool fast, neg, known;
    __host__ __device__ ChainSolver(const CurveAxis& ax, float step, float t1) : a(ax) {
        fast = fast_axis(T, ax) && fast_t(t1);
        neg = fast && step < 0.0f;
        an = neg ? negated(ax) : ax;
        known = false;  // the next crossing starts on the side the chain comes from
    }
    __host__ __device__ bool common(float t_min) const { return T != kLine && fast && known && fast_t(t_min); }
    __host__ __device__ float next_common(float line, float t_min, float t1) {
        const float cn = neg ? -line : line;
        return bisect_below<T, true>(an, neg ? next_up(cn) : cn, cn, t_min, t1, &known);
    }
    __host__ __device__ float next(float line, float t_min, float t1) {
        if (T == kLine) return solve_line(a, line, t_min, t1);
        if (!fast) return bisect_axis<T>(a, line, t_min, t1);
        bool run = true, from = true;
        const bool fb = fast_t(t_min);  // t1 was checked with the axis
        if (!(fb && known)) {
            const float vt0 = eval_axis(T, t_min, a.c[0], a.c[1], a.c[2], a.c[3]);
            run = vt0 != line;
            from = neg ? vt0 > line : vt0 < line;
        }
        float ts = t_min;
        known = fb;  // a start on the line itself is a crossing with no fallback
        if (run) {
            if (from) {
                const float cn = neg ? -line : line;
                ts = bisect_below<T>(an, neg ? next_up(cn) : cn, cn, t_min, t1, &known);
                known = known && fb;
            } else {
                ts = bisect_axis<T>(a, line, t_min, t1);
                known = false;
            }
        }
        return ts;
    }
};

// Bucket of an entry inside its path's (path, cell row) bucket range.  pr = (first row >= 1,
// number of rows): rows first..first+nrows-1 come first, then the off-canvas bucket, then row 0,
// which is the order of the reference's signed key comparison.  A row outside the path's bound
// (impossible for finite input: every curve point lies inside its control points' hull) raises
// *err and is clamped so that no write goes out of range.
__device__ inline uint32_t bucket_local(int32_t key, int2 pr, int32_t* err) {
    if ((uint32_t)key == kSentinelKey) return (uint32_t)pr.y;
    const int y = ((key >> 16) & 0xFFFF) - 0x7FFF;
    const int r = y >> 1;
    if (r == 0) return (uint32_t)pr.y + 1u;
    int l = r - pr.x;
    if (l < 0 || l >= pr.y) {
        atomicOr(err, 1);
        if (pr.y == 0) return 0u;
        l = l < 0 ? 0 : pr.y - 1;
    }
    return (uint32_t)l;
}

// Runs of equal keys among consecutive valid lanes of a wave: a lane heads a run when it is
// valid and the lane below it is invalid or holds another key.  Returns the ballot of heads.
__device__ inline unsigned long long run_heads(uint32_t key, bool valid, unsigned long long vmask) {
    const int lane = threadIdx.x & 63;
    const uint32_t prev = (uint32_t)__shfl_up((int)key, 1);
    const bool prev_valid = lane > 0 && ((vmask >> (lane - 1)) & 1ull);
    return __ballot(valid && (!prev_valid || prev != key));
}

// Length of the run headed by `lane`: up to the next head or the next invalid lane.
__device__ inline int run_length(int lane, unsigned long long heads, unsigned long long vmask) {
    const unsigned long long stops = (heads | ~vmask) & ~((2ull << lane) - 1ull);
    return (stops ? __ffsll((long long)stops) - 1 : 64) - lane;
}

// Wave-aggregated atomicAdd on keyed counters: every run of equal keys takes consecutive slots
// with one atomic issued by its head, all heads at once (entries come in runs: consecutive
// crossings of a curve mostly share their cell row).  Slots of one key are handed out in no
// particular order.  Every lane of the wave must call it.
__device__ inline int32_t wave_slot(int32_t* __restrict__ cnt, uint32_t key, bool valid) {
    const int lane = threadIdx.x & 63;
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long heads = run_heads(key, valid, vmask);
    int32_t base = 0;
    if ((heads >> lane) & 1ull) base = atomicAdd(&cnt[key], run_length(lane, heads, vmask));
    const unsigned long long mine = heads & ((2ull << lane) - 1ull);
    const int h = (valid && mine) ? 63 - __clzll((long long)mine) : lane;
    base = __shfl(base, h);
    return base + (lane - h);
}

// Same runs for counters whose result is not needed.
__device__ inline void wave_count(int32_t* __restrict__ cnt, uint32_t key, bool valid) {
    const int lane = threadIdx.x & 63;
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long heads = run_heads(key, valid, vmask);
    if ((heads >> lane) & 1ull) atomicAdd(&cnt[key], run_length(lane, heads, vmask));
}

// gen_fragment.comp:90-224: cell key of the curve segment [t0, t1] from its end points (pf at t0,
// pl at t1; callers share the evaluation of a point between two consecutive segments): the key
// of its midpoint's cell, or the off-canvas sentinel when t0 >= t1 or the cell row is off the
// canvas, and its winding change across the cell's centre row.
__device__ inline void segment_cell_pts(bool forward, float pfx, float pfy, float plx, float ply, int w, int h,
                                        uint32_t& yx, int& swn) {
    yx = kSentinelKey;
    swn = 0;
    if (!forward) return;
    const int raw_x = wmul(f2i_rd(((pfx + plx) * 0.5f) / (float)kFrag), kFrag);
    const int raw_y = wmul(f2i_rd(((pfy + ply) * 0.5f) / (float)kFrag), kFrag);
    const int pos_x = min(max(raw_x, -kFrag), (w & ~1) + kFrag);
    const int pos_y = min(max(raw_y, -kFrag), (h & ~1) + kFrag);
    const uint32_t y_shift = (uint32_t)(pos_y + 0x7FFF) << 16;
    const uint32_t x_shift = (uint32_t)(pos_x + 0x7FFF) & 0xFFFFu;
    if ((uint32_t)raw_y < (uint32_t)h) yx = y_shift | x_shift;
    const float fwy = (float)(pos_y + 1);
    if (pfy == ply) swn = 0;
    else if (pfy < fwy && fwy <= ply) swn = -1;
    else if (ply < fwy && fwy <= pfy) swn = 1;
}

}  // namespace spr
