// scene.cpp -- RVG parsing, SoA flattening, RVG writing and the synthetic generators.
#include "scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace spr {

namespace {

// The container model of src/core/vg/vg_container.h:21-87: points, curves owning their points
// (endpoints duplicated), paths owning a contiguous curve range.
struct Container {
    std::vector<float> pts;           // 2 per point
    std::vector<uint32_t> curve_begin;  // first point per curve
    std::vector<uint32_t> curve_type;
    std::vector<float> curve_weight;  // per curve (1 except arcs)
    std::vector<uint32_t> path_curve_begin;
    std::vector<uint32_t> rule;
    std::vector<float> color;         // 4 per path
    std::vector<float> opacity;
    std::vector<uint32_t> paint;      // per path: kSolidPaint or an index into paints
    std::vector<Paint> paints;
    std::vector<Stop> stops;
    int path_index = -1;

    uint32_t n_points() const { return (uint32_t)(pts.size() / 2); }

    // VGContainer::newPath, vg_container.h:49-56
    void new_path() {
        ++path_index;
        path_curve_begin.push_back((uint32_t)curve_type.size());
        rule.push_back(0);
        color.insert(color.end(), {0.0f, 0.0f, 0.0f, 1.0f});
        opacity.push_back(0.0f);
        paint.push_back(kSolidPaint);
    }
    // VGContainer::newCurve + addCurve, vg_container.h:58-86 (which stores no points for the
    // QUADRIC and ARC types it declares; here they own 3 points, an arc also its weight)
    void add_curve(uint32_t type, const float* p, int npts, float weight = 1.0f) {
        curve_begin.push_back(n_points());
        curve_type.push_back(type);
        curve_weight.push_back(weight);
        for (int i = 0; i < npts; ++i) { pts.push_back(p[2 * i]); pts.push_back(p[2 * i + 1]); }
    }
};

// loadVG's flattening, scanline_rasterizer.cpp:67-134.
Scene flatten(const Container& c, uint32_t width, uint32_t height) {
    Scene s;
    s.width = width; s.height = height;
    const uint32_t n_paths = (uint32_t)(c.path_index + 1);
    const uint32_t n_curves = (uint32_t)c.curve_type.size();
    const uint32_t n_points = c.n_points();
    for (uint32_t pi = 0; pi < n_paths; ++pi) {
        uint32_t cb = c.path_curve_begin[pi];
        uint32_t ce = pi + 1 != n_paths ? c.path_curve_begin[pi + 1] : n_curves;
        s.fill.push_back(pack_fill(c.color[4 * pi], c.color[4 * pi + 1], c.color[4 * pi + 2],
                                   c.color[4 * pi + 3], c.opacity[pi]));
        s.rule.push_back(c.rule[pi]);
        for (uint32_t ci = cb; ci < ce; ++ci) {
            uint32_t pb = c.curve_begin[ci];
            uint32_t pe = ci + 1 != n_curves ? c.curve_begin[ci + 1] : n_points;
            s.curve_path.push_back(pi);
            s.curve_pos.push_back(pb);
            s.curve_type.push_back(c.curve_type[ci]);
            s.curve_weight.push_back(c.curve_weight[ci]);
            for (uint32_t p = pb; p < pe; ++p) {
                s.pos.push_back(c.pts[2 * p]);
                s.pos.push_back(c.pts[2 * p + 1]);
                s.pos_path.push_back(pi);
            }
        }
    }
    bool arcs = false, paints = false;
    for (uint32_t t : s.curve_type) arcs = arcs || t == 0x13u;
    for (uint32_t q : c.paint) paints = paints || q != kSolidPaint;
    if (!arcs) s.curve_weight.clear();
    if (paints) {
        s.path_paint.assign(c.paint.begin(), c.paint.begin() + n_paths);
        s.paints = c.paints;
        s.stops = c.stops;
    }
    return s;
}

void replace_sep(std::string& s) {
    for (auto& ch : s) if (ch == ',' || ch == ':') ch = ' ';
}

// read_point of rvg.cpp:188-198: one token "x,y" (a trailing ':' allowed).
bool read_point(std::istream& in, float& x, float& y) {
    std::string t;
    if (!(in >> t)) return false;
    replace_sep(t);
    x = 0.0f; y = 0.0f;
    return std::sscanf(t.c_str(), "%f %f", &x, &y) == 2;
}

// PCG32 (O'Neill 2014, pcg32_random_r with the default multiplier).
struct Pcg32 {
    uint64_t state = 0, inc = 0;
    explicit Pcg32(uint64_t seed, uint64_t seq = 0x5eed) {
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    double uniform() { return (double)next() * (1.0 / 4294967296.0); }  // [0, 1)
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    int range(int lo, int hi) { return lo + (int)(next() % (uint32_t)(hi - lo + 1)); }
    double normal() {  // Box-Muller
        double u1 = (next() + 1.0) * (1.0 / 4294967297.0), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

uint32_t rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    uint32_t c = (r & 255u) | ((g & 255u) << 8) | ((b & 255u) << 16) | ((a & 255u) << 24);
    return (c & 0xFF000000u) ? c : 0u;
}

struct Builder {
    Scene s;
    void begin_path(uint32_t rule, uint32_t fill) { s.rule.push_back(rule); s.fill.push_back(fill); }
    void cubic(float x0, float y0, float x1, float y1, float x2, float y2, float x3, float y3) {
        uint32_t p = s.n_paths() - 1;
        s.curve_pos.push_back(s.n_points());
        s.curve_type.push_back(4);
        s.curve_path.push_back(p);
        const float v[8] = {x0, y0, x1, y1, x2, y2, x3, y3};
        for (int i = 0; i < 4; ++i) {
            s.pos.push_back(v[2 * i]); s.pos.push_back(v[2 * i + 1]); s.pos_path.push_back(p);
        }
    }
    // Degree elevation of a quadratic (p0, q, p1) to a cubic, done once on the host:
    // C1 = P0 + 2/3 (Q - P0), C2 = P1 + 2/3 (Q - P1).
    void quad(float x0, float y0, float qx, float qy, float x1, float y1) {
        const float k = 2.0f / 3.0f;
        cubic(x0, y0, x0 + k * (qx - x0), y0 + k * (qy - y0), x1 + k * (qx - x1), y1 + k * (qy - y1), x1, y1);
    }
};

// Random closed cubic scribbles with control points uniform in the canvas.
Scene gen_scribbles(Pcg32& r, uint32_t n_paths, uint32_t W) {
    Builder b;
    const float fw = (float)W;
    for (uint32_t p = 0; p < n_paths; ++p) {
        int nseg = r.range(8, 32);
        uint32_t cr = r.range(0, 255), cg = r.range(0, 255), cb = r.range(0, 255);
        b.begin_path(0, rgba(cr, cg, cb, 255));
        float x0 = (float)(r.uniform() * fw), y0 = (float)(r.uniform() * fw);
        float sx = x0, sy = y0;
        for (int k = 0; k < nseg; ++k) {
            float x1 = (float)(r.uniform() * fw), y1 = (float)(r.uniform() * fw);
            float x2 = (float)(r.uniform() * fw), y2 = (float)(r.uniform() * fw);
            float x3 = (float)(r.uniform() * fw), y3 = (float)(r.uniform() * fw);
            if (k == nseg - 1) { x3 = sx; y3 = sy; }
            b.cubic(x0, y0, x1, y1, x2, y2, x3, y3);
            x0 = x3; y0 = y3;
        }
    }
    b.s.width = W; b.s.height = W;
    return b.s;
}

// One closed quadratic contour around (cx, cy) with radii (rx, ry); ccw = +1 or -1.
void glyph_contour(Builder& b, Pcg32& r, int n, double cx, double cy, double rx, double ry, int dir) {
    std::vector<double> ang(n), rad(n);
    for (int k = 0; k < n; ++k) {
        ang[k] = dir * 6.283185307179586 * (k + 0.35 * (r.uniform() - 0.5)) / n;
        rad[k] = r.uniform(0.75, 1.0);
    }
    for (int k = 0; k < n; ++k) {
        int k1 = (k + 1) % n;
        double a0 = ang[k], a1 = k1 == 0 ? ang[0] + dir * 6.283185307179586 : ang[k1];
        double am = 0.5 * (a0 + a1), rm = r.uniform(0.7, 1.25);
        float x0 = (float)(cx + rx * rad[k] * std::cos(a0)), y0 = (float)(cy + ry * rad[k] * std::sin(a0));
        float x1 = (float)(cx + rx * rad[k1] * std::cos(ang[k1])), y1 = (float)(cy + ry * rad[k1] * std::sin(ang[k1]));
        float qx = (float)(cx + rx * rm * std::cos(am)), qy = (float)(cy + ry * rm * std::sin(am));
        b.quad(x0, y0, qx, qy, x1, y1);
    }
}

// Glyph-like text page: small closed quadratic paths on a jittered row grid, even-odd fill,
// 30% with one counter-wound inner hole.
Scene gen_glyphs(Pcg32& r, uint32_t n_paths, uint32_t W) {
    Builder b;
    const double fw = W;
    int cols = (int)std::ceil(std::sqrt((double)n_paths));
    double pitch = fw / (cols + 1);
    for (uint32_t p = 0; p < n_paths; ++p) {
        int row = (int)(p / cols), col = (int)(p % cols);
        double hgt = std::exp(r.uniform(std::log(8.0), std::log(48.0)));
        double wid = hgt * r.uniform(0.45, 0.85);
        double cx = (col + 1) * pitch + pitch * 0.3 * (r.uniform() - 0.5);
        double cy = (row + 1) * pitch + pitch * 0.3 * (r.uniform() - 0.5);
        cx = std::min(std::max(cx, wid * 0.7 + 1.0), fw - wid * 0.7 - 1.0);
        cy = std::min(std::max(cy, hgt * 0.7 + 1.0), fw - hgt * 0.7 - 1.0);
        int n = r.range(10, 40);
        bool hole = r.uniform() < 0.3;
        uint32_t c = r.range(0, 80);
        b.begin_path(1, rgba(c, r.range(0, 80), r.range(0, 80), 255));
        if (hole) {
            int nh = std::max(4, n / 3);
            glyph_contour(b, r, n - nh, cx, cy, wid * 0.5, hgt * 0.5, +1);
            double k = r.uniform(0.3, 0.5);
            glyph_contour(b, r, nh, cx, cy, wid * 0.5 * k, hgt * 0.5 * k, -1);
        } else {
            glyph_contour(b, r, n, cx, cy, wid * 0.5, hgt * 0.5, +1);
        }
    }
    b.s.width = W; b.s.height = W;
    return b.s;
}

// Illustration: overlapping closed cubic paths; segment count and box size log-normal
// (sigma 1), per-path affine (scale 0.5-4, uniform rotation) applied to the control points,
// 50/50 nonzero/even-odd, alpha 0.3-1.0.  Every control point stays inside the canvas.
Scene gen_illustration(Pcg32& r, uint32_t n_paths, uint32_t W) {
    Builder b;
    const double fw = W;
    const double seg_median = 24.0, box_median = fw / 32.0;
    for (uint32_t p = 0; p < n_paths; ++p) {
        int nseg = (int)std::lround(seg_median * std::exp(r.normal()));
        nseg = std::min(std::max(nseg, 4), 400);
        double box = box_median * std::exp(r.normal());
        double scale = r.uniform(0.5, 4.0), rot = r.uniform(0.0, 6.283185307179586);
        double half = 0.5 * box * scale * 1.4142135623730951;  // radius of the transformed box
        double lim = 0.5 * fw - 2.0;
        if (half > lim) { box *= lim / half; half = lim; }
        double cx = r.uniform(half + 1.0, fw - half - 1.0), cy = r.uniform(half + 1.0, fw - half - 1.0);
        uint32_t rule = r.next() & 1u;
        uint32_t alpha = (uint32_t)(r.uniform(0.3, 1.0) * 255.0);
        b.begin_path(rule, rgba(r.range(0, 255), r.range(0, 255), r.range(0, 255), alpha));
        double cs = std::cos(rot) * scale, sn = std::sin(rot) * scale;
        auto pt = [&](float& x, float& y) {
            double u = (r.uniform() - 0.5) * box, v = (r.uniform() - 0.5) * box;
            x = (float)(cx + cs * u - sn * v);
            y = (float)(cy + sn * u + cs * v);
        };
        float x0, y0;
        pt(x0, y0);
        float sx = x0, sy = y0;
        for (int k = 0; k < nseg; ++k) {
            float x1, y1, x2, y2, x3, y3;
            pt(x1, y1); pt(x2, y2); pt(x3, y3);
            if (k == nseg - 1) { x3 = sx; y3 = sy; }
            b.cubic(x0, y0, x1, y1, x2, y2, x3, y3);
            x0 = x3; y0 = y3;
        }
    }
    b.s.width = W; b.s.height = W;
    return b.s;
}

}  // namespace

uint32_t pack_fill(float r, float g, float b, float a, float opacity) {
    float c[4] = {r, g, b, a * opacity};
    uint32_t out = 0;
    for (int i = 0; i < 4; ++i) {
        float v = c[i] * 255.0f;
        int iv = (v != v) ? 0 : (v <= 0.0f ? 0 : (v >= 255.0f ? 255 : (int)v));
        out |= (uint32_t)iv << (8 * i);
    }
    return (out & 0xFF000000u) ? out : 0u;
}

namespace {

// 2x3 affine map in double: x' = m0 x + m1 y + m2, y' = m3 x + m4 y + m5.
struct Affine {
    double m[6] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0};
    void apply(double x, double y, double& ox, double& oy) const {
        ox = m[0] * x + m[1] * y + m[2];
        oy = m[3] * x + m[4] * y + m[5];
    }
    // this map after `first`
    Affine after(const Affine& first) const {
        Affine r;
        const double* a = m;
        const double* b = first.m;
        r.m[0] = a[0] * b[0] + a[1] * b[3]; r.m[1] = a[0] * b[1] + a[1] * b[4]; r.m[2] = a[0] * b[2] + a[1] * b[5] + a[2];
        r.m[3] = a[3] * b[0] + a[4] * b[3]; r.m[4] = a[3] * b[1] + a[4] * b[4]; r.m[5] = a[3] * b[2] + a[4] * b[5] + a[5];
        return r;
    }
    bool inverse(Affine& out) const {
        const double det = m[0] * m[4] - m[1] * m[3];
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
        const double id = 1.0 / det;
        out.m[0] = m[4] * id; out.m[1] = -m[1] * id; out.m[3] = -m[3] * id; out.m[4] = m[0] * id;
        out.m[2] = -(out.m[0] * m[2] + out.m[1] * m[5]);
        out.m[5] = -(out.m[3] * m[2] + out.m[4] * m[5]);
        return true;
    }
};

// The numbers of one token after ':', ',', '(', ')', '[' and ']' become blanks.
std::vector<double> numbers_of(std::string t) {
    for (auto& ch : t)
        if (ch == ',' || ch == ':' || ch == '(' || ch == ')' || ch == '[' || ch == ']') ch = ' ';
    std::vector<double> v;
    std::istringstream iss(t);
    double x;
    while (iss >> x) v.push_back(x);
    return v;
}

// A transform token: "dyn_identity" or "dyn_affine([a,b,c],[d,e,f])" (read on until the closing
// parenthesis if the writer put blanks inside).
Affine read_transform(std::istream& in, std::string t) {
    Affine xf;
    if (t == "dyn_identity") return xf;
    if (t.compare(0, 10, "dyn_affine") != 0) throw std::runtime_error("rvg: unsupported transform '" + t + "'");
    std::string all = t.substr(10), more;
    while (all.find(')') == std::string::npos && in >> more) all += " " + more;
    const std::vector<double> v = numbers_of(all);
    if (v.size() != 6) throw std::runtime_error("rvg: bad dyn_affine '" + all + "'");
    for (int k = 0; k < 6; ++k) {
        if (!std::isfinite(v[k])) throw std::runtime_error("rvg: non-finite dyn_affine");
        xf.m[k] = v[k];
    }
    return xf;
}

// A colour token rgba(r,g,b,a) or rgb(r,g,b), as the solid paint of rvg.cpp:236-250 reads it.
void read_color(const std::string& t, float* col) {
    if (t.compare(0, 4, "rgba") == 0 && t.size() >= 6) {
        std::string v = t.substr(5, t.length() - 6);
        replace_sep(v);
        std::istringstream iss(v);
        iss >> col[0] >> col[1] >> col[2] >> col[3];
    } else if (t.compare(0, 3, "rgb") == 0 && t.size() >= 6) {
        std::string v = t.substr(4, t.length() - 6);  // same span as rvg.cpp:244
        replace_sep(v);
        std::istringstream iss(v);
        iss >> col[0] >> col[1] >> col[2];
    } else {
        throw std::runtime_error("rvg: bad colour '" + t + "'");
    }
}

// The arc from p0 with homogeneous control point (hx, hy, w) to p2, appended as arcs of weight
// >= 1/2 with Euclidean control points: an arc of smaller weight (w = 0 is a half ellipse, its
// control point at infinity) is split at t = 1/2 into two arcs of weight sqrt((1 + w) / 2), whose
// control points are (p0 + h) / (1 + w) and (h + p2) / (1 + w) and whose common end point is
// (p0 + 2h + p2) / (2 + 2w).
void add_arc(Container& c, double x0, double y0, double hx, double hy, double w, double x2, double y2,
             int depth = 0) {
    if (!(w > -1.0) || !std::isfinite(w) || !std::isfinite(hx) || !std::isfinite(hy))
        throw std::runtime_error("rvg: arc weight must be finite and greater than -1");
    if (w >= 0.5) {
        const float p[6] = {(float)x0, (float)y0, (float)(hx / w), (float)(hy / w), (float)x2, (float)y2};
        c.add_curve(0x13u, p, 3, (float)w);
        return;
    }
    if (depth >= 16) throw std::runtime_error("rvg: arc weight too close to -1");
    const double s = 1.0 + w, w2 = std::sqrt(0.5 * s);
    const double mx = (x0 + 2.0 * hx + x2) / (2.0 * s), my = (y0 + 2.0 * hy + y2) / (2.0 * s);
    add_arc(c, x0, y0, w2 * (x0 + hx) / s, w2 * (y0 + hy) / s, w2, mx, my, depth + 1);
    add_arc(c, mx, my, w2 * (hx + x2) / s, w2 * (hy + y2) / s, w2, x2, y2, depth + 1);
}

}  // namespace

// RVG::load -> parse_header + parse_paths, src/core/vg/rvg.cpp:9-255, plus what that parser
// skips: Q (quadratic) and A (rational quadratic "x1,y1,w1 x2,y2", the control point in
// homogeneous form) segments, dyn_affine path, paint and scene transforms, and linear and radial
// gradient paints ("linear gradient (x1,y1) (x2,y2)" or "radial gradient (cx,cy) (fx,fy) r",
// then "dyn_ramp" offset:colour stops, a spread pad|repeat|reflect|transparent and the paint's
// transform, gradient space to scene).
Scene load_rvg(const std::string& path) {
    std::ifstream fin(path);
    if (!fin.is_open()) throw std::runtime_error("cannot open \"" + path + "\"");
    Container c;
    std::string t;
    float vx, vy, vz, vw, wx, wy, wz, ww;
    // parse_header, rvg.cpp:33-83
    if (!(fin >> t) || t != "viewport") throw std::runtime_error("rvg: expected 'viewport'");
    if (!read_point(fin, vx, vy) || !read_point(fin, vz, vw)) throw std::runtime_error("rvg: bad viewport");
    if (!(fin >> t) || t != "window") throw std::runtime_error("rvg: expected 'window'");
    if (!read_point(fin, wx, wy) || !read_point(fin, wz, ww)) throw std::runtime_error("rvg: bad window");
    if (!(fin >> t) || t != "scene") throw std::runtime_error("rvg: expected 'scene'");
    if (!(fin >> t)) throw std::runtime_error("rvg: truncated scene");
    const Affine scene_xf = read_transform(fin, t);
    (void)vx; (void)vy; (void)wx; (void)wy; (void)wz; (void)ww;

    // parse_paths, rvg.cpp:85-255
    long last_path_vertex_number = -1;
    while (fin >> t) {
        if (t.size() >= 2 && t[0] == '/' && t[1] == '/') { std::getline(fin, t); continue; }
        if (last_path_vertex_number != (long)c.n_points()) {
            last_path_vertex_number = (long)c.n_points();
            c.new_path();
        }
        if (t != "1") throw std::runtime_error("rvg: expected '1' at the start of an element, got '" + t + "'");
        if (!(fin >> t) || t != "element") throw std::runtime_error("rvg: expected 'element'");
        const int pi = c.path_index;
        if (!(fin >> t)) throw std::runtime_error("rvg: truncated element");
        if (t == "ofill") c.rule[pi] = 1;
        else if (t == "nzfill") c.rule[pi] = 0;
        else throw std::runtime_error("rvg: unsupported fill '" + t + "'");
        if (!(fin >> t) || t != "dyn_concrete") throw std::runtime_error("rvg: expected 'dyn_concrete'");
        float dx, dy;
        for (int i = 0; i < 3; ++i)
            if (!read_point(fin, dx, dy)) throw std::runtime_error("rvg: bad dyn_concrete point");
        float p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t path_begin = c.n_points();
        uint32_t contour_begin = c.n_points();
        bool closed = false;
        bool have_tok = false;
        while (fin >> t) {
            if (t != "fL" && t.length() != 1) { have_tok = true; break; }
            char cmd = t[0];
            if (std::string("MZLCQA").find(cmd) == std::string::npos && t != "fL") {
                if (std::string("mzlcqaHhVvSsTt").find(cmd) != std::string::npos)
                    throw std::runtime_error(std::string("rvg: unsupported path command '") + cmd + "'");
                have_tok = true;
                break;
            }
            if (t == "fL") {  // rvg.cpp:285-289
                if (!read_point(fin, p[0], p[1])) throw std::runtime_error("rvg: bad point");
                continue;
            }
            switch (cmd) {
                case 'M':
                    if (!read_point(fin, p[0], p[1])) throw std::runtime_error("rvg: bad point");
                    contour_begin = c.n_points();
                    closed = false;
                    break;
                case 'L':
                    if (!read_point(fin, p[2], p[3])) throw std::runtime_error("rvg: bad point");
                    c.add_curve(2, p, 2);
                    p[0] = p[2]; p[1] = p[3];
                    break;
                case 'Q':
                    if (!read_point(fin, p[2], p[3]) || !read_point(fin, p[4], p[5]))
                        throw std::runtime_error("rvg: bad point");
                    c.add_curve(3, p, 3);
                    p[0] = p[4]; p[1] = p[5];
                    break;
                case 'C':
                    if (!read_point(fin, p[2], p[3]) || !read_point(fin, p[4], p[5]) || !read_point(fin, p[6], p[7]))
                        throw std::runtime_error("rvg: bad point");
                    c.add_curve(4, p, 4);
                    p[0] = p[6]; p[1] = p[7];
                    break;
                case 'A': {
                    std::string h;
                    if (!(fin >> h)) throw std::runtime_error("rvg: truncated arc");
                    std::vector<double> hv = numbers_of(h);
                    float ex, ey;
                    if (hv.size() != 3 || !read_point(fin, ex, ey)) throw std::runtime_error("rvg: bad arc '" + h + "'");
                    // the values as the file's floats, the split in double
                    add_arc(c, p[0], p[1], (float)hv[0], (float)hv[1], (float)hv[2], ex, ey);
                    p[0] = ex; p[1] = ey;
                    break;
                }
                case 'Z':
                    closed = true;
                    break;
            }
        }
        // implicit closing line of the last contour, rvg.cpp:210-223 (Z adds no segment)
        if (!closed && c.n_points() > contour_begin) {
            float fx = c.pts[2 * contour_begin], fy = c.pts[2 * contour_begin + 1];
            float lx = c.pts[c.pts.size() - 2], ly = c.pts[c.pts.size() - 1];
            if (fx != lx || fy != ly) {
                float q[4] = {lx, ly, fx, fy};
                c.add_curve(2, q, 2);
            }
        }
        if (!have_tok) throw std::runtime_error("rvg: truncated element");
        // the path's transform (and the scene's) applied to its points: exact for every curve
        // type under an affine map, an arc keeping its weight
        const Affine path_xf = scene_xf.after(read_transform(fin, t));
        const bool moved = path_xf.m[0] != 1.0 || path_xf.m[1] != 0.0 || path_xf.m[2] != 0.0 ||
                           path_xf.m[3] != 0.0 || path_xf.m[4] != 1.0 || path_xf.m[5] != 0.0;
        if (moved) {
            for (uint32_t q = path_begin; q < c.n_points(); ++q) {
                double ox, oy;
                path_xf.apply(c.pts[2 * q], c.pts[2 * q + 1], ox, oy);
                c.pts[2 * q] = (float)ox;
                c.pts[2 * q + 1] = (float)oy;
            }
        }
        if (!(fin >> t) || t != "dyn_paint") throw std::runtime_error("rvg: expected 'dyn_paint'");
        float opacity;
        if (!(fin >> opacity)) throw std::runtime_error("rvg: bad opacity");
        c.opacity[pi] = opacity;
        if (!(fin >> t)) throw std::runtime_error("rvg: truncated paint");
        float* col = &c.color[4 * pi];
        if (t == "solid") {
            if (!(fin >> t)) throw std::runtime_error("rvg: truncated paint");
            read_color(t, col);
            continue;
        }
        if (t != "linear" && t != "radial") throw std::runtime_error("rvg: unsupported paint '" + t + "'");
        Paint g{};
        g.kind = t == "linear" ? kPaintLinear : kPaintRadial;
        g.opacity = opacity;
        std::string a, b;
        if (!(fin >> t) || t != "gradient" || !(fin >> a) || !(fin >> b)) throw std::runtime_error("rvg: bad gradient");
        const std::vector<double> pa = numbers_of(a), pb = numbers_of(b);
        if (pa.size() != 2 || pb.size() != 2) throw std::runtime_error("rvg: bad gradient points");
        double cx = pa[0], cy = pa[1], fx = pb[0], fy = pb[1], r = 0.0;
        if (g.kind == kPaintRadial) {
            if (!(fin >> r) || !(r > 0.0)) throw std::runtime_error("rvg: bad gradient radius");
            // a focus on or outside the circle moves inside it (SVG 1.1 moves it onto the circle)
            const double d = std::hypot(fx - cx, fy - cy);
            if (d > 0.9999 * r) { fx = cx + (fx - cx) * (0.9999 * r / d); fy = cy + (fy - cy) * (0.9999 * r / d); }
            g.geom[0] = (float)cx; g.geom[1] = (float)cy; g.geom[2] = (float)fx; g.geom[3] = (float)fy;
            g.geom[4] = (float)r;
        } else {
            g.geom[0] = (float)cx; g.geom[1] = (float)cy; g.geom[2] = (float)fx; g.geom[3] = (float)fy;
        }
        if (!(fin >> t) || t != "dyn_ramp") throw std::runtime_error("rvg: expected 'dyn_ramp'");
        g.first_stop = (uint32_t)c.stops.size();
        while (fin >> t) {
            if (t == "pad" || t == "repeat" || t == "reflect" || t == "transparent") break;
            const size_t k = t.find(':');
            if (k == std::string::npos) throw std::runtime_error("rvg: bad ramp stop '" + t + "'");
            Stop st{};
            st.offset = std::strtof(t.substr(0, k).c_str(), nullptr);
            float sc[4] = {0.0f, 0.0f, 0.0f, 1.0f};
            read_color(t.substr(k + 1), sc);
            st.r = sc[0]; st.g = sc[1]; st.b = sc[2]; st.a = sc[3];
            c.stops.push_back(st);
        }
        g.n_stops = (uint32_t)c.stops.size() - g.first_stop;
        if (g.n_stops == 0) throw std::runtime_error("rvg: empty ramp");
        g.spread = t == "pad" ? kSpreadPad : t == "repeat" ? kSpreadRepeat : t == "reflect" ? kSpreadReflect
                                                                               : kSpreadTransparent;
        if (!(fin >> t)) throw std::runtime_error("rvg: truncated gradient");
        Affine inv;
        if (!scene_xf.after(read_transform(fin, t)).inverse(inv))
            throw std::runtime_error("rvg: singular paint transform");
        for (int k = 0; k < 6; ++k) g.xform[k] = (float)inv.m[k];
        const Stop& s0 = c.stops[g.first_stop];
        col[0] = s0.r; col[1] = s0.g; col[2] = s0.b; col[3] = s0.a;  // the path's representative colour
        c.paint[pi] = (uint32_t)c.paints.size();
        c.paints.push_back(g);
    }
    if (c.path_index < 0) throw std::runtime_error("rvg: no paths");
    float fwid = vz, fhgt = vw;
    uint32_t W = fwid > 0 && fwid < 65536 ? (uint32_t)std::ceil(fwid) : 1024;
    uint32_t H = fhgt > 0 && fhgt < 65536 ? (uint32_t)std::ceil(fhgt) : 1024;
    return flatten(c, W, H);
}

void write_rvg(const Scene& s, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write \"" + path + "\"");
    std::fprintf(f, "viewport 0,0 %u,%u\nwindow 0,0 %u,%u\nscene dyn_identity\n", s.width, s.height,
                 s.width, s.height);
    const uint32_t C = s.n_curves();
    uint32_t c = 0;
    for (uint32_t p = 0; p < s.n_paths(); ++p) {
        std::fprintf(f, "  1 element %s dyn_concrete 0,0 0,0 0,0:", s.rule[p] ? "ofill" : "nzfill");
        float lx = 0, ly = 0;
        bool first = true;
        for (; c < C && s.curve_path[c] == p; ++c) {
            const float* q = &s.pos[2 * (size_t)s.curve_pos[c]];
            if (first || q[0] != lx || q[1] != ly) {
                std::fprintf(f, " M %.9g,%.9g", q[0], q[1]);
                first = false;
            }
            const uint32_t t = s.curve_type[c];
            if (t == 2) {
                std::fprintf(f, " L %.9g,%.9g", q[2], q[3]);
                lx = q[2]; ly = q[3];
            } else if (t == 3) {
                std::fprintf(f, " Q %.9g,%.9g %.9g,%.9g", q[2], q[3], q[4], q[5]);
                lx = q[4]; ly = q[5];
            } else if (t == 0x13u) {
                const double w = s.curve_weight[c];
                std::fprintf(f, " A %.17g,%.17g,%.9g %.9g,%.9g", w * q[2], w * q[3], w, q[4], q[5]);
                lx = q[4]; ly = q[5];
            } else {
                std::fprintf(f, " C %.9g,%.9g %.9g,%.9g %.9g,%.9g", q[2], q[3], q[4], q[5], q[6], q[7]);
                lx = q[6]; ly = q[7];
            }
        }
        const uint32_t pp = s.path_paint.empty() ? kSolidPaint : s.path_paint[p];
        if (pp == kSolidPaint) {
            uint32_t col = s.fill[p];
            auto ch = [&](int i) { return (((col >> (8 * i)) & 255u) + 0.5) / 255.0; };
            std::fprintf(f, " Z dyn_identity dyn_paint 1 solid rgba(%.9g,%.9g,%.9g,%.9g)\n", ch(0), ch(1), ch(2),
                         ch(3));
            continue;
        }
        const Paint& g = s.paints[pp];
        std::fprintf(f, " Z dyn_identity dyn_paint %.9g ", g.opacity);
        if (g.kind == kPaintLinear)
            std::fprintf(f, "linear gradient (%.9g,%.9g) (%.9g,%.9g) dyn_ramp", g.geom[0], g.geom[1], g.geom[2],
                         g.geom[3]);
        else
            std::fprintf(f, "radial gradient (%.9g,%.9g) (%.9g,%.9g) %.9g dyn_ramp", g.geom[0], g.geom[1],
                         g.geom[2], g.geom[3], g.geom[4]);
        for (uint32_t k = 0; k < g.n_stops; ++k) {
            const Stop& st = s.stops[g.first_stop + k];
            std::fprintf(f, " %.9g:rgba(%.9g,%.9g,%.9g,%.9g)", st.offset, st.r, st.g, st.b, st.a);
        }
        static const char* spreads[4] = {"pad", "repeat", "reflect", "transparent"};
        Affine inv, fwd;
        for (int k = 0; k < 6; ++k) inv.m[k] = g.xform[k];
        if (!inv.inverse(fwd)) throw std::runtime_error("write_rvg: singular paint transform");
        std::fprintf(f, " %s dyn_affine([%.17g,%.17g,%.17g],[%.17g,%.17g,%.17g])\n", spreads[g.spread & 3u],
                     fwd.m[0], fwd.m[1], fwd.m[2], fwd.m[3], fwd.m[4], fwd.m[5]);
    }
    std::fclose(f);
}

Scene generate(int kind, uint64_t seed, uint32_t n_paths, uint32_t canvas) {
    Pcg32 r(seed);
    switch (kind) {
        case 0: return gen_scribbles(r, n_paths ? n_paths : 64, canvas ? canvas : 1024);
        case 1: return gen_glyphs(r, n_paths ? n_paths : 20000, canvas ? canvas : 2048);
        case 2: return gen_illustration(r, n_paths ? n_paths : 5000, canvas ? canvas : 4096);
        default: throw std::runtime_error("unknown generator kind");
    }
}

Scene subset(const Scene& s, uint32_t first, uint32_t count) {
    if ((uint64_t)first + count > s.n_paths()) throw std::runtime_error("subset out of range");
    Scene o;
    o.width = s.width; o.height = s.height;
    for (uint32_t p = first; p < first + count; ++p) { o.rule.push_back(s.rule[p]); o.fill.push_back(s.fill[p]); }
    if (!s.path_paint.empty()) {
        o.path_paint.assign(s.path_paint.begin() + first, s.path_paint.begin() + first + count);
        o.paints = s.paints;
        o.stops = s.stops;
    }
    const uint32_t C = s.n_curves();
    for (uint32_t c = 0; c < C; ++c) {
        uint32_t p = s.curve_path[c];
        if (p < first || p >= first + count) continue;
        uint32_t pb = s.curve_pos[c];
        uint32_t pe = c + 1 < C ? s.curve_pos[c + 1] : s.n_points();
        o.curve_pos.push_back(o.n_points());
        o.curve_type.push_back(s.curve_type[c]);
        if (!s.curve_weight.empty()) o.curve_weight.push_back(s.curve_weight[c]);
        o.curve_path.push_back(p - first);
        for (uint32_t q = pb; q < pe; ++q) {
            o.pos.push_back(s.pos[2 * q]); o.pos.push_back(s.pos[2 * q + 1]);
            o.pos_path.push_back(p - first);
        }
    }
    return o;
}

std::string validate(const float* pos, const uint32_t* pos_path, uint32_t P, const uint32_t* cpos,
                     const uint32_t* ctype, const uint32_t* cpath, uint32_t C,
                     const uint32_t* rule, uint32_t Np) {
    (void)pos; (void)rule;
    if ((P && (!pos || !pos_path)) || (C && (!cpos || !ctype || !cpath)) || (Np && !rule))
        return "null input array";
    for (uint32_t i = 0; i < P; ++i)
        if (pos_path[i] >= Np) return "pos_path_idx out of range";
    for (uint32_t c = 0; c < C; ++c) {
        uint32_t t = ctype[c];
        if (t != 2 && t != 3 && t != 4 && t != 0x13u)
            return "curve_type must be 2 (line), 3 (quadratic), 4 (cubic) or 0x13 (arc)";
        if ((uint64_t)cpos[c] + (t & 7) > P) return "curve_pos_map points past the point array";
        if (cpath[c] >= Np) return "curve_path_idx out of range";
        if (c && cpath[c] < cpath[c - 1]) return "curve_path_idx must be non-decreasing";
    }
    return std::string();
}

std::string validate_extras(const uint32_t* ctype, uint32_t C, const float* cweight, uint32_t Np,
                            const uint32_t* path_paint, const Paint* paints, uint32_t n_paints,
                            const Stop* stops, uint32_t n_stops) {
    for (uint32_t c = 0; c < C; ++c) {
        if (ctype[c] != 0x13u) continue;
        if (!cweight) return "arcs need curve weights";
        const float w = cweight[c];
        if (!(w > 0.0f) || !std::isfinite(w)) return "an arc weight must be finite and > 0";
    }
    if ((n_paints && !paints) || (n_stops && !stops)) return "null paint or stop array";
    if (path_paint)
        for (uint32_t p = 0; p < Np; ++p)
            if (path_paint[p] != kSolidPaint && path_paint[p] >= n_paints) return "path_paint out of range";
    auto fin = [](float v) { return std::isfinite(v); };
    for (uint32_t k = 0; k < n_paints; ++k) {
        const Paint& g = paints[k];
        if (g.kind != kPaintLinear && g.kind != kPaintRadial) return "paint kind must be 1 (linear) or 2 (radial)";
        if (g.spread > kSpreadTransparent) return "paint spread must be 0..3";
        if (g.n_stops == 0 || (uint64_t)g.first_stop + g.n_stops > n_stops) return "paint stops out of range";
        if (!fin(g.opacity)) return "non-finite paint opacity";
        for (int i = 0; i < 6; ++i) if (!fin(g.xform[i])) return "non-finite paint transform";
        for (int i = 0; i < 5; ++i) if (!fin(g.geom[i])) return "non-finite paint geometry";
        for (uint32_t i = 0; i < g.n_stops; ++i) {
            const Stop& st = stops[g.first_stop + i];
            if (!fin(st.offset) || !fin(st.r) || !fin(st.g) || !fin(st.b) || !fin(st.a)) return "non-finite stop";
            if (i && st.offset < stops[g.first_stop + i - 1].offset) return "stop offsets must not decrease";
        }
        if (g.kind == kPaintRadial) {
            // the per-pixel solve needs a < 0 (the focus strictly inside the circle), in float
            const float dx = g.geom[0] - g.geom[2], dy = g.geom[1] - g.geom[3], r = g.geom[4];
            const float a = (dx * dx + dy * dy) - r * r;
            if (!(r > 0.0f) || !(a < 0.0f)) return "a radial gradient needs r > 0 and its focus inside the circle";
        }
    }
    return std::string();
}

}  // namespace spr
