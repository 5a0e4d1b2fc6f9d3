// scene.cpp -- RVG parsing, SoA flattening, RVG writing and the synthetic generators.
#include "scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
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
    std::vector<uint32_t> path_curve_begin;
    std::vector<uint32_t> rule;
    std::vector<float> color;         // 4 per path
    std::vector<float> opacity;
    int path_index = -1;

    uint32_t n_points() const { return (uint32_t)(pts.size() / 2); }

    // VGContainer::newPath, vg_container.h:49-56
    void new_path() {
        ++path_index;
        path_curve_begin.push_back((uint32_t)curve_type.size());
        rule.push_back(0);
        color.insert(color.end(), {0.0f, 0.0f, 0.0f, 1.0f});
        opacity.push_back(0.0f);
    }
    // VGContainer::newCurve + addCurve, vg_container.h:58-86
    void add_curve(uint32_t type, const float* p, int npts) {
        curve_begin.push_back(n_points());
        curve_type.push_back(type);
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
            for (uint32_t p = pb; p < pe; ++p) {
                s.pos.push_back(c.pts[2 * p]);
                s.pos.push_back(c.pts[2 * p + 1]);
                s.pos_path.push_back(pi);
            }
        }
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

// RVG::load -> parse_header + parse_paths, src/core/vg/rvg.cpp:9-255.
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
    if (!(fin >> t) || t != "dyn_identity") throw std::runtime_error("rvg: only 'scene dyn_identity' is supported");
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
        uint32_t contour_begin = c.n_points();
        bool closed = false;
        bool have_tok = false;
        while (fin >> t) {
            if (t != "fL" && t.length() != 1) { have_tok = true; break; }
            char cmd = t[0];
            if (std::string("MZLC").find(cmd) == std::string::npos && t != "fL") {
                if (std::string("mzlcAaQq").find(cmd) != std::string::npos)
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
                case 'C':
                    if (!read_point(fin, p[2], p[3]) || !read_point(fin, p[4], p[5]) || !read_point(fin, p[6], p[7]))
                        throw std::runtime_error("rvg: bad point");
                    c.add_curve(4, p, 4);
                    p[0] = p[6]; p[1] = p[7];
                    break;
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
        if (!have_tok || t != "dyn_identity")
            throw std::runtime_error("rvg: unsupported path transform '" + t + "'");
        if (!(fin >> t) || t != "dyn_paint") throw std::runtime_error("rvg: expected 'dyn_paint'");
        float opacity;
        if (!(fin >> opacity)) throw std::runtime_error("rvg: bad opacity");
        c.opacity[pi] = opacity;
        if (!(fin >> t) || t != "solid") throw std::runtime_error("rvg: unsupported paint '" + t + "'");
        if (!(fin >> t)) throw std::runtime_error("rvg: truncated paint");
        float* col = &c.color[4 * pi];
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
        float cx0 = 0, cy0 = 0;
        for (; c < C && s.curve_path[c] == p; ++c) {
            const float* q = &s.pos[2 * (size_t)s.curve_pos[c]];
            if (first || q[0] != lx || q[1] != ly) {
                std::fprintf(f, " M %.9g,%.9g", q[0], q[1]);
                cx0 = q[0]; cy0 = q[1];
                first = false;
            }
            if (s.curve_type[c] == 2) {
                std::fprintf(f, " L %.9g,%.9g", q[2], q[3]);
                lx = q[2]; ly = q[3];
            } else {
                std::fprintf(f, " C %.9g,%.9g %.9g,%.9g %.9g,%.9g", q[2], q[3], q[4], q[5], q[6], q[7]);
                lx = q[6]; ly = q[7];
            }
        }
        (void)cx0; (void)cy0;
        uint32_t col = s.fill[p];
        auto ch = [&](int i) { return (((col >> (8 * i)) & 255u) + 0.5) / 255.0; };
        std::fprintf(f, " Z dyn_identity dyn_paint 1 solid rgba(%.9g,%.9g,%.9g,%.9g)\n", ch(0), ch(1), ch(2), ch(3));
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
    const uint32_t C = s.n_curves();
    for (uint32_t c = 0; c < C; ++c) {
        uint32_t p = s.curve_path[c];
        if (p < first || p >= first + count) continue;
        uint32_t pb = s.curve_pos[c];
        uint32_t pe = c + 1 < C ? s.curve_pos[c + 1] : s.n_points();
        o.curve_pos.push_back(o.n_points());
        o.curve_type.push_back(s.curve_type[c]);
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
        if (t != 2 && t != 4) return "curve_type must be 2 (line) or 4 (cubic)";
        if ((uint64_t)cpos[c] + (t & 7) > P) return "curve_pos_map points past the point array";
        if (cpath[c] >= Np) return "curve_path_idx out of range";
        if (c && cpath[c] < cpath[c - 1]) return "curve_path_idx must be non-decreasing";
    }
    return std::string();
}

}  // namespace spr
