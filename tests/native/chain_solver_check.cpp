// Host check of the lane chain solver (ChainSolver in device_common.hpp, the code k_chain_solve
// runs per lane) against the reference's bisection per crossing (bisect_axis,
// make_intersection_1.comp:377-437), bit for bit.  Chains: monotone cubics and quadratics in
// both directions, with fractional, integer and even-integer control values (values landing
// exactly on grid lines, starts on a line, line 0), first grid line after the piece start as
// piece_grid makes it, starts at t = 0 and inside (0, 1/4).  Prints the crossing count, the
// mismatches and the paths taken; exits non-zero on a mismatch.
#include "device_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

using namespace spr;

template <uint32_t T>
static void reference_chain(const CurveAxis& a, float t0, float t1, float line0, float step, int n, int side,
                            float* out) {
    float tprev = t0, line = line0;
    for (int k = 0; k < n; ++k) {
        const float v = bitsf((fbits(bisect_axis<T>(a, line, tprev, t1)) & 0xFFFFFFFCu) | (uint32_t)side);
        out[k] = v;
        tprev = v;
        line += step;
    }
}

template <uint32_t T>
static void solver_chain(const CurveAxis& a, float t0, float t1, float line0, float step, int n, int side,
                         float* out, long* fast_chains) {
    ChainSolver<T> s(a, step, t1);
    *fast_chains += s.fast ? 1 : 0;
    float tprev = t0, line = line0;
    for (int k = 0; k < n; ++k) {
        // both forms the kernel takes: the straight-line one in the usual case, next() otherwise
        const float ts = s.common(tprev) ? s.next_common(line, tprev, t1) : s.next(line, tprev, t1);
        const float v = bitsf((fbits(ts) & 0xFFFFFFFCu) | (uint32_t)side);
        out[k] = v;
        tprev = v;
        line += step;
    }
}

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 100000;
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    static float r[4096], q[4096];
    long total = 0, bad = 0, chains = 0, fast = 0, on_line = 0;
    for (int it = 0; it < iters; ++it) {
        const int kind = it % 4;
        const float span = kind == 0 ? 40.f : (kind == 1 ? 600.f : 8.f);
        CurveAxis a;
        for (int i = 0; i < 4; ++i) {
            float v = U(rng) * span;
            if (kind == 2) v = std::floor(v);
            if (kind == 3) v = 2.f * std::floor(U(rng) * 6.f);
            a.c[i] = v;
        }
        std::sort(a.c, a.c + 4);  // sorted Bernstein coefficients: a monotone piece
        const bool down = (it / 4) % 2;
        if (down) { std::swap(a.c[0], a.c[3]); std::swap(a.c[1], a.c[2]); }
        const bool quad = it % 7 == 0;
        if (quad) a.c[3] = 0.f;
        const float p0 = a.c[0], p1 = quad ? a.c[2] : a.c[3];
        const float step = down ? -2.f : 2.f;
        float line0 = down ? 2.f * std::floor(p0 / 2.f) : 2.f * std::floor(p0 / 2.f) + 2.f;
        if (down && line0 == p0) {
            if (it % 3) line0 -= 2.f;
            else ++on_line;
        }
        int n = 0;
        for (float l = line0; down ? l >= p1 : l <= p1; l += step) ++n;
        if (n <= 0 || n > 4000) continue;
        const int side = it & 1;
        const float t0 = it % 5 == 0 ? 0.f : bitsf(fbits(0.25f * U(rng)) | 2u);
        ++chains;
        if (quad) {
            reference_chain<kQuad>(a, t0, 1.0f, line0, step, n, side, r);
            solver_chain<kQuad>(a, t0, 1.0f, line0, step, n, side, q, &fast);
        } else {
            reference_chain<kCubic>(a, t0, 1.0f, line0, step, n, side, r);
            solver_chain<kCubic>(a, t0, 1.0f, line0, step, n, side, q, &fast);
        }
        for (int k = 0; k < n; ++k) {
            ++total;
            if (fbits(r[k]) != fbits(q[k])) {
                if (bad < 5)
                    std::printf("mismatch: chain %d crossing %d %s %s: %a vs %a\n", it, k, quad ? "quad" : "cubic",
                                down ? "down" : "up", r[k], q[k]);
                ++bad;
            }
        }
    }
    std::printf("chains %ld fast %ld starts_on_line %ld crossings %ld mismatches %ld\n", chains, fast, on_line, total, bad);
    return bad != 0 ? 1 : 0;
}
