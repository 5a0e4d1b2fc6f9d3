"""Pins the CPU oracle (oracle/spr_oracle.cpp) without trusting it.

Each oracle stage is checked against something it does not share code with: closed forms,
float64 mathematics, brute force on small inputs, an independent numpy formulation of a
stage's definition, invariants of closed curves, and the reference's own record fixtures
(workdir/test_data.csv, workdir/test_data3.csv).  A plausible slip in the oracle (a swapped
axis, an off-by-one on the grid, a sign error in the winding, a transposed image) fails at
least one test here.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import ROOT, RVG_DIR
import geometry as G

GOLDEN = ROOT / "tests" / "golden"


def f32(x):
    return np.float32(x)


def frame_of(orc, arrays, w, h, **kw):
    return orc.Frame(arrays, w, h, **kw)


def per_curve_counts(fr):
    c = fr.get("counts")
    return np.diff(c)


# --- transform_pos.comp:33-85 ------------------------------------------------------------------

def test_transform_identity_and_dyadic_mvp(orc):
    rng = np.random.default_rng(0)
    polys = [rng.uniform(8, 200, size=(7, 2)).round(3) for _ in range(3)]
    a = G.polygon_scene(polys, [0, 1, 0], [0xFF0000FF, 0xFF00FF00, 0xFFFF0000])
    fr = frame_of(orc, a, 256, 256)
    assert np.array_equal(fr.get("tpos").reshape(-1, 2), a["pos"])
    # scale by 2, translate by (3, -5), homogeneous w = 2: every step is exact in float32
    m = np.array([[4, 0, 0, 6], [0, 4, 0, -10], [0, 0, 1, 0], [0, 0, 0, 2]], dtype=np.float32)
    fr = frame_of(orc, a, 256, 256, mvp=m)
    exp = (a["pos"].astype(np.float64) * 2 + np.array([3.0, -5.0])).astype(np.float32)
    assert np.array_equal(fr.get("tpos").reshape(-1, 2), exp)


def test_transform_general_mvp_matches_float64(orc):
    rng = np.random.default_rng(1)
    a = G.polygon_scene([rng.uniform(0, 300, size=(50, 2))], [0], [0xFF000000])
    m = rng.normal(size=(4, 4)).astype(np.float32)
    m[3] = [0.001, -0.002, 0.0, 1.0]
    fr = frame_of(orc, a, 300, 300, mvp=m)
    p = a["pos"].astype(np.float64)
    hom = np.c_[p, np.zeros(len(p)), np.ones(len(p))] @ m.astype(np.float64).T
    exp = hom[:, :2] / hom[:, 3:4]
    np.testing.assert_allclose(fr.get("tpos").reshape(-1, 2), exp, rtol=1e-5, atol=1e-3)


REGION_BITS = {  # (x class, y class) -> bit; class 0: < 0, 1: inside, 2: >= size
    (0, 0): 0x10000000, (1, 0): 0x01000000, (2, 0): 0x00100000,
    (0, 1): 0x00010000, (1, 1): 0x10000001, (2, 1): 0x00001000,
    (0, 2): 0x00000100, (1, 2): 0x00000010, (2, 2): 0x00000001,
}


def test_visibility_is_the_half_plane_rule(orc):
    """A path is culled iff all of its points lie strictly beyond one canvas edge."""
    cells = list(REGION_BITS)
    for subset in range(1 << 9):
        chosen = [cells[i] for i in range(9) if subset >> i & 1]
        mask = 0
        for c in chosen:
            mask |= REGION_BITS[c]
        culled = (all(cx == 0 for cx, _ in chosen) or all(cx == 2 for cx, _ in chosen)
                  or all(cy == 0 for _, cy in chosen) or all(cy == 2 for _, cy in chosen))
        assert orc.path_visible(mask) == (not culled), (subset, hex(mask))


def test_transform_outcodes_match_regions(orc):
    W, H = 100, 80
    coord = {0: -5.0, 1: 40.0, 2: 130.0}
    coord_y = {0: -5.0, 1: 40.0, 2: 90.0}
    polys = [[(coord[cx], coord_y[cy]), (coord[cx], coord_y[cy]), (coord[cx], coord_y[cy])]
             for (cx, cy) in REGION_BITS]
    a = G.polygon_scene(polys, [0] * 9, [0xFF000000] * 9)
    fr = frame_of(orc, a, W, H)
    vis = fr.get("visible")
    for p, key in enumerate(REGION_BITS):
        assert vis[p] == np.int32(np.uint32(REGION_BITS[key])), key
    # boundary classes: x = 0 is inside, x = W is outside
    a = G.polygon_scene([[(0.0, 0.0)] * 3, [(float(W), float(H))] * 3], [0, 0], [0, 0])
    vis = frame_of(orc, a, W, H).get("visible")
    assert vis[0] == np.int32(np.uint32(0x10000001)) and vis[1] == 0x00000001


# --- make_intersection_0.comp: solveQuadEquation, monotone cuts, count --------------------------

def test_solve_quad_matches_float64_roots(orc):
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(3000):
        a, b, c = (f32(v) for v in rng.normal(size=3) * rng.choice([0.01, 1, 100], size=3))
        r0, r1 = orc.solve_quad(a, b, c)
        disc = float(b) ** 2 - 4 * float(a) * float(c)
        if disc <= 0:
            # B^2 - AC in float32 may round to either side of zero only when it is tiny
            if disc < -1e-3 * float(b) ** 2:
                assert r0 == 0 and r1 == 0
            continue
        true = np.roots([float(a), float(b), float(c)]).real
        for r in (r0, r1):
            err = np.min(np.abs(true - float(r)) / np.maximum(1.0, np.abs(true)))
            assert err < 2e-4, (a, b, c, r0, r1, true)
        assert abs(float(r0) - float(r1)) > 0 or abs(true[0] - true[1]) < 1e-3 * max(1, abs(true[0]))
        checked += 1
    assert checked > 1000
    # linear case and the double-root / no-root cases
    assert orc.solve_quad(0.0, 2.0, -3.0) == (f32(1.5), f32(1.5))
    assert orc.solve_quad(1.0, 2.0, 1.0) == (f32(0.0), f32(0.0))
    assert orc.solve_quad(1.0, 0.0, 1.0) == (f32(0.0), f32(0.0))
    r = sorted(orc.solve_quad(1.0, -3.0, 2.0))
    assert r == [f32(1.0), f32(2.0)]


def _deriv_roots(p):
    """Roots in (0, 1) of the derivative of a cubic Bezier coordinate (float64)."""
    p0, p1, p2, p3 = (float(v) for v in p)
    a = 3 * (-p0 + 3 * p1 - 3 * p2 + p3)
    b = 6 * (p0 - 2 * p1 + p2)
    c = 3 * (p1 - p0)
    if abs(a) < 1e-12:
        return [] if abs(b) < 1e-12 else [r for r in [-c / b] if 0 < r < 1]
    d = b * b - 4 * a * c
    if d <= 0:
        return []
    s = np.sqrt(d)
    return [r for r in ((-b - s) / (2 * a), (-b + s) / (2 * a)) if 0 < r < 1]


def test_monotone_cuts_are_the_extrema_and_q3_quirk(orc):
    rng = np.random.default_rng(3)
    unsorted_fourth = 0
    for _ in range(4000):
        xs = rng.uniform(0, 100, 4).astype(np.float32)
        ys = rng.uniform(0, 100, 4).astype(np.float32)
        n, q = orc.curve_cuts(4, xs, ys)
        roots = sorted(_deriv_roots(xs) + _deriv_roots(ys))
        # roots within 1e-4 of 0 or 1, or pairs closer than 1e-4, are decided by rounding
        fragile = any(r < 1e-4 or r > 1 - 1e-4 for r in roots) or any(np.diff(roots) < 1e-4)
        if fragile:
            continue
        assert n == len(roots), (xs, ys, q, roots)
        np.testing.assert_allclose(np.sort(q[:n]), roots, atol=2e-4)
        assert np.all(q[n:] == 0)
        if n <= 3:
            assert np.all(np.diff(q[:n]) >= 0)
        else:
            # the reference's 4-cut insertion step compares the 4th cut with itself, so the
            # first three are sorted and the 4th stays where it was discovered (the second
            # y extremum, i.e. the larger one in the order the roots are found)
            assert np.all(np.diff(q[:3]) >= 0)
            y_roots = sorted(_deriv_roots(ys))
            assert abs(q[3] - y_roots[0]) < 2e-4 or abs(q[3] - y_roots[1]) < 2e-4
            unsorted_fourth += q[3] < q[2]
    assert unsorted_fourth > 10, "the 4-cut quirk must leave some 4th cuts out of order"
    # lines are never cut; invisible curves are not cut either
    assert orc.curve_cuts(2, [0, 10, 0, 0], [0, 5, 0, 0])[0] == 0
    assert orc.curve_cuts(4, [0, 100, -100, 0], [0, 0, 0, 0], visible=False)[0] == 0


def test_monotone_pieces_by_sampling(orc):
    rng = np.random.default_rng(4)
    for _ in range(300):
        xs = rng.uniform(0, 100, 4).astype(np.float32)
        ys = rng.uniform(0, 100, 4).astype(np.float32)
        n, q = orc.curve_cuts(4, xs, ys)
        if n == 4:
            continue
        ts = np.concatenate([[0.0], q[:n], [1.0]])
        for t0, t1 in zip(ts[:-1], ts[1:]):
            tt = np.linspace(t0, t1, 200)[3:-3]
            pts = G.bezier(np.c_[xs, ys], tt)
            for ax in range(2):
                d = np.diff(pts[:, ax])
                assert np.all(d >= -1e-6) or np.all(d <= 1e-6)


def even_in(lo, hi, top):
    """#{even k : lo < k <= hi, 0 <= k <= top} for lo <= hi."""
    ks = np.arange(0, top + 1, 2)
    return int(np.sum((ks > lo) & (ks <= hi)))


def test_count_lines_brute_force(orc):
    rng = np.random.default_rng(5)
    W, H = 97, 64
    segs = []
    for _ in range(400):
        p0 = rng.uniform(-40, 140, 2).round(rng.integers(0, 3))
        p1 = rng.uniform(-40, 140, 2).round(rng.integers(0, 3))
        segs.append((p0, p1))
    segs.append((np.array([4.0, 4.0]), np.array([10.0, 10.0])))  # starts exactly on grid lines
    segs.append((np.array([10.0, 3.0]), np.array([4.0, 3.0])))   # horizontal, right to left
    polys = [[tuple(p0), tuple(p1)] for p0, p1 in segs]
    a = G.polygon_scene(polys, [0] * len(polys), [0] * len(polys), close=False)
    fr = frame_of(orc, a, W, H)
    cnt = per_curve_counts(fr)
    vis = fr.get("visible")
    xmax, ymax = (W & ~1) + 2, (H & ~1) + 2
    for c, (p0, p1) in enumerate(segs):
        if not orc.path_visible(int(vis[c])):
            assert cnt[c] == 0
            continue
        # the piece end is the curve evaluated at t = 1, p0 + 1 * (p1 - p0) in float32, which
        # can miss p1 by one ulp (make_intersection_0.comp:388-391)
        e = [np.float32(p0[ax]) + np.float32(1) * (np.float32(p1[ax]) - np.float32(p0[ax])) for ax in (0, 1)]
        x0, x1 = sorted((float(np.float32(p0[0])), float(e[0])))
        y0, y1 = sorted((float(np.float32(p0[1])), float(e[1])))
        exp = 1 + even_in(x0, x1, xmax) + even_in(y0, y1, ymax)
        assert cnt[c] == exp, (c, p0, p1, cnt[c], exp)


def test_count_cubics_float64(orc):
    """Count = sum over monotone pieces of 1 + grid lines crossed on each axis."""
    rng = np.random.default_rng(6)
    W, H = 200, 160
    pts, checked = [], 0
    for _ in range(300):
        pts.append(rng.uniform(-20, 220, size=(4, 2)))
    pos = np.concatenate(pts).astype(np.float32)
    a = dict(pos=pos, pos_path=np.repeat(np.arange(300), 4).astype(np.uint32),
             curve_pos=(np.arange(300) * 4).astype(np.uint32), curve_type=np.full(300, 4, np.uint32),
             curve_path=np.arange(300, dtype=np.uint32), rule=np.zeros(300, np.uint32),
             fill=np.zeros(300, np.uint32))
    fr = frame_of(orc, a, W, H)
    cnt = per_curve_counts(fr)
    vis = fr.get("visible")
    cuts = fr.get("cuts").reshape(-1, 5)
    for c in range(300):
        if not orc.path_visible(int(vis[c])):
            assert cnt[c] == 0
            continue
        n = int(cuts[c, 4:5].view(np.uint32)[0])
        ts = np.concatenate([[0.0], cuts[c, :n].astype(np.float64), [1.0]])
        ends = G.bezier(pos[4 * c:4 * c + 4], ts)
        if np.any(np.abs(ends / 2 - np.round(ends / 2)) < 1e-3):
            continue  # a piece end within 1e-3 of a grid line: decided by float32 rounding
        exp = 0
        for e0, e1 in zip(ends[:-1], ends[1:]):
            exp += 1 + even_in(min(e0[0], e1[0]), max(e0[0], e1[0]), (W & ~1) + 2)
            exp += even_in(min(e0[1], e1[1]), max(e0[1], e1[1]), (H & ~1) + 2)
        assert cnt[c] == exp, c
        checked += 1
    assert checked > 150


def test_float2int_rd(orc):
    for v in [0.0, 0.5, 1.0, 1.9999, 2.0, 7.25, 1e6 + 0.5]:
        assert orc.lib().oracle_float2int_rd(v) == int(np.floor(v))
    for v in [-0.5, -1.5, -2.25, -1000.75]:
        assert orc.lib().oracle_float2int_rd(v) == int(np.floor(v))
    # the reference's float2int_rd subtracts 1 before truncating, so negative integers go one
    # lower than floor (make_intersection_1.comp:76-79)
    for v in [-1.0, -2.0, -64.0]:
        assert orc.lib().oracle_float2int_rd(v) == int(v) - 1


# --- make_intersection_1.comp: the written crossings ---------------------------------------------

def _line_scene(segs):
    polys = [[tuple(p0), tuple(p1)] for p0, p1 in segs]
    return G.polygon_scene(polys, [0] * len(polys), [0] * len(polys), close=False)


def test_write_lines_exact_crossings(orc):
    """For a line the crossings are t = (k - x0) / (x1 - x0) on every even grid line k between
    the ends (and the same for y), written in increasing t after the start entry."""
    rng = np.random.default_rng(7)
    W, H = 120, 90
    segs = [(rng.uniform(1, 119, 2), rng.uniform(1, 89, 2)) for _ in range(150)]
    segs = [(np.array([p0[0], p1[0]]), np.array([p0[1], p1[1]])) for p0, p1 in segs]
    a = _line_scene(segs)
    fr = frame_of(orc, a, W, H)
    inter = fr.get("inter").reshape(-1, 2)
    counts = fr.get("counts")
    for c, (p0, p1) in enumerate(segs):
        e = inter[counts[c]:counts[c + 1]]
        assert np.all(e[:, 0] == c)
        t = (e[:, 1].view(np.uint32) & np.uint32(0xFFFFFFFC)).view(np.float32).astype(np.float64)
        assert t[0] == 0.0
        assert np.all(np.diff(t) >= 0)
        exp = []
        for ax in range(2):
            lo, hi = sorted((float(np.float32(p0[ax])), float(np.float32(p1[ax]))))
            d = float(np.float32(p1[ax])) - float(np.float32(p0[ax]))
            for k in range(0, 200, 2):
                if lo < k <= hi and d != 0:
                    exp.append((k - float(np.float32(p0[ax]))) / d)
        np.testing.assert_allclose(t[1:], np.sort(exp), atol=1e-5)


def test_write_cubic_crossings_land_on_grid(orc):
    s = None
    import vkscanlinepr_amd as V
    s = V.Scene.generate(0, 21, 6, 256)
    a = s.arrays()
    fr = frame_of(orc, a, 256, 256)
    inter = fr.get("inter").reshape(-1, 2)
    pos = a["pos"].astype(np.float64)
    bits = inter[:, 1].view(np.uint32)
    t = (bits & np.uint32(0xFFFFFFFC)).view(np.float32).astype(np.float64)
    side = bits & 3
    good = total = 0
    for i in range(len(inter)):
        if side[i] > 1 or t[i] == 0.0:
            continue  # piece starts carry cut flags; t = 0 is a curve start
        c = inter[i, 0]
        k = int(a["curve_type"][c])
        pt = G.bezier(pos[a["curve_pos"][c]:a["curve_pos"][c] + k], [t[i]])[0]
        v = pt[int(side[i])]
        total += 1
        good += abs(v / 2 - round(v / 2)) * 2 < 0.02
    assert total > 1000 and good / total > 0.995


def _monotone_cubics(rng, n, span=64.0):
    """n cubics whose x runs monotonically over about `span` px (control x non-decreasing)."""
    out = []
    for _ in range(n):
        x0 = rng.uniform(0.0, 200.0)
        xs = np.sort(x0 + rng.uniform(0.0, span, 4))
        ys = rng.uniform(0.0, 200.0, 4)
        out.append((xs.astype(np.float32), ys.astype(np.float32)))
    return out


def test_write_cubic_bisection_runs_24_halvings(orc):
    """CUBIC_ITERATION_NUMBER = 24 (make_intersection_1.comp:7), pinned by a closed form: on the
    bracket [0, 1] every midpoint is a dyadic rational float32 holds exactly, and the result is the
    last midpoint, (t0 + t1) / 2 with t1 - t0 = 2^-(K-1) after K - 1 halvings: t * 2^K is an odd
    integer.  A 23-step (or 20-step) bisection leaves t * 2^24 even; 25 steps leave it fractional.
    The float64 root lies within the last bracket, up to the float32 evaluation's rounding."""
    rng = np.random.default_rng(24)
    odd = checked = 0
    worst = 0.0
    for xs, ys in _monotone_cubics(rng, 400):
        lo, hi = float(xs[0]), float(xs[3])
        k0 = int(np.ceil(lo / 2.0)) * 2
        for c in range(k0, int(hi), 2):
            if not lo < c < hi:
                continue
            t = orc.solve_crossing(4, xs, ys, 0, float(c), 0.0, 1.0)
            if abs(float(G.bezier(np.stack([xs, ys], 1), [float(t)])[0][0]) - c) > 1.0:
                continue  # the 1 px fallback (pinned below)
            m = float(t) * 2.0 ** 24
            assert m == np.floor(m), (t, c)  # a multiple of 2^-24: at least 24 halvings
            checked += 1
            odd += int(m) % 2 == 1
            # float64 root of x(t) = c on the monotone curve, by bisection to 1e-15
            a, b = 0.0, 1.0
            ctrl = np.stack([xs, ys], 1).astype(np.float64)
            for _ in range(60):
                mid = 0.5 * (a + b)
                if G.bezier(ctrl, [mid])[0][0] < c:
                    a = mid
                else:
                    b = mid
            worst = max(worst, abs(float(t) - a) * 2.0 ** 24)
    assert checked > 2000
    assert odd == checked  # exactly 24: a 23-step or 20-step bisection gives even multiples
    # the last bracket is 2^-24 wide, but the float32 sign test near the root is only as good as
    # the evaluation's rounding (a few ulp of a ~200 px coordinate) over the local slope, which is
    # small near a monotone cubic's flat ends: the distance to the float64 root is a sanity bound
    assert worst < 2.0 ** 8, worst


def test_write_cubic_fallback_beyond_one_pixel(orc):
    """make_intersection_1.comp:430-433: when the last midpoint's value misses the grid line by more
    than 1 px, the solve returns the bracket start t_min; within 1 px it keeps the midpoint."""
    xs = np.array([0.0, 3.0, 7.0, 10.0], np.float32)  # x runs 0 -> 10 monotonically
    ys = np.array([0.0, 1.0, 2.0, 3.0], np.float32)
    # no crossing in the bracket: the search runs to t1 and ends 40 px away -> t_min exactly
    assert orc.solve_crossing(4, xs, ys, 0, 50.0, 0.25, 1.0) == np.float32(0.25)
    # 0.9 px beyond the curve's end: within 1 px, the last midpoint (just below t1) is kept
    t = orc.solve_crossing(4, xs, ys, 0, 10.9, 0.25, 1.0)
    assert 1.0 - 1e-6 < t < 1.0
    # 1.1 px beyond: the fallback
    assert orc.solve_crossing(4, xs, ys, 0, 11.1, 0.25, 1.0) == np.float32(0.25)
    # the line at the bracket start itself: t_min without any search
    assert orc.solve_crossing(4, xs, ys, 0, 0.0, 0.0, 1.0) == np.float32(0.0)


def test_write_corner_crossings_are_or_merged(orc):
    """A diagonal through grid corners crosses x = k and y = k at the same t.  The reference
    ORs the side bits of equal-t neighbours into both entries (make_intersection_1.comp:
    400-416), so every crossing entry carries bits 0|1 = 1 and they come in equal pairs."""
    a = _line_scene([(np.array([1.0, 1.0]), np.array([11.0, 11.0]))])
    fr = frame_of(orc, a, 64, 64)
    v = fr.get("inter").reshape(-1, 2)[:, 1].view(np.uint32)
    assert v[0] == 0  # the curve start, t = 0
    cross = v[1:]
    assert len(cross) == 10  # x = 2..10 and y = 2..10 on even lines: 5 + 5
    assert np.all(cross & 3 == 1)
    assert np.all(cross[0::2] == cross[1::2])
    t = (cross[0::2] & np.uint32(0xFFFFFFFC)).view(np.float32)
    np.testing.assert_allclose(t, (np.arange(2, 11, 2) - 1.0) / 10.0, atol=1e-6)


def test_write_piece_start_crossing_keeps_cut_flag(orc):
    """x(t) of the cubic (0, 16, 16, 0) peaks at exactly x = 12 at t = 1/2, an even grid line.
    The second monotone piece starts there (cut flag 2: the piece end is an integer x) and its
    first x crossing is found at that same t, so the OR merge gives both entries the flag
    bits 2 | 0 = 2 (make_intersection_1.comp:400-416)."""
    a = dict(pos=np.array([[0, 1], [16, 1], [16, 1], [0, 1]], np.float32), pos_path=np.zeros(4, np.uint32),
             curve_pos=np.array([0], np.uint32), curve_type=np.array([4], np.uint32),
             curve_path=np.array([0], np.uint32), rule=np.array([0], np.uint32), fill=np.array([0], np.uint32))
    v = frame_of(orc, a, 32, 32).get("inter").reshape(-1, 2)[:, 1].view(np.uint32)
    assert len(v) == 1 + 6 + 1 + 6  # start, x = 2..12, piece start, x = 12..2
    t = (v & np.uint32(0xFFFFFFFC)).view(np.float32)
    at_half = np.nonzero(np.abs(t - 0.5) < 1e-6)[0]
    assert len(at_half) == 2 and np.all(v[at_half] & 3 == 2) and v[at_half[0]] == v[at_half[1]]
    assert np.all(v[(np.abs(t - 0.5) >= 1e-6)] & 3 == 0)


def test_winding_sign_known_answers(orc):
    """gen_fragment.comp:196-205: a fragment that crosses its cell row's middle line y+1
    going up (towards larger y) changes the winding by -1, going down by +1, and a
    fragment that stays on one side (or is horizontal) by 0."""
    for (y0, y1), exp in (((0.5, 1.5), -1), ((1.5, 0.5), 1), ((0.5, 0.9), 0), ((1.0, 1.0), 0)):
        a = _line_scene([(np.array([4.5, y0]), np.array([4.5, y1]))])
        fr = frame_of(orc, a, 16, 16)
        wn = fr.get("wn")
        assert wn.sum() == exp, (y0, y1, wn)


# --- gen_fragment.comp ---------------------------------------------------------------------------

def scenes_small():
    import vkscanlinepr_amd as V
    out = [("cubic_loops", G.cubic_loops(30, 12, 256, 256), 256, 256)]
    for kind, seed, n, w in [(0, 31, 6, 256), (1, 32, 40, 256), (2, 33, 30, 320)]:
        a = V.Scene.generate(kind, seed, n, w).arrays()
        out.append((f"kind{kind}", a, a["width"], a["height"]))
    a = V.Scene.load_rvg(RVG_DIR / "test.rvg").arrays()
    out.append(("test.rvg", a, a["width"], a["height"]))
    return out


@pytest.mark.parametrize("name,arrays,w,h", scenes_small())
def test_fragment_keys_segments_and_closed_winding(orc, name, arrays, w, h):
    fr = frame_of(orc, arrays, w, h)
    inter = fr.get("inter").reshape(-1, 2)
    key = fr.get("key")
    wn = fr.get("wn")
    seg = fr.get("seg")
    path = fr.get("path")
    N = len(inter)
    assert len(key) == N == len(wn)
    # keys: sentinel or (y + 0x7FFF) << 16 | (x + 0x7FFF) with even y in the canvas
    live = key != np.int32(np.uint32(0xFFFEFFFE))
    ku = key[live].view(np.uint32)
    y = (ku >> 16).astype(np.int64) - 0x7FFF
    x = (ku & 0xFFFF).astype(np.int64) - 0x7FFF
    assert np.all((y >= 0) & (y < h) & (y % 2 == 0))
    assert np.all((x >= -2) & (x <= (w & ~1) + 2) & (x % 2 == 0))
    assert set(np.unique(wn)) <= {-1, 0, 1}
    # segment offsets: first entry of each path in the intersection list
    cp = arrays["curve_path"][inter[:, 0]].astype(np.int64)
    assert np.array_equal(path, cp)
    exp_seg = np.searchsorted(cp, np.arange(len(arrays["rule"]) + 1), side="left")
    exp_seg[-1] = N
    assert np.array_equal(seg, exp_seg)
    # every path is a set of closed contours, so each cell row's middle line is crossed up and
    # down equally often: the winding changes of a path sum to zero (for paths inside the
    # canvas whose cubics avoid the 4-cut case, which traverses part of a curve twice)
    sums = np.bincount(cp, weights=wn, minlength=len(arrays["rule"]))
    reg = G.regular_paths(arrays, w, h)
    assert np.all(sums[reg] == 0), np.nonzero(sums * reg)[0][:10]


# --- naive_seg_sort_pairs.comp + shuffle + winding scan ------------------------------------------

@pytest.mark.parametrize("name,arrays,w,h", scenes_small())
def test_sort_and_winding_independent(orc, name, arrays, w, h):
    fr = frame_of(orc, arrays, w, h)
    key, wn, seg = fr.get("key"), fr.get("wn"), fr.get("seg")
    skey, sidx, winding = fr.get("skey"), fr.get("sidx"), fr.get("winding")
    exp_idx = np.arange(len(key), dtype=np.int32)
    for s in range(len(seg) - 1):
        b, e = seg[s], seg[s + 1]
        order = np.lexsort((np.arange(b, e), key[b:e]))  # signed key, then index
        exp_idx[b:e] = b + order
    assert np.array_equal(sidx, exp_idx)
    assert np.array_equal(skey, key[exp_idx])
    assert np.array_equal(winding[:-1], np.concatenate([[0], np.cumsum(wn[exp_idx])[:-1]]))
    assert winding[-1] == wn.sum()


# --- mark + gen_merged + draw, end to end against float64 winding numbers -----------------------

def _star(cx, cy, r, k=5):
    ang = np.pi / 2 + np.arange(k) * 4 * np.pi / k
    return np.c_[cx + r * np.cos(ang), cy + r * np.sin(ang)]


def _square(x0, y0, s, ccw=True):
    v = np.array([[x0, y0], [x0 + s, y0], [x0 + s, y0 + s], [x0, y0 + s]], dtype=np.float64)
    return v if ccw else v[::-1]


def analytic_cases():
    cases = []
    # pentagram: the centre has winding 2 -> filled under nonzero, empty under even-odd
    for rule in (G.NONZERO, G.EVEN_ODD):
        cases.append((f"star_rule{rule}", G.polygon_scene([_star(101.3, 97.7, 80.2)], [rule], [0xFF2040C0]), 200, 200))
    # square with a hole: same orientation vs opposite orientation, both rules
    for rule in (G.NONZERO, G.EVEN_ODD):
        for ccw in (True, False):
            outer = _square(20.5, 30.25, 150.0)
            inner = _square(70.75, 80.5, 50.0, ccw=ccw)
            pos = np.concatenate([outer, outer[:1], inner, inner[:1]]).astype(np.float32)
            a = G.polygon_scene([outer], [rule], [0xFF00A000])
            # one path, two contours
            a["pos"] = pos
            a["pos_path"] = np.zeros(len(pos), np.uint32)
            a["curve_pos"] = np.array([0, 1, 2, 3, 5, 6, 7, 8], np.uint32)
            a["curve_type"] = np.full(8, 2, np.uint32)
            a["curve_path"] = np.zeros(8, np.uint32)
            cases.append((f"hole_rule{rule}_ccw{int(ccw)}", a, 211, 203))
    # overlapping opaque polygons: the later path wins
    rng = np.random.default_rng(8)
    polys = [rng.uniform(10, 180, size=(6, 2)) for _ in range(5)]
    cols = [0xFF0000FF, 0xFF00FF00, 0xFFFF0000, 0xFF00FFFF, 0xFFFF00FF]
    cases.append(("overlap_polys", G.polygon_scene(polys, [0, 1, 0, 1, 0], cols), 190, 190))
    return cases


@pytest.mark.parametrize("name,arrays,w,h", analytic_cases())
def test_image_matches_winding_numbers(orc, name, arrays, w, h):
    fr = frame_of(orc, arrays, w, h)
    img = fr.image()
    exp, far = G.expected_overwrite_image(arrays, w, h)
    assert far.sum() > 0.5 * w * h
    bad = np.any(img != exp, axis=2) & far
    assert not bad.any(), (name, np.argwhere(bad)[:5])


def test_image_matches_winding_numbers_generated(orc, spr):
    glyphs = spr.Scene.generate(1, 42, 12, 256).arrays()
    glyphs = G.select_paths(glyphs, G.regular_paths(glyphs, 256, 256))
    glyphs["fill"] = (glyphs["fill"] | np.uint32(0xFF000000)).astype(np.uint32)  # opaque
    for kind, a in (("loops", G.cubic_loops(41, 8, 256, 256)), ("glyphs", glyphs)):
        assert len(a["rule"]) >= 3
        fr = frame_of(orc, a, 256, 256)
        exp, far = G.expected_overwrite_image(a, 256, 256)
        bad = np.any(fr.image() != exp, axis=2) & far
        assert far.sum() > 10000
        assert not bad.any(), (kind, np.argwhere(bad)[:5])


def test_star_centre_depends_on_rule(orc):
    for rule, filled in ((G.NONZERO, True), (G.EVEN_ODD, False)):
        a = G.polygon_scene([_star(100.0, 100.0, 80.0)], [rule], [0xFF0000FF])
        img = frame_of(orc, a, 200, 200).image()
        centre = img[200 - 1 - 100, 100]
        assert (tuple(centre) == (255, 0, 0, 255)) == filled


def test_orientation_and_rule_invariances(orc, spr):
    """Reversing every contour flips every winding sign: the image must not change."""
    a = G.cubic_loops(51, 6, 256, 256)
    rev = dict(a)
    pos = a["pos"].copy()
    # reverse point order inside each path (curves reversed accordingly)
    for p in range(len(a["rule"])):
        idx = np.nonzero(a["pos_path"] == p)[0]
        pos[idx] = pos[idx[::-1]]
    rev["pos"] = pos
    i0 = frame_of(orc, a, 256, 256).image()
    i1 = frame_of(orc, rev, 256, 256).image()
    diff = np.any(i0 != i1, axis=2)
    exp, far = G.expected_overwrite_image(dict(a, fill=a["fill"] | np.uint32(0xFF000000)), 256, 256)
    assert not (diff & far).any()


@pytest.mark.parametrize("name,arrays,w,h", scenes_small())
def test_record_invariants(orc, name, arrays, w, h):
    fr = frame_of(orc, arrays, w, h)
    rec = fr.get("records").reshape(-1, 4)
    skey, path = fr.get("skey"), fr.get("path")
    live = skey != np.int32(np.uint32(0xFFFEFFFE))
    ku = skey.view(np.uint32)
    ys = (ku >> 16).astype(np.int64) - 0x7FFF
    xs = (ku & 0xFFFF).astype(np.int64) - 0x7FFF
    inside = live & (xs >= 0) & (xs < w) & (ys >= 0) & (ys < h)
    # one cell record per distinct in-canvas (path, cell)
    n_cells = len(np.unique(np.c_[path[inside], skey[inside]], axis=0))
    cells = rec[:, 3] != 0
    assert cells.sum() == n_cells
    assert np.all(rec[cells, 1] == 2)
    assert np.array_equal(rec[cells, 3], np.arange(1, n_cells + 1))
    assert np.all(rec[~cells, 1] > 0)
    _spans_bridge_cells(rec, w)
    # gen_merged_fragment_and_span.comp:64-100 writes the cell of fragment f first and the span
    # that ends at that cell right after it
    x = rec[:, 0] & 0xFFFF
    for i in np.nonzero(~cells)[0]:
        assert i > 0 and cells[i - 1] and rec[i - 1, 0] >> 16 == rec[i, 0] >> 16
        assert x[i] + rec[i, 1] == x[i - 1]


def _spans_bridge_cells(rec, width):
    """Each span (index 0) fills the gap between two cells of its row with its own colour:
    a cell ends where it starts (unless the span was clamped at x = 0) and a cell starts where
    it ends (unless that cell lies right of the canvas, where cells are not emitted)."""
    cells = rec[:, 3] != 0
    y, x = rec[:, 0] >> 16, rec[:, 0] & 0xFFFF
    have = set(zip(y[cells].tolist(), x[cells].tolist(), rec[cells, 2].tolist()))
    n = 0
    for i in np.nonzero(~cells)[0]:
        if rec[i, 1] >= 1200:
            continue  # full-width background rows of the recorded fixture
        assert x[i] == 0 or (y[i], x[i] - 2, rec[i, 2]) in have, i
        assert x[i] + rec[i, 1] >= width or (y[i], x[i] + rec[i, 1], rec[i, 2]) in have, i
        n += 1
    return n


# --- the draw stage against the reference's own record fixtures ----------------------------------

def _paint(records, w, h):
    img = np.full((h, w, 4), 255, dtype=np.uint8)
    for yx, width, fill, _ in records:
        x, y = int(yx) & 0xFFFF, int(yx) >> 16
        col = np.frombuffer(np.int32(fill).tobytes(), dtype=np.uint8)
        for yy in (y, y + 1):
            if 0 <= yy < h:
                img[h - 1 - yy, max(x, 0):min(x + int(width), w)] = col
    return img


def _fixture(name):
    if name == "test_data.csv":
        return np.loadtxt(GOLDEN / name, delimiter=",", dtype=np.int64).astype(np.int32)
    return np.load(GOLDEN / name)["records"]


@pytest.mark.parametrize("name", ["test_data.csv", "test_data3.npz"])
def test_reference_record_fixture_structure(name):
    """The reference's recorded outputs have the record layout the oracle emits: cells are
    2 wide and numbered 1, 2, 3, ... in output order; spans carry index 0 and bridge two
    cells of their row.  (These lists were recorded with the span placed between its two
    cells, where the shader in workdir/shaders writes it after the cell that ends it; the
    two orders draw the same pixels because a span never overlaps a cell of its path.)"""
    rec = _fixture(name)
    cells = rec[:, 3] != 0
    assert np.all(rec[cells, 1] == 2)
    assert np.array_equal(rec[cells, 3], np.arange(1, cells.sum() + 1))
    assert _spans_bridge_cells(rec, 1200) > 100


@pytest.mark.parametrize("name", ["test_data.csv", "test_data3.npz"])
def test_draw_reference_fixture(orc, name):
    """The reference draws these lists on a 1200x1024 target (scanlinepr.vert:37)."""
    rec = _fixture(name)
    img = orc.draw(rec, 1200, 1024)
    assert np.array_equal(img, _paint(rec, 1200, 1024))


def test_draw_src_over_closed_forms(orc):
    # opaque colour replaces, alpha 0 leaves white, 50% black over white
    r = np.array([[(4 << 16) | 10, 6, 0xFF336699 - (1 << 32), 1]], dtype=np.int64).astype(np.int32)
    img = orc.draw(r, 32, 16, compose=1)
    assert tuple(img[16 - 1 - 4, 10]) == (0x99, 0x66, 0x33, 0xFF)
    assert tuple(img[16 - 1 - 5, 15]) == (0x99, 0x66, 0x33, 0xFF)
    assert tuple(img[16 - 1 - 4, 16]) == (255, 255, 255, 255)
    r = np.array([[(2 << 16) | 0, 4, 0x00123456, 1]], dtype=np.int32)
    assert np.all(orc.draw(r, 8, 8, compose=1) == 255)
    r = np.array([[(0 << 16) | 0, 2, 0x80000000 - (1 << 32), 1]], dtype=np.int64).astype(np.int32)
    px = orc.draw(r, 4, 4, compose=1)[3, 0]
    a = 128 / 255
    assert tuple(px[:3]) == (int((1 - a) * 255 + 0.5),) * 3 and px[3] == 255
    # two records: order matters for colour, each applied once
    r = np.array([[(0 << 16) | 0, 2, 0x80FF0000 - (1 << 32), 1],
                  [(0 << 16) | 0, 2, 0x800000FF - (1 << 32), 2]], dtype=np.int64).astype(np.int32)
    px = orc.draw(r, 2, 2, compose=1)[1, 0].astype(np.float64)
    d = np.ones(4)
    for s in (np.array([0, 0, 1.0]), np.array([1.0, 0, 0])):
        d[:3] = s * a + d[:3] * (1 - a)
        d[3] = a + d[3] * (1 - a)
    assert np.all(np.abs(px - np.floor(d * 255 + 0.5)) <= 1)


# --- committed golden vectors ------------------------------------------------------------------

def test_oracle_reproduces_golden_vectors(orc, spr):
    import make_golden
    g = np.load(GOLDEN / "oracle_golden.npz")
    for name, arrays, w, h, mvp, compose in make_golden.cases(spr):
        fr = frame_of(orc, arrays, w, h, mvp=mvp, compose=compose)
        for stage in make_golden.STAGES:
            k = f"{name}/{stage}"
            if k in g:
                assert np.array_equal(fr.get(stage), g[k]), k
        assert make_golden.digest(fr.image()) == str(g[f"{name}/image_sha256"]), name
