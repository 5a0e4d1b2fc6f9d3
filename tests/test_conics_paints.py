"""Pins of the quadratic / arc curves and gradient paints (CPU).

The reference declares QUADRIC and ARC but leaves their shader cases TODO, and draws solid
colours only, so these parts of the oracle restate the library's documented definitions
(include/spr.h).  They are pinned here against what the mathematics fixes, computed in float64
independently of the oracle: points of loaded arcs on their circles and ellipses, stationary
points against a bounded float64 minimisation, crossings against the closed-form root of
N(t) - c D(t) inside the last bisection bracket, and gradient colours worked out by hand."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

QUAD, ARC = 3, 0x13


def _rvg(tmp_path, body, paint="solid rgba(0,0,0,1)", xf="dyn_identity"):
    f = tmp_path / "s.rvg"
    f.write_text("viewport 0,0 400,400\nwindow 0,0 400,400\nscene dyn_identity\n"
                 f"  1 element nzfill dyn_concrete 0,0 0,0 0,0: {body} {xf} dyn_paint 1 {paint}\n")
    return f


def _arcs(a):
    """(P0, P1, P2, w) of every stored arc, float64."""
    out = []
    for c in np.flatnonzero(a["curve_type"] == ARC):
        b = a["curve_pos"][c]
        p = a["pos"][b:b + 3].astype(np.float64)
        out.append((p[0], p[1], p[2], float(a["curve_weight"][c])))
    return out


def _rat(p0, p1, p2, w, t):
    """Textbook rational quadratic Bezier with Euclidean control point p1 of weight w."""
    t = np.asarray(t, dtype=np.float64)[:, None]
    b0, b1, b2 = (1 - t) ** 2, 2 * t * (1 - t) * w, t ** 2
    return (b0 * p0 + b1 * p1 + b2 * p2) / (b0 + b1 + b2)


T = np.linspace(0.0, 1.0, 101)
R2 = math.sqrt(0.5)


@pytest.mark.parametrize("body,n_arcs,sweep", [
    # quarter circle of radius 100 about the origin: control point (100,100) of weight cos 45
    (f"M 100,0 A {100 * R2:.9f},{100 * R2:.9f},{R2:.9f} 0,100", 1, 0.5 * math.pi),
    # half circle, control point at infinity (w = 0): split into two quarter arcs
    ("M 100,0 A 0,100,0 -100,0", 2, math.pi),
    # the major arc from +60 to -60 degrees (homogeneous control negated): 240 degrees, two arcs
    (f"M 50,{100 * math.sin(math.pi / 3):.9f} A -100,0,-0.5 50,{-100 * math.sin(math.pi / 3):.9f}", 2,
     4 * math.pi / 3),
])
def test_loaded_arcs_lie_on_the_circle(spr, tmp_path, body, n_arcs, sweep):
    a = spr.Scene.load_rvg(_rvg(tmp_path, body)).arrays()
    arcs = _arcs(a)
    assert len(arcs) == n_arcs
    assert all(0.5 <= w <= 1.0 for *_, w in arcs)
    total = 0.0
    for k, (p0, p1, p2, w) in enumerate(arcs):
        pts = _rat(p0, p1, p2, w, T)
        r = np.hypot(pts[:, 0], pts[:, 1])
        assert np.abs(r - 100.0).max() < 2e-3, (k, np.abs(r - 100.0).max())
        if k:
            assert np.allclose(arcs[k - 1][2], p0, atol=1e-4)  # the pieces chain
        ang = np.unwrap(np.arctan2(pts[:, 1], pts[:, 0]))
        total += abs(ang[-1] - ang[0])
    assert abs(total - sweep) < 1e-4


def test_path_transform_maps_arcs_exactly(spr, tmp_path):
    """dyn_affine([2,0,50],[0,1,20]): the half circle becomes a half ellipse of semi-axes 200
    and 100 about (50, 20); the weights are unchanged."""
    a = spr.Scene.load_rvg(_rvg(tmp_path, "M 100,0 A 0,100,0 -100,0", xf="dyn_affine([2,0,50],[0,1,20])")).arrays()
    for p0, p1, p2, w in _arcs(a):
        pts = _rat(p0, p1, p2, w, T)
        e = ((pts[:, 0] - 50.0) / 200.0) ** 2 + ((pts[:, 1] - 20.0) / 100.0) ** 2
        assert np.abs(e - 1.0).max() < 4e-5
        assert abs(w - R2) < 1e-6


def test_loaded_quadratic_keeps_its_points(spr, tmp_path):
    a = spr.Scene.load_rvg(_rvg(tmp_path, "M 10,20 Q 150,300 290,20 Z")).arrays()
    c = np.flatnonzero(a["curve_type"] == QUAD)
    assert len(c) == 1
    b = a["curve_pos"][c[0]]
    assert np.array_equal(a["pos"][b:b + 3], np.array([[10, 20], [150, 300], [290, 20]], np.float32))
    assert a["curve_weight"] is None  # no arc in the scene


def test_oracle_conic_evaluation(orc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        x = rng.uniform(-500, 500, 3).astype(np.float32)
        w = np.float32(rng.uniform(0.5, 1.5))
        t = np.float32(rng.uniform(0, 1))
        # arc: the textbook rational form (control point x1 of weight w)
        ref = _rat(np.float64(x[0]), np.float64(x[1]), np.float64(x[2]), float(w), [float(t)])[0]
        got = orc.eval_axis(ARC, float(t), [x[0], x[1], x[2], w])
        assert abs(got - ref) <= 1e-4 * (1 + abs(ref))
        # quadratic: the Bernstein polynomial
        refq = (1 - t) ** 2 * x[0] + 2 * t * (1 - t) * x[1] + t * t * x[2]
        gq = orc.eval_axis(QUAD, float(t), [x[0], x[1], x[2], 0])
        assert abs(gq - float(refq)) <= 1e-4 * (1 + abs(float(refq)))
    # the arc's midpoint in closed form: (x0 + 2 w x1 + x2) / (2 + 2 w)
    mid = orc.eval_axis(ARC, 0.5, [0.0, 100.0, 40.0, 0.6])
    assert abs(mid - (0 + 2 * 0.6 * 100 + 40) / 3.2) < 1e-4


def _stationary(f, lo, hi):
    r = minimize_scalar(f, bounds=(lo, hi), method="bounded", options={"xatol": 1e-11})
    return r.x


def test_oracle_conic_cuts_at_the_stationary_points(orc):
    # circular arc from -30 to 120 degrees about the origin, radius 100: x is extreme at 0
    # degrees, y at 90 degrees
    a0, a1 = math.radians(-30), math.radians(120)
    half = 0.5 * (a1 - a0)
    w = math.cos(half)
    mid = 0.5 * (a0 + a1)
    p0 = np.array([100 * math.cos(a0), 100 * math.sin(a0)])
    p2 = np.array([100 * math.cos(a1), 100 * math.sin(a1)])
    p1 = np.array([100 / w * math.cos(mid), 100 / w * math.sin(mid)])
    xs = np.array([p0[0], p1[0], p2[0], w], np.float32)
    ys = np.array([p0[1], p1[1], p2[1], 0.0], np.float32)
    n, q = orc.curve_cuts(ARC, xs, ys)
    assert n == 2
    tx = _stationary(lambda t: -_rat(p0, p1, p2, w, [t])[0, 0], 0, 1)
    ty = _stationary(lambda t: -_rat(p0, p1, p2, w, [t])[0, 1], 0, 1)
    assert np.allclose(sorted([tx, ty]), q[:2], atol=2e-5), (tx, ty, q)
    assert q[2] == 0 and q[3] == 0
    # the two extrema sit at 0 and 90 degrees
    for t, ang in zip(sorted([tx, ty]), (0.0, 90.0)):
        pt = _rat(p0, p1, p2, w, [t])[0]
        assert abs(math.degrees(math.atan2(pt[1], pt[0])) - ang) < 1e-3
    # a quadratic: t = (p0 - p1) / (p0 - 2 p1 + p2) per axis, ascending
    xs = np.array([0, 200, 50, 0], np.float32)
    ys = np.array([0, -80, 120, 0], np.float32)
    n, q = orc.curve_cuts(QUAD, xs, ys)
    ref = sorted([(0 - 200) / (0 - 400 + 50), (0 + 80) / (0 + 160 + 120)])
    assert n == 2 and np.allclose(q[:2], ref, atol=1e-6)
    # monotone curves have no cut; an invisible curve none either
    n, _ = orc.curve_cuts(QUAD, np.array([0, 1, 2, 0], np.float32), np.array([0, 1, 3, 0], np.float32))
    assert n == 0
    n, _ = orc.curve_cuts(ARC, xs, ys, visible=False)
    assert n == 0


def _conic_root(x0, x1, x2, w, c, lo, hi):
    """Root in [lo, hi] of N(t) - c D(t) = 0 (a quadratic in t), in float64."""
    # Bernstein coefficients b_i = w_i (x_i - c) with weights (1, w, 1)
    b0, b1, b2 = x0 - c, w * (x1 - c), x2 - c
    A, B, C = b0 - 2 * b1 + b2, 2 * (b1 - b0), b0
    roots = [-C / B] if abs(A) < 1e-300 else list(np.roots([A, B, C]).real)
    roots = [r for r in roots if lo - 1e-9 <= r <= hi + 1e-9]
    assert len(roots) == 1
    return roots[0]


@pytest.mark.parametrize("kind", [QUAD, ARC])
def test_oracle_conic_crossings_run_24_halvings(orc, kind):
    """Quadratics and arcs take the cubic's 24-step bisection: on [0, 1] the result is the last
    midpoint, an odd multiple of 2^-24 (a 23- or 20-step bisection gives even multiples), and it
    lies next to the closed-form float64 root of N(t) - c D(t) = 0."""
    rng = np.random.default_rng(7 + kind)
    worst = 0.0
    for _ in range(400):
        x0, x2 = 0.0, float(rng.uniform(50, 400))
        x1 = float(rng.uniform(0.05, 0.95)) * x2  # monotone in x for w > 0
        w = 1.0 if kind == QUAD else float(rng.uniform(0.5, 1.0))
        xs = np.array([x0, x1, x2, w if kind == ARC else 0.0], np.float32)
        ys = np.zeros(4, np.float32)
        c = float(np.float32(rng.uniform(1, x2 - 1)))
        t = float(orc.solve_crossing(kind, xs, ys, 0, c, 0.0, 1.0))
        m = t * 2.0 ** 24
        assert m == np.floor(m) and int(m) % 2 == 1, (t, xs, c)
        ref = _conic_root(float(xs[0]), float(xs[1]), float(xs[2]), float(np.float32(w)), c, 0.0, 1.0)
        worst = max(worst, abs(t - ref) * 2.0 ** 24)
    # within a few final brackets: the float32 sign test near the root is as good as the
    # evaluation's rounding over the local slope
    assert worst < 16, worst


# ---- gradient paints: colours worked out by hand from the definitions in include/spr.h

def _paint(spr, kind, spread, geom, stops, opacity=1.0, xform=(1, 0, 0, 0, 1, 0)):
    p = np.zeros(1, dtype=spr.PAINT_DTYPE)
    p[0] = (kind, spread, 0, len(stops), opacity, list(xform), list(geom) + [0] * (5 - len(geom)))
    return p, np.array(stops, np.float32)


def _rgba(c):
    return [(c >> (8 * i)) & 255 for i in range(4)]


RB = [[0.0, 1, 0, 0, 1], [1.0, 0, 0, 1, 1]]  # red to blue


def test_linear_gradient_known_values(spr, orc):
    p, st = _paint(spr, spr.PAINT_LINEAR, spr.SPREAD_PAD, [0, 0, 100, 0], RB)
    # pixel 24: centre 24.5, t = 0.245: red 0.755 * 255 = 192.5 -> 192, blue 62.5 -> 62
    assert _rgba(orc.paint(p, st, 24, 7)) == [192, 0, 62, 255]
    assert _rgba(orc.paint(p, st, -10, 0)) == [255, 0, 0, 255]   # pad below 0: first stop
    assert _rgba(orc.paint(p, st, 150, 0)) == [0, 0, 255, 255]   # pad above 1: last stop
    # opacity scales alpha only: 0.5 * 255 = 127.5 -> 127
    p2, _ = _paint(spr, spr.PAINT_LINEAR, spr.SPREAD_PAD, [0, 0, 100, 0], RB, opacity=0.5)
    assert _rgba(orc.paint(p2, st, 24, 0)) == [192, 0, 62, 127]
    # the paint transform: scene -> gradient space halves x, so pixel 49 (centre 49.5) gives
    # t = 0.2475: red 0.7525 * 255 = 191.89 -> 191, blue 63.11 -> 63
    p3, _ = _paint(spr, spr.PAINT_LINEAR, spr.SPREAD_PAD, [0, 0, 100, 0], RB, xform=(0.5, 0, 0, 0, 0.5, 0))
    assert _rgba(orc.paint(p3, st, 49, 0)) == [191, 0, 63, 255]
    # the inverse MVP: a view shifted right by 10 shows scene x = 24.5 at pixel 34
    m = np.eye(4, dtype=np.float32)
    m[0, 3] = 10.0
    assert _rgba(orc.paint(p, st, 34, 7, mvp=m)) == [192, 0, 62, 255]
    # equal end points: the last stop's colour, whatever the spread
    p4, _ = _paint(spr, spr.PAINT_LINEAR, spr.SPREAD_REPEAT, [5, 5, 5, 5], RB)
    assert _rgba(orc.paint(p4, st, 0, 0)) == [0, 0, 255, 255]


def test_gradient_spreads(spr, orc):
    # pixel 124: t = 1.245
    for spread, want in [(spr.SPREAD_PAD, [0, 0, 255, 255]),
                         (spr.SPREAD_REPEAT, [192, 0, 62, 255]),      # 0.245
                         (spr.SPREAD_REFLECT, [62, 0, 192, 255]),     # 2 - 1.245 = 0.755
                         (spr.SPREAD_TRANSPARENT, [0, 0, 0, 0])]:
        p, st = _paint(spr, spr.PAINT_LINEAR, spread, [0, 0, 100, 0], RB)
        assert _rgba(orc.paint(p, st, 124, 0)) == want, spread
    # reflect below 0: t = -0.755 -> 0.755 (m = 1.245, 2 - m)
    p, st = _paint(spr, spr.PAINT_LINEAR, spr.SPREAD_REFLECT, [0, 0, 100, 0], RB)
    assert _rgba(orc.paint(p, st, -76, 0)) == [62, 0, 192, 255]


def test_multi_stop_ramp(spr, orc):
    stops = [[0.0, 0, 0, 0, 1], [0.5, 1, 0, 0, 1], [1.0, 1, 1, 1, 1]]
    p, st = _paint(spr, spr.PAINT_LINEAR, spr.SPREAD_PAD, [0.5, 0, 100.5, 0], stops)
    # pixel 75: t = 0.75, halfway between the red and the white stops: green/blue 127.5 -> 127
    assert _rgba(orc.paint(p, st, 75, 0)) == [255, 127, 127, 255]
    # pixel 25: t = 0.25, halfway between black and red
    assert _rgba(orc.paint(p, st, 25, 0)) == [127, 0, 0, 255]


def test_radial_gradient_known_values(spr, orc):
    bw = [[0.0, 0, 0, 0, 1], [1.0, 1, 1, 1, 1]]
    # focus at the centre: t = |g - c| / r.  Pixel (69, 49), centre (69.5, 49.5), c = (50, 50),
    # r = 40: t = hypot(19.5, 0.5) / 40 = 0.48766 -> 124.35 -> 124
    p, st = _paint(spr, spr.PAINT_RADIAL, spr.SPREAD_PAD, [50, 50, 50, 50, 40], bw)
    assert _rgba(orc.paint(p, st, 69, 49)) == [124, 124, 124, 255]
    assert _rgba(orc.paint(p, st, 50, 50)) == [4, 4, 4, 255]  # t = hypot(.5, .5) / 40: 4.51 -> 4
    # focus off the centre: c = (0.5, 0.5), f = (50.5, 0.5), r = 100.  The point 25 left of c
    # lies on the circle of centre f + t (c - f) and radius 100 t for t = 1/2:
    # |-25 - 50 + 50 t| = 100 t.  Pixel (-25, 0) has that centre: 127.5 -> 127
    p, st = _paint(spr, spr.PAINT_RADIAL, spr.SPREAD_PAD, [0.5, 0.5, 50.5, 0.5, 100], bw)
    assert _rgba(orc.paint(p, st, -25, 0)) == [127, 127, 127, 255]
    # the focus itself is t = 0, the circle point on the ray beyond c is t = 1
    assert _rgba(orc.paint(p, st, 50, 0)) == [0, 0, 0, 255]
    assert _rgba(orc.paint(p, st, -100, 0)) == [255, 255, 255, 255]


def _arc_scene(**kw):
    base = dict(pos=np.array([[0, 0], [5, 9], [10, 0], [10, 0], [0, 0]], np.float32), pos_path=[0] * 5,
                curve_pos=[0, 3], curve_type=[ARC, 2], curve_path=[0, 0], rule=[0], fill=[0xFF000000],
                width=16, height=16)
    base.update(kw)
    return base


def test_scene_create_validates_weights_and_paints(spr):
    with pytest.raises(spr.SprError, match="arcs need curve weights"):
        spr.Scene.create(**_arc_scene())
    for w in (0.0, -1.0, np.nan, np.inf):
        with pytest.raises(spr.SprError, match="arc weight must be finite and > 0"):
            spr.Scene.create(**_arc_scene(curve_weight=[w, 1.0]))
    s = spr.Scene.create(**_arc_scene(curve_weight=[0.7, 1.0]))
    assert s.arrays()["curve_weight"][0] == np.float32(0.7)
    stops = np.array([[0, 1, 0, 0, 1], [1, 0, 0, 1, 1]], np.float32)

    def paint(kind=1, spread=0, first=0, n=2, geom=(0, 0, 10, 0, 0), xform=(1, 0, 0, 0, 1, 0)):
        p = np.zeros(1, dtype=spr.PAINT_DTYPE)
        p[0] = (kind, spread, first, n, 1.0, list(xform), list(geom))
        return p

    ok = dict(curve_weight=[0.7, 1.0], path_paint=[0], stops=stops)
    spr.Scene.create(**_arc_scene(**ok, paints=paint()))
    bad = [
        (dict(paints=paint(kind=3)), "paint kind"),
        (dict(paints=paint(spread=4)), "paint spread"),
        (dict(paints=paint(n=3)), "paint stops out of range"),
        (dict(paints=paint(n=0)), "paint stops out of range"),
        (dict(paints=paint(kind=2, geom=(0, 0, 20, 0, 10))), "focus inside the circle"),
        (dict(paints=paint(kind=2, geom=(0, 0, 0, 0, 0))), "r > 0"),
        (dict(paints=paint(xform=(np.nan, 0, 0, 0, 1, 0))), "non-finite paint transform"),
        (dict(paints=paint(), stops=stops[::-1].copy()), "stop offsets must not decrease"),
        (dict(paints=paint(), path_paint=[1]), "path_paint out of range"),
    ]
    for kw, msg in bad:
        args = dict(ok)
        args.update(kw)
        with pytest.raises(spr.SprError, match=msg):
            spr.Scene.create(**_arc_scene(**args))
    # solid paths may be marked SOLID_PAINT next to painted ones
    s = spr.Scene.create(**_arc_scene(curve_weight=[0.7, 1.0], path_paint=[spr.SOLID_PAINT], paints=paint(),
                                      stops=stops))
    assert s.arrays()["path_paint"][0] == spr.SOLID_PAINT


def test_write_rvg_round_trip_keeps_conics_and_paints(spr, tmp_path):
    """car.rvg (arcs, gradients, affine paint transforms) written and read back: the same
    curves, weights, stops and paints (transforms inverted twice, to float rounding)."""
    from conftest import RVG_DIR
    s = spr.Scene.load_rvg(RVG_DIR / "car.rvg")
    f = tmp_path / "car2.rvg"
    s.write_rvg(f)
    a, b = s.arrays(), spr.Scene.load_rvg(f).arrays()
    assert np.array_equal(a["curve_type"], b["curve_type"])
    assert np.array_equal(a["rule"], b["rule"])
    assert np.array_equal(a["path_paint"], b["path_paint"])
    np.testing.assert_allclose(b["pos"], a["pos"], rtol=0, atol=1e-3)
    arc = a["curve_type"] == ARC
    np.testing.assert_allclose(b["curve_weight"][arc], a["curve_weight"][arc], rtol=1e-6)
    np.testing.assert_array_equal(b["stops"], a["stops"])
    pa, pb = a["paints"], b["paints"]
    for k in ("kind", "spread", "first_stop", "n_stops"):
        assert np.array_equal(pa[k], pb[k]), k
    np.testing.assert_allclose(pb["xform"], pa["xform"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(pb["geom"], pa["geom"], rtol=1e-6)
