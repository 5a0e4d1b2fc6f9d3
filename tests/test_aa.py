"""Pins of the anti-aliased coverage (CPU): the oracle's 64-sample cell masks and its
coverage-weighted draw.

The reference leaves the boundary cells' "comb-like multisampling" stencil mask TODO
(gen_fragment.comp:202, scanlinepr.vert:48-50), so the oracle restates the library's documented
definition (include/spr.h, spr_opts.aa).  It is pinned here independently: on polygon scenes
(lines only, so the chord polygon is the polygon itself) every sample bit must equal the
nonzero / even-odd rule of the float64 winding number of the polygon at that sample, and a
straight edge gives the closed-form sample counts."""
from __future__ import annotations

import numpy as np
import pytest

ALL = np.uint64(0xFFFFFFFFFFFFFFFF)
SX = np.array([((7 * j) % 16 + 0.5) / 16 for j in range(16)])
SY = np.array([(j + 0.5) / 16 for j in range(16)])


def _polygon_scene(spr, rng, n_paths, canvas):
    pos, pos_path, cpos, ctype, cpath, rule, fill, polys = [], [], [], [], [], [], [], []
    for p in range(n_paths):
        n = int(rng.integers(3, 10))
        c = rng.uniform(0.2 * canvas, 0.8 * canvas, 2)
        r = rng.uniform(0.05 * canvas, 0.3 * canvas)
        ang = np.sort(rng.uniform(0, 2 * np.pi, n)) if p % 3 else rng.uniform(0, 2 * np.pi, n)  # some self-intersect
        rad = r * rng.uniform(0.3, 1.0, n)
        v = np.stack([c[0] + rad * np.cos(ang), c[1] + rad * np.sin(ang)], 1).astype(np.float32)
        polys.append(v.astype(np.float64))
        for k in range(n):
            cpos.append(len(pos_path))
            ctype.append(2)
            cpath.append(p)
            for q in (v[k], v[(k + 1) % n]):
                pos.append(q)
                pos_path.append(p)
        rule.append(p % 2)
        fill.append(0xFF000000 | (p * 7919 + 12345) & 0xFFFFFF)  # distinct, opaque
    scene = spr.Scene.create(np.array(pos, np.float32), pos_path, cpos, ctype, cpath, rule, fill,
                             width=canvas, height=canvas)
    return scene, polys, rule, fill


def _winding(poly, x, y):
    """Winding number of a closed polygon about (x, y), float64 crossing rule."""
    a, b = poly, np.roll(poly, -1, 0)
    up = (a[:, 1] <= y) & (y < b[:, 1])
    dn = (b[:, 1] <= y) & (y < a[:, 1])
    cr = (b[:, 0] - a[:, 0]) * (y - a[:, 1]) - (x - a[:, 0]) * (b[:, 1] - a[:, 1])
    return int(((up & (cr > 0)).sum()) - ((dn & (cr < 0)).sum()))


def _edge_distance(poly, x, y):
    a, b = poly, np.roll(poly, -1, 0)
    d = b - a
    t = np.clip(((x - a[:, 0]) * d[:, 0] + (y - a[:, 1]) * d[:, 1]) / np.maximum((d ** 2).sum(1), 1e-30), 0, 1)
    px, py = a[:, 0] + t * d[:, 0], a[:, 1] + t * d[:, 1]
    return float(np.min(np.hypot(px - x, py - y)))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cell_masks_match_point_in_polygon(spr, orc, seed):
    rng = np.random.default_rng(seed)
    canvas = 160
    scene, polys, rule, fill = _polygon_scene(spr, rng, 8, canvas)
    a = scene.arrays()
    f = orc.Frame(a, canvas, canvas, compose=1, aa=16)
    rec = f.get("records").reshape(-1, 4)
    cm = f.get("cmask")
    assert len(cm) == len(rec) > 0
    path_of = {int(np.uint32(c)): p for p, c in enumerate(fill)}
    checked = ambiguous = cells = 0
    for k, r in enumerate(rec):
        if r[3] == 0:  # a span: full coverage
            assert cm[k] == ALL
            continue
        cells += 1
        x, y = int(r[0] & 0xFFFF), int(r[0] >> 16)
        p = path_of[int(np.uint32(r[2]))]
        for py in range(2):
            for px in range(2):
                for j in range(16):
                    xs, ys = x + px + SX[j], y + py + SY[j]
                    if _edge_distance(polys[p], xs, ys) < 1e-3:
                        ambiguous += 1
                        continue
                    wn = _winding(polys[p], xs, ys)
                    inside = (wn & 1) != 0 if rule[p] == 1 else wn != 0
                    bit = (int(cm[k]) >> ((py * 2 + px) * 16 + j)) & 1
                    assert bit == int(inside), (seed, k, p, px, py, j, wn)
                    checked += 1
    assert cells > 100 and checked > 50 * cells and ambiguous < checked // 1000 + 5


def test_straight_edge_sample_counts(spr, orc):
    """A square whose left edge is x = 10.3: in the edge's pixel column 10 the samples with
    x offset (k + 1/2)/16 >= 0.3 are inside, k = 5..15: 11 of 16; columns 11+ are whole; an
    opaque red square drawn over white gives green = blue = (1 - 11/16) * 255 + 1/2 -> 80."""
    x0, x1, y0, y1 = 10.3, 40.0, 8.0, 30.0
    pts = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    pos, pos_path, cpos = [], [], []
    for k in range(4):
        cpos.append(len(pos_path))
        pos += [pts[k], pts[(k + 1) % 4]]
        pos_path += [0, 0]
    scene = spr.Scene.create(np.array(pos, np.float32), pos_path, cpos, [2] * 4, [0] * 4, [0], [0xFF0000FF],
                             width=64, height=48)
    f = orc.Frame(scene.arrays(), 64, 48, compose=1, aa=16)
    img = f.image()
    row = lambda yy: img[48 - 1 - yy]  # canvas row yy is image row h-1-yy
    for yy in (12, 13, 20):
        assert row(yy)[10].tolist() == [255, 80, 80, 255], yy
        assert row(yy)[11].tolist() == [255, 0, 0, 255]
        assert row(yy)[9].tolist() == [255, 255, 255, 255]
    # the binary frame paints the whole cell column 10..11 (no partial coverage)
    fb = orc.Frame(scene.arrays(), 64, 48, compose=1)
    assert fb.image()[48 - 1 - 12][10].tolist() == [255, 0, 0, 255]
    # bottom edge y = 8.0 on a cell boundary: row 8 whole, row 7 empty
    assert row(8)[20].tolist() == [255, 0, 0, 255] and row(7)[20].tolist() == [255, 255, 255, 255]


def test_aa_draw_weights_colour_by_coverage(spr, orc):
    """Every pixel of an anti-aliased frame of one opaque path equals the closed form of the
    src-over blend with the pixel's covered-sample share."""
    rng = np.random.default_rng(5)
    scene, polys, rule, fill = _polygon_scene(spr, rng, 1, 96)
    f = orc.Frame(scene.arrays(), 96, 96, compose=1, aa=16)
    img = f.image().astype(np.int32)
    rec = f.get("records").reshape(-1, 4)
    cm = f.get("cmask")
    col = [(fill[0] >> (8 * i)) & 255 for i in range(3)]
    for k, r in enumerate(rec):
        if r[3] == 0:
            continue
        x, y = int(r[0] & 0xFFFF), int(r[0] >> 16)
        for py in range(2):
            for px in range(2):
                n = bin((int(cm[k]) >> ((py * 2 + px) * 16)) & 0xFFFF).count("1")
                cov = np.float32(n) / np.float32(16)
                for ch in range(3):
                    s = np.float32(col[ch]) / np.float32(255) * np.float32(1) * cov
                    v = s + np.float32(1) * (np.float32(1) - cov)
                    assert img[96 - 1 - (y + py), x + px, ch] == int(v * np.float32(255) + np.float32(0.5))
