"""Independent float64 geometry used by the CPU tests to pin the oracle.

Nothing here shares code with the oracle: curves are evaluated in Bernstein form in float64,
winding numbers come from a crossing-number test on densely sampled polylines.
"""
from __future__ import annotations

import numpy as np

NONZERO, EVEN_ODD = 0, 1


def bezier(ctrl: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Cubic (4 control points) or line (2) in Bernstein form, float64. ctrl: (k, 2)."""
    c = np.asarray(ctrl, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)[:, None]
    if len(c) == 2:
        return (1 - t) * c[0] + t * c[1]
    s = 1 - t
    return s ** 3 * c[0] + 3 * s * s * t * c[1] + 3 * s * t * t * c[2] + t ** 3 * c[3]


def polygon_scene(polys, rules, fills, close=True):
    """Scene arrays (the loadVG SoA) from polygons given as (n, 2) vertex lists; each polygon
    is one path made of line curves (the closing edge included)."""
    pos, pos_path, cpos, ctype, cpath = [], [], [], [], []
    for p, poly in enumerate(polys):
        v = np.asarray(poly, dtype=np.float32)
        if close:
            v = np.concatenate([v, v[:1]])
        base = len(pos)
        pos.extend(v.tolist())
        pos_path.extend([p] * len(v))
        for i in range(len(v) - 1):
            cpos.append(base + i)
            ctype.append(2)
            cpath.append(p)
    return dict(pos=np.asarray(pos, np.float32).reshape(-1, 2), pos_path=np.asarray(pos_path, np.uint32),
                curve_pos=np.asarray(cpos, np.uint32), curve_type=np.asarray(ctype, np.uint32),
                curve_path=np.asarray(cpath, np.uint32), rule=np.asarray(rules, np.uint32),
                fill=np.asarray(fills, np.uint32))


def path_polylines(arrays: dict, step: float = 0.25):
    """Per path: list of dense polylines (one per curve), sample spacing <= step pixels."""
    pos = np.asarray(arrays["pos"], dtype=np.float64).reshape(-1, 2)
    out = [[] for _ in range(len(arrays["rule"]))]
    for c in range(len(arrays["curve_type"])):
        k = int(arrays["curve_type"][c])
        ctrl = pos[int(arrays["curve_pos"][c]):int(arrays["curve_pos"][c]) + k]
        length = np.sum(np.linalg.norm(np.diff(ctrl, axis=0), axis=1))
        n = max(2, int(np.ceil(length / step)) + 1)
        out[int(arrays["curve_path"][c])].append(bezier(ctrl, np.linspace(0.0, 1.0, n)))
    return out


def winding_numbers(polylines, px: np.ndarray, py: np.ndarray) -> np.ndarray:
    """Winding number of the closed curve set at points (px, py) (crossing-number rule)."""
    wn = np.zeros(px.shape, dtype=np.int64)
    for pl in polylines:
        a, b = pl[:-1], pl[1:]
        for i0 in range(0, len(a), 512):
            ax, ay = a[i0:i0 + 512, 0][:, None], a[i0:i0 + 512, 1][:, None]
            bx, by = b[i0:i0 + 512, 0][:, None], b[i0:i0 + 512, 1][:, None]
            X, Y = px[None, :], py[None, :]
            left = (bx - ax) * (Y - ay) - (X - ax) * (by - ay)
            up = (ay <= Y) & (by > Y) & (left > 0)
            down = (by <= Y) & (ay > Y) & (left < 0)
            wn += up.sum(axis=0) - down.sum(axis=0)
    return wn


def expected_overwrite_image(arrays: dict, w: int, h: int, margin: float = 3.0):
    """Overwrite composition of opaque paths at pixel centres, and a mask of the pixels that
    lie farther than `margin` from every path boundary (where the rasteriser's 2x2 boundary
    cells cannot reach).  Canvas row y is image row h-1-y."""
    lines = path_polylines(arrays)
    jj, ii = np.meshgrid(np.arange(w), np.arange(h))
    px = (jj + 0.5).reshape(-1).astype(np.float64)
    py = ((h - 1 - ii) + 0.5).reshape(-1).astype(np.float64)
    img = np.full((h * w, 4), 255, dtype=np.uint8)
    near = np.zeros((h, w), dtype=bool)
    r = int(np.ceil(margin)) + 1
    for p, pls in enumerate(lines):
        if not pls:
            continue
        wn = winding_numbers(pls, px, py)
        inside = (wn != 0) if int(arrays["rule"][p]) == NONZERO else ((wn & 1) != 0)
        col = np.frombuffer(np.uint32(arrays["fill"][p]).tobytes(), dtype=np.uint8)
        img[inside] = col
        pts = np.concatenate(pls)
        cx = np.floor(pts[:, 0]).astype(np.int64)
        cy = np.floor(pts[:, 1]).astype(np.int64)
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                x, y = cx + dx, cy + dy
                ok = (x >= 0) & (x < w) & (y >= 0) & (y < h)
                near[h - 1 - y[ok], x[ok]] = True
    return img.reshape(h, w, 4), ~near


def derivative_root_count(ctrl) -> int:
    """Number of interior extrema (both axes) of a cubic, float64: 4 means the reference's
    4-cut case, whose unsorted last cut doubles part of the curve."""
    c = np.asarray(ctrl, dtype=np.float64)
    n = 0
    for ax in range(2):
        p0, p1, p2, p3 = c[:, ax]
        a = 3 * (-p0 + 3 * p1 - 3 * p2 + p3)
        b = 6 * (p0 - 2 * p1 + p2)
        cc = 3 * (p1 - p0)
        if abs(a) < 1e-12:
            n += int(abs(b) > 1e-12 and 0 < -cc / b < 1)
            continue
        d = b * b - 4 * a * cc
        if d > 0:
            s = np.sqrt(d)
            n += sum(0 < r < 1 for r in ((-b - s) / (2 * a), (-b + s) / (2 * a)))
    return n


def regular_paths(arrays: dict, w: int, h: int) -> np.ndarray:
    """Paths whose control points all lie inside the canvas and whose cubics have fewer than
    four extrema: for them the rasteriser's output is exactly the winding-number fill."""
    pos = np.asarray(arrays["pos"], dtype=np.float64).reshape(-1, 2)
    ok = np.ones(len(arrays["rule"]), dtype=bool)
    for c in range(len(arrays["curve_type"])):
        k = int(arrays["curve_type"][c])
        ctrl = pos[int(arrays["curve_pos"][c]):int(arrays["curve_pos"][c]) + k]
        p = int(arrays["curve_path"][c])
        if np.any(ctrl < 1) or np.any(ctrl[:, 0] > w - 1) or np.any(ctrl[:, 1] > h - 1):
            ok[p] = False
        if k == 4 and derivative_root_count(ctrl) >= 4:
            ok[p] = False
    return ok


def select_paths(arrays: dict, keep: np.ndarray) -> dict:
    """The scene restricted to the paths where keep is True (indices renumbered)."""
    keep = np.asarray(keep, dtype=bool)
    new_id = np.cumsum(keep) - 1
    pk = keep[arrays["pos_path"]]
    ck = keep[arrays["curve_path"]]
    new_pos_idx = np.cumsum(pk) - 1
    return dict(pos=np.asarray(arrays["pos"]).reshape(-1, 2)[pk],
                pos_path=new_id[arrays["pos_path"][pk]].astype(np.uint32),
                curve_pos=new_pos_idx[arrays["curve_pos"][ck]].astype(np.uint32),
                curve_type=arrays["curve_type"][ck].astype(np.uint32),
                curve_path=new_id[arrays["curve_path"][ck]].astype(np.uint32),
                rule=arrays["rule"][keep].astype(np.uint32), fill=arrays["fill"][keep].astype(np.uint32))


def cubic_loops(seed: int, n_paths: int, w: int, h: int, curves=(3, 9)) -> dict:
    """Closed contours of cubic curves inside the canvas (one contour per path, consecutive
    curves share end points), each cubic drawn again until it has fewer than four extrema.
    Rules alternate nonzero / even-odd; colours are opaque."""
    rng = np.random.default_rng(seed)
    pos, pos_path, cpos, ctype, cpath = [], [], [], [], []
    for p in range(n_paths):
        k = int(rng.integers(curves[0], curves[1] + 1))
        cx, cy = rng.uniform(0.3 * w, 0.7 * w), rng.uniform(0.3 * h, 0.7 * h)
        rad = rng.uniform(0.1, 0.28) * min(w, h)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        knots = np.c_[cx + rad * np.cos(ang), cy + rad * np.sin(ang)]
        base = len(pos)
        pts = [knots[0]]
        for i in range(k):
            a, b = knots[i], knots[(i + 1) % k]
            while True:
                c1 = a + rng.uniform(-0.5, 0.5, 2) * rad
                c2 = b + rng.uniform(-0.5, 0.5, 2) * rad
                ctrl = np.array([a, c1, c2, b])
                if derivative_root_count(ctrl) < 4 and np.all(ctrl >= 2) and np.all(ctrl[:, 0] <= w - 2) \
                        and np.all(ctrl[:, 1] <= h - 2):
                    break
            pts.extend([c1, c2, b])
            cpos.append(base + 3 * i)
            ctype.append(4)
            cpath.append(p)
        pos.extend(pts)
        pos_path.extend([p] * len(pts))
    fills = [0xFF000000 | int(c) for c in rng.integers(0, 1 << 24, n_paths)]
    return dict(pos=np.asarray(pos, np.float32).reshape(-1, 2), pos_path=np.asarray(pos_path, np.uint32),
                curve_pos=np.asarray(cpos, np.uint32), curve_type=np.asarray(ctype, np.uint32),
                curve_path=np.asarray(cpath, np.uint32),
                rule=np.asarray([p & 1 for p in range(n_paths)], np.uint32),
                fill=np.asarray(fills, np.uint32))
