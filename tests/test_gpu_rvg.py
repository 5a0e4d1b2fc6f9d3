"""GPU parity on the full RVG scenes: quadratic and arc curves (the CurveType values the
reference declares but leaves TODO in its shaders), dyn_affine path and paint transforms, and
linear and radial gradient paints, against the CPU oracle.

car.rvg (arcs that are half ellipses, split at load into quarter arcs, 236 gradients, affine
transforms), chord.rvg (9,312 circular arcs), drops.rvg and embrace.rvg (radial and linear
gradients under affine paint transforms) are the reference's own scenes
(workdir/input/rvg).  Every integer stage and the overwrite image must be identical; the
src-over image too (both accumulate in fp32 in the same order)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import RVG_DIR

pytestmark = pytest.mark.gpu

SCENES = ["car.rvg", "chord.rvg", "drops.rvg", "embrace.rvg"]
STAGES = ["tpos", "visible", "cuts", "counts", "inter", "key", "wn", "seg", "skey", "sidx"]


def _compare_stages(r, f):
    for st in STAGES:
        g, o = r.export(st), f.get(st)
        assert g.shape == o.shape, (st, g.shape, o.shape)
        if g.dtype == np.float32:
            bad = np.flatnonzero(g.view(np.uint32) != o.view(np.uint32))
        else:
            bad = np.flatnonzero(g != o)
        assert bad.size == 0, (st, bad[:8], g[bad[:8]], o[bad[:8]])
    N = len(f.get("inter")) // 2
    assert np.array_equal(r.export("winding")[: N + 1], f.get("winding")[: N + 1])
    assert np.array_equal(r.export("flags"), f.get("flags"))
    assert np.array_equal(r.export("records"), f.get("records"))


def _image_diff(img, ref):
    return np.abs(img.astype(np.int16) - ref.astype(np.int16)).max(-1)


@pytest.mark.parametrize("compose", [0, 1])
@pytest.mark.parametrize("name", SCENES)
def test_rvg_scene_matches_oracle(spr, orc, name, compose):
    scene = spr.Scene.load_rvg(RVG_DIR / name)
    a = scene.arrays()
    w, h = a["width"], a["height"]
    n_arcs = int((a["curve_type"] == spr.CURVE_ARC).sum())
    print(f"[rvg] {name} compose={compose} {w}x{h} paths={scene.n_paths} arcs={n_arcs} "
          f"paints={len(a['paints'])}", flush=True)
    f = orc.Frame(a, w, h, compose=compose, threads=8)
    r = spr.Rasterizer(w, h, compose=compose, debug=True)
    r.load_scene(scene)
    img = r.render()
    _compare_stages(r, f)
    d = _image_diff(img, f.image())
    assert d.max() == 0, (name, int(d.max()), int((d > 0).sum()))
    r.close()


@pytest.mark.parametrize("name", SCENES)
def test_rvg_production_frame_matches_oracle(spr, orc, name):
    """Frames without debug export (fused mark + emit, 32-bit sort words) and a second frame of
    the same handle."""
    scene = spr.Scene.load_rvg(RVG_DIR / name)
    a = scene.arrays()
    w, h = a["width"], a["height"]
    ref = orc.Frame(a, w, h, compose=1, threads=8).image()
    r = spr.Rasterizer(w, h, compose=1)
    r.load_scene(scene)
    for frame in range(2):
        print(f"[rvg-prod] {name} frame={frame}", flush=True)
        assert _image_diff(r.render(), ref).max() == 0, (name, frame)
    r.close()


def _mvp(angle, scale, tx, ty):
    c, s = np.cos(angle) * scale, np.sin(angle) * scale
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 1], m[0, 3] = c, -s, tx
    m[1, 0], m[1, 1], m[1, 3] = s, c, ty
    return m


@pytest.mark.parametrize("name", ["car.rvg", "drops.rvg"])
def test_rvg_under_mvp(spr, orc, name):
    """A rotated, scaled and shifted view: arcs keep their weights under the affine map and the
    gradients follow the geometry through the inverse MVP."""
    scene = spr.Scene.load_rvg(RVG_DIR / name)
    a = scene.arrays()
    w, h = a["width"], a["height"]
    m = _mvp(0.3, 1.7, -120.0, -80.0)
    print(f"[rvg-mvp] {name}", flush=True)
    f = orc.Frame(a, w, h, mvp=m, compose=1, threads=8)
    r = spr.Rasterizer(w, h, compose=1, debug=True)
    r.load_scene(scene)
    r.set_mvp(m)
    img = r.render()
    _compare_stages(r, f)
    assert _image_diff(img, f.image()).max() == 0
    r.close()


@pytest.mark.parametrize("compose", [0, 1])
def test_rvg_batch_matches_oracle(spr, orc, compose):
    """The four scenes plus a generated document in one batch (arc weights and paints
    concatenated, paint indices renumbered): every document equals its own oracle frame, on the
    eager frames and on the graph replays."""
    scenes = [spr.Scene.load_rvg(RVG_DIR / n) for n in SCENES] + [spr.Scene.generate(2, 5, 40, 300)]
    canv = [(s.arrays()["width"], s.arrays()["height"]) for s in scenes[:-1]] + [(300, 200)]
    refs = [orc.Frame(s.arrays(), cw, ch, compose=compose, threads=8).image() for s, (cw, ch) in zip(scenes, canv)]
    r = spr.Rasterizer(64, 64, compose=compose)
    r.load_batch(scenes, canv)
    for frame in range(4):
        print(f"[rvg-batch] compose={compose} frame={frame}", flush=True)
        r.render(readback=False)
        for d in range(len(scenes)):
            assert _image_diff(r.batch_image(d), refs[d]).max() == 0, (frame, d)
    r.close()


def test_rvg_learned_frame_matches_oracle(spr, orc):
    """The learned single-pass frame on the arc scene (its features and polygon bound read
    three-point curves) and its graph replays."""
    train = spr.Scene.generate(2, 4, 1500, 1024)
    scene = spr.Scene.load_rvg(RVG_DIR / "chord.rvg")
    a = scene.arrays()
    w, h = a["width"], a["height"]
    ref = orc.Frame(a, w, h, compose=1, threads=8).image()
    r = spr.Rasterizer(w, h, compose=1, mode=spr.MODE_LEARNED)
    r.learned_train(train, epochs=10, chunk_paths=500)
    r.load_scene(scene)
    for frame in range(4):
        print(f"[rvg-learned] frame={frame}", flush=True)
        assert _image_diff(r.render(), ref).max() == 0, frame
    r.close()


def test_scene_create_with_conics_and_paints(spr, orc):
    """A scene built from arrays (spr_scene_create): a quadratic, an arc of weight 0.6 and a
    three-quarter circle split into arcs, each path with another paint kind and spread."""
    pos, pos_path, cpos, ctype, cpath = [], [], [], [], []
    weights = []

    def curve(path, t, pts, wgt=1.0):
        cpos.append(len(pos_path))
        ctype.append(t)
        cpath.append(path)
        weights.append(wgt)
        for p in pts:
            pos.append(p)
            pos_path.append(path)

    # path 0: a quadratic lens, closed by a line
    curve(0, spr.CURVE_QUAD, [(40.0, 60.0), (150.0, 300.0), (260.0, 60.0)])
    curve(0, spr.CURVE_LINE, [(260.0, 60.0), (40.0, 60.0)])
    # path 1: a circle of radius 90 at (350, 250) from four quarter arcs (w = cos 45 deg)
    cx, cy, rr, wq = 350.0, 250.0, 90.0, float(np.cos(np.pi / 4))
    ang = [0.0, np.pi / 2, np.pi, 1.5 * np.pi, 2 * np.pi]
    for k in range(4):
        p0 = (cx + rr * np.cos(ang[k]), cy + rr * np.sin(ang[k]))
        p2 = (cx + rr * np.cos(ang[k + 1]), cy + rr * np.sin(ang[k + 1]))
        mid = 0.5 * (ang[k] + ang[k + 1])
        pc = (cx + rr / wq * np.cos(mid), cy + rr / wq * np.sin(mid))
        curve(1, spr.CURVE_ARC, [p0, pc, p2], wq)
    # path 2: a flat arc of weight 0.6 closed by a line, nonzero
    curve(2, spr.CURVE_ARC, [(100.0, 400.0), (200.0, 520.0), (300.0, 400.0)], 0.6)
    curve(2, spr.CURVE_LINE, [(300.0, 400.0), (100.0, 400.0)])
    paints = np.zeros(3, dtype=spr.PAINT_DTYPE)
    stops = np.array([[0.0, 1, 0, 0, 1], [0.5, 0, 1, 0, 0.6], [1.0, 0, 0, 1, 1],
                      [0.0, 1, 1, 0, 1], [1.0, 0, 0.5, 0.5, 0.8]], dtype=np.float32)
    ident = [1, 0, 0, 0, 1, 0]
    paints[0] = (spr.PAINT_LINEAR, spr.SPREAD_REFLECT, 0, 3, 0.9, ident, [40, 60, 120, 100, 0])
    paints[1] = (spr.PAINT_RADIAL, spr.SPREAD_PAD, 3, 2, 1.0, ident, [350, 250, 380, 260, 90])
    paints[2] = (spr.PAINT_LINEAR, spr.SPREAD_REPEAT, 0, 3, 1.0, [0.5, 0, 0, 0, 0.5, 0], [50, 200, 100, 260, 0])
    fill = np.array([0xFF0000FF, 0xFF00FF00, 0x80FF0000], dtype=np.uint32)
    scene = spr.Scene.create(np.array(pos, np.float32), pos_path, cpos, ctype, cpath, [0, 1, 0], fill,
                             width=512, height=600, curve_weight=weights, path_paint=[0, 1, 2],
                             paints=paints, stops=stops)
    a = scene.arrays()
    for compose in (0, 1):
        print(f"[scene-create] compose={compose}", flush=True)
        f = orc.Frame(a, 512, 600, compose=compose)
        r = spr.Rasterizer(512, 600, compose=compose, debug=True)
        r.load_scene(scene)
        img = r.render()
        _compare_stages(r, f)
        assert _image_diff(img, f.image()).max() == 0
        r.close()
