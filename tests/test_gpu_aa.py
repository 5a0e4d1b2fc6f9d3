"""GPU parity of the anti-aliased mode (spr_opts.aa = 16): every record's 64-sample coverage
mask (k_cell_cover) and the coverage-weighted composite against the oracle, bit for bit, on the
reference's RVG scenes (arcs and gradients included), generated scenes, a transformed view and
an odd-sized canvas; plus the option's validation."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import RVG_DIR

pytestmark = pytest.mark.gpu

GEN = {  # name: (kind, seed, paths, canvas)
    "illustration": (2, 3, 150, 1024),
    "glyphs": (1, 2, 600, 1024),
    "scribbles_odd": (0, 11, 24, 777),
}


def _scene(spr, name):
    if name.endswith(".rvg"):
        s = spr.Scene.load_rvg(RVG_DIR / name)
        a = s.arrays()
        return s, a["width"], a["height"]
    k, seed, n, c = GEN[name]
    return spr.Scene.generate(k, seed, n, c), c, c


@pytest.mark.parametrize("name", ["tiger.rvg", "car.rvg", "chord.rvg", "drops.rvg"] + list(GEN))
def test_aa_coverage_and_image_match_oracle(spr, orc, name):
    scene, w, h = _scene(spr, name)
    a = scene.arrays()
    print(f"[aa] {name} {w}x{h} paths={scene.n_paths}", flush=True)
    f = orc.Frame(a, w, h, compose=1, threads=8, aa=16)
    r = spr.Rasterizer(w, h, compose=1, aa=16)
    r.load_scene(scene)
    for frame in range(2):
        img = r.render()
        cm = r.export("coverage")
        ref = f.get("cmask")
        assert cm.shape == ref.shape, (cm.shape, ref.shape)
        bad = np.flatnonzero(cm != ref)
        assert bad.size == 0, (frame, bad[:8], cm[bad[:8]], ref[bad[:8]])
        d = np.abs(img.astype(np.int16) - f.image().astype(np.int16)).max(-1)
        assert d.max() == 0, (frame, int(d.max()), int((d > 0).sum()))
    st = r.stats()
    assert st.n_cells + st.n_spans == len(ref)
    # the binary frame of the same scene differs only on boundary cells
    assert (cm != np.uint64(0xFFFFFFFFFFFFFFFF)).sum() > 0
    r.close()


def test_aa_under_mvp(spr, orc):
    scene, w, h = _scene(spr, "car.rvg")
    m = np.eye(4, dtype=np.float32)
    c, s = np.cos(0.4) * 1.3, np.sin(0.4) * 1.3
    m[0, 0], m[0, 1], m[0, 3] = c, -s, -60.0
    m[1, 0], m[1, 1], m[1, 3] = s, c, -140.0
    print("[aa-mvp] car", flush=True)
    f = orc.Frame(scene.arrays(), w, h, mvp=m, compose=1, threads=8, aa=16)
    r = spr.Rasterizer(w, h, compose=1, aa=16)
    r.load_scene(scene)
    r.set_mvp(m)
    img = r.render()
    assert np.array_equal(r.export("coverage"), f.get("cmask"))
    assert np.array_equal(img, f.image())
    r.close()


def test_aa_option_validation(spr):
    with pytest.raises(spr.SprError, match="aa must be 0 or 16"):
        spr.Rasterizer(64, 64, compose=1, aa=8)
    with pytest.raises(spr.SprError, match="compose src-over and the plain mode"):
        spr.Rasterizer(64, 64, compose=0, aa=16)
    with pytest.raises(spr.SprError, match="compose src-over and the plain mode"):
        spr.Rasterizer(64, 64, compose=1, aa=16, mode=spr.MODE_LEARNED)
    r = spr.Rasterizer(64, 64, compose=1, aa=16)
    with pytest.raises(spr.SprError, match="without anti-aliased coverage"):
        r.load_batch([spr.Scene.generate(0, 1, 4, 64)], [(64, 64)])
    r.close()
    r = spr.Rasterizer(64, 64, compose=1)
    r.load_scene(spr.Scene.generate(0, 1, 4, 64))
    r.render()
    with pytest.raises(spr.SprError, match="anti-aliased frames"):
        r.export("coverage")
    r.close()
