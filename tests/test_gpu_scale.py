"""GPU parity at the configured sizes, against the CPU oracle.

- The benched frame: the 4096^2 illustration (5,000 paths, ~49M intersections, src-over), in the
  production form (32-bit sort words, mark + emit fused into the band sort) and in the debug form
  (every stage exported and compared bit for bit).
- The 20,000-glyph page at 2048^2 (even-odd, overwrite): every stage bit for bit, and the
  production frame.
- The learned single-pass frame on the illustration at 4096^2, its predictor trained on 50,000
  paths: identical to the plain frame and to the oracle.
- One constructed long cubic (more than 2,000 crossings per axis), compared on every stage.

Each frame also reports which kernel routes it took (wave-solved chains, workgroup-merged pieces,
long and oversized sort bands), so the parity below is known to cover them.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
DEBUG_STAGES = ["tpos", "visible", "cuts", "counts", "inter", "key", "wn", "seg", "skey", "sidx"]


def _close_images(img, oi, exact):
    if exact:
        assert np.array_equal(img, oi)
        return
    diff = np.abs(img.astype(np.int16) - oi.astype(np.int16)).max(-1)
    assert (diff <= 1).mean() >= 0.999, (diff <= 1).mean()


def _check_debug_frame(spr, f, scene, w, h, compose):
    """Every exported stage of a debug frame against the oracle frame f."""
    r = spr.Rasterizer(w, h, compose=compose, debug=True)
    r.load_scene(scene)
    img = r.render()
    for st in DEBUG_STAGES:
        g, o = r.export(st), f.get(st)
        assert g.shape == o.shape, (st, g.shape, o.shape)
        if g.dtype == np.float32:
            assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), st
        else:
            bad = np.flatnonzero(g != o)
            assert bad.size == 0, (st, bad[:10], g[bad[:10]], o[bad[:10]])
        del g, o
    N = len(f.get("skey"))
    assert np.array_equal(r.export("winding")[: N + 1], f.get("winding")[: N + 1])
    assert np.array_equal(r.export("flags"), f.get("flags"))
    assert np.array_equal(r.export("records"), f.get("records"))
    _close_images(img, f.image(), compose == 0)
    st = r.stats()
    r.close()
    return st


def _check_production_frame(spr, f, scene, w, h, compose):
    r = spr.Rasterizer(w, h, compose=compose)
    r.load_scene(scene)
    img = r.render()
    st = r.stats()
    rec = f.get("records").reshape(-1, 4)  # cells carry their 1-based index in .w, spans 0
    assert st.n_intersections == len(f.get("skey"))
    assert st.n_cells == int((rec[:, 3] != 0).sum())
    assert st.n_spans == int((rec[:, 3] == 0).sum())
    _close_images(img, f.image(), compose == 0)
    r.close()
    return st, img


def _routes(st):
    return dict(wave_chains=st.n_wave_chains, heavy_pieces=st.n_heavy_pieces, long_bands=st.n_long_bands,
                big_bands=st.n_big_bands)


@pytest.fixture(scope="module")
def illustration(spr, orc):
    scene = spr.Scene.generate(2, 3, 5000, 4096)
    print(f"[scale] illustration_4096: oracle on {THREADS} threads", flush=True)
    f = orc.Frame(scene.arrays(), 4096, 4096, compose=1, threads=THREADS)
    return scene, f


def test_illustration_4096_production_frame(spr, illustration):
    """The frame bench.py times, against the oracle: intersection, cell and span counts exact,
    and at least 99.9% of the pixels within 1/255 (src-over)."""
    scene, f = illustration
    st, img = _check_production_frame(spr, f, scene, 4096, 4096, 1)
    routes = _routes(st)
    print(f"[scale] illustration_4096 production: N={st.n_intersections} routes={routes}", flush=True)
    assert st.n_intersections > 40_000_000
    # the frame exercises the wave-solved chains, the workgroup-merged pieces, the long bands and
    # the oversized ones (sorted and emitted on the second stream beside the wave sorts)
    assert routes["wave_chains"] > 0 and routes["heavy_pieces"] > 0 and routes["long_bands"] > 0
    assert routes["big_bands"] > 0
    # src-over accumulates in fp32 in path order on both sides: identical in practice
    assert np.array_equal(img, f.image())


def test_illustration_4096_debug_frame_every_stage(spr, illustration):
    """The same scene with every stage exported: each one bit for bit."""
    scene, f = illustration
    st = _check_debug_frame(spr, f, scene, 4096, 4096, 1)
    print(f"[scale] illustration_4096 debug: routes={_routes(st)}", flush=True)
    assert st.n_wave_chains > 0 and st.n_heavy_pieces > 0


def test_glyph_page_2048_full(spr, orc):
    """20,000 glyphs (30% with a counter-wound hole), even-odd, 2048^2: every stage, then the
    production frame."""
    scene = spr.Scene.generate(1, 2, 20000, 2048)
    assert scene.n_paths == 20000
    print(f"[scale] glyphs_2048: oracle on {THREADS} threads", flush=True)
    f = orc.Frame(scene.arrays(), 2048, 2048, compose=0, threads=THREADS)
    _check_debug_frame(spr, f, scene, 2048, 2048, 0)
    st, _ = _check_production_frame(spr, f, scene, 2048, 2048, 0)
    print(f"[scale] glyphs_2048 production: N={st.n_intersections} routes={_routes(st)}", flush=True)


def test_learned_illustration_4096_identical(spr, illustration):
    """The learned single-pass frame at the benched size, its predictor trained on the GPU on
    50,000 paths of the same generator (another seed): identical to the plain frame (hence to
    the oracle), on the eager frames and on the captured-graph replays."""
    scene, f = illustration
    r = spr.Rasterizer(4096, 4096, compose=1, mode=spr.MODE_LEARNED)
    print("[scale] learned: training on 50,000 paths", flush=True)
    rep = r.learned_train(spr.Scene.generate(2, 4, 50000, 4096), chunk_paths=5000)
    assert rep["samples"] == 50000
    r.load_scene(scene)
    p = spr.Rasterizer(4096, 4096, compose=1)
    p.load_scene(scene)
    ip = p.render()
    sp = p.stats()
    p.close()
    assert np.array_equal(ip, f.image())
    for frame in range(4):  # eager (budgets), eager + capture, replays
        il = r.render()
        assert np.array_equal(il, ip), frame
        sl = r.stats()
        assert (sl.n_intersections, sl.n_cells, sl.n_spans) == (sp.n_intersections, sp.n_cells, sp.n_spans)
    print(f"[scale] learned: overflow paths {sl.n_overflow_paths}, allocated {sl.alloc_intersections} "
          f"for {sl.n_intersections}", flush=True)
    assert sl.alloc_intersections > 0
    r.close()


def _long_cubic_scene():
    """Two closed paths on a 4096^2 canvas: a near-diagonal cubic across the whole canvas (one
    monotone piece crossing about 2,000 vertical and 2,000 horizontal grid lines) closed by a
    line, and an S-shaped cubic with four interior extrema closed the same way."""
    pos = np.array([
        [20.3, 30.7], [1400.2, 1350.9], [2700.6, 2750.1], [4050.4, 4060.2],  # cubic 0
        [4050.4, 4060.2], [20.3, 30.7],                                     # line back
        [100.5, 3900.25], [6000.0, 3000.0], [-2000.0, 1000.0], [3950.75, 150.5],  # cubic 1
        [3950.75, 150.5], [100.5, 3900.25],                                 # line back
    ], np.float32)
    pos_path = np.array([0] * 6 + [1] * 6, np.uint32)
    curve_pos = np.array([0, 4, 6, 10], np.uint32)
    curve_type = np.array([4, 2, 4, 2], np.uint32)
    curve_path = np.array([0, 0, 1, 1], np.uint32)
    rule = np.array([0, 1], np.uint32)
    fill = np.array([0xFF2040C0, 0xFF10A020], np.uint32)
    return dict(pos=pos, pos_path=pos_path, curve_pos=curve_pos, curve_type=curve_type,
                curve_path=curve_path, rule=rule, fill=fill)


def test_long_cubic_every_stage(spr, orc):
    a = _long_cubic_scene()
    f = orc.Frame(a, 4096, 4096, threads=THREADS)
    N = len(f.get("skey"))
    assert N > 6000
    for debug in (True, False):
        r = spr.Rasterizer(4096, 4096, debug=debug)
        r.load_paths(a["pos"], a["pos_path"], a["curve_pos"], a["curve_type"], a["curve_path"], a["rule"],
                     a["fill"])
        img = r.render()
        st = r.stats()
        print(f"[scale] long cubic debug={debug}: N={st.n_intersections} routes={_routes(st)}", flush=True)
        assert st.n_wave_chains >= 2
        assert st.n_intersections == N
        if debug:
            for stage in DEBUG_STAGES + ["flags", "records"]:
                g, o = r.export(stage), f.get(stage)
                if g.dtype == np.float32:
                    assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), stage
                else:
                    assert np.array_equal(g, o), stage
            assert np.array_equal(r.export("winding")[: N + 1], f.get("winding")[: N + 1])
        assert np.array_equal(img, f.image())
        r.close()
