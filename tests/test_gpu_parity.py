"""GPU parity: every stage of the HIP pipeline against the CPU oracle, bit for bit.

Integer stages (counts, intersections, keys, sorted order, winding numbers, flags, records) and
the overwrite image must be identical; the src-over image is compared within 1/255 on 99.9% of
pixels (it is also expected to be identical, as both accumulate in fp32 in the same order).
"""
import sys

import numpy as np
import pytest

from conftest import RVG_DIR

pytestmark = pytest.mark.gpu

CASES = {
    # name: (generator kind, seed, n_paths, canvas, compose)
    "scribbles_A": (0, 1, 64, 1024, 0),
    "glyphs_small": (1, 2, 2000, 2048, 0),
    "illustration_small": (2, 3, 200, 4096, 1),
    "illustration_overwrite": (2, 7, 120, 2048, 0),
    "scribbles_odd_canvas": (0, 11, 24, 777, 0),
}


def _scene(spr, name):
    if name.startswith("rvg:"):
        return spr.Scene.load_rvg(RVG_DIR / name[4:]), 0
    kind, seed, n, canvas, compose = CASES[name]
    return spr.Scene.generate(kind, seed, n, canvas), compose


def _frames(spr, orc, name):
    scene, compose = _scene(spr, name)
    a = scene.arrays()
    w, h = a["width"], a["height"]
    print(f"[parity] case={name} w={w} h={h} paths={scene.n_paths} curves={scene.n_curves}", flush=True)
    f = orc.Frame(a, w, h, compose=compose, threads=8)
    r = spr.Rasterizer(w, h, compose=compose, debug=True)
    r.load_scene(scene)
    img = r.render()
    return a, f, r, img


@pytest.mark.parametrize("name", list(CASES) + ["rvg:tiger.rvg", "rvg:test.rvg", "rvg:reschart.rvg"])
def test_all_stages_bit_exact(spr, orc, name):
    a, f, r, img = _frames(spr, orc, name)
    N = len(f.get("inter")) // 2
    assert N > 0
    for st in ["tpos", "visible", "cuts", "counts", "inter", "key", "wn", "seg", "skey", "sidx"]:
        g, o = r.export(st), f.get(st)
        assert g.shape == o.shape, (st, g.shape, o.shape)
        if g.dtype == np.float32:
            assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), st
        else:
            bad = np.flatnonzero(g != o)
            assert bad.size == 0, (st, bad[:10], g[bad[:10]], o[bad[:10]])
    w = r.export("winding")
    assert np.array_equal(w[: N + 1], f.get("winding")[: N + 1])
    assert np.array_equal(r.export("flags"), f.get("flags"))
    assert np.array_equal(r.export("records"), f.get("records"))
    st = r.stats()
    assert st.n_intersections == N
    assert st.n_cells + st.n_spans == len(f.get("records")) // 4
    oi = f.image()
    if CASES.get(name, (0, 0, 0, 0, 0))[4] == 0:
        assert np.array_equal(img, oi)
    else:
        diff = np.abs(img.astype(np.int16) - oi.astype(np.int16)).max(-1)
        assert (diff <= 1).mean() >= 0.999
        assert (diff == 0).mean() > 0.99


@pytest.mark.parametrize("name", list(CASES) + ["rvg:tiger.rvg"])
def test_production_frames_match_oracle(spr, orc, name):
    """Frames without debug export sort 32-bit words without the entry index and fuse mark + emit
    into the band sort (each band writes its own records and bucket ranges): the intersection,
    cell and span counts and the image are still the oracle's, and the sorted keys and winding
    numbers, which such frames never store, cannot be exported."""
    scene, compose = _scene(spr, name)
    a = scene.arrays()
    w, h = a["width"], a["height"]
    print(f"[production] case={name} w={w} h={h}", flush=True)
    f = orc.Frame(a, w, h, compose=compose, threads=8)
    r = spr.Rasterizer(w, h, compose=compose)
    r.load_scene(scene)
    img = r.render()
    N = len(f.get("skey"))
    # production frames keep neither the fragment keys nor the sorted ones: the piece kernels store
    # each entry's bucket and narrow sort word instead
    for stage in ("key", "wn", "skey", "winding"):
        with pytest.raises(RuntimeError):
            r.export(stage)
    st = r.stats()
    assert st.n_intersections == N
    if name == "scribbles_A":
        # few long chains: each launch wave-solves its costliest classes (none reaches 896 crossings)
        assert st.n_wave_chains > 0
    rec = f.get("records").reshape(-1, 4)  # cells carry their 1-based index in .w, spans 0
    assert st.n_cells == int((rec[:, 3] != 0).sum())
    assert st.n_spans == int((rec[:, 3] == 0).sum())
    oi = f.image()
    if compose == 0:
        assert np.array_equal(img, oi)
    else:
        diff = np.abs(img.astype(np.int16) - oi.astype(np.int16)).max(-1)
        assert (diff <= 1).mean() >= 0.999
        assert (diff == 0).mean() > 0.99


def test_injected_intersections_integer_stages(spr, orc):
    """Integer stages on identical intersection inputs: the oracle's intersections injected."""
    scene, _ = _scene(spr, "scribbles_A")
    a = scene.arrays()
    f = orc.Frame(a, a["width"], a["height"], threads=8)
    r = spr.Rasterizer(a["width"], a["height"], debug=True)
    r.load_scene(scene)
    inter = f.get("inter").copy()
    r.inject_intersections(inter)
    img = r.render()
    for st in ["inter", "key", "wn", "seg", "skey", "sidx", "flags", "records"]:
        assert np.array_equal(r.export(st), f.get(st)), st
    assert np.array_equal(img, f.image())
    # a perturbed intersection list (every other t nudged by one ulp) still matches the oracle
    # fed the same list
    pert = inter.reshape(-1, 2).copy()
    pert[::2, 1] = np.maximum(pert[::2, 1] - 4, 0)
    f2 = orc.Frame(a, a["width"], a["height"], inter_override=pert, threads=8)
    r.inject_intersections(pert)
    img2 = r.render()
    for st in ["key", "wn", "skey", "sidx", "flags", "records"]:
        assert np.array_equal(r.export(st), f2.get(st)), st
    assert np.array_equal(img2, f2.image())
    r.inject_intersections(None)


def test_injected_intersections_production_frame(spr, orc):
    """Injected intersections in a frame without debug (narrow sort words, mark + emit fused into
    the band sort): the fragment stage then stores keys rather than the packed bucket and sort
    word of generated entries, and the image, counts and keys are the oracle's."""
    scene, _ = _scene(spr, "scribbles_A")
    a = scene.arrays()
    f = orc.Frame(a, a["width"], a["height"], threads=8)
    r = spr.Rasterizer(a["width"], a["height"])
    r.load_scene(scene)
    r.inject_intersections(f.get("inter").copy())
    img = r.render()
    assert np.array_equal(img, f.image())
    assert np.array_equal(r.export("key"), f.get("key"))
    assert np.array_equal(r.export("wn"), f.get("wn"))
    st = r.stats()
    rec = f.get("records").reshape(-1, 4)
    assert st.n_cells == int((rec[:, 3] != 0).sum())
    assert st.n_spans == int((rec[:, 3] == 0).sum())
    r.inject_intersections(None)
    assert np.array_equal(r.render(), img)  # generated again: packed entries, same image
    with pytest.raises(RuntimeError):
        r.export("key")
    r.close()


def test_repeat_frames_identical(spr):
    scene, _ = _scene(spr, "illustration_overwrite")
    a = scene.arrays()
    r = spr.Rasterizer(a["width"], a["height"])
    r.load_scene(scene)
    i1 = r.render()
    i2 = r.render()
    i3 = r.render()
    assert np.array_equal(i1, i2) and np.array_equal(i2, i3)


def test_mvp_and_culling(spr, orc):
    """A non-identity MVP (scale + translate) that pushes some paths off-canvas."""
    scene, _ = _scene(spr, "scribbles_A")
    a = scene.arrays()
    m = np.array([[1.7, 0.2, 0, -300.0], [-0.1, 1.3, 0, 150.0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)
    f = orc.Frame(a, 1024, 1024, mvp=m, threads=8)
    r = spr.Rasterizer(1024, 1024, debug=True)
    r.load_scene(scene)
    r.set_mvp(m)
    img = r.render()
    assert np.array_equal(r.export("tpos").view(np.uint32), f.get("tpos").view(np.uint32))
    assert np.array_equal(r.export("visible"), f.get("visible"))
    assert np.array_equal(r.export("records"), f.get("records"))
    assert np.array_equal(img, f.image())


def test_empty_and_offcanvas(spr, orc):
    scene, _ = _scene(spr, "scribbles_A")
    a = scene.arrays()
    r = spr.Rasterizer(256, 256, debug=True)
    # everything translated far outside the canvas: every path culled, white image
    m = np.eye(4, dtype=np.float32)
    m[0, 3] = 5000.0
    r.load_scene(scene)
    r.set_mvp(m)
    img = r.render()
    assert (img == 255).all()
    assert r.stats().n_intersections == 0
    # zero paths
    r.load_paths(np.zeros((0, 2)), [], [], [], [], [], [])
    img = r.render()
    assert (img == 255).all()


def test_golden_vectors(spr):
    """The HIP path against the committed golden vectors (oracle outputs on fixed scenes)."""
    import make_golden
    g = np.load(make_golden.HERE / "oracle_golden.npz")
    for name, arrays, w, h, mvp, compose in make_golden.cases(spr):
        print(f"[golden] case={name}", flush=True)
        r = spr.Rasterizer(w, h, compose=compose, debug=True)
        r.load_paths(arrays["pos"], arrays["pos_path"], arrays["curve_pos"], arrays["curve_type"],
                     arrays["curve_path"], arrays["rule"], arrays["fill"])
        if mvp is not None:
            r.set_mvp(mvp)
        img = r.render()
        for stage in make_golden.STAGES:
            k = f"{name}/{stage}"
            if k in g:
                got = r.export(stage)[: len(g[k])]
                assert np.array_equal(got, g[k]), k
        assert make_golden.digest(img) == str(g[f"{name}/image_sha256"]), name
        r.close()


@pytest.mark.parametrize("knobs", [
    dict(band_window_log=10), dict(scan_shape=1), dict(scan_shape=2), dict(scan_shape=3),
    dict(chain_blocks_per_cu=4, merge_blocks_per_cu=16), dict(chain_blocks_per_cu=16, merge_blocks_per_cu=4),
    dict(band_window_log=10, scan_shape=3, chain_blocks_per_cu=16, merge_blocks_per_cu=16),
])
def test_tuning_knobs_keep_output(spr, orc, knobs):
    """Every launch knob the autotuner can pick gives the oracle's frame: 1024-entry band windows
    (bands up to 2048 words), each large-scan tile shape, other persistent-kernel occupancies."""
    for name in ("illustration_small", "glyphs_small"):
        scene, compose = _scene(spr, name)
        a = scene.arrays()
        w, h = a["width"], a["height"]
        print(f"[knobs] {knobs} case={name}", flush=True)
        f = orc.Frame(a, w, h, compose=compose, threads=8)
        r = spr.Rasterizer(w, h, compose=compose)
        r.set_tuning(**knobs)
        r.load_scene(scene)
        img = r.render()
        st = r.stats()
        assert st.n_intersections == len(f.get("skey"))
        if compose == 0:
            assert np.array_equal(img, f.image())
        else:
            diff = np.abs(img.astype(np.int16) - f.image().astype(np.int16)).max(-1)
            assert (diff <= 1).mean() >= 0.999
        r.close()
    r = spr.Rasterizer(64, 64)
    with pytest.raises(spr.SprError):
        r.set_tuning(band_window_log=11)
    r.close()
