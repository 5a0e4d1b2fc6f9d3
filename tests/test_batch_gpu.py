"""GPU tests of batched documents and captured-graph replays.

Every document of a batch must come out bit-identical to rendering that document alone with the
plain method (each document keeps its own canvas and its own winding numbers), on the eager
frames and on every replay of the captured graph, with eager work of another handle and host
synchronisations between the replays."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DOCS = [  # (kind, seed, n_paths, width, height)
    (0, 101, 12, 256, 256), (1, 102, 150, 512, 384), (2, 103, 60, 777, 500), (0, 104, 5, 300, 1024),
    (2, 105, 200, 1024, 1024), (1, 106, 40, 256, 256), (2, 107, 1, 64, 64), (1, 108, 300, 640, 480),
    (0, 109, 30, 512, 512), (2, 110, 90, 2048, 1536),
]


def _plain(spr, scene, w, h, compose):
    p = spr.Rasterizer(w, h, compose=compose)
    p.load_scene(scene)
    img = p.render()
    p.close()
    return img


@pytest.mark.parametrize("compose", [0, 1])
def test_batch_matches_single_documents(spr, compose):
    scenes = [spr.Scene.generate(k, s, n, max(w, h)) for k, s, n, w, h in DOCS]
    ref = [_plain(spr, sc, w, h, compose) for sc, (_, _, _, w, h) in zip(scenes, DOCS)]
    r = spr.Rasterizer(64, 64, compose=compose)
    r.load_batch(scenes, [(w, h) for _, _, _, w, h in DOCS])
    other = spr.Rasterizer(256, 256)
    other.load_scene(spr.Scene.generate(2, 7, 50, 256))
    for frame in range(5):  # eager (budgets), eager + capture, then graph replays
        print(f"[batch] compose={compose} frame={frame}", flush=True)
        r.render(readback=False)
        for d, (_, _, _, w, h) in enumerate(DOCS):
            assert np.array_equal(r.batch_image(d), ref[d]), (frame, d)
        other.render()  # eager work of another handle between replays
    st = r.stats()
    assert st.n_paths == sum(n for _, _, n, _, _ in DOCS)
    r.close()
    other.close()


def test_learned_mode_graph_replays(spr):
    """A learned scene is captured after two frames; every replay equals the plain image."""
    train = spr.Scene.generate(2, 4, 1500, 512)
    r = spr.Rasterizer(512, 512, compose=1, mode=spr.MODE_LEARNED)
    r.learned_train(train, epochs=10, chunk_paths=500)
    scene = spr.Scene.generate(2, 12, 150, 512)
    ref = _plain(spr, scene, 512, 512, 1)
    r.load_scene(scene)
    for frame in range(5):
        print(f"[learned-graph] frame={frame}", flush=True)
        assert np.array_equal(r.render(), ref), frame
        _plain(spr, scene, 512, 512, 0)  # eager work between replays
    r.close()


@pytest.mark.parametrize("compose", [0, 1])
def test_batch_documents_match_oracle(spr, orc, compose):
    """Each document of a batch (square and non-square canvases, one single-path document)
    against the CPU oracle rendering it alone on its own canvas."""
    import os
    threads = max(1, min(16, os.cpu_count() or 1))
    scenes = [spr.Scene.generate(k, s, n, max(w, h)) for k, s, n, w, h in DOCS]
    r = spr.Rasterizer(64, 64, compose=compose)
    r.load_batch(scenes, [(w, h) for _, _, _, w, h in DOCS])
    for frame in range(3):  # the first frame counts (budget), then eager + capture, then a replay
        print(f"[batch-oracle] compose={compose} frame={frame}", flush=True)
        r.render(readback=False)
    assert r.stats().learned_fallbacks == 1  # the first frame had no budget: counted, rendered again
    for d, (sc, (_, _, _, w, h)) in enumerate(zip(scenes, DOCS)):
        f = orc.Frame(sc.arrays(), w, h, compose=compose, threads=threads)
        img, oi = r.batch_image(d), f.image()
        if compose == 0:
            assert np.array_equal(img, oi), d
        else:
            diff = np.abs(img.astype(np.int16) - oi.astype(np.int16)).max(-1)
            assert (diff <= 1).mean() >= 0.999, d
    r.close()


def test_autotune_keeps_output(spr):
    scenes = [spr.Scene.generate(k, s, n, max(w, h)) for k, s, n, w, h in DOCS[:5]]
    ref = [_plain(spr, sc, w, h, 0) for sc, (_, _, _, w, h) in zip(scenes, DOCS[:5])]
    r = spr.Rasterizer(64, 64)
    r.load_batch(scenes, [(w, h) for _, _, _, w, h in DOCS[:5]])
    rep = r.autotune(reps=2)
    print(f"[autotune] {rep}", flush=True)
    assert rep["n_tried"] == 9 and rep["best_ms"] <= rep["default_ms"] + 1e-9
    assert rep["chain_blocks_per_cu"] in (4, 8, 16) and rep["merge_blocks_per_cu"] in (4, 8, 16)
    assert rep["scan_shape"] in (0, 1, 2, 3) and rep["band_window_log"] in (9, 10)
    for _ in range(3):
        r.render(readback=False)
        for d in range(5):
            assert np.array_equal(r.batch_image(d), ref[d])
    r.close()


def test_autotune_plain_handle(spr, orc):
    """The autotuner on a plain (count-then-write) handle times eager frames; the output stays the
    oracle's."""
    scene = spr.Scene.generate(2, 3, 200, 2048)
    a = scene.arrays()
    r = spr.Rasterizer(2048, 2048, compose=1)
    r.load_scene(scene)
    rep = r.autotune(reps=2)
    print(f"[autotune-plain] {rep}", flush=True)
    assert rep["n_tried"] == 9
    f = orc.Frame(a, 2048, 2048, compose=1, threads=8)
    assert np.array_equal(r.render(), f.image())
    r.close()
