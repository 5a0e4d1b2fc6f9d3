"""GPU parity of the oversized-band route: a (path, cell row) bucket of more than one sort tile
(2048 entries) is sorted through global memory, and in production frames its cells and spans
are emitted from the sorted keys by a kernel of its own.  A zigzag path confined to two cell rows
puts several thousand fragments into each of its buckets; production frames (narrow words,
mark + emit fused into the sort) and debug frames (separate mark + emit) must both give the
oracle's record counts and image, and the debug records must be the oracle's bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 1024, 256


def _zigzag_scene():
    """Path 0: 8000 line segments between y = 101.1 and 102.9 from x = 4 to 1004, closed by one
    line back; path 1: a rectangle crossing the same rows.  Curves own their points."""
    xs = np.linspace(4.0, 1004.0, 8001, dtype=np.float64)
    ys = np.where(np.arange(8001) % 2 == 0, 101.1, 102.9)
    zig = np.stack([xs, ys], 1)
    loop = np.concatenate([zig, zig[:1]])  # back to the start
    rect = np.array([[300.0, 60.0], [700.0, 60.0], [700.0, 180.0], [300.0, 180.0], [300.0, 60.0]])
    pos, pos_path, cpos, ctype, cpath = [], [], [], [], []
    for p, pts in enumerate((loop, rect)):
        for a, b in zip(pts[:-1], pts[1:]):
            cpos.append(len(pos))
            ctype.append(2)
            cpath.append(p)
            pos += [a, b]
            pos_path += [p, p]
    return {
        "pos": np.asarray(pos, np.float32), "pos_path": np.asarray(pos_path, np.uint32),
        "curve_pos": np.asarray(cpos, np.uint32), "curve_type": np.asarray(ctype, np.uint32),
        "curve_path": np.asarray(cpath, np.uint32), "rule": np.asarray([0, 1], np.uint32),
        "fill": np.asarray([0xC0FF4020, 0x8020A0FF], np.uint32), "width": W, "height": H,
    }


def _load(r, a):
    r.load_paths(a["pos"], a["pos_path"], a["curve_pos"], a["curve_type"], a["curve_path"], a["rule"],
                 a["fill"])


@pytest.mark.parametrize("compose", [0, 1])
def test_oversized_buckets_match_oracle(spr, orc, compose):
    a = _zigzag_scene()
    print(f"[big-bands] compose={compose}", flush=True)
    f = orc.Frame(a, W, H, compose=compose, threads=8)
    # the scene does exercise the route: some bucket of path 0 holds more than 2048 entries
    skey, seg = f.get("skey"), f.get("seg")
    rows = (skey[seg[0]:seg[1]].astype(np.int64) >> 16) & 0xFFFF
    assert np.bincount(rows).max() > 2048
    rec = f.get("records").reshape(-1, 4)  # cells carry their 1-based index in .w, spans 0
    n_cells, n_spans = int((rec[:, 3] != 0).sum()), int((rec[:, 3] == 0).sum())
    oi = f.image()
    for debug in (False, True):
        r = spr.Rasterizer(W, H, compose=compose, debug=debug)
        _load(r, a)
        img = r.render()
        st = r.stats()
        assert st.n_intersections == len(skey)
        assert (st.n_cells, st.n_spans) == (n_cells, n_spans), f"debug={debug}"
        assert np.array_equal(img, oi), f"debug={debug}"
        if debug:
            assert np.array_equal(r.export("records"), f.get("records"))
            assert np.array_equal(r.export("skey"), skey)
        r.close()
