"""GPU record replay (the reference's MOCK_DATA path) and PNG output.

Record lists injected with spr_debug_inject(SPR_STAGE_RECORDS) are drawn by the GPU in list order;
the result is compared with the oracle's draw stage: the reference's own recorded lists
(workdir/test_data.csv, test_data3.csv, on the reference's 1200x1024 target), random lists in both
composite modes, and the records of an oracle frame (which must give that frame's image)."""
from __future__ import annotations

import zlib

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
GOLDEN = ROOT / "tests" / "golden"


def _fixture(name):
    if name == "test_data.csv":
        return np.loadtxt(GOLDEN / name, delimiter=",", dtype=np.int64).astype(np.int32)
    return np.load(GOLDEN / name)["records"]


@pytest.mark.parametrize("name", ["test_data.csv", "test_data3.npz"])
@pytest.mark.parametrize("compose", [0, 1])
def test_replay_reference_record_files(spr, orc, name, compose):
    rec = _fixture(name)
    r = spr.Rasterizer(1200, 1024, compose=compose)
    r.inject_records(rec)
    print(f"[replay] {name} compose={compose} records={len(rec)}", flush=True)
    img = r.render()
    assert np.array_equal(img, orc.draw(rec, 1200, 1024, compose=compose))
    st = r.stats()
    assert st.n_cells == int((rec[:, 3] != 0).sum()) and st.n_cells + st.n_spans == len(rec)
    r.close()


@pytest.mark.parametrize("compose", [0, 1])
def test_replay_random_records(spr, orc, compose):
    """Overlapping records anywhere (also off the canvas and across column chunks), each colour
    with its own alpha: the order of application decides every pixel."""
    rng = np.random.default_rng(5 + compose)
    w, h, n = 2500, 300, 3000
    x = rng.integers(-50, w + 50, n)
    y = rng.integers(-3, h + 3, n)
    rec = np.stack([(y << 16) | (x & 0xFFFF), rng.integers(1, 1500, n),
                    rng.integers(0, 1 << 32, n, dtype=np.int64), rng.integers(0, 5, n)], 1).astype(np.int64)
    rec = rec.astype(np.uint32).view(np.int32).reshape(-1, 4)
    r = spr.Rasterizer(w, h, compose=compose)
    r.inject_records(rec)
    assert np.array_equal(r.render(), orc.draw(rec, w, h, compose=compose))
    r.close()


def test_replay_frame_records_give_the_frame(spr, orc):
    """The records of a whole oracle frame replayed: the frame's image (src-over included)."""
    scene = spr.Scene.generate(2, 3, 120, 1024)
    a = scene.arrays()
    f = orc.Frame(a, 1024, 1024, compose=1, threads=8)
    rec = f.get("records").reshape(-1, 4)
    r = spr.Rasterizer(1024, 1024, compose=1)
    r.load_scene(scene)
    r.inject_records(rec)
    assert np.array_equal(r.render(), f.image())
    r.inject_records(None)  # cleared: the scene again
    assert np.array_equal(r.render(), f.image())
    assert r.stats().n_intersections == len(f.get("skey"))
    r.close()


def _png_pixels(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        assert int.from_bytes(data[pos + 8 + n:pos + 12 + n], "big") == zlib.crc32(typ + body)
        if typ == b"IHDR":
            w, h = int.from_bytes(body[:4], "big"), int.from_bytes(body[4:8], "big")
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(h, 4 * w + 1)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 4)


def test_write_png_of_a_frame(spr, tmp_path):
    from conftest import RVG_DIR
    r = spr.Rasterizer(512, 512)
    r.load_rvg(RVG_DIR / "tiger.rvg")
    img = r.render()
    path = tmp_path / "tiger.png"
    r.write_png(path)
    assert np.array_equal(_png_pixels(path), img)
    r.close()
