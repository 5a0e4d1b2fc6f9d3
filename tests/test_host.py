"""Host-side checks that need no GPU: the C ABI library loads and exports every entry point
that include/spr.h declares, and the scene front end (RVG parser, generators, subsetting,
writer) behaves as the reference's loader does (VkScanlinePR/src/core/vg/rvg.cpp:9-255,
VkScanlinePR/src/core/scanline/scanline_rasterizer.cpp:67-189)."""
from __future__ import annotations

import re

import numpy as np
import pytest

from conftest import ROOT, RVG_DIR
import geometry as G

HEADER = ROOT / "include" / "spr.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b(spr_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    assert "spr_create" in names and "spr_render_frame" in names and len(names) >= 20


def test_library_exports_every_declared_function(spr):
    L = spr.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(set(spr.EXPORTS)) == sorted(set(declared_functions()) & set(spr.EXPORTS))
    assert set(declared_functions()) <= set(spr.EXPORTS), set(declared_functions()) - set(spr.EXPORTS)


def _rvg_counts(path):
    """Paths, explicit L/C curves and implicit closing lines of an RVG file, counted from its
    text: the last contour of an element that is not closed by Z and does not end at its
    start gets a closing line (rvg.cpp:210-223)."""
    n_paths = n_lines = n_cubics = n_close = 0
    for line in path.read_text().splitlines():
        if " element " not in line:
            continue
        n_paths += 1
        body = line.split(":", 1)[1].split("dyn_identity")[0].split()
        closed, start, cur = False, None, None
        i = 0
        while i < len(body):
            tok = body[i]
            if tok == "M":
                start = cur = body[i + 1]; closed = False; i += 2
            elif tok == "L":
                n_lines += 1; cur = body[i + 1]; i += 2
            elif tok == "C":
                n_cubics += 1; cur = body[i + 3]; i += 4
            elif tok == "Z":
                closed = True; i += 1
            else:
                i += 1
        if not closed and start is not None:
            sx, sy = (np.float32(v) for v in start.split(","))
            cx, cy = (np.float32(v) for v in cur.split(","))
            n_close += (sx != cx) or (sy != cy)
    return n_paths, n_lines, n_cubics, n_close


@pytest.mark.parametrize("name", ["test", "tiger", "reschart"])
def test_rvg_fixture_counts(spr, name):
    f = RVG_DIR / f"{name}.rvg"
    a = spr.Scene.load_rvg(f).arrays()
    n_paths, n_lines, n_cubics, n_close = _rvg_counts(f)
    assert len(a["rule"]) == n_paths
    assert int(np.sum(a["curve_type"] == 4)) == n_cubics
    assert int(np.sum(a["curve_type"] == 2)) == n_lines + n_close
    assert np.all(np.diff(a["curve_path"].astype(np.int64)) >= 0)
    assert np.all(a["curve_pos"] + a["curve_type"] <= len(a["pos_path"]))
    # each curve's points belong to its path
    for c in range(0, len(a["curve_type"]), 97):
        k = a["curve_type"][c]
        assert np.all(a["pos_path"][a["curve_pos"][c]:a["curve_pos"][c] + k] == a["curve_path"][c])


def test_rvg_known_values(spr):
    a = spr.Scene.load_rvg(RVG_DIR / "test.rvg").arrays()
    assert (a["width"], a["height"]) == (765, 990)
    # first element: nzfill rgba(1,1,1,1) rectangle M 0,0 L 765,0 L 765,990 L 0,990 L 0,0 Z
    assert a["rule"][0] == 0 and a["fill"][0] == 0xFFFFFFFF
    assert a["rule"][1] == 1 and a["fill"][1] == 0xFF000000
    first = a["curve_pos"][a["curve_path"] == 0]
    pts = a["pos"][first]
    assert np.array_equal(pts[:4], np.array([[0, 0], [765, 0], [765, 990], [0, 990]], np.float32))
    # first curve of the second element: L from 74.2583,912.88 to 74.2359,915.133
    c = np.nonzero(a["curve_path"] == 1)[0][0]
    assert a["curve_type"][c] == 2
    np.testing.assert_array_equal(a["pos"][a["curve_pos"][c]:a["curve_pos"][c] + 2],
                                  np.array([[74.2583, 912.88], [74.2359, 915.133]], np.float32))


@pytest.mark.parametrize("body,msg", [
    ("M 0,0 A 1,1 0 0 1 5,5 Z", "bad arc '1,1'"),  # SVG's arc syntax, not RVG's (x1,y1,w1 x2,y2)
    ("M 0,0 A 1,1,-1 5,5 Z", "arc weight must be finite and greater than -1"),
    ("M 0,0 q 1,1 2,2 Z", "unsupported path command 'q'"),
    ("M 0,0 l 1,1 Z", "unsupported path command 'l'"),
])
def test_rvg_rejects_unsupported_commands(spr, tmp_path, body, msg):
    f = tmp_path / "bad.rvg"
    f.write_text("viewport 0,0 10,10\nwindow 0,0 10,10\nscene dyn_identity\n"
                 f"  1 element nzfill dyn_concrete 0,0 0,0 0,0: {body} dyn_identity dyn_paint 1 solid rgba(0,0,0,1)\n")
    with pytest.raises(spr.SprError, match=re.escape(msg)):
        spr.Scene.load_rvg(f)


def test_rvg_rejects_gradients_and_missing_files(spr, tmp_path):
    f = tmp_path / "grad.rvg"
    f.write_text("viewport 0,0 10,10\nwindow 0,0 10,10\nscene dyn_identity\n"
                 "  1 element nzfill dyn_concrete 0,0 0,0 0,0: M 0,0 L 5,5 Z dyn_identity dyn_paint 1 "
                 "lineargradient pad 0,0 1,1 ramp 0 rgba(0,0,0,1) 1 rgba(1,1,1,1)\n")
    with pytest.raises(spr.SprError, match="unsupported paint"):
        spr.Scene.load_rvg(f)
    with pytest.raises(spr.SprError):
        spr.Scene.load_rvg(tmp_path / "does_not_exist.rvg")


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_generators_are_deterministic(spr, kind):
    a = spr.Scene.generate(kind, 7, 50, 512).arrays()
    b = spr.Scene.generate(kind, 7, 50, 512).arrays()
    c = spr.Scene.generate(kind, 8, 50, 512).arrays()
    for k in ("pos", "curve_type", "curve_path", "rule", "fill"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["pos"][:10], c["pos"][:10])
    assert len(a["rule"]) == 50 and a["width"] == a["height"] == 512
    assert set(np.unique(a["curve_type"])) <= {2, 4}
    assert np.all(np.diff(a["curve_path"].astype(np.int64)) >= 0)
    if kind == 1:
        assert np.all(a["rule"] == 1)
    if kind == 0:
        assert np.all((a["pos"] >= 0) & (a["pos"] <= 512))


def test_subset_matches_path_selection(spr):
    s = spr.Scene.generate(2, 9, 40, 512)
    sub = s.subset(11, 17).arrays()
    keep = np.zeros(40, dtype=bool)
    keep[11:28] = True
    exp = G.select_paths(s.arrays(), keep)
    for k in ("pos", "pos_path", "curve_pos", "curve_type", "curve_path", "rule", "fill"):
        assert np.array_equal(np.asarray(sub[k]).reshape(np.asarray(exp[k]).shape), exp[k]), k
    with pytest.raises(spr.SprError):
        s.subset(30, 20)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_write_rvg_round_trip(spr, tmp_path, kind):
    s = spr.Scene.generate(kind, 12, 30, 300)
    f = tmp_path / "s.rvg"
    s.write_rvg(f)
    a, b = s.arrays(), spr.Scene.load_rvg(f).arrays()
    for k in ("curve_type", "curve_path", "rule", "fill"):
        assert np.array_equal(a[k], b[k]), k
    # curve control points are identical (%.9g round-trips float32)
    for c in range(len(a["curve_type"])):
        k = a["curve_type"][c]
        assert np.array_equal(a["pos"][a["curve_pos"][c]:a["curve_pos"][c] + k],
                              b["pos"][b["curve_pos"][c]:b["curve_pos"][c] + k]), c


def test_cpp_adapter_header_compiles_and_links(spr, tmp_path):
    """include/spr_rasterizer.hpp (the C++ face with the reference's method set) compiles with
    strict warnings and links against libspr.so; no GPU call is made."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "t.cpp"
    src.write_text('#include "spr_rasterizer.hpp"\n'
                   "int main() { spr::Rasterizer r; spr_stats s{}; (void)s;\n"
                   "  try { r.render(); } catch (const std::runtime_error&) { return 0; }\n"
                   "  return 1; }\n")
    lib_dir = spr.LIB_PATH.parent
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(src),
                    f"-L{lib_dir}", "-lspr", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_png_encoder_round_trip(spr, tmp_path):
    """spr_write_png_rgba: the PNG decodes with Python's zlib (which also checks the chunk CRCs
    and the stream's Adler-32) to the same pixels, for sizes that cross the 65535-byte stored
    deflate blocks."""
    import zlib
    rng = np.random.default_rng(3)
    for w, h in ((1, 1), (7, 3), (300, 200)):
        img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        p = tmp_path / f"img_{w}x{h}.png"
        spr.write_png(p, img)
        data = p.read_bytes()
        assert data[:8] == b"\x89PNG\r\n\x1a\n"
        pos, idat = 8, b""
        while pos < len(data):
            n = int.from_bytes(data[pos:pos + 4], "big")
            typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
            assert int.from_bytes(data[pos + 8 + n:pos + 12 + n], "big") == zlib.crc32(typ + body)
            if typ == b"IHDR":
                assert (int.from_bytes(body[:4], "big"), int.from_bytes(body[4:8], "big")) == (w, h)
                assert body[8:] == bytes([8, 6, 0, 0, 0])
            idat += body if typ == b"IDAT" else b""
            pos += 12 + n
        raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(h, 4 * w + 1)
        assert np.all(raw[:, 0] == 0) and np.array_equal(raw[:, 1:].reshape(h, w, 4), img)
    with pytest.raises(spr.SprError):
        spr.write_png(tmp_path / "no_dir" / "x.png", np.zeros((2, 2, 4), np.uint8))
