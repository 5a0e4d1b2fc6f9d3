"""Writes the committed golden vectors of tests/golden/.

  oracle_golden.npz   stage outputs of the CPU oracle on small fixed scenes (the RVG fixtures
                      and seeded synthetic scenes), so that a change of the oracle and the HIP
                      path's agreement with it are both visible against fixed data
  test_data3.npz      the reference's recorded output-record list workdir/test_data3.csv as
                      int32 (yx, width, fill, frag_index) rows; test_data.csv is kept verbatim

Run from the repository root:  python tests/golden/make_golden.py [--reference-root DIR]
"""
from __future__ import annotations

import argparse
import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
STAGES = ("counts", "inter", "skey", "sidx", "winding", "records")


def digest(img: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest()


def cases(V):
    """(name, arrays, width, height, mvp, compose) of every golden case."""
    out = []
    t = V.Scene.load_rvg(HERE / "rvg" / "test.rvg").arrays()
    out.append(("test_rvg", t, t["width"], t["height"], None, 0))
    m = np.array([[0.75, 0.05, 0, 40.0], [-0.05, 0.75, 0, 60.0], [0, 0, 1, 0], [0.0001, 0, 0, 1.0]],
                 dtype=np.float32)
    out.append(("test_rvg_mvp", t, t["width"], t["height"], m, 0))
    s = V.Scene.generate(0, 61, 8, 256).arrays()
    out.append(("scribbles_256_srcover", s, 256, 256, None, 1))
    g = V.Scene.generate(1, 62, 40, 256).arrays()
    out.append(("glyphs_256", g, 256, 256, None, 0))
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference-root", default="/root/reference")
    args = ap.parse_args()
    sys.path.insert(0, str(ROOT))
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    import vkscanlinepr_amd as V

    data = {}
    for name, arrays, w, h, mvp, compose in cases(V):
        fr = O.Frame(arrays, w, h, mvp=mvp, compose=compose)
        for st in STAGES:
            if name.startswith("test_rvg") or st in ("counts", "records"):
                data[f"{name}/{st}"] = fr.get(st)
        data[f"{name}/image_sha256"] = np.array(digest(fr.image()))
    np.savez_compressed(HERE / "oracle_golden.npz", **data)

    csv = Path(args.reference_root) / "workdir" / "test_data3.csv"
    if csv.exists():
        rec = np.loadtxt(csv, delimiter=",", dtype=np.int64).astype(np.int32)
        np.savez_compressed(HERE / "test_data3.npz", records=rec)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main()
