"""CPU checks of the learned allocation's host-visible rules.

The learned frame sizes its buffers from the predicted arenas plus an upper bound of every
path's intersections (vkscanlinepr_amd/csrc/k_learned.hip, k_path_features): 1.5 x the control
polygon's |dx| + |dy| sum, x 1.001, + 20 per curve + 2, for every path that is not culled.  The
oracle's exact counts must never exceed it, including the reference's 4-cut order quirk, which
runs one piece backwards over part of the curve."""
from __future__ import annotations

import numpy as np
import pytest

import geometry as G


def _bound(arrays):
    pos = np.asarray(arrays["pos"], np.float64).reshape(-1, 2)
    Np = len(arrays["rule"])
    var = np.zeros(Np)
    ncur = np.zeros(Np)
    for c in range(len(arrays["curve_type"])):
        k = int(arrays["curve_type"][c])
        q = pos[int(arrays["curve_pos"][c]):int(arrays["curve_pos"][c]) + k]
        p = int(arrays["curve_path"][c])
        var[p] += np.abs(np.diff(q, axis=0)).sum()
        ncur[p] += 1
    return 1.5 * var * 1.001 + 20 * ncur + 2


@pytest.mark.parametrize("kind,seed,n,canvas", [(0, 1, 64, 1024), (2, 3, 150, 1024), (1, 2, 300, 512),
                                                (2, 9, 80, 256)])
def test_polygon_bound_holds(orc, spr, kind, seed, n, canvas):
    a = spr.Scene.generate(kind, seed, n, canvas).arrays()
    fr = orc.Frame(a, canvas, canvas)
    counts = np.diff(fr.get("counts"))
    per_path = np.bincount(a["curve_path"], weights=counts, minlength=n)
    vis = np.array([orc.path_visible(int(v)) for v in fr.get("visible")])
    b = np.where(vis, _bound(a), 0)
    assert np.all(per_path <= b), np.nonzero(per_path > b)[0][:5]
    assert np.all(per_path[~vis] == 0)


def test_polygon_bound_holds_for_quirk_curves(orc):
    """Cubics with four interior extrema (the unsorted 4th cut) at many scales."""
    rng = np.random.default_rng(11)
    ctrl = []
    while len(ctrl) < 200:
        q = rng.uniform(0, 200, size=(4, 2))
        if G.derivative_root_count(q) == 4:
            ctrl.append(q)
    pos = np.concatenate(ctrl).astype(np.float32)
    m = len(ctrl)
    a = dict(pos=pos, pos_path=np.repeat(np.arange(m), 4).astype(np.uint32),
             curve_pos=(np.arange(m) * 4).astype(np.uint32), curve_type=np.full(m, 4, np.uint32),
             curve_path=np.arange(m, dtype=np.uint32), rule=np.zeros(m, np.uint32), fill=np.zeros(m, np.uint32))
    fr = orc.Frame(a, 256, 256)
    counts = np.diff(fr.get("counts"))
    assert np.all(counts <= _bound(a))
