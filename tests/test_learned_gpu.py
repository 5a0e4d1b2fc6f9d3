"""GPU tests of the learned single-pass allocation: the predictor's HIP kernels against the CPU
oracle (oracle/mlp_oracle.py), training on the GPU, and frames whose output must be identical
to the plain method's whatever the predictions are."""
from __future__ import annotations

import numpy as np
import pytest

import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _random_model(seed):
    rng = np.random.default_rng(seed)
    flat = rng.normal(scale=0.25, size=M.N_PARAMS).astype(np.float32)
    mean = rng.normal(size=16).astype(np.float32)
    inv_std = rng.uniform(0.5, 2.0, size=16).astype(np.float32)
    return flat, mean, inv_std


def test_mfma_inference_matches_oracle(spr):
    flat, mean, inv_std = _random_model(0)
    r = spr.Rasterizer(64, 64)
    r.mlp_set_params(np.concatenate([flat, mean, inv_std]))
    rng = np.random.default_rng(1)
    for n in (1, 15, 16, 17, 1000):
        X = (rng.normal(size=(n, 16)) / inv_std + mean).astype(np.float32)
        got = r.mlp_predict(X)
        ref = M.forward_bf16(flat, mean, inv_std, X)
        # same bf16 operands; only the fp32 accumulation order differs
        assert np.max(np.abs(got - ref)) < 2e-2 * (1 + np.max(np.abs(ref))), n
        y64, _ = M.forward_f64(flat, mean, inv_std, X)
        assert np.max(np.abs(got - y64)) < 0.05 * (1 + np.max(np.abs(y64))), n
    r.close()


def test_training_gradients_match_oracle(spr):
    flat, mean, inv_std = _random_model(2)
    r = spr.Rasterizer(64, 64)
    r.mlp_set_params(np.concatenate([flat, mean, inv_std]))
    rng = np.random.default_rng(3)
    X = (rng.normal(size=(300, 16)) / inv_std + mean).astype(np.float32)
    T = rng.normal(size=(300, 2)).astype(np.float32)
    g, loss = r.mlp_gradients(X, T)
    gr, lr = M.grads_f64(flat, mean, inv_std, X, T)
    assert abs(loss - lr) < 1e-4 * (1 + lr)
    scale = np.max(np.abs(gr))
    assert np.max(np.abs(g - gr)) < 1e-4 * scale
    for name, _ in M.SIZES:  # every block of the layout is covered
        a, b = M.unpack(g)[name], M.unpack(gr)[name]
        assert np.max(np.abs(a - b)) <= 1e-4 * scale + 1e-6, name
    r.close()


@pytest.fixture(scope="module")
def trained(spr, tmp_path_factory):
    """A predictor trained on the GPU on 3,000 illustration paths (seed 4) at 1024^2."""
    r = spr.Rasterizer(1024, 1024, compose=spr.COMPOSE_SRC_OVER, mode=spr.MODE_LEARNED)
    train = spr.Scene.generate(2, 4, 3000, 1024)
    print("[learned] training on 3000 paths", flush=True)
    rep = r.learned_train(train, epochs=40, chunk_paths=1000)
    path = tmp_path_factory.mktemp("model") / "mlp.bin"
    r.learned_save(path)
    return r, rep, path


def _plain_and_learned(spr, r_learned, scene, w, compose, mvp=None):
    p = spr.Rasterizer(w, w, compose=compose)
    p.load_scene(scene)
    r_learned.load_scene(scene)
    if mvp is not None:
        p.set_mvp(mvp)
        r_learned.set_mvp(mvp)
    ip = p.render()
    il = r_learned.render()
    return p, ip, il


def test_training_reduces_loss(trained):
    r, rep, _ = trained
    assert rep["samples"] == 3000 and rep["steps"] > 50
    assert rep["val_loss"] < 0.25 * rep["first_loss"]
    assert rep["val_loss"] < 0.5


def test_learned_frame_identical_to_plain(spr, trained):
    r, _, _ = trained
    scene = spr.Scene.generate(2, 3, 400, 1024)
    p, ip, il = _plain_and_learned(spr, r, scene, 1024, spr.COMPOSE_SRC_OVER)
    assert np.array_equal(ip, il)
    sp, sl = p.stats(), r.stats()
    assert sp.n_intersections == sl.n_intersections and sp.n_cells == sl.n_cells and sp.n_spans == sl.n_spans
    assert sl.alloc_intersections > 0
    # a second frame runs without any host readback and is identical too
    il2 = r.render()
    assert np.array_equal(ip, il2)
    p.close()
    # full records (debug handles): identical too
    pd = spr.Rasterizer(1024, 1024, compose=spr.COMPOSE_SRC_OVER, debug=True)
    qd = spr.Rasterizer(1024, 1024, compose=spr.COMPOSE_SRC_OVER, mode=spr.MODE_LEARNED, debug=True)
    qd.mlp_set_params(r.mlp_params())
    pd.load_scene(scene)
    qd.load_scene(scene)
    pd.render(readback=False)
    qd.render(readback=False)
    assert np.array_equal(pd.export("records"), qd.export("records"))
    pd.close()
    qd.close()


def test_forced_overflow_and_generous_arenas(spr, trained):
    r, _, path = trained
    scene = spr.Scene.generate(1, 5, 300, 1024)
    p = spr.Rasterizer(1024, 1024)
    p.load_scene(scene)
    ip = p.render()
    for margin, slack in ((-20.0, 0), (4.0, 64), (0.0, 0)):
        q = spr.Rasterizer(1024, 1024, mode=spr.MODE_LEARNED)
        q.learned_load(path)
        q.learned_set_margin(margin, slack)
        q.load_scene(scene)
        print(f"[learned] margin={margin} slack={slack}", flush=True)
        assert np.array_equal(q.render(), ip), margin
        assert np.array_equal(q.render(), ip), margin
        st = q.stats()
        if margin < -10:
            assert st.n_overflow_paths > 0.9 * scene.n_paths * 0.5  # every visible path overflowed
        q.close()
    p.close()


def test_margin_change_after_capture(spr, trained):
    """A new margin after the frame was captured as a graph: the next frames use it (the graph
    and the budgets are rebuilt), and stay identical to the plain frame."""
    r, _, path = trained
    scene = spr.Scene.generate(2, 3, 400, 1024)
    p = spr.Rasterizer(1024, 1024, compose=spr.COMPOSE_SRC_OVER)
    p.load_scene(scene)
    ip = p.render()
    p.close()
    q = spr.Rasterizer(1024, 1024, compose=spr.COMPOSE_SRC_OVER, mode=spr.MODE_LEARNED)
    q.learned_load(path)
    q.load_scene(scene)
    for frame in range(3):  # eager, eager + capture, replay
        assert np.array_equal(q.render(), ip), frame
    n0 = q.stats().n_overflow_paths
    visible = int(np.sum(q.export("features").reshape(-1, 16)[:, 12] > 0))
    assert n0 < visible
    for margin, slack in ((-20.0, 0), (4.0, 64)):
        print(f"[learned] margin change to {margin}", flush=True)
        q.learned_set_margin(margin, slack)
        for frame in range(3):
            assert np.array_equal(q.render(), ip), (margin, frame)
            n = q.stats().n_overflow_paths
            if margin < -10:
                assert n >= 0.95 * visible  # every visible path overflows its (empty) capacity
            else:
                assert n <= n0
    q.close()


@pytest.mark.skip(reason="this case (a learned frame whose budget was exceeded, so that the device "
                         "dropped paths and the host redid the frame with the plain method) faulted "
                         "the GPU twice; that code path (the per-path drop and the overflow pool) is "
                         "removed, a frame over budget is now culled before any entry is written "
                         "(test_over_budget_frame_renders_plain), and this case stays off the GPU")
def test_budget_growth_falls_back_to_plain(spr, trained):
    r, _, path = trained
    scene = spr.Scene.generate(2, 6, 200, 1024)
    q = spr.Rasterizer(1024, 1024, mode=spr.MODE_LEARNED)
    q.learned_load(path)
    q.learned_set_margin(-20.0, 0)  # everything goes to the overflow pool
    q.load_scene(scene)
    q.set_mvp(np.diag([0.25, 0.25, 1, 1]).astype(np.float32))  # small first frame: small budgets
    q.render(readback=False)
    big = np.diag([2.0, 2.0, 1, 1]).astype(np.float32)
    q.set_mvp(big)
    il = q.render()
    assert q.stats().learned_fallbacks >= 1
    p = spr.Rasterizer(1024, 1024)
    p.load_scene(scene)
    p.set_mvp(big)
    assert np.array_equal(p.render(), il)
    # the grown budgets hold the next frame without a fallback
    n = q.stats().learned_fallbacks
    assert np.array_equal(q.render(), il)
    assert q.stats().learned_fallbacks == n
    q.close()
    p.close()


def test_over_budget_frame_renders_plain(spr, trained):
    """The allocation is predicted for the scene at a quarter of its scale; the identity MVP then
    needs about four times the entries.  That frame exceeds the allocation: the device culls it,
    the host renders it with the plain method (identical output) and grows the allocation, and
    the next frames are learned frames again (eager, captured, replayed), identical too."""
    r, _, path = trained
    scene = spr.Scene.generate(2, 6, 200, 1024)
    q = spr.Rasterizer(1024, 1024, mode=spr.MODE_LEARNED)
    q.learned_load(path)
    q.load_scene(scene)
    p = spr.Rasterizer(1024, 1024)
    p.load_scene(scene)
    small = np.diag([0.25, 0.25, 1, 1]).astype(np.float32)
    q.set_mvp(small)
    p.set_mvp(small)
    assert np.array_equal(q.render(), p.render())
    n_small = p.stats().n_intersections
    assert q.stats().learned_fallbacks == 0
    eye = np.eye(4, dtype=np.float32)
    q.set_mvp(eye)
    p.set_mvp(eye)
    ip = p.render()
    assert p.stats().n_intersections > 2 * n_small
    print("[learned] full-scale frame after a quarter-scale prediction", flush=True)
    assert np.array_equal(q.render(), ip)
    n = q.stats().learned_fallbacks
    assert n == 1
    for frame in range(3):  # eager on the grown allocation, eager + capture, replay
        assert np.array_equal(q.render(), ip), frame
        assert q.stats().learned_fallbacks == n
    q.close()
    p.close()


def test_mvp_changes_replay_the_captured_frame(spr, trained):
    """Interactive use: a new MVP every frame.  Still frames replay a graph without the prediction;
    the first new MVP captures the graph with it, which reads the MVP at each replay, so the pans,
    a rotation and zooms out after it are replays of that graph (no new capture), each identical
    to the plain frame with that MVP; a zoom in beyond the allocation falls back to the plain
    method once, and the frames after it are captured and replayed again."""
    r, _, path = trained
    scene = spr.Scene.generate(2, 9, 300, 1024)
    q = spr.Rasterizer(1024, 1024, compose=spr.COMPOSE_SRC_OVER, mode=spr.MODE_LEARNED)
    q.learned_load(path)
    q.learned_set_margin(2.0, 0)  # capacities of 4x the prediction: room for the pans and zooms out
    q.load_scene(scene)
    p = spr.Rasterizer(1024, 1024, compose=spr.COMPOSE_SRC_OVER)
    p.load_scene(scene)

    def view(scale, angle, dx, dy):
        c, s_ = scale * np.cos(angle), scale * np.sin(angle)
        return np.array([[c, -s_, 0, dx], [s_, c, 0, dy], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)

    # the scene shrunk to 15%: every view below fits the predicted allocation (4x margin)
    q.set_mvp(view(0.15, 0.0, 100.0, 100.0))
    p.set_mvp(view(0.15, 0.0, 100.0, 100.0))
    for _ in range(3):  # eager with the prediction, eager + capture (no prediction), replay
        q.render(readback=False)
    st = q.stats()
    assert st.graph_captures == 1
    mvps = [view(0.15, 0.0, 107.5, 96.75), view(0.15, 0.0, 60.0, 300.0), view(0.15, 0.1, 500.0, 400.0),
            view(0.13, 0.3, 200.0, 250.0), view(0.1, -0.2, 700.0, 90.0), view(0.15, 0.0, 100.0, 100.0)]
    for k, m in enumerate(mvps):
        q.set_mvp(m)
        p.set_mvp(m)
        print(f"[mvp] frame {k}", flush=True)
        ip = p.render()
        assert np.array_equal(q.render(), ip), k
        st = q.stats()
        # the first new MVP: eager + capture of the frame with the prediction; then its replays
        assert st.learned_fallbacks == 0 and st.graph_captures == 2, k
        p.render(readback=False)  # eager work of another handle between replays
    n_fb = q.stats().learned_fallbacks
    big = np.eye(4, dtype=np.float32)  # the full-size scene: about 6.7 times the entries
    q.set_mvp(big)
    p.set_mvp(big)
    ip = p.render()
    assert np.array_equal(q.render(), ip)
    assert q.stats().learned_fallbacks == n_fb + 1  # over the allocation: the plain method
    caps = q.stats().graph_captures
    for k in range(3):
        assert np.array_equal(q.render(), ip), k
    assert q.stats().graph_captures == caps + 1 and q.stats().learned_fallbacks == n_fb + 1
    q.close()
    p.close()


def test_save_load_bit_exact(spr, trained):
    r, _, path = trained
    q = spr.Rasterizer(64, 64)
    q.learned_load(path)
    assert np.array_equal(q.mlp_params(), r.mlp_params())
    X = np.random.default_rng(7).normal(size=(100, 16)).astype(np.float32)
    assert np.array_equal(q.mlp_predict(X), r.mlp_predict(X))
    q.close()


def test_trained_network_beats_the_polygon_bound(spr, trained):
    """What the network gains over the method without it: the capacity that needs no
    prediction is a bound from the control polygon (a monotone piece crosses at most
    |dx|/2 + |dy|/2 + 2 grid lines).  The trained network allocates far less at a small
    overflow rate."""
    r, _, _ = trained
    scene = spr.Scene.generate(2, 8, 500, 1024)
    r.load_scene(scene)
    r.render(readback=False)
    exact = r.export("path_counts").reshape(-1, 2)[:, 0].astype(np.float64)
    feat = r.export("features").reshape(-1, 16)
    pred = r.export("pred").reshape(-1, 2)[:, 0]
    vis = feat[:, 12] > 0
    caps = np.where(vis, np.ceil(np.exp2(pred)) - 1, 0)
    # Bezier curves vary no more than their control polygons; a monotone piece crosses at most
    # |dx|/2 + 1 vertical and |dy|/2 + 1 horizontal grid lines (plus its start entry), and the
    # <= 5 pieces cover the curve at most three times (the 4-cut order quirk): the learned
    # frame's own bound, 1.5 (sum |dx| + sum |dy|) x 1.001 + 20 per curve + 2
    half_dx = np.exp2(feat[:, 5].astype(np.float64)) - 1
    half_dy = np.exp2(feat[:, 6].astype(np.float64)) - 1
    n_curves = np.round(np.exp2(feat[:, 0].astype(np.float64)) - 1)
    bound = np.where(vis, np.ceil(3.0 * (half_dx + half_dy) * 1.001) + 20 * n_curves + 2, 0)
    bad = np.nonzero(bound < exact)[0]
    for i in bad[:8]:
        print(f"[learned] path {i}: exact {exact[i]:.0f} bound {bound[i]:.0f} curves {n_curves[i]:.0f} "
              f"half_dx {half_dx[i]:.1f} half_dy {half_dy[i]:.1f} vis {vis[i]} f12 {feat[i, 12]}", flush=True)
    assert bad.size == 0  # the bound never overflows...
    over_bound = bound.sum() / exact.sum() - 1
    st = r.stats()
    over_mlp = st.alloc_intersections / exact.sum() - 1
    overflow_rate = st.n_overflow_paths / max(1, int(vis.sum()))
    print(f"[learned] over-allocation: bound {over_bound:.2f}, mlp {over_mlp:.2f}; "
          f"overflow rate {overflow_rate:.3f}; raw prediction total {caps.sum() / exact.sum():.2f}x", flush=True)
    assert over_mlp < 0.5 * over_bound  # ...but costs far more memory
    assert overflow_rate < 0.2
