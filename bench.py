"""Benchmark: one frame of the scanline path-rasterisation pipeline on one MI355X.

A step is one frame: device-resident SoA input (after load) -> final RGBA8 image in HBM, with
every count, allocation and host synchronisation of the chosen allocation mode inside the timed
region (image readback excluded).  Default workload: the illustration scene (5,000 paths of
4-400 cubic segments, log-normal sizes, per-path affine, mixed fill rules, alpha 0.3-1.0,
src-over composite into 4096x4096 RGBA8), synthetic from a seeded generator.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 exits: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import vkscanlinepr_amd as V  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Stage boundaries recorded by the library on its stream (stage_ms[k] = events k -> k+1).  The
# hot kernels (chain_solve, band_sort, winding_scan, mark_emit, composite, ...) sit alone between
# two events, so these are their device times.
STAGE_NAMES = ["transform", "count", "scans+sync", "chain_prep", "chain_solve", "piece_fragments",
               "gen_fragment", "bucket_scatter", "band_sort", "winding_scan", "mark_emit",
               "composite"]

WORKLOADS = {
    # name: (generator kind, seed, n_paths, canvas, compose)
    "illustration_4096": (2, 3, 5000, 4096, V.COMPOSE_SRC_OVER),
    "glyphs_2048": (1, 2, 20000, 2048, V.COMPOSE_OVERWRITE),
    "scribbles_1024": (0, 1, 64, 1024, V.COMPOSE_OVERWRITE),
}


# Frames that export no stage, on canvases up to 32764 wide (every workload here), fuse mark +
# emit into the band sort: the band_sort stage then holds both, the mark_emit stage only brackets
# two events, and the sort is rated with the bytes of every reference stage it replaces.
FUSED_MARK_EMIT = True


def stage_bytes(P, C, Np, N, R, w, h):
    """Algorithmic HBM bytes per stage: every input read once and every output written once, in
    the reference's data types; a kernel that fuses several reference stages counts the sum of
    their bytes.  transform 20 B/point + 4 B/path; count 72 B/curve; intersection write
    72 B/curve + 8 B/intersection and fragment keys 24 B/intersection + 48 B/curve, split into
    the solve (48 B/curve of curve data in, one 4 B t out per intersection) and the piece merge
    fused with the keys (the rest: 72 B/curve + 28 B/intersection); sort 16 B/entry + the
    winding gather 12 B/entry + the winding scan 8 B/entry (bucket scatter and band sort
    together); mark 20 B/entry + flag scan 2 x 8 B/entry + compaction 24 B/entry + 16 B/record
    (one fused kernel); composite 16 B/record + 4 B/pixel.  The scans before the host sync count
    the reference's scan of the per-curve counts (4 B in, 4 B out per curve).  The chain prep and
    order, which have no reference stage, count their per-curve I/O: the curve data the count pass
    reads (72 B) and its offset (4 B) in, five 24 B piece descriptors and ten class bytes out, the
    class bytes read twice by the order (226 B/curve; the 32 B descriptor and 4 B order slot of
    each non-empty chain are left out, so the rate is a lower bound)."""
    return {
        "transform": 20 * P + 4 * Np,
        "count": 72 * C,
        "scans+sync": 8 * (C + 1),
        "chain_prep": 226 * C,
        "chain_solve": 48 * C + 4 * N,
        "piece_fragments": 72 * C + 28 * N,
        "sort": 36 * N,
        "mark_emit": 60 * N + 16 * R,
        "composite": 16 * R + 4 * w * h,
    }


def fold_mark_emit(stages, algo, io, N, R):
    """Mark + emit fused into the sort: one stage, timed as the sort plus what is left between the
    mark_emit events, rated with the algorithmic bytes of both, and with the fused kernel's own
    I/O (key, index and winding change in, 12 B/entry; records out, 16 B/record)."""
    stages["sort"] += stages["mark_emit"]
    algo["sort"] += algo.pop("mark_emit")
    io.pop("mark_emit")
    io["sort"] = 12 * N + 16 * R


def stage_io_bytes(P, C, Np, N, R, w, h):
    """Each kernel's own inputs read once and outputs written once, in the reference's data types
    (no bytes for the intermediate buffers a fused kernel never materialises): piece merge with
    keys 72 B/curve + t in, key, winding change and path out (16 B/intersection); sort with the
    winding gather and scan: key, index, winding change in and key, index, winding number out
    (24 B/entry); mark and emit: key, winding number, path in (12 B/entry) and 16 B/record out;
    the scans before the host sync: per-curve counts and per-path bucket counts, 8 B each."""
    return {
        "transform": 20 * P + 4 * Np,
        "count": 72 * C,
        "scans+sync": 8 * (C + 1) + 8 * (Np + 1),
        "chain_prep": 226 * C,
        "chain_solve": 48 * C + 4 * N,
        "piece_fragments": 72 * C + 16 * N,
        "sort": 24 * N,
        "mark_emit": 12 * N + 16 * R,
        "composite": 16 * R + 4 * w * h,
    }


# The intersection solve is bound by FP32 vector issue, not HBM: each crossing of a cubic costs 24
# bisection steps, each a midpoint (add, mul), a cubic evaluated as six lerps of a sub, a mul and
# an add, and the sign test (sub): 21 operations, no FMA (bit parity with the reference needs
# every rounding).  Line crossings take one reciprocal product and are not counted.
SOLVE_FLOP_PER_CUBIC_CROSSING = 24 * 21
# MI355X published FP32 vector peak (FMA counted as two operations, packed FP32 as two lanes), and
# the rate this kernel can reach without FMA or packed math (bit parity rules out both):
# 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz.
FP32_PEAK_TFLOPS = 157.3
VALU_NOFMA_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def cubic_crossings(counts, cuts, curve_type):
    """Crossings solved by bisection: each cubic's entries minus its pieces' start entries (a
    monotone piece holds its start plus its x and y grid crossings)."""
    n = np.diff(counts.astype(np.int64))
    n_cuts = cuts.reshape(-1, 5)[:, 4].view(np.uint32).astype(np.int64)
    cubic = curve_type == 4
    return int(np.maximum(n[cubic] - (n_cuts[cubic] + 1), 0).sum())


def rate_stages(stages, algo, io, traffic):
    """Per stage: device time, the algorithmic rate (every reference stage a kernel replaces read and
    written once in the reference's types; effective for fused kernels), the
    rate on the kernels' own I/O and the measured HBM traffic rate, each as a fraction of the
    8 TB/s peak.  Times are those of a frame with every kernel on one stream."""
    table = {}
    for k, ms in stages.items():
        row = {"ms": round(ms, 4)}
        for label, nbytes in (("algo", algo.get(k)), ("io", io.get(k)), ("traffic", (traffic or {}).get(k))):
            if nbytes is None or ms <= 0:
                row[f"{label}_GB/s"] = row[f"{label}_frac"] = None
            else:
                gbs = nbytes / (ms * 1e-3) / 1e9
                row[f"{label}_GB/s"] = round(gbs, 1)
                row[f"{label}_frac"] = round(gbs / HBM_PEAK_GBS, 4)
        table[k] = row
    return table


def cpu_baseline(kind, seed, n_paths, canvas, compose, budget_s=20.0):
    """The CPU oracle (C++ restatement of the reference shaders) on a bounded prefix of the
    same scene, all host threads (capped at 16)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    threads = min(16, os.cpu_count() or 1)
    scene = V.Scene.generate(kind, seed, n_paths, canvas)
    k = max(1, min(n_paths, 50))
    best = None
    while True:
        sub = scene.subset(0, k)
        a = sub.arrays()
        t0 = time.perf_counter()
        O.Frame(a, canvas, canvas, compose=compose, threads=threads)
        dt = time.perf_counter() - t0
        best = (k, dt)
        if dt > budget_s / 4 or k >= n_paths:
            break
        k = min(n_paths, int(k * max(2.0, budget_s / 4 / max(dt, 1e-3))))
    k, dt = best
    return {"value": k / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"first {k} of {n_paths} paths of the same scene, {canvas}x{canvas}, "
                      f"one frame in {dt:.2f} s"}


def documents(seed=5, n_docs=256):
    """Batched-documents workload: n_docs documents, path counts log-uniform in [10, 20000],
    square canvases 2^U{8..12} (256^2 .. 4096^2), generators round-robin (scribbles, glyphs,
    illustration).  Scribble documents keep the scribble workload's 64 paths: scribble paths
    span the canvas, and 20,000 of them would be ~1e9 intersections in one document."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(10), np.log(20000), n_docs)).astype(int)
    cv = 2 ** rng.integers(8, 13, n_docs)
    out = []
    for d in range(n_docs):
        kind = d % 3
        paths = min(int(n[d]), 64) if kind == 0 else int(n[d])
        out.append((V.Scene.generate(kind, seed * 1000 + d, paths, int(cv[d])), int(cv[d])))
    return out


def bench_documents(a) -> dict:
    """256 documents rendered in groups, each group one batch frame replayed as one captured
    graph, against the same documents rendered one by one with the plain two-pass method."""
    docs = documents()
    # the sequential plain renders (the baseline, inputs resident) also give each document's
    # exact intersection count
    plain = {}
    seq_ms = 0.0
    counts = []
    for sc, c in docs:
        if c not in plain:
            plain[c] = V.Rasterizer(c, c, device=0)
        p = plain[c]
        p.load_scene(sc)
        p.render(readback=False)
        p.synchronize()
        t1 = time.perf_counter()
        p.render(readback=False)
        p.synchronize()
        seq_ms += (time.perf_counter() - t1) * 1e3
        counts.append(int(p.stats().n_intersections))
    for p in plain.values():
        p.close()
    # groups of consecutive documents whose counts fit one frame's budget (int32 entry offsets,
    # with the budget's headroom): each group is one batch, one graph launch per frame
    limit = 9.0e8
    groups, cur, acc = [], [], 0
    for i, n in enumerate(counts):
        if cur and acc + n > limit:
            groups.append(cur)
            cur, acc = [], 0
        cur.append(i)
        acc += n
    if cur:
        groups.append(cur)
    # identity against plain renders on a spread of documents (every kind, the smallest and the
    # largest canvas, the most intersections)
    check = sorted({0, 1, 2, len(docs) - 1, int(np.argmax(counts)), int(np.argmin(counts)),
                    int(np.argmax([c for _, c in docs])), int(np.argmin([c for _, c in docs]))})
    batched_ms = 0.0
    tune = None
    same = True
    for gi, g in enumerate(groups):
        r = V.Rasterizer(64, 64, device=0)
        r.load_batch([docs[i][0] for i in g], [(docs[i][1], docs[i][1]) for i in g])
        for _ in range(max(a.warmup, 3)):  # count and render, eager + capture, replay
            r.render(readback=False)
        if gi == 0 and a.autotune:
            tune = r.autotune(reps=2)
            for _ in range(3):
                r.render(readback=False)
        t0 = time.perf_counter()
        for _ in range(a.steps):
            r.render(readback=False)  # one graph launch; waits for the group's counters
        batched_ms += (time.perf_counter() - t0) / a.steps * 1e3
        for j, i in enumerate(g):
            if i in check:
                sc, c = docs[i]
                p = V.Rasterizer(c, c, device=0)
                p.load_scene(sc)
                same &= bool(np.array_equal(r.batch_image(j), p.render()))
                p.close()
        r.close()
    n_paths = int(sum(sc.n_paths for sc, _ in docs))
    return {
        "metric": "documents/s, 256 batched documents (one captured graph per group) vs sequential plain renders",
        "value": len(docs) / (batched_ms * 1e-3), "unit": "documents/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": batched_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+int32", "data": "synthetic (seeded generators)",
        "config": {"workload": "documents_256", "documents": len(docs), "paths": n_paths,
                   "intersections": int(sum(counts)), "groups": len(groups),
                   "group_sizes": [len(g) for g in groups],
                   "sequential_plain_ms": seq_ms, "speedup_vs_sequential": seq_ms / batched_ms,
                   "images_identical_to_plain": same, "checked_documents": check, "autotune": tune},
    }


def measured_traffic(workload, mode):
    """HBM bytes per launch of each stage's kernels, from the committed rocprofv3 FETCH_SIZE +
    WRITE_SIZE summary of this workload (already scaled by the copy-kernel calibration), or an
    empty dict when no such measurement is committed."""
    f = ROOT / "profiles" / f"traffic_{workload}_{mode}.json"
    if not f.exists():
        return {}
    return {k: d.get("bytes_per_launch") for k, d in json.loads(f.read_text()).get("stages", {}).items()}


def time_frames(r, steps, warmup):
    """Wall-clock ms per frame (render + the frame's counters back on the host) and the mean
    device time of the frames (HIP events on the handle's stream)."""
    for _ in range(warmup):
        r.render(readback=False)
    r.synchronize()
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render(readback=False)
        dev.append(r.stats().frame_ms)  # waits for the frame
    r.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, float(np.mean(dev))


def stage_table(r, steps, sizes, workload, mode):
    """Per-stage device times of frames that run every kernel on the main stream (HIP events
    between the stages), rated by algorithmic bytes, own I/O and measured traffic."""
    r.set_stage_timing(True)
    acc = np.zeros(16)
    for _ in range(steps):
        r.render(readback=False)
        acc += np.array(r.stats().stage_ms[:16])
    serial_ms = float(r.stats().frame_ms)
    r.set_stage_timing(False)
    stages, algo, io = fold_stages(acc[: len(STAGE_NAMES)] / steps, sizes)
    return rate_stages(stages, algo, io, measured_traffic(workload, mode)), serial_ms, algo, io


def fold_stages(stage_ms, sizes):
    """The per-event stage times (STAGE_NAMES order) folded into the rows of the table, with the
    algorithmic and own-I/O bytes of each row."""
    P, C, Np, N, R, w, h = sizes
    stages = dict(zip(STAGE_NAMES, (float(x) for x in stage_ms)))
    # the reference's sort (with its winding gather and scan) is the bucket scatter, the band sort
    # and the winding scan folded into it
    stages["sort"] = stages.pop("bucket_scatter") + stages.pop("band_sort") + stages.pop("winding_scan")
    # the separate fragment-key kernel runs only on injected intersections: generated ones get
    # their keys inside the piece kernels, so this row would only time two adjacent events
    stages.pop("gen_fragment")
    algo = stage_bytes(P, C, Np, N, R, w, h)
    io = stage_io_bytes(P, C, Np, N, R, w, h)
    if FUSED_MARK_EMIT:
        fold_mark_emit(stages, algo, io, N, R)
        stages.pop("mark_emit")
    return stages, algo, io


def bench_scan(a) -> dict:
    """The standalone exclusive scan (the reference's naive_scan.comp) on n int32 in HBM, timed
    with HIP events on the stream it runs on; 8 algorithmic bytes per element."""
    from vkscanlinepr_amd import hip as H

    n = a.scan_n
    x = np.random.default_rng(6).integers(0, 1 << 10, n, dtype=np.int32)
    xd, yd = H.DeviceBuffer(4 * n), H.DeviceBuffer(4 * (n + 1))
    xd.upload(x)
    st = H.Stream()
    e0, e1 = H.Event(), H.Event()
    r = V.Rasterizer(64, 64, device=0)
    for _ in range(a.warmup):
        r.scan_i32(xd.ptr, yd.ptr, n, st.handle)
    st.synchronize()
    t0 = time.perf_counter()
    e0.record(st)
    for _ in range(a.steps):
        r.scan_i32(xd.ptr, yd.ptr, n, st.handle)
    e1.record(st)
    st.synchronize()
    wall_ms = (time.perf_counter() - t0) / a.steps * 1e3
    ms = e0.elapsed_ms(e1) / a.steps
    # the timed output against numpy (the sums stay below 2^31 here)
    ok = bool(np.array_equal(yd.download(np.int32, n + 1),
                             np.concatenate([[0], np.cumsum(x, dtype=np.int64)]).astype(np.int32)))
    r.close()
    gbs = 8 * n / (ms * 1e-3) / 1e9
    return {
        "metric": "exclusive scan GB/s (int32, 8 B per element) vs 8 TB/s HBM peak",
        "value": gbs, "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic (seeded uniform 0..1023)",
        "config": {"workload": "scan", "n": n, "device_ms_per_scan": ms, "matches_numpy_cumsum": ok},
        "roofline": {"bound": "hbm", "kernel": "k_scan_i32", "achieved": gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": 8 * n, "kernel_ms": ms},
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="illustration_4096",
                    choices=sorted(WORKLOADS) + ["documents_256", "scan"])
    ap.add_argument("--autotune", action="store_true", help="documents_256: autotune the first group")
    ap.add_argument("--mode", default="both", choices=["both", "plain", "learned"],
                    help="both: the plain two-pass frame (value) and the learned single-pass frame")
    ap.add_argument("--train-paths", type=int, default=50000,
                    help="learned: paths of the training scene (same generator, seed 4)")
    ap.add_argument("--scan-n", type=int, default=100_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-aa", action="store_true", help="skip the anti-aliased frames of the same scene")
    ap.add_argument("--stages", action="store_true", help="print per-stage times to stderr")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 1 GiB (the reference kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        import torch
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel(), "copies": 5}))
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: frames render on one GPU; several GPUs would each render a band of image "
              "rows (or whole documents) and gather the bands once, which is not implemented.",
              file=sys.stderr)
        return 2

    if a.workload == "documents_256":
        print(json.dumps(bench_documents(a)))
        return 0
    if a.workload == "scan":
        print(json.dumps(bench_scan(a)))
        return 0
    kind, seed, n_paths, canvas, compose = WORKLOADS[a.workload]
    scene = V.Scene.generate(kind, seed, n_paths, canvas)
    primary = "learned" if a.mode == "learned" else "plain"
    learned = None
    rl = None
    if a.mode in ("both", "learned"):
        # the predictor is trained here, on the GPU, on paths of the same generator (another
        # seed) counted by the plain method; not part of any timed region
        rl = V.Rasterizer(canvas, canvas, compose=compose, mode=V.MODE_LEARNED, device=0)
        train = V.Scene.generate(kind, 4, a.train_paths, canvas)
        learned = {"train": rl.learned_train(train, chunk_paths=5000)}
        del train
        rl.load_scene(scene)
    rp = None
    if a.mode in ("both", "plain"):
        rp = V.Rasterizer(canvas, canvas, compose=compose, mode=V.MODE_PLAIN, device=0)
        rp.load_scene(scene)
        plain_ms, plain_dev = time_frames(rp, a.steps, a.warmup)
    r = rp if rp is not None else rl
    if rl is not None:
        learned_ms, learned_dev = time_frames(rl, a.steps, a.warmup)
    ms_per_step, dev_ms = (plain_ms, plain_dev) if primary == "plain" else (learned_ms, learned_dev)
    st = r.stats()
    P, C, Np = st.n_points, st.n_curves, st.n_paths
    N, R = st.n_intersections, st.n_cells + st.n_spans
    sizes = (P, C, Np, N, R, canvas, canvas)
    table, serial_ms, algo, io = stage_table(r, a.steps, sizes, a.workload, primary)
    if a.stages:
        for k, v in table.items():
            print(f"  {k:16s} {v['ms']:8.3f} ms  algo {v['algo_GB/s'] or 0:8.1f}  io {v['io_GB/s'] or 0:8.1f}  "
                  f"traffic {v['traffic_GB/s'] or 0:8.1f} GB/s", file=sys.stderr)
    traffic = measured_traffic(a.workload, primary)
    # the intersection solve: FP32 vector work of the cubic crossings
    xc = cubic_crossings(r.export("counts"), r.export("cuts"), scene.arrays()["curve_type"])
    solve_ms = table["chain_solve"]["ms"]
    tflops = SOLVE_FLOP_PER_CUBIC_CROSSING * xc / (solve_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_chain_solve (both launches)", "achieved": tflops,
            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS,
            "peak_no_fma_no_packed": VALU_NOFMA_PEAK_TOPS, "frac_no_fma_no_packed": tflops / VALU_NOFMA_PEAK_TOPS,
            "traffic": traffic.get("chain_solve"), "algorithmic_flop": SOLVE_FLOP_PER_CUBIC_CROSSING * xc,
            "cubic_crossings": xc, "kernel_ms": solve_ms}
    # the longest HBM-bound stage, rated on its own I/O; the algorithmic figure (every reference
    # stage a fused kernel replaces, buffers it never touches included) is labelled apart
    hb = max((k for k in table if k in io and k != "chain_solve"), key=lambda k: table[k]["ms"])
    hb_ms = table[hb]["ms"]
    roof_hbm = {"bound": "hbm", "kernel": hb, "achieved": io[hb] / (hb_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": io[hb] / (hb_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "io_bytes": io[hb],
                "traffic": traffic.get(hb), "kernel_ms": hb_ms,
                "algorithmic_equivalent_GB/s": algo[hb] / (hb_ms * 1e-3) / 1e9,
                "algorithmic_equivalent_bytes": algo[hb]}
    out = {
        "metric": "ms/frame & Mpaths/s at 4096^2 (learned 1-pass vs plain 2-pass); HBM GB/s",
        "value": Np / (ms_per_step * 1e-3) / 1e6,
        "unit": "Mpaths/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+int32", "data": "synthetic (seeded generator)",
        "config": {"workload": a.workload, "mode": primary, "paths": Np, "curves": C,
                   "points": P, "intersections": N, "records": R, "canvas": canvas,
                   "compose": "src_over" if compose else "overwrite",
                   "device_ms_per_frame": dev_ms, "serialized_frame_ms": serial_ms,
                   "routes": {"wave_chains": st.n_wave_chains, "heavy_pieces": st.n_heavy_pieces,
                              "long_bands": st.n_long_bands, "big_bands": st.n_big_bands}},
        "roofline": roof,
        "roofline_hbm": roof_hbm,
        "stages": table,
    }
    if rp is not None and compose == V.COMPOSE_SRC_OVER and not a.no_aa:
        # the same frame with anti-aliased coverage (16 samples per pixel in the boundary cells)
        ra = V.Rasterizer(canvas, canvas, compose=compose, mode=V.MODE_PLAIN, device=0, aa=V.AA_SAMPLES)
        ra.load_scene(scene)
        aa_ms, aa_dev = time_frames(ra, a.steps, a.warmup)
        cov = ra.export("coverage")
        out["aa"] = {"samples_per_pixel": V.AA_SAMPLES, "ms_per_step": aa_ms, "device_ms_per_frame": aa_dev,
                     "Mpaths/s": Np / (aa_ms * 1e-3) / 1e6,
                     "partial_cells": int((cov != np.uint64(0xFFFFFFFFFFFFFFFF)).sum()), "records": int(len(cov))}
        ra.close()
        del cov
    if rl is not None:
        sl = rl.stats()
        il = rl.render()
        visible = int(np.sum(rl.export("features").reshape(-1, 16)[:, 12] > 0))
        learned.update({
            "ms_per_step": learned_ms, "device_ms_per_frame": learned_dev,
            "Mpaths/s": Np / (learned_ms * 1e-3) / 1e6,
            "overflow_paths": int(sl.n_overflow_paths), "visible_paths": visible,
            "overflow_rate": sl.n_overflow_paths / max(1, visible),
            "alloc_intersections": int(sl.alloc_intersections),
            "over_allocation": sl.alloc_intersections / max(1, N) - 1.0,
            "fallback_frames": int(sl.learned_fallbacks),
        })
        if rp is not None:
            learned["plain_ms_per_step"] = plain_ms
            learned["speedup_vs_plain"] = plain_ms / learned_ms
            learned["image_identical_to_plain"] = bool(np.array_equal(rp.render(), il))
            out["config"]["plain_ms_per_step"] = plain_ms
            out["config"]["learned_ms_per_step"] = learned_ms
            out["config"]["learned_speedup_vs_plain"] = plain_ms / learned_ms
        # interactive view: a new MVP every frame (a 2-px pan, so the scene's counts stay put and the
        # captured frame is replayed with the matrix copied from pinned memory inside the graph)
        pan = np.eye(4, dtype=np.float32)
        for k in range(a.warmup):
            pan[0, 3] = 2.0 * (k + 1)
            rl.set_mvp(pan)
            rl.render(readback=False)
        rl.synchronize()
        t0 = time.perf_counter()
        for k in range(a.steps):
            pan[0, 3] = 2.0 * ((k + a.warmup) % 16 + 1)
            rl.set_mvp(pan)
            rl.render(readback=False)
        rl.synchronize()
        pan_ms = (time.perf_counter() - t0) * 1e3 / a.steps
        ls = rl.stats()
        learned["interactive_pan"] = {"ms_per_step": pan_ms, "graph_captures": int(ls.graph_captures),
                                      "fallback_frames": int(ls.learned_fallbacks)}
        rl.set_mvp(np.eye(4, dtype=np.float32))
        out["learned"] = learned
        rl.close()
    if rp is not None:
        rp.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(kind, seed, n_paths, canvas, compose)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
