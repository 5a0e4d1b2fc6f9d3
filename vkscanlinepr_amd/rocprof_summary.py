"""Turns rocprofv3 CSV output into the short summaries kept under profiles/.

    python -m vkscanlinepr_amd.rocprof_summary kernels STATS_CSV TITLE COMMAND > profiles/X.md
    python -m vkscanlinepr_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR COPY_FETCH_DIR \
        COPY_WRITE_DIR WORKLOAD MODE COMMAND > profiles/traffic_WORKLOAD_MODE.json

`traffic` averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel, scales each by the
reading of a plain 1 GiB device-to-device copy (bytes moved / counter reading), and sums the
kernels of each pipeline stage into HBM bytes per launch of that stage.
"""
from __future__ import annotations

import collections
import csv
import json
import re
import sys
from pathlib import Path

# pipeline stage -> kernels (names without namespace) whose traffic it sums
STAGE_KERNELS = {
    "transform": ["k_transform"],
    "count": ["k_count"],
    "chain_prep": ["k_chain_prep", "k_chain_order"],
    "chain_solve": ["k_chain_solve"],
    "piece_fragments": ["k_piece_small", "k_piece_fragments", "k_piece_waves", "k_piece_heavy"],
    "gen_fragment": ["k_gen_fragment"],
    "bucket_scatter": ["k_bucket_scatter", "k_band_flags", "k_band_emit"],
    "band_sort": ["k_band_classify2", "k_sort_bands_wave", "k_sort_big_bands", "k_emit_big_bands",
                  "k_band_totals"],
    # with mark + emit fused into the band sort (frames that export no stage), the whole sort
    "sort": ["k_bucket_scatter", "k_band_flags", "k_band_emit", "k_band_classify2", "k_sort_bands_wave",
             "k_sort_big_bands", "k_emit_big_bands", "k_band_totals", "k_mark_emit"],
    "mark_emit": ["k_mark_emit"],
    "cell_cover": ["k_cell_cover"],
    "composite": ["k_composite"],
    "scan": ["k_scan_i32"],
}


def kernel_name(full: str, templated: bool = False) -> str:
    s = re.sub(r"\(anonymous namespace\)::", "", full)
    s = re.sub(r"^void ", "", s)
    s = s.split("(")[0]
    args = ""
    if "<" in s:
        args = "<" + s.split("<", 1)[1]
        s = s.split("<")[0]
    s = s.split("::")[-1]
    return s + re.sub(r"spr::", "", args) if templated else s


def _per_kernel(dirpath: Path, counter: str):
    rows = list(csv.DictReader(open(next(Path(dirpath).glob("*counter_collection.csv")))))
    agg = collections.defaultdict(list)
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        agg[(kernel_name(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return agg


def traffic(fetch_dir, write_dir, cfetch_dir, cwrite_dir, workload, mode, command):
    one_gib_kib = float(1 << 20)
    cf = _per_kernel(cfetch_dir, "FETCH_SIZE")
    cw = _per_kernel(cwrite_dir, "WRITE_SIZE")
    copy_f = [v for (k, _), vs in cf.items() if "copyBuffer" in k for v in vs]
    copy_w = [v for (k, _), vs in cw.items() if "copyBuffer" in k for v in vs]
    f_scale = one_gib_kib / (sum(copy_f) / len(copy_f))
    w_scale = one_gib_kib / (sum(copy_w) / len(copy_w))
    fe = _per_kernel(fetch_dir, "FETCH_SIZE")
    wr = _per_kernel(write_dir, "WRITE_SIZE")
    kern = collections.defaultdict(lambda: [0.0, 0.0])
    for (k, g), vs in fe.items():
        kern[k][0] += sum(vs) / len(vs) * f_scale * 1024
    for (k, g), vs in wr.items():
        kern[k][1] += sum(vs) / len(vs) * w_scale * 1024
    stages = {}
    for st, ks in STAGE_KERNELS.items():
        have = [k for k in ks if k in kern]
        if not have:
            continue
        fb = sum(kern[k][0] for k in have)
        wb = sum(kern[k][1] for k in have)
        stages[st] = {"kernels": have, "fetch_bytes": round(fb), "write_bytes": round(wb),
                      "bytes_per_launch": round(fb + wb)}
    return {"workload": workload, "mode": mode, "command": command,
            "calibration": {"copy_bytes": 1 << 30, "fetch_scale": f_scale, "write_scale": w_scale,
                            "note": "bytes moved by a 1 GiB device-to-device copy / counter reading"},
            "stages": stages}


def kernels(stats_csv, title, command):
    rows = list(csv.DictReader(open(stats_csv)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = [f"# {title}", "", "Command (run from the repository root on one MI355X, after "
           "`python -m vkscanlinepr_amd.build`):", "", f"    {command}", "",
           "rocprofv3 kernel-trace summary (per-kernel device time):", "",
           "| kernel | calls | avg us | min us | max us | % |", "|---|---:|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        out.append(f"| {kernel_name(r['Name'], True)} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{100 * float(r['TotalDurationNs']) / total:.1f} |")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    if sys.argv[1] == "traffic":
        print(json.dumps(traffic(*sys.argv[2:9]), indent=1))
    elif sys.argv[1] == "kernels":
        sys.stdout.write(kernels(*sys.argv[2:5]))
    else:
        raise SystemExit(__doc__)
