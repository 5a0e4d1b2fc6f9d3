"""Builds libspr.so (gfx950 HIP kernels + C ABI) in-tree, and the CPU oracle used by the tests.

    python -m vkscanlinepr_amd.build            # both
    python -m vkscanlinepr_amd.build --lib      # only libspr.so

hipcc cross-compiles for gfx950 without a GPU.  Every HIP source is built with
-ffp-contract=off (no fused multiply-add: the float stages keep the reference shaders' rounding)
and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libspr.so"
ORACLE_SRC = ROOT / "oracle" / "spr_oracle.cpp"
ORACLE_LIB = ROOT / "oracle" / "build" / "libspr_oracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
ORACLE_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                "-fopenmp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    objs = [BUILD / (s.name + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", str(src), "-o", str(obj)])

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    ORACLE_LIB.parent.mkdir(parents=True, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([cxx, *ORACLE_FLAGS, str(ORACLE_SRC), "-o", str(tmp)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="build only libspr.so")
    ap.add_argument("--oracle", action="store_true", help="build only the CPU oracle")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    if not a.oracle:
        print("built", build_lib(a.j))
    if not a.lib:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
