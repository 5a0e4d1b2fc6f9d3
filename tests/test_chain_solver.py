"""The lane chain solver that k_chain_solve runs (ChainSolver, device_common.hpp) against the
reference's bisection per crossing (bisect_axis, make_intersection_1.comp:377-437), bit for bit,
on the host: tests/native/chain_solver_check.cpp built with hipcc for the CPU.

The solver's short form (one compare per bisection step, the negated axis for chains stepping
down, no start evaluation after a crossing without fallback) must agree with the reference on
every crossing, including values that land exactly on a grid line, where the reference's sign
test counts fl(c - c) = +0 as above the line."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc is not installed")
    exe = tmp_path_factory.mktemp("chain_solver") / "chain_solver_check"
    cmd = [HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           f"-I{ROOT / 'vkscanlinepr_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "chain_solver_check.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def test_chain_solver_matches_reference_bisection(checker):
    r = subprocess.run([str(checker), "60000"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:]
    words = r.stdout.split()
    stats = dict(zip(words[::2], map(int, words[1::2])))
    assert stats["mismatches"] == 0
    assert stats["crossings"] > 1_000_000 and stats["fast"] > 0.9 * stats["chains"]
    assert stats["starts_on_line"] > 100
