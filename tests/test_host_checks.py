"""Host routines that replace standard-library calls of the reference, checked against those calls:
the head-of-shuffle used for GOSS draws and bin sub-samples (csrc/host/shuffle_head.h) against
std::mt19937 and std::shuffle (reference booster.cpp:572, binning.cpp:95), and the value sort of
the bin edges (csrc/host/value_sort.h) against std::sort (reference binning.cpp:15).  Each C++
check program is compiled with the host compiler and run; it prints each failure."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _run_check(tmp_path, name):
    exe = tmp_path / name
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "tidegbm_amd" / "csrc" / "host"),
                    "-o", str(exe), str(ROOT / "tests" / "cpp" / f"{name}.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-4000:]
    assert r.stdout.strip().splitlines()[-1].startswith("ok"), r.stdout[-4000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs the host C++ compiler")
def test_shuffle_head_matches_std_shuffle(tmp_path):
    _run_check(tmp_path, "shuffle_head_check")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs the host C++ compiler")
def test_value_sort_matches_std_sort(tmp_path):
    _run_check(tmp_path, "value_sort_check")
