"""Build the native libraries in-tree.

* tidegbm_amd/lib/libtgbm_hip.so   — gfx950 HIP kernels + thin device C ABI (hipcc)
* tidegbm_amd/lib/libtidegbm_amd.so — C++ host booster + drop-in tide_* C API (g++), linked to
  the HIP library through the thin ABI
* oracle/liboracle.so              — the CPU restatement used by the tests (checker only)
* oracle/_ref/libtide_ref.so       — the reference core built from its own sources, when the
  reference tree is present (checker only)

Run: python tidegbm_amd/build.py (a plain script: it must not import the package, whose
__init__ loads the library being built)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "tidegbm_amd"
LIB = PKG / "lib"
DEV_SRC = ["kernels_fit.hip", "session.hip", "forest.hip", "serve.hip", "goss.hip"]
HOST_SRC = ["booster.cpp", "serialize.cpp", "c_api.cpp", "batch.cpp", "server.cpp", "tuner.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: reference formulas must round like its x86-64 build (no FMA).  The
# packed-fp32 feature is disabled for every HIP object (hipcc's host pass warns that it does
# not know the feature; harmless).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # the first kernel arguments arrive in SGPRs at wave launch (no dependent load of the
             # task-descriptor pointer before each kernel's first read)
             "-mllvm", "-amdgpu-kernarg-preload-count=16"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-Wall", "-Wno-unused-function"]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(map(str, cmd))}")
    return "\n".join(lines)


def build_device():
    LIB.mkdir(parents=True, exist_ok=True)
    srcs = [str(PKG / "csrc" / "device" / s) for s in DEV_SRC if (PKG / "csrc" / "device" / s).exists()]
    out = LIB / "libtgbm_hip.so"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", str(out) + ".tmp", *srcs])
    os.replace(str(out) + ".tmp", out)
    return out


def build_host():
    srcs = [str(PKG / "csrc" / "host" / s) for s in HOST_SRC if (PKG / "csrc" / "host" / s).exists()]
    out = LIB / "libtidegbm_amd.so"
    _run(["g++", *HOST_FLAGS, "-shared", "-o", str(out) + ".tmp", *srcs,
          f"-L{LIB}", "-ltgbm_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(str(out) + ".tmp", out)
    return out


def build_oracle():
    _run(["make", "-s", "-f", str(ROOT / "oracle" / "Makefile")])
    if Path("/root/reference/cpp/src/booster.cpp").exists():
        _run(["make", "-s", "-f", str(ROOT / "oracle" / "Makefile.ref")])


def build(oracle: bool = True):
    build_device()
    build_host()
    if oracle and shutil.which("make"):
        build_oracle()


if __name__ == "__main__":
    build()
    print("built:", ", ".join(p.name for p in sorted(LIB.glob("*.so"))))
