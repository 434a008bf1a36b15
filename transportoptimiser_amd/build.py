"""Builds libroadmc.so (HIP, gfx950) in-tree and the CPU oracle library.

``python -m transportoptimiser_amd.build`` compiles both.  The shared
libraries stay inside the source tree (``transportoptimiser_amd/lib`` and
``oracle/_build``) so they travel with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libroadmc.so"
ORACLE_DIR = REPO / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for device code: hipcc -O3 would
# otherwise emit v_pk_{add,mul,fma}_f32 on its own.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{REPO / 'include'}", f"-I{CSRC}",
          "-Wall", "-Wno-unused-result"]


# Per-source flags.  The road builder and the grid surrogate follow the CPU
# statement's FP64 rounding (separate multiplies and adds), so they are
# compiled without multiply-add contraction; hipcc's default for device code
# contracts across statements and ignores `#pragma clang fp contract`.
EXTRA_FLAGS = {
    "road_kernels.hip": ["-ffp-contract=off"],
    "roadmc_road.cpp": ["-ffp-contract=off"],
    "knn_kernels.hip": ["-ffp-contract=off"],
    "roadmc_knn.cpp": ["-ffp-contract=off"],
    "rovs_kernels.hip": ["-ffp-contract=off"],
}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _sources() -> list[pathlib.Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def build_lib(jobs: int = 8) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in _sources():
        obj = OBJ / (src.name + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        cmds.append([HIPCC, f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS,
                     *EXTRA_FLAGS.get(src.name, []), *lang, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs]])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> pathlib.Path:
    (ORACLE_DIR / "_build").mkdir(exist_ok=True)
    srcs = sorted(ORACLE_DIR.glob("*.c"))
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c99",
          "-Wall", f"-I{ORACLE_DIR}", "-o", str(tmp), *[str(s) for s in srcs], "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_all(jobs: int = 8) -> None:
    build_oracle()
    build_lib(jobs)


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print(f"built {LIB} and {ORACLE_LIB}")
