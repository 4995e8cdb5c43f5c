"""Build recipe for the gfx950 HIP library ``libaknn.so`` (in-tree, next to this file).

``__graft_entry__.build()`` runs it; so does ``python _build.py`` from this folder.
Every HIP unit is compiled by hipcc for gfx950 only, in parallel, then linked into one
shared object exporting the C ABI declared in ``include/aknn.h``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                     # the package folder holding csrc/
CSRC = os.path.join(ROOT, "csrc")
BUILD_DIR = os.path.join(ROOT, "_build")
LIB_PATH = os.path.join(PKG_DIR, "libaknn.so")
SOURCES = [
    "aknn_abi.hip",
    "aknn_knn.hip",
    "aknn_knn_f32_aniso.hip",
    "aknn_knn_f32_euclid.hip",
    "aknn_knn_f64_aniso.hip",
    "aknn_knn_f64_euclid.hip",
    "aknn_agg.hip",
    "aknn_points.hip",
    "aknn_tiles.hip",
    "aknn_sort.hip",
]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libaknn.so)")


def compile_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-function",
        # packed FP32 math stays off (hipcc would otherwise emit v_pk_* from float code)
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fno-gpu-rdc",
        "-fvisibility=hidden",
    ]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP unit and link ``libaknn.so``; returns the library path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([cc] + compile_flags() + ["-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, fut in zip(cmds, [ex.submit(_run, c) for c in cmds]):
            if verbose:
                print(" ".join(cmd), flush=True)
            fut.result()
    tmp = LIB_PATH + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
