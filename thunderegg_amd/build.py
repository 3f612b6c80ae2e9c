"""Builds libthunderegg_amd.so in-tree with hipcc for gfx950.

Usage: python -m thunderegg_amd.build   (or thunderegg_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libthunderegg_amd.so")

SOURCES = [
    "cabi.cpp",
    "gmg.cpp",
    "batch.cpp",
    "tune.cpp",
    "forest.cpp",
    "timer.cpp",
    "kernels/ghost.hip",
    "kernels/stencil.hip",
    "kernels/transfer.hip",
    "kernels/blas.hip",
    "kernels/dft.hip",
    "kernels/dft_mfma.hip",
    "kernels/fused2d.hip",
    "kernels/tail2d.hip",
    "kernels/patch_bicgstab.hip",
    "kernels/krylov.hip",
    "kernels/cycle3d.hip",
    "kernels/bc.hip",
    "kernels/noalloc.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed-FP32 (v_pk_*_f32) instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(HERE, "..", "include"),
    "-I" + CSRC,
]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = _obj(src)
    cmd = [HIPCC] + FLAGS + (["-x", "hip"] if src.endswith(".hip") else []) + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn:
        print(f"[{src}]\n{warn}", file=sys.stderr)
    return obj


def _obj(src: str) -> str:
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def _up_to_date(src: str) -> bool:
    obj = _obj(src)
    if not os.path.exists(obj):
        return False
    deps = [os.path.join(CSRC, src)]
    for root in (CSRC, os.path.join(HERE, "..", "include")):
        for dp, _, files in os.walk(root):
            deps += [os.path.join(dp, f) for f in files if f.endswith(".h")]
    return os.path.getmtime(obj) > max(os.path.getmtime(d) for d in deps)


def build(jobs: int | None = None, incremental: bool = False) -> str:
    """Compile every source (incremental=True: only those older than their object
    or any header) and link the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    todo = [s for s in srcs if not (incremental and _up_to_date(s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [_obj(s) for s in srcs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(incremental="--incremental" in sys.argv))
