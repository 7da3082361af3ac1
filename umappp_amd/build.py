"""Builds libumappp_amd.so (HIP, gfx950) in-tree with hipcc.

Run as ``python -m umappp_amd.build`` or through ``__graft_entry__.build()``.
Every HIP translation unit is compiled with ``-ffp-contract=off`` (the union
weights and the epoch schedule must round like separate IEEE multiplies and
adds) and with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libumappp_amd.so"

SOURCES = ["graph_kernels.hip", "layout_kernels.hip", "sell_kernels.hip", "capi.hip"]
HOST_SOURCES = ["host_api.cpp"]
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in res.stdout.splitlines()
        if "packed-fp32-ops" not in ln and "is not a recognized feature for this target" not in ln
    )
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = OBJ / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = OBJ / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([CXX, "-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", f"-I{ROOT / 'include'}",
                     "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for fut in [ex.submit(_run, c) for c in cmds]:
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
