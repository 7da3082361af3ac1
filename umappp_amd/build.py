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
# floor probes of the layout epoch (used by bench.py only, not part of the library)
PROBES_SRC = "probes.hip"
PROBES_LIB = PKG / "libumx_probes.so"
TORCH_DIR = PKG / "_torch_build"
TORCH_NAME = "umappp_amd_torch"

SOURCES = ["graph_kernels.hip", "layout_kernels.hip", "sell_kernels.hip", "loss_kernels.hip", "spectral.hip", "knn.hip", "capi.hip"]
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
    probes_tmp = PROBES_LIB.with_suffix(".so.tmp")
    cmds.append([HIPCC, *FLAGS, "-shared", str(CSRC / PROBES_SRC), "-o", str(probes_tmp)])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for fut in [ex.submit(_run, c) for c in cmds]:
            fut.result()
    os.replace(probes_tmp, PROBES_LIB)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libumappp_amd.so", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_torch(verbose)
    return LIB


def build_torch(verbose: bool = False) -> Path:
    """Builds the libtorch extension (umappp_amd/_torch_build/umappp_amd_torch.so)
    holding the autograd loss.  It calls libumappp_amd's C ABI, whose symbols
    it takes from the library that umappp_amd loads first (RTLD_GLOBAL)."""
    import ctypes

    import torch  # noqa: F401  (the extension links against this torch)
    from torch.utils import cpp_extension

    ctypes.CDLL(str(LIB), mode=ctypes.RTLD_GLOBAL)  # the extension's umx_* symbols
    TORCH_DIR.mkdir(exist_ok=True)
    for old in (TORCH_DIR / "torch_loss.o", TORCH_DIR / f"{TORCH_NAME}.so"):  # always compile afresh
        if old.exists():
            old.unlink()
    cpp_extension.load(name=TORCH_NAME, sources=[str(CSRC / "torch_loss.cpp")],
                       extra_include_paths=[str(ROOT / "include")], extra_cflags=["-O2"],
                       build_directory=str(TORCH_DIR), with_cuda=False, verbose=verbose, is_python_module=True)
    out = TORCH_DIR / f"{TORCH_NAME}.so"
    if not out.exists():
        raise RuntimeError(f"torch extension was not produced at {out}")
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
