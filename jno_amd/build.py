"""Build recipe of the jno_fem HIP library (gfx950): hipcc, in-tree, one shared object.

    python -m jno_amd.build          # -> jno_amd/libjno_fem.so
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB = PKG / "libjno_fem.so"
OBJDIR = PKG / "_build"

SOURCES = ["common.hip", "mesh.hip", "pattern.hip", "plan.hip", "assemble_p1.hip", "sparse.hip", "pcg.hip", "pcg_static.hip", "p2.hip", "ebe.hip", "ebe_box.hip", "batch.hip", "transient.hip"]

HIPFLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-Wall",
    "-Wno-unused-result",
    # products and sums are rounded as written: bitwise-symmetric element matrices, exact zeros
    # of axis-aligned geometry, and numpy-identical lattice coordinates depend on it
    "-ffp-contract=off",
    # packed-FP32 instructions are kept out of all device code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the jno_fem HIP library cannot be built")


def _compile(src: str) -> Path:
    obj = OBJDIR / (Path(src).stem + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
