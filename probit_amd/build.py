"""Build libprobit_amd.so (gfx950) in-tree with hipcc.

Usage: ``python -m probit_amd.build`` (also called by ``__graft_entry__.build``).
The shared library lands next to this file so it travels with the source tree.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libprobit_amd.so"
BUILD_DIR = HERE.parent / "_build"

SOURCES = ["gram.hip", "lik.hip", "chol.hip", "solve.hip", "vec.hip", "grad.hip", "sample.hip",
           "api.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return exe


def _compile(src: Path, obj: Path) -> None:
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")


def build(verbose: bool = True) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    objs = []
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)), 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = []
        for s in SOURCES:
            src = CSRC / s
            obj = BUILD_DIR / (src.stem + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, src, obj))
        for f in futs:
            f.result()
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
