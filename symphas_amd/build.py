"""Build the gfx950 HIP library (libsps_fd.so) in-tree with hipcc.

Every HIP source is compiled for gfx950 only, with packed FP32 instructions
disabled; the reference-order kernels additionally with -ffp-contract=off so
that their FP64 arithmetic rounds exactly like the CPU SolverFT.
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
OBJ = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libsps_fd.so"

SOURCES = {
    "api.hip": [],
    "ref_kernels.hip": ["-ffp-contract=off"],
    "fused2d.hip": [],
    "fused3d.hip": [],
    "reduce.hip": [],
    "batch.hip": [],
    "nn.hip": [],
    "tune.hip": [],
    "textio.hip": [],
    "multi_ref.hip": ["-ffp-contract=off"],
    "multi.hip": [],
    "ckpt.hip": [],
    "regional.hip": [],
}

COMMON = [
    "-O3",
    "--offload-arch=gfx950",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library needs ROCm's hipcc")


def _compile(src: str, extra: list[str]) -> Path:
    obj = OBJ / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-6000:]}")
    return obj


def _stale(src: str) -> bool:
    """An object is stale when older than its source or than any header."""
    obj = OBJ / (Path(src).stem + ".o")
    if not obj.exists():
        return True
    deps = [CSRC / src, *CSRC.glob("*.hpp"), *(PKG.parent / "include").glob("*.h")]
    return obj.stat().st_mtime < max(d.stat().st_mtime for d in deps)


def build(verbose: bool = True, stale_only: bool = False) -> Path:
    """Compile every source (stale_only: only those whose object is older
    than the source or a header) and link the library."""
    OBJ.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    jobs = min(8, len(SOURCES))

    def one(kv):
        src, extra = kv
        if stale_only and not _stale(src):
            return OBJ / (Path(src).stem + ".o")
        return _compile(src, extra)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, SOURCES.items()))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)
