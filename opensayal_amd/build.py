"""Builds the solver library ``libopensayal_amd.so`` in-tree with hipcc for gfx950.

Every translation unit is compiled with ``-ffp-contract=off`` (no fused
multiply-add, so float arithmetic rounds after each operation as the
reference's source states) and without fast-math (division and square root
stay correctly rounded).  ``-packed-fp32-ops`` is switched off for device code.

Usage: ``python -m opensayal_amd.build [--incremental]`` (the full build is
also called by ``__graft_entry__.build``).
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libopensayal_amd.so"
CLI = PKG / "fluid-simulator"
OBJ = PKG / "_build"
ROOT = PKG.parent
CXX = os.environ.get("CXX", "g++")

SOURCES = [
    "json_value.cpp",
    "config.cpp",
    "fluid_host.cpp",
    "batch_host.cpp",
    "kernels_step.hip",
    "kernels_project.hip",
    "kernels_batch.hip",
    "kernels_view.hip",
    "noalloc.hip",
    "arith_check.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-ffp-contract=off", "-fno-fast-math",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc",
]


def _headers_mtime() -> float:
    hs = list(CSRC.glob("*.hpp")) + list((ROOT / "include").glob("*.h"))
    return max(h.stat().st_mtime for h in hs)


def _up_to_date(src: Path, out: Path) -> bool:
    """The object exists and is newer than its source and every header."""
    if not out.exists():
        return False
    t = out.stat().st_mtime
    return t > src.stat().st_mtime and t > _headers_mtime()


def _compile(src: str, incremental: bool = False) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    path = CSRC / src
    if incremental and _up_to_date(path, out):
        return out
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", str(path), "-o", str(out)]
    else:
        cmd = [HIPCC, *COMMON, "-c", str(path), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return out


def build(verbose: bool = True, incremental: bool = False) -> Path:
    """Compiles every source (with ``incremental``, only those whose object
    is older than the source or a header) and links the shared library;
    returns its path."""
    OBJ.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o in objs]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    build_cli(verbose)
    return LIB


def build_cli(verbose: bool = True) -> Path:
    """Links the headless fluid-simulator program against the library; it
    has no device code, so the host compiler builds it."""
    tmp = CLI.with_suffix(".tmp")
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", "-o", str(tmp),
           str(CSRC / "cli.cpp"), str(CSRC / "json_value.cpp"),
           f"-I{ROOT / 'include'}", f"-L{LIB.parent}", "-lopensayal_amd",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, CLI)
    if verbose:
        print(f"built {CLI}", file=sys.stderr)
    return CLI


def clean() -> None:
    if OBJ.exists():
        shutil.rmtree(OBJ)
    if LIB.exists():
        LIB.unlink()


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv[1:])
