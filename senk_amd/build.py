"""Build recipe for libsenk_amd.so (HIP kernels for gfx950 + host C++ + C ABI).

Compiles every source under senk_amd/csrc with hipcc into senk_amd/_build/*.o (in
parallel) and links senk_amd/libsenk_amd.so in-tree, so the library travels with the
source tree; then builds the command-line workflow senk_amd/senk_main (host C++ over the
library, senk_amd/cli/senk_main.cpp).  Run as ``python -m senk_amd.build`` or through
__graft_entry__.build().
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
BUILD = PKG / "_build"
LIB = PKG / "libsenk_amd.so"
CLI_SRC = PKG / "cli" / "senk_main.cpp"
CLI = PKG / "senk_main"
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("SENK_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: kernels keep the reference's separate multiply and add, so the
# order-preserving kernels reproduce the reference's bits.
# -packed-fp32-ops off: no v_pk_*_f32 generation.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    errs = [ln for ln in proc.stderr.splitlines() if "warning" in ln and "packed-fp32" not in ln]
    for ln in errs:
        print(ln, file=sys.stderr)
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")), 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *DEVICE, "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


def build_cli() -> Path:
    """The command-line workflow, linked against the in-tree library (rpath $ORIGIN)."""
    tmp = CLI.with_suffix(".tmp")
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(CLI_SRC), f"-L{PKG}", "-lsenk_amd",
           "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build of {CLI.name} failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True)
