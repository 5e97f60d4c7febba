"""Builds libdemsph.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m this_studio_amd.build  (or build() from Python).
Every .hip file under csrc/ is compiled to an object in _build/ and linked
into this_studio_amd/libdemsph.so; the C++ example programs under
examples/ (host code over the DEMSolver / WCSPHSolver facade) are built
into examples/bin/ against that library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libdemsph.so"
EXAMPLES = PKG / "examples"
EXAMPLES_BIN = EXAMPLES / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DEMSPH_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for all device code: their forms have
# produced silently wrong results on this hardware pool.
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-pthread", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err:
        print(err, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    shutil.move(str(tmp), str(LIB))
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


def build_examples(verbose: bool = True) -> list[Path]:
    """Compiles examples/*.cpp (host-only C++) against libdemsph.so."""
    EXAMPLES_BIN.mkdir(parents=True, exist_ok=True)
    srcs = sorted(EXAMPLES.glob("*.cpp"))

    def one(src: Path) -> Path:
        exe = EXAMPLES_BIN / src.stem
        cmd = [HIPCC, "-O2", "-std=c++17", "-I", str(ROOT / "include"), "-I", str(EXAMPLES), str(src), "-o", str(exe),
               "-L", str(PKG), "-ldemsph", "-Wl,-rpath,$ORIGIN/../.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed for {src.name}:\n{r.stdout}\n{r.stderr}")
        return exe

    with cf.ThreadPoolExecutor(max_workers=min(8, max(1, len(srcs)))) as ex:
        exes = list(ex.map(one, srcs))
    if verbose:
        print(f"built {len(exes)} example programs in {EXAMPLES_BIN}")
    return exes


if __name__ == "__main__":
    build()
    build_examples()
