"""Build libmcgpu.so (HIP, gfx950) in-tree with hipcc.

Usage: ``python -m mcgpu.build`` from ``gpu-monte-carlo_amd/`` or
``mcgpu.build.build()``.  Each ``csrc/*.hip`` is compiled to an object in
parallel, then linked into ``mcgpu/libmcgpu.so``.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent                      # gpu-monte-carlo_amd/
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
BUILD_DIR = ROOT / "_build"
LIB_PATH = PKG_DIR / "libmcgpu.so"

ARCH = os.environ.get("MCGPU_ARCH", "gfx950")

# -ffp-contract=off keeps every float expression in source order (bitwise
# agreement with the CPU oracle); packed FP32 ops are disabled on this target.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-munsafe-fp-atomics",
]


# Per-file additions.  The persistent batched kernel runs short loops whose
# induction variables LLVM's loop strength reduction turns into a dozen
# pointer increments each (VGPRs 136 -> 82 without it, fewer instructions).
FILE_FLAGS = {
    "mc_batch.hip": ["-mllvm", "-disable-lsr"],
    "mc_pack.hip": ["-mllvm", "-disable-lsr"],
}


def _hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; ROCm is required to build libmcgpu.so")
    return path


def sources():
    return sorted(CSRC.glob("*.hip"))


def _compile(src, verbose):
    obj = BUILD_DIR / (src.stem + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, *FILE_FLAGS.get(src.name, []), "-I", str(INCLUDE), "-I", str(CSRC), "-c", str(src),
           "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose=False, jobs=None):
    """Compile every HIP source and link libmcgpu.so; returns its path."""
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = sources()
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
