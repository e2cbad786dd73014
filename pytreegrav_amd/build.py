"""Compile the gfx950 HIP library ``pytreegrav_amd/libptg_hip.so`` with hipcc.

Each translation unit is compiled to an object in parallel, then linked into one shared library.
The build/moment unit is compiled with -ffp-contract=off so Morton keys, node geometry and FP64
moments are IEEE-exact; the walk and direct-sum units allow contraction.  Packed FP32 instructions
are disabled for every unit.

    python -m pytreegrav_amd.build          # build (always recompiles)
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libptg_hip.so")
ARCH = os.environ.get("PTG_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-command-line-argument",
]
UNITS = {
    "build.hip": ["-ffp-contract=off"],
    "walk.hip": [],
    "direct.hip": [],
    "learned.hip": [],
    "colden.hip": [],
}


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return exe


def _compile(unit, extra, verbose):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJDIR, unit.replace(".hip", ".o"))
    cmd = [hipcc(), "-c", src, "-o", obj] + COMMON + extra
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    """Compile every unit for the target architecture and link libptg_hip.so (always rebuilds)."""
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], kv[1], verbose), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
