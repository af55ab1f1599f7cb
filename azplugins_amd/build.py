"""Build libazp.so (gfx950) from azplugins_amd/csrc with hipcc.

Each translation unit is compiled to an object under ``azplugins_amd/_build``
in parallel and linked into ``azplugins_amd/lib/libazp.so``.  Translation
units whose arithmetic must match the host oracle bit for bit (neighbor
predicate, DPD pair terms) are compiled with ``-ffp-contract=off``.

Usage: python -m azplugins_amd.build
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libazp.so")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # packed FP32 (v_pk_*_f32) stays off on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

NO_CONTRACT = {"k_nlist.hip", "k_tile_walk.hip", "k_dpd.hip", "k_integrate.hip", "k_wall.hip"}


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if src in NO_CONTRACT:
        cmd.insert(1, "-ffp-contract=off")
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(
        line for line in r.stderr.splitlines() if "packed-fp32-ops" not in line
    )
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{errs}")
    return obj, errs


def build(verbose=False):
    """Compile every HIP source and link libazp.so (always recompiles)."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = []
    for obj, errs in results:
        objs.append(obj)
        if verbose and errs.strip():
            print(errs, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
