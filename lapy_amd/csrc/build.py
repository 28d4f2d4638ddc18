"""Build liblapyhip.so for gfx950 with hipcc (no cmake).

Every .hip source is compiled to an object with its own flags, then linked
into lapy_amd/_lib/liblapyhip.so.  The assembly and curvature sources are
compiled with -ffp-contract=off: their element arithmetic must not fuse
a*b+c so that it reproduces numpy's rounding (the curvature's dihedral angles
are arccos of dot products near 1, where a one-ulp difference in the dot
product is amplified).
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(OUT_DIR, "liblapyhip.so")
OBJ_DIR = os.path.join(ROOT, "_build", "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-Wno-unused-result",
]
SOURCES = {
    "lh_core.hip": [],
    "lh_assembly.hip": ["-ffp-contract=off"],
    "lh_sparse.hip": [],
    "lh_dense.hip": [],
    "lh_mfma.hip": [],
    "lh_rr.hip": [],
    "lh_poly.hip": [],
    "lh_reorder.hip": [],
    "lh_eigs.hip": [],
    "lh_pcg.hip": [],
    "lh_mpnn.hip": [],
    "lh_rrlib.hip": [],
    "lh_curv.hip": ["-ffp-contract=off"],
}
# rocSOLVER (and its rocBLAS) only for the Rayleigh-Ritz of wide trial bases
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def _compile(name: str, extra: list[str]) -> str:
    src = os.path.join(HERE, name)
    obj = os.path.join(OBJ_DIR, name.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, *extra, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, *LIBS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
