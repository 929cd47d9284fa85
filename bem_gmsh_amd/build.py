"""Build libbemhip.so in-tree with hipcc for gfx950.

    python -m bem_gmsh_amd.build            # compile csrc/*.hip -> bem_gmsh_amd/libbemhip.so

Every translation unit is compiled with packed-FP32 instructions disabled.
assemble.hip and elastic.hip are compiled with -ffp-contract=off (FP64 parity path: no fused
multiply-add where the reference's FP64 arithmetic has none).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libbemhip.so")
ARCH = os.environ.get("BEMHIP_ARCH", "gfx950")

SOURCES = {
    "api.hip": [],
    "assemble.hip": ["-ffp-contract=off"],
    "lu.hip": ["-mllvm", "-pragma-unroll-threshold=200000"],
    "solve.hip": [],
    "adjoint.hip": [],
    "gmres.hip": [],
    "fmm.hip": [],
    "elastic.hip": ["-ffp-contract=off"],
}

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-Wno-unused-value", "-munsafe-fp-atomics"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)

    def compile_one(item):
        src, extra = item
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [cc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(7, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES.items()))
    tmp = OUT + ".tmp"
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
