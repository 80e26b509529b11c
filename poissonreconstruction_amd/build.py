"""Build libprx.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Usage: python -m poissonreconstruction_amd.build   (or build_lib() from Python)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libprx.so")
ARCH = os.environ.get("PRX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["k_grid.hip", "k_points.hip", "k_stencil.hip", "solver.hip", "reconstruct.hip", "measure.hip", "k_adjoint.hip", "batch.hip", "k_mc.hip", "k_mg.hip", "noalloc.hip", "persist.hip"]

# -packed-fp32-ops: hipcc -O3 may form packed FP32 instructions from scalar
# float code; they are kept out of every kernel.
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: libprx needs ROCm's hipcc")


def build_lib(verbose: bool = False) -> str:
    """Compile every source to an object in a temp dir, then link libprx.so."""
    cc = hipcc()
    tmp = tempfile.mkdtemp(prefix="prx_build_")
    try:
        procs = []
        objs = []
        for src in SOURCES:
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            objs.append(obj)
            cmd = [cc, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        failed = []
        for src, p in procs:
            out, _ = p.communicate()
            text = out.decode(errors="replace")
            if p.returncode != 0:
                failed.append((src, text))
            elif verbose and text.strip():
                print(text)
        if failed:
            msg = "\n".join(f"--- {s}\n{t}" for s, t in failed)
            raise RuntimeError(f"hipcc failed:\n{msg}")
        out_tmp = LIB + ".tmp"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out_tmp, *objs]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
        os.replace(out_tmp, LIB)
        return LIB
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv))
