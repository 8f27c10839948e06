"""Build libvmcx.so (HIP, gfx950) in-tree with hipcc.

    python -m vmc_torch_amd.build        # or __graft_entry__.build()

The library is placed next to this file so that it travels with the source
tree; the Python host code loads it from there and nowhere else.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvmcx.so")
SOURCES = ["capi.hip", "slater.hip", "hubbard.hip", "sr.hip", "chol.hip", "minres.hip", "nnbf.hip", "fpeps.hip", "bmps.hip", "proposal.hip", "syrk.hip", "symmetry.hip", "noalloc.hip"]
ARCH = os.environ.get("VMCX_OFFLOAD_ARCH", "gfx950")


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return exe


def compile_flags():
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # never emit packed FP32 (v_pk_*_f32) instructions
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def build(verbose=True):
    """Compile every source to an object, then link libvmcx.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmds.append([hipcc(), *compile_flags(), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    # one hipcc per source, at most 8 at a time
    jobs = max(1, min(8, os.cpu_count() or 1, len(cmds)))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(compile_one, cmds))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocblas"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
