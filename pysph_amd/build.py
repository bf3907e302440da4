"""Build libsph_mi355x.so for gfx950 with hipcc (no GPU needed).

The library is written in-tree next to this file so that it travels with the
source tree.  `python -m pysph_amd.build` (or __graft_entry__.build()) runs it.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsph_mi355x.so")
SOURCES = ["sph_abi.hip", "sph_bin.hip", "sph_passes.hip", "sph_fast.hip", "sph_integrate.hip",
           "sph_batch.hip", "sph_mlp.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False, out=LIB, defines=()):
    """Compile every source and link `out`; `defines` are extra -D macros
    (tuning constants) for variant builds."""
    objdir = os.path.join(HERE, "_obj", os.path.basename(out).replace(".so", ""))
    os.makedirs(objdir, exist_ok=True)
    extra = ["-D" + d for d in defines]
    cmds = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmds.append([HIPCC] + [f for f in FLAGS if f != "-shared"] + extra +
                    ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the sources compile independently: one hipcc per source, in parallel
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        for f in [ex.submit(subprocess.check_call, cmd) for cmd in cmds]:
            f.result()
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
