"""Build recipe for libdnlg.so (HIP kernels + C ABI) for gfx950.

Run as ``python build.py`` from anywhere; the library lands next to this
file so that dnlPython and the tests load it in-tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/kernels.hip", "csrc/elem_fp64.hip", "csrc/elem_mixed.hip", "csrc/history.hip", "csrc/batch.hip",
           "csrc/api.hip", "csrc/server.hip", "csrc/energy.hip", "csrc/autotune.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 (v_pk_*_f32) code generation: scalar FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fused within one expression only, decided by the front
    # end: with fusion left to the back end ("fast"), which products get
    # fused follows the IR the vectorizers leave, and that depends on the
    # register budget a kernel's amdgpu_waves_per_eu sets -- the same source
    # then rounded differently under two schedules
    "-ffp-contract=on",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build(verbose=False):
    """Each source to an object (in parallel), then one shared library."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(HERE, "libdnlg.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src).replace(".hip", ".o"))
        cmd = [HIPCC, *compile_flags, "-c", os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    jobs = int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(SOURCES)))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
