"""Build recipe for libdnlg.so (HIP kernels + C ABI) for gfx950.

Run as ``python build.py`` from anywhere; the library lands next to this
file so that dnlPython and the tests load it in-tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/kernels.hip", "csrc/history.hip", "csrc/batch.hip", "csrc/api.hip", "csrc/server.hip",
           "csrc/energy.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 (v_pk_*_f32) code generation: scalar FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build(verbose=False):
    out = os.path.join(HERE, "libdnlg.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *[os.path.join(HERE, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
