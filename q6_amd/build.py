"""Compiles the HIP engine for gfx950 into q6_amd/libq6nb.so (in-tree)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/qnb_kernels.hip", "csrc/qnb_api.hip", "csrc/qnb_batch.hip", "csrc/qnb_adjoint.hip",
           "csrc/qnb_qbonded.hip", "csrc/qnb_fep.hip", "csrc/qnb_exc.hip"]
OUT = os.path.join(HERE, "libq6nb.so")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    cmd = ["hipcc"] + FLAGS + ["-o", OUT + ".tmp"] + [os.path.join(HERE, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
