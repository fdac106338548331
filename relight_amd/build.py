"""Builds librelight_amd.so in-tree with hipcc for gfx950."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librelight_amd.so")
SOURCES = ["rl_kernels.hip", "rl_photon.hip", "rl_radiosity.hip", "rl_post.hip", "rl_raster.hip", "rl_runtime.cpp", "rl_bvh.cpp", "rl_obj.cpp"]

# -ffp-contract=off keeps host (lumel setup) and device (ray setup, shading) arithmetic rounding
# as written; -packed-fp32-ops keeps the compiler from emitting packed fp32 instructions.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Wno-unused-value", "-Wno-unused-command-line-argument",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print(" ".join(cmd))
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
