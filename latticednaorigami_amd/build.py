"""Build libldo_hip.so for gfx950 in-tree (latticednaorigami_amd/lib/)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libldo_hip.so")
SOURCES = ["sampler.hip", "mlp.hip", "plan.cpp", "tables.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # identical FP64 arithmetic to the reference's x86 build: no FMA fusion
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


DRIVER = os.path.join(LIB_DIR, "ldo_run")


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the reference-compatible driver: host C++ over the C ABI
    cmd = ["g++", "-O2", "-std=c++17", "-o", DRIVER + ".tmp",
           os.path.join(CSRC, "driver", "ldo_run.cpp"), "-L" + LIB_DIR, "-lldo_hip",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
