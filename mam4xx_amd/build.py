"""Build recipe of the HIP library (gfx950) -- called by __graft_entry__.build().

Produces ``mam4xx_amd/lib/libmam4x.so`` in-tree (git-ignored; it travels to
the GPU box with the source tree).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmam4x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["mam4x.hip", "chain_consts.cpp"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # packed-FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the FP64 drop-in performs the reference's IEEE operations one by one
    # (no a*b+c fusion), so it differs from the CPU build only in libm bits
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def build(verbose=False, extra_flags=()):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC] + HIP_FLAGS + list(extra_flags) + ["-o", LIB + ".tmp"] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra_flags=sys.argv[1:])
