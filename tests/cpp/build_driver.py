"""Build recipe of the GPU test drivers: the C++ host-interface driver
(tests/cpp/adapter_driver.cpp -> tests/_build/libadapter_driver.so, linked
against mam4xx_amd/lib/libmam4x.so, build that first), the division
check kernel (tests/cpp/div_check.hip -> tests/_build/libdivcheck.so) and the
VALU-rate probe (tests/cpp/valu_rates.hip -> tests/_build/libvalurates.so).
Called by __graft_entry__.build(); the GPU tests load the results and fail
when they are missing.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(ROOT, "tests", "_build", "libadapter_driver.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-o", OUT + ".tmp", os.path.join(HERE, "adapter_driver.cpp"),
           "-L" + os.path.join(ROOT, "mam4xx_amd", "lib"), "-lmam4x",
           "-Wl,-rpath,$ORIGIN/../../mam4xx_amd/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    div = os.path.join(os.path.dirname(OUT), "libdivcheck.so")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
           "-o", div + ".tmp", os.path.join(HERE, "div_check.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(div + ".tmp", div)
    rates = os.path.join(os.path.dirname(OUT), "libvalurates.so")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
           "-o", rates + ".tmp", os.path.join(HERE, "valu_rates.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(rates + ".tmp", rates)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
