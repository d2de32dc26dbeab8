"""The oracle and the CPU build of the kernel's cell code under
AddressSanitizer + UndefinedBehaviorSanitizer and under the reference test
suite's floating-point traps (FE_DIVBYZERO | FE_INVALID | FE_OVERFLOW,
src/validation/validation.hpp:52, src/tests/testing.cpp:214-218).

tests/cpp/sanitize_driver.cpp runs the oracle's step, the cell code's FP64,
mixed and tangent builds and its reverse sweep (both SOA storages) on a
scene, in its own process: a sanitizer report or a trapped exception ends
that process with a non-zero status.  Scenes cover every cloud category,
a pre-cloud-chemistry state, negative inputs (clipped by the chain), empty
modes, one and three condensation sub-steps.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mam4xx_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
EXE = os.path.join(BUILD, "sanitize_driver")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _have_asan():
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    return out.returncode == 0 and os.path.isabs(out.stdout.strip())


@pytest.fixture(scope="module")
def driver():
    if not _have_asan():
        pytest.skip("gcc without libasan")
    os.makedirs(BUILD, exist_ok=True)
    # per-process intermediate names, then an atomic rename: concurrent test
    # processes never write a file another one is running
    tag = ".%d" % os.getpid()
    obj = os.path.join(BUILD, "mam4_oracle_san%s.o" % tag)
    subprocess.run(["gcc", "-c", "-O1", "-g", "-std=gnu11", "-fPIC", "-ffp-contract=off",
                    "-Wno-unknown-pragmas"] + SAN +
                   ["-o", obj, os.path.join(ROOT, "oracle", "mam4_oracle.c")], check=True)
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
                    "-Wno-unknown-pragmas"] + SAN +
                   ["-o", EXE + tag, os.path.join(HERE, "cpp", "sanitize_driver.cpp"),
                    os.path.join(ROOT, "mam4xx_amd", "csrc", "chain_consts.cpp"), obj, "-lm"],
                   check=True)
    os.remove(obj)
    os.replace(EXE + tag, EXE)
    return EXE


def _write(sc, prefix, precld=None):
    for k in ("state", "temperature", "pressure", "height", "qv", "cloud_fraction", "pblh",
              "dgncur_a", "dgncur_awet", "wetdens"):
        np.ascontiguousarray(sc[k], dtype=np.float64).tofile("%s.%s.bin" % (prefix, k))
    if precld is not None:
        np.ascontiguousarray(precld, dtype=np.float64).tofile(prefix + ".precld.bin")


def _scene(seed):
    sc = scenes.generate(12, 24, seed=seed)
    rng = np.random.default_rng(seed)
    sc["cloud_fraction"][:, :3] = 1.0
    sc["cloud_fraction"][:, 3:6] = rng.uniform(0.05, 0.95, (24, 3))
    sc["state"][:, :2, 6] *= -1.0         # negative inputs, clipped by the chain
    sc["state"][8:29, :, 7] = 0.0         # empty interstitial modes
    sc["state"][:4, :, 7] = 0.0
    return sc


@pytest.mark.parametrize("ns,precld,traps", [(1, False, True), (1, True, True),
                                             (3, True, True), (1, False, False)],
                         ids=["plain-traps", "precld-traps", "3sub-precld-traps", "plain"])
def test_clean_under_sanitizers(driver, tmp_path, ns, precld, traps):
    sc = _scene(60 + ns)
    pre = None
    if precld:
        rng = np.random.default_rng(61)
        pre = sc["state"] * rng.uniform(0.5, 1.0, sc["state"].shape)
    prefix = str(tmp_path / "scene")
    _write(sc, prefix, pre)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    out = subprocess.run([driver, prefix, "12", "24", str(ns), "1" if traps else "0"],
                         capture_output=True, text=True, env=env, timeout=600)
    print(out.stdout, out.stderr[-3000:])
    assert out.returncode == 0, (out.returncode, out.stderr[-3000:])
    assert "oracle mismatches 0, non-finite or inconsistent 0" in out.stdout


def test_traps_are_armed(driver, tmp_path):
    """The driver's trap set-up does end a process that divides by zero."""
    import signal
    prefix = str(tmp_path / "scene")
    _write(_scene(70), prefix)
    out = subprocess.run([driver, prefix, "12", "24", "1", "2"], capture_output=True, text=True,
                         timeout=120)
    # AddressSanitizer's handler reports the signal and exits with 1
    assert out.returncode == -signal.SIGFPE or (out.returncode == 1 and
                                                "AddressSanitizer: FPE" in out.stderr), \
        (out.returncode, out.stdout, out.stderr[-500:])
