"""Helpers for the tests of the C++ host layer (include/openfea_amd_cae.hpp):
build and run its driver tests/cpp/implicit_pipeline and read what it wrote."""
import os
import subprocess

import numpy as np

from conftest import REPO

DRIVER = os.path.join(REPO, "tests", "cpp", "implicit_pipeline")
CAE = os.path.join(REPO, "tests", "cpp", "cae_process")


def driver():
    """Path of the driver, built (make) only if it is not there yet."""
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", os.path.join(REPO, "tests", "cpp")], check=True)
    return DRIVER


def run_cae(deck, prefix, *args, timeout=120):
    """tests/cpp/cae_process (CAE_process: NLFEA Newton or explicit)."""
    if not os.path.exists(CAE):
        subprocess.run(["make", "-C", os.path.join(REPO, "tests", "cpp")], check=True)
    cmd = [CAE, deck, str(prefix)] + [str(a) for a in args]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def run(deck, prefix, mode=None, timeout=120):
    cmd = [driver(), deck, str(prefix)] + ([mode] if mode else [])
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def summary(stdout):
    line = [ln for ln in stdout.splitlines() if ln.startswith("summary ")][-1]
    return {k: v for k, v in (t.split("=", 1) for t in line.split()[1:])}


def load(prefix, name, dtype):
    return np.fromfile(f"{prefix}.{name}.{'i32' if dtype == np.int32 else 'f64'}", dtype=dtype)
