"""Shared test setup: the gpu marker, import paths and library fixtures."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PACKAGE = os.path.join(REPO, "kami-hair-solver_amd")
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "libkhs_oracle.so")

for p in (PACKAGE, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


def _ensure_oracle():
    if not os.path.exists(ORACLE_LIB):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    return ORACLE_LIB


@pytest.fixture(scope="session")
def oracle_lib():
    import khs_native

    return khs_native.load_library(_ensure_oracle())


@pytest.fixture(scope="session")
def gpu_lib():
    import khs_native

    if not os.path.exists(khs_native.GPU_LIBRARY):
        raise RuntimeError(
            "HIP library missing: run __graft_entry__.build() first; "
            + khs_native.GPU_LIBRARY
        )
    return khs_native.load_library(khs_native.GPU_LIBRARY)
