"""The HIP library (when built) and the oracle load and export every entry
point the headers declare. No GPU work is done."""
import ctypes
import os
import re

import pytest

import khs_native as kn

HEADERS = [os.path.join(os.path.dirname(__file__), "..", "include", h)
           for h in ("kami_hair_solver.h", "kami_hair_solver_ext.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        names += re.findall(r"KHS_API\s+[^;(]*?\b(khs\w+)\s*\(", open(h).read())
    return names


def test_headers_declare_reference_abi():
    names = declared_functions()
    reference = [n for n in names if n.startswith("khs") and not n.startswith("khsx")]
    assert len(reference) == 33
    assert set(reference) == set(kn.ABI_FUNCTIONS)
    assert set(n for n in names if n.startswith("khsx")) == set(kn.EXT_FUNCTIONS)


def test_oracle_exports_everything(oracle_lib):
    for name in declared_functions():
        assert hasattr(oracle_lib, name), name


def test_hip_library_exports_everything():
    if not os.path.exists(kn.GPU_LIBRARY):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(kn.GPU_LIBRARY)
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.khsGetAbiVersion() == 5
