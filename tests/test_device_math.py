"""CPU check of the HIP library's element math (device_math.hpp compiled for
the host) against the oracle's dual-number strains and closest points."""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_device_math_matches_oracle():
    src = os.path.join(HERE, "native", "check_device_math.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check")
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, src])
        out = json.loads(subprocess.check_output([exe]).decode())
    # strains and closest points repeat the oracle's operations exactly
    assert out["strain_rel"] == 0.0
    assert out["geometry_abs"] == 0.0
    # the analytic Jacobian agrees with forward-mode differentiation to the
    # accuracy the latter reaches near the exp-map series threshold
    assert out["jacobian_rel"] < 1e-8
