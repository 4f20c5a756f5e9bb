"""The reference's scenario tests run against the CPU oracle.

These pin the oracle's model build, Newton loop, contact, sweep and
checkpoint logic to the assertions of the reference's own test suite
(tests/test_solver.cpp:18-343, tests/test_c_api.c:11-30).
"""
import pytest

import scenarios


@pytest.mark.parametrize("case", scenarios.ALL, ids=[f.__name__ for f in scenarios.ALL])
def test_reference_scenario_on_oracle(oracle_lib, case):
    case(oracle_lib)


@pytest.mark.parametrize("case", scenarios.EXTENSIONS, ids=[f.__name__ for f in scenarios.EXTENSIONS])
def test_extension_scenario_on_oracle(oracle_lib, case):
    case(oracle_lib)
