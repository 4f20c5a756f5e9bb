"""The bake driver (.khc cache, resumable bakes) against the CPU oracle."""
import pytest

import bake_cases


@pytest.mark.parametrize("case", bake_cases.ALL, ids=[f.__name__ for f in bake_cases.ALL])
def test_bake_on_oracle(oracle_lib, tmp_path, case):
    case(oracle_lib, tmp_path)
