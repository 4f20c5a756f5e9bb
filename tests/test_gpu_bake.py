"""The bake driver (.khc cache, resumable bakes over the device rollback
ring) against the HIP library."""
import numpy as np
import pytest

import bake
import bake_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", bake_cases.ALL, ids=[f.__name__ for f in bake_cases.ALL])
def test_bake_on_gpu(gpu_lib, tmp_path, case):
    print(f"[gpu] bake case {case.__name__}", flush=True)
    case(gpu_lib, tmp_path)


def test_gpu_cache_matches_oracle_cache(gpu_lib, oracle_lib, tmp_path):
    """Every frame record of the GPU's cache is within 1e-7 m of the
    oracle's."""
    print("[gpu] bake cache against the oracle", flush=True)
    paths = {}
    for name, lib in (("gpu", gpu_lib), ("oracle", oracle_lib)):
        s = bake_cases.animated_solver(lib)
        paths[name] = tmp_path / f"{name}.khc"
        b = bake.Bake(s, bake_cases.START, bake_cases.END, bake_cases.FPS, paths[name])
        assert b.run(), b.error
        s.close()
    for frame in range(bake_cases.START, bake_cases.END + 1):
        a = bake.read_cache_frame(paths["gpu"], frame)
        r = bake.read_cache_frame(paths["oracle"], frame)
        assert np.max(np.abs(a - r)) < 1e-7, frame
