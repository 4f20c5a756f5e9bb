"""The reference's scenario tests run against the HIP library on the GPU
(tests/test_solver.cpp:18-343, tests/test_c_api.c:11-30 of the reference)."""
import pytest

import khs_native as kn
import scenarios

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", scenarios.ALL, ids=[f.__name__ for f in scenarios.ALL])
def test_reference_scenario_on_gpu(gpu_lib, case):
    print(f"[gpu] scenario {case.__name__}", flush=True)
    case(gpu_lib)


@pytest.mark.parametrize("case", scenarios.EXTENSIONS, ids=[f.__name__ for f in scenarios.EXTENSIONS])
def test_extension_scenario_on_gpu(gpu_lib, case):
    print(f"[gpu] extension scenario {case.__name__}", flush=True)
    case(gpu_lib)


def test_gpu_info(gpu_lib):
    """tests/test_c_api.c:26-29 asks for compute capability >= 12 (an NVIDIA
    sm_120 part); on MI355X the device reports gfx950, i.e. major 9, minor 5."""
    info = kn.gpu_info(gpu_lib)
    assert info.available == 1
    assert info.device_ordinal == 0
    assert info.compute_capability_major == 9 and info.compute_capability_minor == 5
    assert info.total_vram_bytes > 200 * 2**30
    assert b"gfx950" in info.device_name
