"""The allocation-free broad phase (khsxBroadphasePairsNoAlloc) against the
allocating one (khsxBroadphasePairs): equal sorted pair sets on two calls in
a row, the second with fewer pairs, so the second must not see what the
first left in the shared staging array; capacities above the staging limit
are refused."""
import numpy as np
import pytest

import grooms
import khs_native as kn

pytestmark = pytest.mark.gpu


def _pulled(solver, lo, hi, seed):
    q, v = solver.get_state()
    x = q.reshape(-1, 6)
    _, _, _, flags = solver.node_model()
    rng = np.random.default_rng(seed)
    for k in np.nonzero((flags & 1) == 0)[0]:
        x[k, :3] *= (0.098 + rng.uniform(lo, hi)) / np.linalg.norm(x[k, :3])
    solver.check(solver.set_state(q, v, q))


def test_no_alloc_broadphase_matches_allocating_one(gpu_lib):
    print("[gpu] allocation-free broad phase against the allocating one", flush=True)
    scene = grooms.small_sphere_scene()
    s = grooms.build_solver(gpu_lib, scene)
    sizes = []
    for lo, hi, seed in ((5e-5, 1.6e-3, 1), (8e-4, 1.6e-3, 2)):
        _pulled(s, lo, hi, seed)
        want = s.broadphase_pairs()
        got = s.broadphase_pairs_no_alloc()
        print(f"pairs allocating {len(want)} allocation-free {len(got)}")
        assert len(want) > 0
        assert np.array_equal(got, want)
        sizes.append(len(want))
    assert sizes[1] < sizes[0]
    # a capacity above the staging limit is refused with a message
    count = kn.C.c_uint64(0)
    big = np.zeros((4, 2), dtype=np.uint32)
    rc = gpu_lib.khsxBroadphasePairsNoAlloc(s.handle, kn._ptr(big), kn.KHSX_NO_ALLOC_PAIRS + 1, kn.C.byref(count))
    assert rc == kn.KHS_ERROR_INVALID_ARGUMENT
    assert "stages at most" in s.last_error()
    # a capacity below the total fills what fits and reports the total
    short = np.zeros((5, 2), dtype=np.uint32)
    rc = gpu_lib.khsxBroadphasePairsNoAlloc(s.handle, kn._ptr(short), 5, kn.C.byref(count))
    assert rc == kn.KHS_ERROR_INVALID_ARGUMENT and count.value == sizes[1]
    s.close()
