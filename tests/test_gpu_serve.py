"""The bake server on the GPU: uneven jobs served together give bit for bit
the result of each job baked alone (jobs share no device state and every
reduction has a fixed order), and the admission estimate covers what each
job holds on the device."""
import numpy as np
import pytest

import serve

pytestmark = pytest.mark.gpu


def _uneven_jobs():
    spec = [(0, 300, 6), (1, 1200, 4), (2, 600, 8)]  # (seed, strands, frames)
    return [serve.Job(job_id=k, arrival=0.0, scene=serve.ellipsoid_groom(seed, strands, subdivisions=4),
                      frames=frames) for k, (seed, strands, frames) in enumerate(spec)]


def test_serving_alone_equals_in_the_mix(gpu_lib):
    print("[gpu] three uneven bake jobs served together, then each alone", flush=True)
    mixed = _uneven_jobs()
    rep = serve.BakeServer(gpu_lib, slots=3, keep_positions=True).run(mixed)
    print(rep)
    assert rep["completed"] == 3 and rep["failed"] == 0
    assert rep["peak_concurrent_jobs"] == 3
    assert rep["estimate_covers_resident"]
    for job in mixed:
        alone = _uneven_jobs()[job.job_id]
        r = serve.BakeServer(gpu_lib, slots=1, keep_positions=True).run([alone])
        assert r["completed"] == 1
        print(f"job {job.job_id}: {job.nodes} nodes, {job.substeps} substeps; alone {alone.substeps}")
        assert alone.nodes == job.nodes and alone.substeps == job.substeps
        assert np.array_equal(alone.positions, job.positions)
