"""The bake server's admission and bookkeeping, on the CPU oracle: jobs
start in arrival order as worker slots and the memory budget allow, a job
larger than the whole budget is refused, and every admitted job completes
with its frames stepped."""
import numpy as np

import serve


def _jobs(n, strands, frames, spacing):
    return [serve.Job(job_id=k, arrival=k * spacing, scene=serve.ellipsoid_groom(k, strands[k % len(strands)],
                                                                                  subdivisions=1),
                      frames=frames) for k in range(n)]


def test_estimate_grows_with_every_input():
    base = serve.estimated_device_bytes(1000, 20, 642, 1280, 10)
    assert serve.estimated_device_bytes(2000, 20, 642, 1280, 10) > base
    assert serve.estimated_device_bytes(1000, 20, 642, 1280, 20) > base
    assert serve.estimated_device_bytes(1000, 20, 642, 2560, 10) > base
    assert serve.estimated_device_bytes(1000, 20, 642, 1280, 10, ring=20) > base


def test_slots_and_budget_bound_admission(oracle_lib):
    jobs = _jobs(5, (3, 5, 4), 2, 0.0)
    one = max(serve.job_bytes(j) for j in jobs)
    # two slots, and a budget that fits two of the largest jobs but not three
    srv = serve.BakeServer(oracle_lib, slots=2, hbm_budget=2.5 * one)
    rep = srv.run(jobs)
    print(rep)
    assert rep["completed"] == 5 and rep["failed"] == 0 and rep["refused"] == 0
    assert rep["peak_concurrent_jobs"] <= 2
    assert rep["peak_admitted_bytes"] <= 2.5 * one
    assert all(j.substeps == 2 * 8 and j.nodes > 0 for j in jobs)
    # first come, first served: admissions follow arrivals
    order = sorted(jobs, key=lambda j: j.admitted)
    assert [j.job_id for j in order] == sorted(j.job_id for j in jobs)
    assert rep["value"] > 0 and rep["bake_latency_s"]["p99"] >= rep["bake_latency_s"]["p50"]


def test_job_above_the_budget_is_refused(oracle_lib):
    jobs = _jobs(2, (3, 40), 1, 0.0)
    srv = serve.BakeServer(oracle_lib, slots=2, hbm_budget=serve.job_bytes(jobs[0]) * 1.01)
    rep = srv.run(jobs)
    assert rep["completed"] == 1 and rep["refused"] == 1
    assert jobs[1].error is not None and "budget" in jobs[1].error
