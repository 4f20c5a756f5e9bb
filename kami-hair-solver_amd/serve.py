"""Bake server: bake jobs of uneven groom sizes arrive over time, are admitted
as worker slots and device memory allow, and run concurrently on one GPU.

This is the Blender add-on's bake worker (blender_extension/kami_hair_solver/
addon.py:727-772, one worker thread stepping a resident animation frame by
frame and keeping a rollback ring of recent frames, :100-103, 642-701) turned
into a request loop:

- Each admitted job gets its own solver: its rod model, collider, every
  animation frame of roots and collider, and a 10-frame device rollback ring
  (khsxSetRollbackRing) stay resident in HBM until the job ends.
- Admission: a queued job starts when a worker slot is free and its
  estimated device bytes fit in the HBM budget next to the running jobs'
  (first come, first served; a job too large for the whole budget is
  refused). A finished job's memory is released before the next admission.
- Each solver runs on its own HIP stream (the library creates one per
  solver) from its own host thread; library calls release the interpreter
  lock, so the jobs' kernels and host round trips overlap on the device.
  Jobs share no device state, so every result is bit-identical to the same
  job baked alone (test_serving_alone_equals_in_the_mix).

The server records, per job, arrival, admission and finish times and every
substep's device time, and reports throughput (node-substeps/s over the
whole window) with p50/p99 substep and bake (arrival to last frame) latency.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

import grooms
import khs_native as kn

LIFTED = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)
RING_FRAMES = 10  # the add-on's rollback ring (addon.py:100-103)


# ------------------------------------------------------------------ jobs
@dataclass
class Job:
    """One bake request: a groom on its own animated ellipsoid collider."""
    job_id: int
    arrival: float              # seconds after the server starts
    scene: grooms.Scene
    frames: int                 # animation frames stepped (1..frames)
    # filled by the server
    admitted: float = math.nan
    finished: float = math.nan
    nodes: int = 0
    substeps: int = 0
    substep_ms: list = field(default_factory=list)
    positions: np.ndarray | None = None
    error: str | None = None
    device_bytes: int = 0


def ellipsoid_groom(seed: int, strands: int, *, median_nodes: float = 40.0, sigma: float = 0.41,
                    axes=None, spin_deg_per_s: float = 30.0, fps: float = 24.0,
                    subdivisions: int = 5) -> grooms.Scene:
    """A groom on a scaled icosphere (the job's own collider), vectorized so
    that large jobs are generated in well under a second: roots uniform on
    the upper part of the ellipsoid, node counts lognormal (median 40, sigma
    0.41, clipped 8..80), 8 mm segments leaving along the surface normal
    with a gentle random curl; body and roots rotate about z."""
    rng = np.random.default_rng(seed)
    if axes is None:
        axes = rng.uniform([0.07, 0.08, 0.10], [0.09, 0.10, 0.12])
    axes = np.asarray(axes, float)
    verts, tris = grooms.icosphere(subdivisions, 1.0)
    verts = verts * axes
    # roots: directions with z > 0.2, mapped onto the ellipsoid
    d = rng.normal(size=(strands * 3, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d[d[:, 2] > 0.2][:strands]
    while d.shape[0] < strands:  # vanishingly rare top-up
        e = rng.normal(size=(strands, 3))
        e /= np.linalg.norm(e, axis=1, keepdims=True)
        d = np.concatenate([d, e[e[:, 2] > 0.2]])[:strands]
    surface = d / np.linalg.norm(d / axes, axis=1, keepdims=True)
    normal = surface / axes ** 2
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    counts = np.clip(np.round(median_nodes * np.exp(sigma * rng.normal(size=strands))), 8, 80).astype(int)
    seg = 0.008
    pts, offsets = [], [0]
    # a curl: the tangent turns about a fixed axis perpendicular to the normal
    side = np.cross(normal, [0.0, 0.0, 1.0])
    side_norm = np.linalg.norm(side, axis=1, keepdims=True)
    side = np.where(side_norm > 1e-6, side / np.maximum(side_norm, 1e-12), [1.0, 0.0, 0.0])
    rate = rng.uniform(0.0, 0.06, strands)
    for s in range(strands):
        k = np.arange(counts[s])
        ang = rate[s] * k
        t = np.cos(ang)[:, None] * normal[s] + np.sin(ang)[:, None] * side[s]
        p = surface[s] + 0.0005 * normal[s] + seg * np.concatenate([[np.zeros(3)], np.cumsum(t[:-1], axis=0)])
        pts.append(p)
        offsets.append(offsets[-1] + counts[s])
    spin = math.radians(spin_deg_per_s) / fps
    return grooms.Scene(points=np.concatenate(pts), offsets=np.array(offsets, dtype=np.uint32), vertices=verts,
                        triangles=tris, desc=dict(substeps=8, maximum_substeps=32),
                        material=dict(friction=0.3, barrier_distance=1.0e-3), frame_dt=1.0 / fps,
                        spin_per_frame=spin)


def poisson_jobs(rate: float, window: float, seed: int, *, strands=(500, 20000), frames=(10, 60),
                 subdivisions: int = 5) -> list[Job]:
    """Arrivals of a Poisson process of `rate` jobs/s over `window` seconds;
    strand counts log-uniform over `strands`, frame counts uniform over
    `frames` (inclusive)."""
    rng = np.random.default_rng(seed)
    jobs, t, k = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > window:
            break
        n = int(round(math.exp(rng.uniform(math.log(strands[0]), math.log(strands[1])))))
        f = int(rng.integers(frames[0], frames[1] + 1))
        jobs.append(Job(job_id=k, arrival=t, scene=ellipsoid_groom(seed * 100003 + k, n, subdivisions=subdivisions),
                        frames=f))
        k += 1
    return jobs


# -------------------------------------------------------------- admission
def estimated_device_bytes(nodes: int, strands: int, vertices: int, triangles: int, frames: int,
                           ring: int = RING_FRAMES) -> int:
    """Upper estimate of a job's resident device memory, from the library's
    buffer layout: per node the state and solver vectors, model, Gauss-Newton
    blocks and Thomas factors, root targets and the 512-entry contact list
    (about 3.9 KB); 48 B per node and 24 B per collider vertex per resident
    animation frame; 96 B per node per rollback slot; about 0.4 KB per
    triangle for the boxes, sort keys and hierarchy; and a fixed 16 MiB for
    small buffers and the sort's temporary storage."""
    per_node = 217 * 8 + 15 * 4 + 512 * 4 + 64
    per_triangle = 12 + 48 + 128 + 4 * 4 + 112 + 8 + 8 + 16 + 40
    per_vertex = 4 * 24
    return int(nodes * per_node + strands * 32 + triangles * per_triangle + vertices * per_vertex
               + frames * (48 * nodes + 24 * vertices) + ring * 96 * nodes + (16 << 20))


def job_bytes(job: Job) -> int:
    # internal nodes are at least the groom's points (more where segments
    # exceed the element length; these grooms' 8 mm segments do not)
    nodes = int(job.scene.points.shape[0])
    return estimated_device_bytes(nodes, job.scene.strand_count, job.scene.vertices.shape[0],
                                  job.scene.triangles.shape[0], job.frames + 1)


# ----------------------------------------------------------------- server
class BakeServer:
    def __init__(self, lib, *, slots: int = 4, hbm_budget: float = 200e9, precision_mode: int = 0,
                 solve_algorithm: int = 0, keep_positions: bool = False, **desc):
        self.lib = lib
        self.slots = slots
        self.budget = int(hbm_budget)
        self.precision_mode = precision_mode
        self.solve_algorithm = solve_algorithm
        self.keep_positions = keep_positions
        self.desc = {**LIFTED, **desc}
        self._lock = threading.Condition()
        self._resident = 0
        self._running = 0
        self.peak_resident = 0
        self.peak_running = 0

    def bake(self, job: Job, t0: float):
        """Builds the job's solver, uploads its animation, steps every frame
        keeping the rollback ring, and releases it."""
        s = None
        try:
            s = grooms.build_solver(self.lib, job.scene, **self.desc)
            s.check(s.set_convergence_mode(1))
            if self.precision_mode:
                s.check(s.set_precision_mode(self.precision_mode))
            if self.solve_algorithm:
                s.check(s.set_solve_algorithm(self.solve_algorithm))
            s.check(s.allocate_animation(job.frames + 1))
            for f in range(job.frames + 1):
                s.check(s.set_root_frame(f, job.scene.frame_points(f)))
                s.check(s.set_collider_frame(f, job.scene.frame_vertices(f)))
            s.check(s.finalize_animation())
            s.check(s.set_rollback_ring(RING_FRAMES))
            job.nodes = s.node_count()
            job.device_bytes = int(s.gpu_stats().resident_bytes)
            for f in range(1, job.frames + 1):
                s.check(s.step_frame(f, job.scene.frame_dt))
                job.substeps += s.step_stats().substeps
                job.substep_ms.extend(s.substep_times().tolist())
            if self.keep_positions:
                job.positions = s.internal_positions()
        except Exception as e:  # reported with the job
            job.error = f"{type(e).__name__}: {e}"
        finally:
            if s is not None:
                s.close()
            job.finished = time.perf_counter() - t0

    def run(self, jobs: list[Job], *, time_scale: float = 1.0) -> dict:
        """Serves `jobs` (arrival times in seconds, multiplied by
        time_scale) and returns the report."""
        jobs = sorted(jobs, key=lambda j: j.arrival)
        queue: list[Job] = []
        workers: list[threading.Thread] = []
        t0 = time.perf_counter()
        pending = list(jobs)
        refused = []

        def finish(job, need):
            with self._lock:
                self._resident -= need
                self._running -= 1
                self._lock.notify_all()

        def worker(job, need):
            try:
                self.bake(job, t0)
            finally:
                finish(job, need)

        while pending or queue:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival * time_scale <= now:
                queue.append(pending.pop(0))
            started = False
            with self._lock:
                while queue:
                    job = queue[0]
                    need = job_bytes(job)
                    if need > self.budget:
                        refused.append(queue.pop(0))
                        job.error = f"needs {need} bytes, above the {self.budget}-byte budget"
                        continue
                    if self._running >= self.slots or self._resident + need > self.budget:
                        break
                    queue.pop(0)
                    self._resident += need
                    self._running += 1
                    self.peak_resident = max(self.peak_resident, self._resident)
                    self.peak_running = max(self.peak_running, self._running)
                    job.admitted = time.perf_counter() - t0
                    th = threading.Thread(target=worker, args=(job, need))
                    th.start()
                    workers.append(th)
                    started = True
                if not started:
                    # wait for a job to finish or the next arrival
                    wait = 0.05
                    if pending:
                        wait = min(wait, max(0.0, pending[0].arrival * time_scale - (time.perf_counter() - t0)))
                    if queue or pending:
                        self._lock.wait(timeout=max(wait, 0.001))
        for th in workers:
            th.join()
        elapsed = time.perf_counter() - t0
        return self.report(jobs, elapsed, time_scale, refused)

    def report(self, jobs, elapsed, time_scale, refused) -> dict:
        done = [j for j in jobs if j.error is None]
        failed = [j for j in jobs if j.error is not None and j not in refused]
        node_substeps = sum(j.nodes * j.substeps for j in done)
        def pct(x):
            x = np.asarray(x, float)
            return ({"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99))} if x.size
                    else {"p50": None, "p99": None})

        sub = [t for j in done for t in j.substep_ms]  # the oracle keeps no substep times
        bake = [j.finished - j.arrival * time_scale for j in done]
        wait = [j.admitted - j.arrival * time_scale for j in done]
        return {
            "value": node_substeps / elapsed,
            "unit": "node-substeps/s",
            "jobs": len(jobs),
            "completed": len(done),
            "failed": len(failed),
            "refused": len(refused),
            "elapsed_s": elapsed,
            "nodes_total": int(sum(j.nodes for j in done)),
            "substep_latency_ms": pct(sub),
            "bake_latency_s": pct(bake),
            "admission_wait_s": pct(wait),
            "slots": self.slots,
            "hbm_budget_bytes": self.budget,
            "peak_admitted_bytes": int(self.peak_resident),
            "peak_concurrent_jobs": int(self.peak_running),
            "estimate_covers_resident": bool(all(job_bytes(j) >= j.device_bytes for j in done)),
            "errors": [j.error for j in failed][:3],
        }
