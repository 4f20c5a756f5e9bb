#!/usr/bin/env python3
"""Benchmark of the incremental-potential hair step on one MI355X.

Workload: the 6,757-strand groom (lognormal node counts, median 36, clipped
8..80; 261,305 nodes) on a head + torso body of 225,184 vertices rotating
30 deg/s, mu = 0.3, barrier 1 mm, 8 substeps per 24 fps frame, run to a
1e-6 relative Newton residual (the reference's residual floors lifted).
A step is one animation frame; inputs (every frame's root targets and
collider) are resident in HBM before the timed region.

Prints one JSON line: node-substeps/s of the whole job, p50/p99 substep
latency, the roofline of the dominant kernel (assembly of gradient and
Gauss-Newton blocks) from HIP events recorded on the library's stream, and
the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "kami-hair-solver_amd"))

import numpy as np  # noqa: E402

import grooms  # noqa: E402
import khs_native as kn  # noqa: E402

LIFTED = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)
HBM_PEAK_GBPS = 8000.0
# Algorithmic bytes of one assembly launch per node: read q, q_old, v
# (144 B), node constants R_ref, m, m_rot (88 B) and element constants
# L0, rest shear, rest curvature (56 B); write the gradient (48 B).
ASSEMBLY_BYTES_PER_NODE = 336


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup_animation(lib, scene, frames, **desc):
    s = grooms.build_solver(lib, scene, **desc)
    s.check(s.set_convergence_mode(1))
    s.check(s.allocate_animation(frames))
    for f in range(frames):
        s.check(s.set_root_frame(f, scene.frame_points(f)))
        s.check(s.set_collider_frame(f, scene.frame_vertices(f)))
    s.check(s.finalize_animation())
    return s


def cpu_baseline(scene, strands, threads):
    oracle_path = os.path.join(REPO, "oracle", "_build", "libkhs_oracle.so")
    lib = kn.load_library(oracle_path)
    sub = grooms.strand_subset(scene, strands)
    s = setup_animation(lib, sub, 2, **LIFTED)
    s.check(s.set_threads(threads))
    t0 = time.perf_counter()
    s.check(s.step_frame(1, scene.frame_dt))
    elapsed = time.perf_counter() - t0
    st = s.step_stats()
    nodes = s.node_count()
    return {
        "value": nodes * st.substeps / elapsed,
        "unit": "node-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {strands} strands ({nodes} nodes) of the same groom and full 225,184-vertex body, "
                  f"1 frame x {st.substeps} substeps to a 1e-6 residual, {elapsed:.1f} s on {threads} threads "
                  f"({platform.processor() or platform.machine()})",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=29, help="timed frames")
    ap.add_argument("--warmup", type=int, default=1, help="untimed frames before the timed ones")
    ap.add_argument("--cpu-strands", type=int, default=4096)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    t = time.perf_counter()
    scene = grooms.head_scene()
    log(f"[bench] groom generated in {time.perf_counter() - t:.1f} s: {scene.strand_count} strands, "
        f"{scene.points.shape[0]} points, {scene.vertices.shape[0]} body vertices")
    if not os.path.exists(kn.GPU_LIBRARY):
        raise SystemExit("HIP library missing; run __graft_entry__.build()")
    lib = kn.load_library(kn.GPU_LIBRARY)
    frames = args.warmup + args.steps + 1
    t = time.perf_counter()
    s = setup_animation(lib, scene, frames, **LIFTED)
    nodes = s.node_count()
    log(f"[bench] built and uploaded {frames} frames in {time.perf_counter() - t:.1f} s; {nodes} nodes")
    for f in range(1, args.warmup + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        st = s.step_stats()
        log(f"[bench] warmup frame {f}: {s.gpu_stats().last_frame_milliseconds:.1f} ms, "
            f"substeps {st.substeps}, newton {st.newton_iterations}")
    s.check(s.set_kernel_timing(True))
    substep_ms, total_substeps, newton, ls_evals = [], 0, 0, 0
    min_gap, penetrations = float("inf"), 0
    t0 = time.perf_counter()
    for f in range(args.warmup + 1, args.warmup + args.steps + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        st = s.step_stats()
        total_substeps += st.substeps
        newton += st.newton_iterations
        ls_evals += st.line_search_evaluations
        min_gap = min(min_gap, st.minimum_gap)
        if st.minimum_gap <= 1e-7:
            penetrations += 1
        substep_ms.extend(s.substep_times().tolist())
    elapsed = time.perf_counter() - t0
    timing = s.kernel_timing()
    log(f"[bench] timed {args.steps} frames in {elapsed:.3f} s; newton {newton}, line-search evals {ls_evals}, "
        f"min gap {min_gap:.3e}")
    value = nodes * total_substeps / elapsed
    asm_avg_ms = timing.assembly_ms / max(timing.assembly_launches, 1)
    achieved = ASSEMBLY_BYTES_PER_NODE * nodes / (asm_avg_ms * 1e-3) / 1e9 if timing.assembly_launches else None
    traffic = None
    traffic_file = os.path.join(REPO, "profiles", "assembly_hbm_traffic.json")
    if os.path.exists(traffic_file):
        traffic = json.load(open(traffic_file)).get("bytes_per_launch")
    result = {
        "metric": "node-substeps/s FP64 vs mixed at 1e-6 residual; p50/p99 substep latency",
        "value": value,
        "unit": "node-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "6,757-strand groom, 261,305 nodes, 225,184-vertex body rotating 30 deg/s, mu 0.3, "
                        "barrier 1 mm, 8 substeps/frame at 24 fps, relative Newton residual 1e-6",
            "form": "plain FP64 step",
            "strands": scene.strand_count,
            "nodes": nodes,
            "collider_triangles": int(scene.triangles.shape[0]),
        },
        "substep_latency_ms": {
            "p50": float(np.percentile(substep_ms, 50)) if substep_ms else None,
            "p99": float(np.percentile(substep_ms, 99)) if substep_ms else None,
            "count": len(substep_ms),
        },
        "newton_iterations_per_substep": newton / max(total_substeps, 1),
        "line_search_evaluations_per_substep": ls_evals / max(total_substeps, 1),
        "minimum_gap_m": min_gap,
        "penetrating_frames": penetrations,
        "kernel_ms": {
            "assembly_avg": asm_avg_ms,
            "assembly_launches": timing.assembly_launches,
            "energy_avg": timing.energy_ms / max(timing.energy_launches, 1),
            "solve_avg": timing.solve_ms / max(timing.solve_launches, 1),
            "collision_total": timing.collision_ms,
            "assembly_total": timing.assembly_ms,
            "energy_total": timing.energy_ms,
            "solve_total": timing.solve_ms,
        },
        "roofline": {
            "kernel": "k_evaluate<gradient> (gradient + Gauss-Newton block assembly)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBPS) if achieved else None,
            "traffic": traffic,
        },
        "reference_derived_node_substeps_per_s": 3.16e5,
    }
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        result["cpu_baseline"] = cpu_baseline(scene, args.cpu_strands, args.cpu_threads)
        log(f"[bench] cpu baseline in {time.perf_counter() - t:.1f} s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
