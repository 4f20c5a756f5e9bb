#!/usr/bin/env python3
"""Benchmark of the incremental-potential hair step on one MI355X.

Workload: the 6,757-strand groom (lognormal node counts, median 36, clipped
8..80; 261,305 nodes) on a head + torso body of 225,184 vertices rotating
30 deg/s, mu = 0.3, barrier 1 mm, 8 substeps per 24 fps frame, run to a
1e-6 relative Newton residual (the reference's residual floors lifted).
A step is one animation frame; inputs (every frame's root targets and
collider) are resident in HBM before the timed region.

Prints one JSON line: node-substeps/s of the whole job, p50/p99 substep
latency, the roofline of the dominant kernel class (by device time, from HIP
events recorded on the library's stream), the algorithmic-byte roofline of
the whole substep, and the CPU oracle timed on a bounded sample of the same
workload.
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
# Algorithmic bytes per node of one launch of each kernel class. Assembly
# and a line-search energy evaluation read q (or q_base and d), q_old, v
# (144 B), node constants R_ref, m, m_rot (88 B) and element constants L0,
# rest shear, rest curvature (56 B) and write 48 B (gradient, or the trial
# state). The block solve's only algorithmic traffic is the direction it
# writes (48 B): the Gauss-Newton blocks are intermediates, and the gradient
# it reads for the slope g.d is the assembly's output, not counted twice.
KERNEL_CLASSES = {
    "assembly": ("k_evaluate<gradient> with its contact list pass (gradient + Gauss-Newton blocks)", 336),
    "energy": ("k_evaluate<line search> with its contact list pass (trial energy)", 336),
    "solve": ("k_block_solve (block-tridiagonal Thomas per strand)", 48),
}


def substep_bytes(nodes, newton, line_search, vertices, triangles, pairs, substeps):
    """Algorithmic HBM bytes of `substeps` substeps: per node 528 B of
    substep begin/end and first assembly, 480 B per Newton iteration, 336 B
    per line-search trial; 72 B per collider vertex (lerp), 132 B per
    triangle (LBVH input), 84 B per candidate pair (indices and vertices)."""
    return (nodes * (528 * substeps + 480 * newton + 336 * line_search)
            + (72 * vertices + 132 * triangles) * substeps + 84 * pairs)


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_config1(threads):
    """The 64-strand sphere scene, one frame at 4 substeps to a 1e-6
    residual, on the oracle with `threads` threads."""
    lib = kn.load_library(os.path.join(REPO, "oracle", "_build", "libkhs_oracle.so"))
    scene = grooms.small_sphere_scene()
    s = grooms.build_solver(lib, scene, **LIFTED)
    s.check(s.set_convergence_mode(1))
    s.check(s.set_threads(threads))
    t0 = time.perf_counter()
    s.check(s.step(scene.frame_dt))
    elapsed = time.perf_counter() - t0
    nodes, substeps = s.node_count(), s.step_stats().substeps
    s.close()
    return {"value": nodes * substeps / elapsed, "unit": "node-substeps/s", "cores": threads,
            "sample": f"64 strands ({nodes} nodes) on the 1,280-triangle icosphere, 1 frame ({substeps} substeps), "
                      f"{elapsed:.2f} s"}


def cpu_baseline(scene, strands, frames, threads):
    oracle_path = os.path.join(REPO, "oracle", "_build", "libkhs_oracle.so")
    lib = kn.load_library(oracle_path)
    sub = grooms.strand_subset(scene, strands) if strands < scene.strand_count else scene
    s = setup_animation(lib, sub, frames + 1, **LIFTED)
    s.check(s.set_threads(threads))
    substeps = 0
    t0 = time.perf_counter()
    for f in range(1, frames + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        substeps += s.step_stats().substeps
    elapsed = time.perf_counter() - t0
    nodes = s.node_count()
    return {
        "value": nodes * substeps / elapsed,
        "unit": "node-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sub.strand_count} strands ({nodes} nodes) of the same groom and the full 225,184-vertex body, "
                  f"frames 1-{frames} ({substeps} substeps) to a 1e-6 residual, {elapsed:.1f} s on {threads} "
                  f"threads ({cpu_model()}, {os.cpu_count()} logical CPUs)",
    }


def run_batched(lib, scene, jobs, warmup, frames, algorithm):
    """`jobs` copies of the workload stepped concurrently, each solver on
    its own host thread and HIP stream (the library call releases the
    interpreter lock). Returns aggregate node-substeps/s over the timed
    frames and the per-job substep latencies."""
    import threading
    solvers = [setup_animation(lib, scene, warmup + frames + 1, **LIFTED) for _ in range(jobs)]
    for s in solvers:
        s.check(s.set_solve_algorithm(algorithm))
    for s in solvers:
        for f in range(1, warmup + 1):
            s.check(s.step_frame(f, scene.frame_dt))
    counts = [0] * jobs
    latencies = [[] for _ in range(jobs)]
    errors = []
    start = threading.Barrier(jobs + 1)

    def work(k):
        s = solvers[k]
        start.wait()
        try:
            for f in range(warmup + 1, warmup + frames + 1):
                s.check(s.step_frame(f, scene.frame_dt))
                counts[k] += s.step_stats().substeps
                latencies[k].extend(s.substep_times().tolist())
        except Exception as e:  # reported after the join
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(jobs)]
    for t in threads:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    for t in threads:
        t.join()
    elapsed = time.perf_counter() - t0
    if errors:
        raise errors[0]
    nodes = solvers[0].node_count()
    lat = np.concatenate([np.asarray(x) for x in latencies])
    for s in solvers:
        s.close()
    return {
        "jobs": jobs,
        "frames_per_job": frames,
        "value": nodes * sum(counts) / elapsed,
        "unit": "node-substeps/s",
        "elapsed_s": elapsed,
        "substep_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "form": "independent jobs on concurrent streams, plain FP64",
    }


def run_mixed(lib, scene, warmup, frames, fp64_positions):
    """The same workload with precision mode 1 (FP32 blocks and block solve,
    FP64 residual, energies, barrier, step bound and convergence test) from
    the same start: node-substeps/s over `frames` frames after `warmup`, the
    Newton work it took, the relative residual its substeps reached, and the
    position error against the FP64 step after the same frames."""
    s = setup_animation(lib, scene, warmup + frames + 1, **LIFTED)
    s.check(s.set_precision_mode(1))
    for f in range(1, warmup + 1):
        s.check(s.step_frame(f, scene.frame_dt))
    substeps = newton = ls_evals = 0
    residuals = []
    t0 = time.perf_counter()
    for f in range(warmup + 1, warmup + frames + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        st = s.step_stats()
        substeps += st.substeps
        newton += st.newton_iterations
        ls_evals += st.line_search_evaluations
        residuals.append(st.relative_residual_norm)
    elapsed = time.perf_counter() - t0
    nodes = s.node_count()
    err = float(np.max(np.abs(s.internal_positions() - fp64_positions))) if fp64_positions is not None else None
    s.close()
    return {
        "value": nodes * substeps / elapsed,
        "unit": "node-substeps/s",
        "frames": frames,
        "form": "FP32 Gauss-Newton blocks and block solve; FP64 assembly, residual, barrier, CCD, line search "
                "and convergence test",
        "newton_iterations_per_substep": newton / max(substeps, 1),
        "line_search_evaluations_per_substep": ls_evals / max(substeps, 1),
        "relative_residual_last_substep_median": float(np.median(residuals)),
        "reaches_1e-6_residual": bool(max(residuals) <= 1e-6),
        "max_position_error_vs_fp64_m": err,
    }


SOLVE_KERNELS = {
    0: "k_block_solve (block-tridiagonal Thomas per strand)",
    1: "k_block_solve_twisted (block-tridiagonal, from both ends)",
    2: "k_block_solve_fused (block-tridiagonal, fused elimination)",
}
SOLVE_ALGORITHMS = {
    0: "block Thomas with the Gauss-Jordan inverse (the reference's)",
    1: "twisted block factorization",
    2: "fused elimination (factors from Gauss-Jordan on [S | U | rhs], no inverse)",
}


def timing_pass(lib, scene, args):
    """The main form's timed frames again, from the same start, with kernel
    timing on: per-class kernel averages (HIP events on the solver's stream)
    and the khsGetGpuStats phase split. The trajectory is deterministic, so
    this is the timed pass's work."""
    s = setup_animation(lib, scene, args.warmup + args.steps + 1, **LIFTED)
    s.check(s.set_solve_algorithm(args.solve_algorithm))
    for f in range(1, args.warmup + 1):
        s.check(s.step_frame(f, scene.frame_dt))
    s.check(s.set_kernel_timing(True))
    phase_ms = {"assembly": 0.0, "collision": 0.0, "optimization": 0.0}
    newton = 0
    t0 = time.perf_counter()
    for f in range(args.warmup + 1, args.warmup + args.steps + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        newton += s.step_stats().newton_iterations
        gs = s.gpu_stats()
        for k in phase_ms:
            phase_ms[k] += getattr(gs, f"last_{k}_milliseconds")
    elapsed = time.perf_counter() - t0
    timing = s.kernel_timing()
    s.close()
    return timing, phase_ms, newton, elapsed


def run_solve_algorithm(lib, scene, warmup, frames, algorithm, reference_positions):
    """The same FP64 workload with another block-solve algorithm (0 block
    Thomas, 1 twisted factorization, 2 fused elimination) from the same
    start: node-substeps/s,
    the solve kernel's average launch, the Newton work, the relative residual
    the substeps reached, and the largest position difference from the main
    form after the same frames (in contact, where substeps stop at the
    iteration cap, rounding differences grow to millimetres)."""
    s = setup_animation(lib, scene, warmup + frames + 1, **LIFTED)
    s.check(s.set_solve_algorithm(algorithm))
    for f in range(1, warmup + 1):
        s.check(s.step_frame(f, scene.frame_dt))
    s.check(s.set_kernel_timing(True))
    substeps = newton = 0
    residuals, lat = [], []
    t0 = time.perf_counter()
    for f in range(warmup + 1, warmup + frames + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        st = s.step_stats()
        substeps += st.substeps
        newton += st.newton_iterations
        residuals.append(st.relative_residual_norm)
        lat.extend(s.substep_times().tolist())
    elapsed = time.perf_counter() - t0
    timing = s.kernel_timing()
    nodes = s.node_count()
    diff = (float(np.max(np.abs(s.internal_positions() - reference_positions)))
            if reference_positions is not None else None)
    s.close()
    return {
        "solve_algorithm": SOLVE_ALGORITHMS[algorithm],
        # with kernel timing on: its event bubbles are in this value
        "value_with_kernel_timing": nodes * substeps / elapsed,
        "unit": "node-substeps/s",
        "frames": frames,
        "ms_per_frame": elapsed * 1e3 / frames,
        "solve_avg_ms": timing.solve_ms / max(timing.solve_launches, 1),
        "solve_launches": timing.solve_launches,
        "substep_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "newton_iterations_per_substep": newton / max(substeps, 1),
        "relative_residual_last_substep_median": float(np.median(residuals)),
        "max_position_difference_vs_main_m": diff,
    }


def run_uneven(lib, warmup, frames, algorithm):
    """The uneven-length groom (50,000 strands of 4..256 nodes, about 6.5M
    nodes, per-strand E and radius) on the rotating head + torso body, plain
    FP64 to the 1e-6 residual: node-substeps/s over `frames` frames after
    `warmup`, with substep latency, Newton work and the block solve's
    average launch time."""
    t = time.perf_counter()
    scene, young, radius = grooms.uneven_scene()
    s = grooms.build_solver(lib, scene, **LIFTED)
    s.check(s.set_solve_algorithm(algorithm))
    s.check(s.set_convergence_mode(1))
    s.check(s.set_strand_materials(young, radius))
    total = warmup + frames + 1
    s.check(s.allocate_animation(total))
    for f in range(total):
        s.check(s.set_root_frame(f, scene.frame_points(f)))
        s.check(s.set_collider_frame(f, scene.frame_vertices(f)))
    s.check(s.finalize_animation())
    nodes = s.node_count()
    log(f"[bench] uneven groom: {scene.strand_count} strands, {nodes} nodes, ready in {time.perf_counter() - t:.1f} s")
    for f in range(1, warmup + 1):
        s.check(s.step_frame(f, scene.frame_dt))
    s.check(s.set_kernel_timing(True))
    substeps = newton = 0
    lat, gaps = [], []
    t0 = time.perf_counter()
    for f in range(warmup + 1, warmup + frames + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        st = s.step_stats()
        substeps += st.substeps
        newton += st.newton_iterations
        gaps.append(st.minimum_gap)
        lat.extend(s.substep_times().tolist())
    elapsed = time.perf_counter() - t0
    timing = s.kernel_timing()
    resident = int(s.gpu_stats().resident_bytes)
    s.close()
    return {
        "value": nodes * substeps / elapsed,
        "unit": "node-substeps/s",
        "workload": f"{scene.strand_count} strands of 4..256 nodes ({nodes} nodes), per-strand E U(1,5) GPa and "
                    f"radius U(30,80) um, 225,184-vertex body rotating 30 deg/s, 8 substeps/frame, relative "
                    f"Newton residual 1e-6, plain FP64",
        "frames": frames,
        "substep_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "newton_iterations_per_substep": newton / max(substeps, 1),
        "minimum_gap_m": float(min(gaps)),
        "solve_avg_ms": timing.solve_ms / max(timing.solve_launches, 1),
        "assembly_avg_ms": timing.assembly_ms / max(timing.assembly_launches, 1),
        "resident_bytes": resident,
    }


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB (reads and writes mib MiB),
    the known kernel against which the PMC traffic counters are scaled. It
    goes through the HIP runtime the library itself links."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")  # the soname the library links
    size = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(size)) != 0:
            raise RuntimeError("hipMalloc failed for the calibration copy")
    ok = hip.hipMemset(a, 1, ctypes.c_size_t(size)) == 0
    ok = ok and hip.hipMemcpy(b, a, ctypes.c_size_t(size), 3) == 0  # hipMemcpyDeviceToDevice
    ok = ok and hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    if not ok:
        raise RuntimeError("calibration copy failed")
    log(f"[bench] calibration copy of {mib} MiB done")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=29, help="timed frames")
    ap.add_argument("--warmup", type=int, default=1, help="untimed frames before the timed ones")
    ap.add_argument("--cpu-strands", type=int, default=6757, help="strands of the CPU sample (all by default)")
    ap.add_argument("--cpu-frames", type=int, default=2, help="frames of the CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-strands-one-thread", type=int, default=400,
                    help="strands of the one-thread CPU sample (frame 1)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this size first (for scaling PMC traffic counters)")
    ap.add_argument("--batch-jobs", type=int, default=4, help="concurrent jobs of the batched form (0: skip)")
    ap.add_argument("--mixed-frames", type=int, default=8, help="timed frames of the mixed-precision form (0: skip)")
    ap.add_argument("--solve-algorithm", type=int, default=2, choices=(0, 1, 2),
                    help="block solve of the main form: 0 block Thomas, 1 twisted factorization, 2 fused elimination")
    ap.add_argument("--other-solve-frames", type=int, default=8,
                    help="timed frames of the main workload with each other solve algorithm (0: skip)")
    ap.add_argument("--uneven-frames", type=int, default=3,
                    help="timed frames of the 50,000-strand uneven-length groom (0: skip)")
    ap.add_argument("--serve-window", type=float, default=20.0, help="arrival window of the served form, s (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=1.0, help="mean job arrivals per second of the served form")
    ap.add_argument("--serve-max-strands", type=int, default=6000, help="largest job of the served form, strands")
    ap.add_argument("--serve-slots", type=int, default=4, help="concurrent jobs of the served form")
    ap.add_argument("--batch-frames", type=int, default=None,
                    help="timed frames per job of the batched form (default: --steps)")
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
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
    frames = args.warmup + args.steps + 1
    t = time.perf_counter()
    s = setup_animation(lib, scene, frames, **LIFTED)
    s.check(s.set_solve_algorithm(args.solve_algorithm))
    nodes = s.node_count()
    log(f"[bench] built and uploaded {frames} frames in {time.perf_counter() - t:.1f} s; {nodes} nodes")
    for f in range(1, args.warmup + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        st = s.step_stats()
        log(f"[bench] warmup frame {f}: {s.gpu_stats().last_frame_milliseconds:.1f} ms, "
            f"substeps {st.substeps}, newton {st.newton_iterations}")
    substep_ms, total_substeps, newton, ls_evals, pairs = [], 0, 0, 0, 0
    min_gap, penetrations = float("inf"), 0
    mixed_last = args.warmup + args.mixed_frames
    other_last = args.warmup + args.other_solve_frames
    main_at_other_last = None
    last_residuals = []  # relative residual of each frame's last substep
    fp64_at_mixed_last = None
    dispatches0, waits0 = s.launch_counts()
    t0 = time.perf_counter()
    for f in range(args.warmup + 1, args.warmup + args.steps + 1):
        s.check(s.step_frame(f, scene.frame_dt))
        if f == mixed_last:
            # the positions the mixed form is compared with (outside the
            # step; the copy costs a few milliseconds of the timed region)
            fp64_at_mixed_last = s.internal_positions()
        if f == other_last:
            main_at_other_last = fp64_at_mixed_last if f == mixed_last else s.internal_positions()
        st = s.step_stats()
        total_substeps += st.substeps
        newton += st.newton_iterations
        ls_evals += st.line_search_evaluations
        pairs += st.contact_candidate_count
        min_gap = min(min_gap, st.minimum_gap)
        if st.minimum_gap <= 1e-7:
            penetrations += 1
        last_residuals.append(st.relative_residual_norm)
        substep_ms.extend(s.substep_times().tolist())
    elapsed = time.perf_counter() - t0
    dispatches1, waits1 = s.launch_counts()
    log(f"[bench] timed {args.steps} frames in {elapsed:.3f} s; newton {newton}, line-search evals {ls_evals}, "
        f"min gap {min_gap:.3e}")
    value = nodes * total_substeps / elapsed
    s.close()
    t = time.perf_counter()
    timing, phase_ms, timing_newton, timing_elapsed = timing_pass(lib, scene, args)
    log(f"[bench] timing pass: {args.steps} frames in {timing_elapsed:.3f} s ({time.perf_counter() - t:.1f} s with "
        f"set-up), newton {timing_newton}")
    totals = {k: (getattr(timing, k + "_ms"), getattr(timing, k + "_launches")) for k in KERNEL_CLASSES}
    dominant = max(totals, key=lambda k: totals[k][0])
    dom_ms, dom_launches = totals[dominant]
    dom_avg_ms = dom_ms / max(dom_launches, 1)
    kernel_name, bytes_per_node = KERNEL_CLASSES[dominant]
    if dominant == "solve":
        kernel_name = SOLVE_KERNELS[args.solve_algorithm]
    achieved = bytes_per_node * nodes / (dom_avg_ms * 1e-3) / 1e9 if dom_launches else None
    traffic = None
    traffic_file = os.path.join(REPO, "profiles", f"{dominant}_hbm_traffic.json")
    if os.path.exists(traffic_file):
        traffic = json.load(open(traffic_file)).get("bytes_per_launch")
    n_vertices, n_triangles = int(scene.vertices.shape[0]), int(scene.triangles.shape[0])
    step_bytes = substep_bytes(nodes, newton, ls_evals, n_vertices, n_triangles, pairs, total_substeps)
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
            "form": "plain FP64 step, " + SOLVE_ALGORITHMS[args.solve_algorithm],
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
        # substeps stop at the 1e-6 relative residual or at the reference's
        # 24-iteration cap; in contact the cap often comes first
        "residual_rule": "relative Newton residual 1e-6 or 24 iterations, whichever first (floors lifted)",
        "last_substep_relative_residual": {
            "median": float(np.median(last_residuals)),
            "max": float(np.max(last_residuals)),
            "frames_reaching_1e-6": int(sum(r <= 1e-6 for r in last_residuals)),
            "frames": len(last_residuals),
        },
        "line_search_evaluations_per_substep": ls_evals / max(total_substeps, 1),
        # GPU dispatches (kernels and device copies) and host waits over the
        # timed frames, everything included (per-substep collider and tree
        # work, rejected line-search trials)
        "dispatches_per_newton_iteration": (dispatches1 - dispatches0) / max(newton, 1),
        "host_waits_per_substep": (waits1 - waits0) / max(total_substeps, 1),
        "host_waits_per_newton_iteration": (waits1 - waits0) / max(newton, 1),
        "minimum_gap_m": min_gap,
        "penetrating_frames": penetrations,
        # khsGetGpuStats phase times summed over the timing pass's frames
        "phase_ms": phase_ms,
        "kernel_timing_pass": {
            "how": "the timed frames again from the same start with kernel timing on (HIP events around each "
                   "kernel class); the headline pass runs without, as each event record idles the GPU ~4 us",
            "newton_iterations": timing_newton,
            "same_work_as_timed_pass": timing_newton == newton,
            "ms_per_step": timing_elapsed * 1e3 / args.steps,
        },
        "kernel_ms": {
            "assembly_avg": timing.assembly_ms / max(timing.assembly_launches, 1),
            "assembly_launches": timing.assembly_launches,
            "energy_avg": timing.energy_ms / max(timing.energy_launches, 1),
            "energy_launches": timing.energy_launches,
            "solve_avg": timing.solve_ms / max(timing.solve_launches, 1),
            "solve_launches": timing.solve_launches,
            "assembly_total": timing.assembly_ms,
            "energy_total": timing.energy_ms,
            "solve_total": timing.solve_ms,
            "collision_total": timing.collision_ms,
        },
        "roofline": {
            "kernel": kernel_name,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBPS) if achieved else None,
            "traffic": traffic,
            "bytes_per_node": bytes_per_node,
            "nodes_per_launch": nodes,
            "avg_launch_ms": dom_avg_ms,
            # the measured HBM traffic over the same launch time: what the
            # kernel actually moves (blocks, factors), against the peak
            "traffic_gbps": traffic / (dom_avg_ms * 1e-3) / 1e9 if traffic and dom_launches else None,
            "traffic_frac": traffic / (dom_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS if traffic and dom_launches else None,
        },
        "substep_roofline": {
            "algorithmic_bytes": step_bytes,
            "achieved": step_bytes / elapsed / 1e9,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": step_bytes / elapsed / 1e9 / HBM_PEAK_GBPS,
        },
        "reference_derived_node_substeps_per_s": 3.16e5,
    }
    if args.other_solve_frames > 0:
        result["other_solve_algorithms"] = []
        for algorithm in (0, 1, 2):
            if algorithm == args.solve_algorithm:
                continue
            t = time.perf_counter()
            other = run_solve_algorithm(lib, scene, args.warmup, args.other_solve_frames, algorithm,
                                        main_at_other_last)
            result["other_solve_algorithms"].append(other)
            log(f"[bench] solve algorithm {algorithm}: {args.other_solve_frames} frames in "
                f"{time.perf_counter() - t:.1f} s, solve {other['solve_avg_ms'] * 1e3:.1f} us")
    if args.mixed_frames > 0:
        t = time.perf_counter()
        result["mixed_precision"] = run_mixed(lib, scene, args.warmup, args.mixed_frames, fp64_at_mixed_last)
        log(f"[bench] mixed form: {args.mixed_frames} frames in {time.perf_counter() - t:.1f} s, "
            f"{result['mixed_precision']['value']:.4g} node-substeps/s")
    if args.batch_jobs > 0:
        t = time.perf_counter()
        result["batched"] = run_batched(lib, scene, args.batch_jobs, args.warmup, args.batch_frames or args.steps,
                                        args.solve_algorithm)
        result["batched"]["solve_algorithm"] = SOLVE_ALGORITHMS[args.solve_algorithm]
        log(f"[bench] batched form: {args.batch_jobs} jobs in {time.perf_counter() - t:.1f} s, "
            f"{result['batched']['value']:.4g} node-substeps/s")
    if args.uneven_frames > 0:
        t = time.perf_counter()
        try:
            result["uneven"] = run_uneven(lib, 1, args.uneven_frames, args.solve_algorithm)
            result["uneven"]["solve_algorithm"] = SOLVE_ALGORITHMS[args.solve_algorithm]
            log(f"[bench] uneven groom in {time.perf_counter() - t:.1f} s, "
                f"{result['uneven']['value']:.4g} node-substeps/s")
        except kn.SolverError as e:  # reported in the line; the headline stands
            result["uneven"] = {"value": None, "error": str(e)}
            log(f"[bench] uneven groom failed: {e}")
    if args.serve_window > 0:
        import serve
        t = time.perf_counter()
        jobs = serve.poisson_jobs(args.serve_rate, args.serve_window, seed=5, strands=(500, args.serve_max_strands),
                                  frames=(10, 30))
        log(f"[bench] serving: {len(jobs)} jobs generated in {time.perf_counter() - t:.1f} s")
        t = time.perf_counter()
        served = serve.BakeServer(lib, slots=args.serve_slots, hbm_budget=200e9,
                                  solve_algorithm=args.solve_algorithm).run(jobs)
        served["arrivals"] = (f"Poisson, {args.serve_rate}/s over {args.serve_window} s; strands log-uniform "
                              f"500..{args.serve_max_strands}, nodes lognormal median 40 (8..80), frames 10..30, "
                              f"each on its own rotating ellipsoid (icosphere subdivision 5)")
        served["form"] = ("continuous admission by slots and a 200 GB HBM budget; each job resident with a "
                          "10-frame rollback ring on its own stream; plain FP64")
        served["solve_algorithm"] = SOLVE_ALGORITHMS[args.solve_algorithm]
        result["served"] = served
        log(f"[bench] served {served['completed']} jobs in {time.perf_counter() - t:.1f} s, "
            f"{served['value']:.4g} node-substeps/s")
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        base = cpu_baseline(scene, args.cpu_strands, args.cpu_frames, args.cpu_threads)
        base["cpu_model"] = cpu_model()
        # one thread on a bounded sample of the same groom (the first
        # strands, the full body, frame 1), and the 64-strand scene in full
        base["one_thread"] = cpu_baseline(scene, args.cpu_strands_one_thread, 1, 1)
        base["sphere_scene"] = {"one_thread": cpu_config1(1), "threads": cpu_config1(args.cpu_threads)}
        result["cpu_baseline"] = base
        log(f"[bench] cpu baseline in {time.perf_counter() - t:.1f} s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
