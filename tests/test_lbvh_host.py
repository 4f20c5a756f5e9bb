"""CPU run of the HIP library's LBVH construction kernels (lbvh_build.hpp
compiled for the host, threads executed one by one with leaves arriving in
three orders): the self-check must be clean and box queries must return
exactly the brute-force overlap set, also on a tree refit with grown boxes
when the leaves are tested against the ungrown ones."""
import json
import os
import subprocess
import tempfile

import numpy as np

import grooms

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def _run(vertices, triangles, queries):
    src = os.path.join(HERE, "native", "check_lbvh.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check")
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(REPO, "include"),
                               "-I", os.path.join(REPO, "kami-hair-solver_amd", "csrc"), "-o", exe, src])
        fmt = lambda row: " ".join("%.17g" % x for x in row)
        lines = [f"{len(vertices)} {len(triangles)}"] + [fmt(v) for v in vertices]
        lines += [" ".join(map(str, t)) for t in triangles] + [str(len(queries))] + [fmt(q) for q in queries]
        out = subprocess.run([exe], input="\n".join(lines), capture_output=True, text=True, check=True)
    return json.loads(out.stdout)


def _queries(n, seed, radius, extent, pad):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n, 3))
    c = c / np.linalg.norm(c, axis=1, keepdims=True) * rng.uniform(0.9 * radius, 1.1 * radius, (n, 1))
    e = rng.uniform(0, extent, (n, 3))
    return np.concatenate([c - e, c + e, np.full((n, 1), pad)], axis=1)


def test_lbvh_kernels_on_icosphere():
    sc = grooms.small_sphere_scene()
    out = _run(sc.vertices, sc.triangles, _queries(1500, 0, 0.098, 0.01, 2.4e-4))
    assert out["box_violations"] == out["chain_violations"] == out["order_violations"] == 0
    assert out["leaf_box_mismatches"] == out["inverted_boxes"] == 0
    assert out["query_mismatches"] == 0 and out["pairs"] > 1000
    assert out["leaf_box_query_mismatches"] == 0


def test_lbvh_kernels_on_duplicate_codes():
    # many triangles share Morton codes: a tiny far triangle fan plus a big one
    v, t = grooms.icosphere(2, 1.0)
    v = np.concatenate([v, v * 1e-9 + 5.0])
    t = np.concatenate([t, t + len(v) // 2]).astype(np.uint32)
    out = _run(v, t, _queries(500, 1, 1.0, 0.2, 1e-3))
    assert out["box_violations"] == out["chain_violations"] == out["order_violations"] == 0
    assert out["query_mismatches"] == 0
    assert out["leaf_box_query_mismatches"] == 0
