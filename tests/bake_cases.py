"""Bake-driver cases run against the oracle on the CPU and the HIP library on
the GPU: the .khc cache of a bake, a cancelled bake resumed from a rewound
frame, and the cache checks of the Blender add-on
(blender_extension/kami_hair_solver/addon.py:580-602, 642-772)."""
import math
import os

import numpy as np
import pytest

import bake
import khs_native as kn

START, END, FPS = 11, 18, 24.0


def animated_solver(lib):
    """Two hanging strands whose roots circle, one fixed root node each."""
    desc = kn.default_desc(lib)
    desc.gravity = kn.KhsVec3(0, -9.81, 0)
    desc.fixed_root_nodes = 1
    s = kn.HairSolver(lib, desc)
    strand = np.array([[0, 0, 0], [0, -0.02, 0], [0, -0.04, 0], [0, -0.06, 0]], float)
    points = np.concatenate([strand, strand + [0.01, 0, 0.003]])
    s.check(s.set_hair(points, [0, 4, 8], kn.default_material(lib)))
    s.check(s.build())
    frames = END - START + 1
    s.check(s.allocate_animation(frames))
    for k in range(frames):
        s.check(s.set_root_frame(k, points + [0.004 * math.sin(0.7 * k), 0.0, 0.003 * (1.0 - math.cos(0.7 * k))]))
        s.check(s.set_collider_frame(k, None))
    s.check(s.finalize_animation())
    return s


def straight_positions(lib):
    """Original-point positions after each scene frame, stepped directly."""
    s = animated_solver(lib)
    out = {}
    for frame in range(START, END + 1):
        s.check(s.step_frame(frame - START, 1.0 / FPS))
        out[frame] = s.original_positions()
    s.close()
    return out


def cache_records_every_frame(lib, tmp_path):
    expected = straight_positions(lib)
    s = animated_solver(lib)
    path = tmp_path / "hair.khc"
    b = bake.Bake(s, START, END, FPS, path, checkpoint_frames=3)
    assert b.run(), b.error
    assert path.exists() and not bake.temporary_path(path).exists()
    raw = path.read_bytes()
    points = s.original_count()
    assert bake.CACHE_HEADER.unpack(raw[:bake.CACHE_HEADER.size]) == (bake.CACHE_MAGIC, START, END, points)
    assert len(raw) == bake.CACHE_HEADER.size + (END - START + 1) * points * 24
    for frame in range(START, END + 1):
        assert np.array_equal(bake.read_cache_frame(path, frame), expected[frame]), frame
    assert bake.read_cache_frame(path, START - 1) is None
    assert bake.read_cache_frame(path, END + 1) is None
    assert np.array_equal(b.final_positions, expected[END])
    assert b.retained == [END - 2, END - 1, END]
    s.close()


def cancelled_bake_resumes(lib, tmp_path):
    """A bake cancelled after frame 15 resumes at frame 14 (rewinding to the
    end of 13) and finishes with the same cache, byte for byte, as an
    uninterrupted bake."""
    reference = tmp_path / "straight.khc"
    s = animated_solver(lib)
    assert bake.Bake(s, START, END, FPS, reference, checkpoint_frames=4).run()
    s.close()

    s = animated_solver(lib)
    path = tmp_path / "resumed.khc"
    b = bake.Bake(s, START, END, FPS, path, checkpoint_frames=4)

    def cancel_after_15(frame):
        if frame == 15:
            b.cancelled = True

    assert not b.run(cancel_after_15)
    assert isinstance(b.error, RuntimeError)
    assert b.failed_frame == 16 and b.completed_frame == 15
    assert not path.exists() and bake.temporary_path(path).exists()
    assert b.resume_frame_range() == (13, 16)
    with pytest.raises(RuntimeError):
        b.rewind(12)  # frame 11's end state has left the ring
    b.rewind(14)
    assert b.completed_frame == 13 and b.retained == [12, 13]
    assert np.array_equal(b.final_positions, bake.read_cache_frame(reference, 13))
    assert np.array_equal(s.original_positions(), b.final_positions)
    b.cancelled = False
    assert b.run(), b.error
    assert path.read_bytes() == reference.read_bytes()
    s.close()


def cache_checks(lib, tmp_path):
    """The resume checks of addon.py:712-723 and a failed frame write."""
    s = animated_solver(lib)
    temporary = tmp_path / "hair.khc.partial"
    points = s.original_count()
    with bake.open_cache_for_calculation(temporary, START, END, points) as f:
        s.check(s.step_frame(0, 1.0 / FPS))
        bake.write_cache_frame(f, s)
    with pytest.raises(RuntimeError, match="header"):
        bake.open_cache_for_calculation(temporary, START, END + 1, points, START)
    with pytest.raises(RuntimeError, match="shorter"):
        bake.open_cache_for_calculation(temporary, START, END, points, START + 1)
    with bake.open_cache_for_calculation(temporary, START, END, points, START) as f:
        assert f.tell() == bake.CACHE_HEADER.size + points * 24
    fd = os.open(temporary, os.O_RDONLY)
    try:
        assert s.write_cache_frame(fd) == kn.KHS_ERROR_INTERNAL
        assert "cache frame" in s.last_error()
    finally:
        os.close(fd)
    assert s.write_cache_frame(-1) == kn.KHS_ERROR_INVALID_ARGUMENT
    s.close()


ALL = [cache_records_every_frame, cancelled_bake_resumes, cache_checks]
