"""Bake driver over the solver library: the .khc hair cache and resumable bakes.

This is the Blender add-on's bake loop without Blender
(blender_extension/kami_hair_solver/addon.py). The cache file format, the
frame order, the temporary file that becomes the cache on success, and the
resume rules are kept. Two things change:

- Each frame record is gathered on the device and written by the library
  (khsxWriteCacheFrame). The add-on converts every point to a Python tuple
  first (addon.py:580-584).
- The states kept for rewinding live in the device rollback ring
  (khsxSetRollbackRing / khsxRollback). The add-on keeps host copies of
  whole checkpoints (addon.py:642-646).

Frames are scene frames `start..end`; the solver's animation frame is
`frame - start`, as in addon.py:743-759.
"""
from __future__ import annotations

import os
import struct
import time
from pathlib import Path

import numpy as np

import khs_native as kn

CACHE_MAGIC = b"KAMIHC1\0"
CACHE_HEADER = struct.Struct("<8sIII")  # magic, start, end, point count (addon.py:23-24)
TEMPORARY_SUFFIX = ".partial"


def temporary_path(path: Path) -> Path:
    """The file a bake writes until it finishes (addon.py:729)."""
    return path.with_suffix(path.suffix + TEMPORARY_SUFFIX)


def open_cache_for_calculation(temporary: Path, start: int, end: int, point_count: int,
                               completed_frame: int | None = None):
    """Opens the cache for writing (addon.py:704-724).

    A fresh bake truncates the file and writes the header. A resumed bake
    (`completed_frame` given) checks that the header matches and that the
    file holds every completed frame. It then cuts the file after the last
    completed frame and leaves the position there.
    """
    if completed_frame is None:
        file = temporary.open("wb")
        file.write(CACHE_HEADER.pack(CACHE_MAGIC, start, end, point_count))
        return file
    file = temporary.open("r+b")
    raw = file.read(CACHE_HEADER.size)
    if len(raw) != CACHE_HEADER.size or CACHE_HEADER.unpack(raw) != (CACHE_MAGIC, start, end, point_count):
        file.close()
        raise RuntimeError("The header of the unfinished cache does not match the bake being resumed.")
    expected = CACHE_HEADER.size + max(0, completed_frame - start + 1) * point_count * 3 * 8
    file.seek(0, os.SEEK_END)
    if file.tell() < expected:
        file.close()
        raise RuntimeError("The unfinished cache is shorter than its completed frames; it cannot be resumed.")
    file.truncate(expected)
    file.seek(expected)
    return file


def write_cache_frame(file, solver: kn.HairSolver) -> None:
    """Appends the solver's current original-point positions as one frame
    record (addon.py:580-584). Python's buffer is flushed first so the
    library's bytes follow everything already written, and the file object
    is moved to the new end afterwards (it caches its position)."""
    file.flush()
    solver.check(solver.write_cache_frame(file.fileno()))
    file.seek(0, os.SEEK_END)


def read_cache_frame(path: Path, frame: int) -> np.ndarray | None:
    """One frame's positions as a (points, 3) array, or None when the file
    has no such frame (addon.py:587-602)."""
    with Path(path).open("rb") as file:
        raw = file.read(CACHE_HEADER.size)
        if len(raw) != CACHE_HEADER.size:
            return None
        magic, start, end, point_count = CACHE_HEADER.unpack(raw)
        if magic != CACHE_MAGIC or frame < start or frame > end:
            return None
        file.seek(CACHE_HEADER.size + (frame - start) * point_count * 24)
        data = file.read(point_count * 24)
        if len(data) != point_count * 24:
            return None
        return np.frombuffer(data, dtype="<f8").reshape(point_count, 3).copy()


class Bake:
    """One bake of scene frames `start..end` into `path`.

    The solver must hold a finalized animation of `end - start + 1` frames.
    `checkpoint_frames` is the add-on's "rewind frames" setting: the most
    recent frames that can be resumed from (default 10, addon.py:100-103).
    """

    def __init__(self, solver: kn.HairSolver, start: int, end: int, fps: float, path, checkpoint_frames: int = 10):
        if end < start:
            raise ValueError("The end frame must not precede the start frame.")
        if checkpoint_frames < 1:
            raise ValueError("At least one frame must be kept for rewinding.")
        self.solver = solver
        self.start, self.end, self.fps = start, end, fps
        self.path = Path(path)
        self.checkpoint_frames = checkpoint_frames
        self.completed_frame = start - 1
        self.failed_frame: int | None = None
        self.error: Exception | None = None
        self.frame_seconds: list[float] = []
        self.retained: list[int] = []  # scene frames whose end state the ring holds
        self.final_positions: np.ndarray | None = None
        self.cancelled = False
        solver.check(solver.set_rollback_ring(checkpoint_frames))

    # ------------------------------------------------------------ running
    def run(self, frame_callback=None) -> bool:
        """Bakes from the first frame not yet completed to the end
        (addon.py:727-772). Returns True when the cache is complete. On
        failure, `error` and `failed_frame` say what happened, and the
        unfinished file is kept so the bake can be resumed."""
        temporary = temporary_path(self.path)
        resume = self.completed_frame >= self.start
        self.error = None
        try:
            with open_cache_for_calculation(temporary, self.start, self.end, self.solver.original_count(),
                                            self.completed_frame if resume else None) as file:
                for frame in range(self.completed_frame + 1, self.end + 1):
                    self.failed_frame = frame
                    if self.cancelled:
                        raise RuntimeError("The bake was cancelled.")
                    began = time.perf_counter()
                    self.solver.check(self.solver.step_frame(frame - self.start, 1.0 / self.fps))
                    write_cache_frame(file, self.solver)
                    self._remember(frame)
                    self.completed_frame = frame
                    self.frame_seconds.append(time.perf_counter() - began)
                    if frame_callback:
                        frame_callback(frame)
            os.replace(temporary, self.path)
            self.failed_frame = None
            self.final_positions = self.solver.original_positions()
            return True
        except Exception as exception:  # noqa: BLE001 - kept for the caller, as the add-on does
            self.error = exception
            return False

    def _remember(self, frame: int) -> None:
        # the device ring keeps the newest `checkpoint_frames` frames
        # (addon.py:642-646)
        self.retained.append(frame)
        del self.retained[:-self.checkpoint_frames]

    # ----------------------------------------------------------- resuming
    def resume_frame_range(self) -> tuple[int, int] | None:
        """First and last scene frames a failed bake can resume from
        (addon.py:649-656)."""
        if not self.retained or self.failed_frame is None:
            return None
        return min(self.retained) + 1, min(self.failed_frame, max(self.retained) + 1)

    def rewind(self, resume_frame: int) -> None:
        """Prepares to resume at `resume_frame` (addon.py:659-688): the
        device state is rolled back to the end of the frame before it. The
        frames after that one are dropped from the retained list and from the
        record, and the next run cuts them from the cache."""
        valid = self.resume_frame_range()
        if valid is None or not valid[0] <= resume_frame <= valid[1]:
            if valid:
                raise RuntimeError(f"Choose a resume frame from {valid[0]} to {valid[1]}.")
            raise RuntimeError("No rewind state is kept.")
        kept = resume_frame - 1
        self.solver.check(self.solver.rollback(kept - self.start))
        self.completed_frame = kept
        self.failed_frame = resume_frame
        self.frame_seconds = self.frame_seconds[:max(0, kept - self.start + 1)]
        self.retained = [f for f in self.retained if f <= kept]
        self.final_positions = read_cache_frame(temporary_path(self.path), kept)
