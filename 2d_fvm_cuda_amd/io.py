"""File output of the reference's RGRP_2D_CUDA.

Text: src/wrappers.cuh:370-397 writes four files ``rho_<t>s.txt``, ``p_<t>s.txt``,
``u_<t>s.txt`` and ``v_<t>s.txt`` in ``dir``, where ``<t>`` is
``std::to_string(t_end)`` (printf "%f") and every value is streamed with the
default ostream precision (printf "%.6g") followed by one space, one line per
mesh row (row 0 first).
"""
from __future__ import annotations

import os

import numpy as np


def write_matrix_txt(dir: str, U: np.ndarray, cTime: float) -> None:
    if U is None:
        U = np.zeros((0, 0, 5))
    tag = "%f" % cTime
    names = ("rho", "p", "u", "v")
    for k, name in enumerate(names):
        path = os.path.join(dir, f"{name}_{tag}s.txt")
        with open(path, "w") as f:
            for row in U:
                f.write("".join("%.6g " % x for x in row[:, k]))
                f.write("\n")


def normalize_u8(a: np.ndarray) -> np.ndarray:
    """Min-max normalisation to 0..255 with rounding to nearest, the
    cv::normalize(..., 0, 255, NORM_MINMAX, CV_8UC1) step of
    src/wrappers.cuh:323-345 (a constant field maps to 0)."""
    a = np.asarray(a, dtype=np.float64)
    lo, hi = float(a.min()), float(a.max())
    if not hi > lo:
        return np.zeros(a.shape, dtype=np.uint8)
    scale = 255.0 / (hi - lo)
    return np.clip(np.rint(a * scale - lo * scale), 0, 255).astype(np.uint8)


def write_frame_pgm(dir: str, rho: np.ndarray, index: int) -> str:
    """Record-mode frame: ``<dir>/<index>.pgm``, the density of every cell
    (ghost cells included, row 0 first) as an 8-bit binary PGM after
    normalize_u8.  The reference writes ``<index>.png`` through OpenCV with
    the viridis colour map applied to the same 8-bit image
    (src/wrappers.cuh:323-345); the grey levels here are that image before
    colouring."""
    os.makedirs(dir, exist_ok=True)
    img = normalize_u8(rho)
    path = os.path.join(dir, f"{index}.pgm")
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (img.shape[1], img.shape[0]))
        f.write(np.ascontiguousarray(img).tobytes())
    return path


def read_pgm(path: str) -> np.ndarray:
    """Read an 8-bit binary PGM written by write_frame_pgm."""
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(b"\n", 3)
    if parts[0] != b"P5":
        raise ValueError(f"{path}: not a binary PGM")
    w, h = (int(x) for x in parts[1].split())
    return np.frombuffer(parts[3], dtype=np.uint8, count=w * h).reshape(h, w)
