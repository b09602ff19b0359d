"""File output of the reference's RGRP_2D_CUDA.

Text: src/wrappers.cuh:370-397 writes four files ``rho_<t>s.txt``, ``p_<t>s.txt``,
``u_<t>s.txt`` and ``v_<t>s.txt`` in ``dir``, where ``<t>`` is
``std::to_string(t_end)`` (printf "%f") and every value is streamed with the
default ostream precision (printf "%.6g") followed by one space, one line per
mesh row (row 0 first).
"""
from __future__ import annotations

import os
import struct
import zlib

import numpy as np

from .viridis import VIRIDIS_U8


def write_matrix_txt(dir: str, U: np.ndarray, cTime: float) -> None:
    if U is None:
        U = np.zeros((0, 0, 5))
    tag = "%f" % cTime
    names = ("rho", "p", "u", "v")
    U = np.asarray(U, dtype=np.float64)
    row_fmt = "%.6g " * U.shape[1] + "\n"  # one format per row: half the time of one per value
    for k, name in enumerate(names):
        path = os.path.join(dir, f"{name}_{tag}s.txt")
        with open(path, "w") as f:
            f.write("".join(row_fmt % tuple(row) for row in U[:, :, k].tolist()))


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


def _viridis_u32() -> np.ndarray:
    lut = np.zeros((256, 4), dtype=np.uint8)
    lut[:, :3] = np.frombuffer(VIRIDIS_U8, dtype=np.uint8).reshape(256, 3)
    return lut.view("<u4").reshape(256)


_VIRIDIS_U32 = _viridis_u32()


def colour_viridis(img: np.ndarray) -> np.ndarray:
    """The 8-bit image through the viridis table (the cv::applyColorMap(...,
    COLORMAP_VIRIDIS) step of src/wrappers.cuh:342): shape + (3,), RGB.
    (One 4-byte table entry per pixel, then the fourth byte dropped: four
    times faster than indexing the 3-byte rows.)"""
    img = np.asarray(img, dtype=np.uint8)
    rgba = _VIRIDIS_U32[img].view(np.uint8).reshape(img.shape + (4,))
    return np.ascontiguousarray(rgba[..., :3])


def _png_chunk(kind: bytes, data: bytes) -> bytes:
    return (struct.pack(">I", len(data)) + kind + data +
            struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF))


def write_frame_png(dir: str, rho: np.ndarray, index: int) -> str:
    """Record-mode frame ``<dir>/<index>.png``: the density of every cell
    (ghost cells included, row 0 first) after normalize_u8, coloured with
    viridis, as an 8-bit RGB PNG -- the file src/wrappers.cuh:323-345 writes
    through OpenCV (zlib level 1, OpenCV's default for PNG)."""
    return write_gray_png(dir, normalize_u8(rho), index)


def write_gray_png(dir: str, gray: np.ndarray, index: int) -> str:
    """``<dir>/<index>.png`` from the frame's normalised 8-bit image (as
    normalize_u8 or the device's fvm2d_download_frame_gray makes it),
    coloured with viridis."""
    os.makedirs(dir, exist_ok=True)
    rgb = colour_viridis(gray)
    h, w = rgb.shape[:2]
    raw = b"".join(b"\x00" + rgb[j].tobytes() for j in range(h))  # filter type 0
    z = zlib.compress(raw, 1)
    idat = 1 << 30  # a chunk holds at most 2^31 - 1 bytes: 1 GiB IDAT chunks
    png = (b"\x89PNG\r\n\x1a\n" +
           _png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
           b"".join(_png_chunk(b"IDAT", z[k:k + idat]) for k in range(0, len(z), idat)) +
           _png_chunk(b"IEND", b""))
    path = os.path.join(dir, f"{index}.png")
    with open(path, "wb") as f:
        f.write(png)
    return path


def read_png(path: str) -> np.ndarray:
    """Read an 8-bit RGB PNG without interlacing whose rows use filter type 0
    (the frames written here and by include/fvm2d_frame.hpp); checks every
    chunk's CRC.  Returns an (h, w, 3) uint8 array."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG")
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        if crc != zlib.crc32(kind + body) & 0xFFFFFFFF:
            raise ValueError(f"{path}: bad CRC in {kind!r}")
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype, _, _, interlace = hdr
    if (depth, ctype, interlace) != (8, 2, 0):
        raise ValueError(f"{path}: not an 8-bit RGB PNG without interlacing")
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(h, 3 * w + 1)
    if np.any(raw[:, 0] != 0):
        raise ValueError(f"{path}: rows with a filter other than type 0")
    return raw[:, 1:].reshape(h, w, 3).copy()
