"""Writes the 8-bit viridis colour table used for record-mode frames.

    python 2d_fvm_cuda_amd/make_viridis.py

The reference colours its frames with OpenCV's COLORMAP_VIRIDIS
(src/wrappers.cuh:323-345), a 256-entry map whose 8-bit table is
round(255 * v) of the viridis data published by its authors (Stefan van
der Walt and Nathaniel Smith, CC0), which matplotlib ships.  This script
reads that data from matplotlib (only here, at generation time) and writes
the rounded table twice: as a C++ array for the drop-in header
(include/fvm2d_viridis.hpp) and as bytes for the Python mirror
(2d_fvm_cuda_amd/viridis.py).  tests/test_record_io.py checks that both
copies agree with each other and, where matplotlib is importable, with
this rule.
"""
from __future__ import annotations

import os

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def table() -> bytes:
    from matplotlib import _cm_listed
    data = _cm_listed._viridis_data
    assert len(data) == 256
    out = bytearray()
    for rgb in data:
        for v in rgb:
            out.append(max(0, min(255, int(round(255.0 * v)))))
    return bytes(out)


def main() -> None:
    t = table()
    rows = []
    for i in range(256):
        r, g, b = t[3 * i:3 * i + 3]
        rows.append(f"{r:3d}, {g:3d}, {b:3d}")
    body = ",\n".join("    " + ", ".join(rows[i:i + 4]) for i in range(0, 256, 4))
    hpp = (
        "// 8-bit viridis colour table (R, G, B) of record-mode frames: round(255 v)\n"
        "// of the viridis data of its authors (CC0), the table OpenCV's\n"
        "// COLORMAP_VIRIDIS applies in the reference's frame writer\n"
        "// (src/wrappers.cuh:323-345).  Generated by 2d_fvm_cuda_amd/make_viridis.py.\n"
        "#pragma once\n\n"
        "namespace fvm2d {\n\n"
        "inline constexpr unsigned char kViridis[256 * 3] = {\n"
        f"{body}}};\n\n"
        "}  // namespace fvm2d\n")
    open(os.path.join(ROOT, "include", "fvm2d_viridis.hpp"), "w").write(hpp)
    hexs = t.hex()
    lines = "\n".join(f'    "{hexs[i:i + 96]}"' for i in range(0, len(hexs), 96))
    py = (
        '"""8-bit viridis colour table (R, G, B) of record-mode frames: round(255 v)\n'
        "of the viridis data of its authors (CC0), the table OpenCV's\n"
        "COLORMAP_VIRIDIS applies in the reference's frame writer\n"
        '(src/wrappers.cuh:323-345).  Generated by make_viridis.py."""\n\n'
        "VIRIDIS_U8 = bytes.fromhex(\n"
        f"{lines}\n)\n")
    open(os.path.join(PKG, "viridis.py"), "w").write(py)


if __name__ == "__main__":
    main()
