"""Text output of the reference's writeFinalResultTo_txt.

src/wrappers.cuh:370-397 writes four files ``rho_<t>s.txt``, ``p_<t>s.txt``,
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
