"""GPU checks of the drop-in front ends: the C++ header driven by a
test2D-style caller (tests/cpp/drop_in_caller.cpp) and the Python mirror,
including record mode (a density frame after the first step and every
gapPerFrame steps)."""
import glob
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cpp", "drop_in_caller")


def _caller():
    if not os.path.exists(CALLER):
        lib_dir = os.path.join(ROOT, "2d_fvm_cuda_amd")
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        CALLER + ".cpp", "-L", lib_dir, "-lfvm2d", "-Wl,-rpath," + lib_dir,
                        "-o", CALLER], check=True)
    return CALLER


def _frames(d):
    return sorted(int(os.path.basename(p)[:-4]) for p in glob.glob(os.path.join(d, "*.pgm")))


def test_python_mirror_record_mode(gpu, problems, tmp_path):
    io = __import__("importlib").import_module("2d_fvm_cuda_amd.io")
    U0 = problems.initialize_4shocks(48)
    r = gpu.RGRP_2D_CUDA(U0)
    r.setTimeLayout(0.02, 0.5)
    r.setRecordModel(True, 4, str(tmp_path))
    U = r.solve()
    n = r.getTimeStagesCount()
    plain, n_plain, _ = gpu.solve(U0, t_end=0.02)
    assert n == n_plain
    assert np.array_equal(U, plain)
    idx = _frames(str(tmp_path))
    assert idx == list(range(0, n, 4))
    for k in idx:
        Uk, nk, _ = gpu.solve(U0, t_end=0.02, max_steps=k + 1)
        assert nk == k + 1
        assert np.array_equal(io.read_pgm(os.path.join(str(tmp_path), f"{k}.pgm")),
                              io.normalize_u8(Uk[..., 0]))


def test_drop_in_caller_matches_library_and_records(gpu, problems, tmp_path):
    io = __import__("importlib").import_module("2d_fvm_cuda_amd.io")
    exe = _caller()
    out = subprocess.run([exe, "48", "0.02", "1"], capture_output=True, text=True, check=True)
    steps, _, rho_sum = out.stdout.split()
    U0 = problems.initialize_4shocks(48)
    ref, n_ref, _ = gpu.solve(U0, t_end=0.02, mode=gpu.MODE_EXACT)
    assert int(steps) == n_ref
    s = 0.0
    for row in ref[..., 0]:
        for c in row:
            s += float(c)  # the caller's sequential sum
    assert float(rho_sum) == s
    rec = str(tmp_path / "frames")
    subprocess.run([exe, "48", "0.02", "1", rec, "3"], capture_output=True, text=True, check=True)
    idx = _frames(rec)
    assert idx == list(range(0, n_ref, 3))
    for k in idx[:4]:
        Uk, _, _ = gpu.solve(U0, t_end=0.02, max_steps=k + 1, mode=gpu.MODE_EXACT)
        assert np.array_equal(io.read_pgm(os.path.join(rec, f"{k}.pgm")), io.normalize_u8(Uk[..., 0]))
