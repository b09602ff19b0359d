"""GPU checks of the drop-in front ends against the oracle: the C++ header
driven by a test2D-style caller (tests/cpp/drop_in_caller.cpp) and the
Python mirror, including record mode (a density frame after the first step
and every gapPerFrame steps, holding the state before that step's ghost
fill, src/wrappers.cuh:253-261) and the text output of
writeFinalResultTo_txt (src/wrappers.cuh:318-321,370-397)."""
import glob
import importlib
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cpp", "drop_in_caller")
io = importlib.import_module("2d_fvm_cuda_amd.io")


def _caller():
    lib_dir = os.path.join(ROOT, "2d_fvm_cuda_amd")
    src = CALLER + ".cpp"
    deps = [src] + glob.glob(os.path.join(ROOT, "include", "*.h*"))
    if not os.path.exists(CALLER) or any(os.path.getmtime(CALLER) < os.path.getmtime(d) for d in deps):
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src,
                        "-L", lib_dir, "-lfvm2d", "-Wl,-rpath," + lib_dir, "-o", CALLER], check=True)
    return CALLER


def _frames(d):
    return sorted(int(os.path.basename(p)[:-4]) for p in glob.glob(os.path.join(d, "*.png")))


@pytest.mark.parametrize("strategy", [0, 1, 2, 3, 4])
def test_frames_equal_oracle_prefill_states(gpu, orc, problems, strategy):
    """fvm2d_download_frame inside a time loop returns, bit for bit in EXACT
    mode, the state the oracle's record mode copies out (after the branch
    average, before the ghost fill) at every step."""
    rng = problems.SplitMix64(31 + strategy)
    ny, nx = 50, 67
    U0 = problems.random_four_quadrant(rng, nx, ny)
    prm = dict(t_end=0.03, strategy=strategy)
    if strategy in (1, 2, 3):  # fast inflow (jets at speed 800): 12 steps
        prm["max_steps"] = 12
    _, n, _, frames, idx = orc.solve_record(U0, gap=1, nthreads=2, **prm)
    assert idx == list(range(n))
    s = gpu.Solver(nx, ny, mode=gpu.MODE_EXACT, **prm)
    s.upload(U0)
    s.begin()
    with pytest.raises(gpu.FVM2DError, match="no step"):
        s.download_frame()
    with pytest.raises(gpu.FVM2DError, match="no step"):
        s.download_frame_gray()
    for k in range(n):
        s.advance(1)
        steps, _, _ = s.sync()
        assert steps == k + 1
        got = s.download_frame()
        assert np.array_equal(got[..., :4], frames[k][..., :4]), k
        # the device's 8-bit image of the same frame
        assert np.array_equal(s.download_frame_gray(), io.normalize_u8(frames[k][..., 0])), k
    s.close()


def test_frame_gray_of_a_constant_density_is_zero(gpu, problems):
    """hi == lo: every pixel 0 (as normalize_u8 and cv::normalize give)."""
    U0 = np.zeros((20, 30, 5))
    U0[..., 0], U0[..., 1] = 1.25, 0.8
    U0[..., 2] = 0.3
    s = gpu.Solver(30, 20, t_end=1.0, max_steps=4)
    s.upload(U0)
    s.begin()
    s.advance(2)
    assert s.sync()[0] == 2
    assert np.all(s.download_frame()[..., 0] == 1.25)
    g = s.download_frame_gray()
    assert g.shape == (20, 30) and g.dtype == np.uint8 and not g.any()
    s.close()


def test_python_mirror_record_mode(gpu, orc, problems, tmp_path):
    U0 = problems.initialize_4shocks(48)
    r = gpu.RGRP_2D_CUDA(U0)
    r.setTimeLayout(0.02, 0.5)
    r.setArithmeticMode(gpu.MODE_EXACT)
    r.setRecordModel(True, 4, str(tmp_path))
    U = r.solve()
    n = r.getTimeStagesCount()
    ref, n_ref, _, frames, idx = orc.solve_record(U0, gap=4, nthreads=2, t_end=0.02)
    assert n == n_ref
    assert np.array_equal(U[..., :4], ref[..., :4])
    assert _frames(str(tmp_path)) == idx == list(range(0, n, 4))
    for k, f in zip(idx, frames):
        assert np.array_equal(io.read_png(os.path.join(str(tmp_path), f"{k}.png")),
                              io.colour_viridis(io.normalize_u8(f[..., 0]))), k


def test_drop_in_caller_matches_oracle_records_and_writes_text(gpu, orc, problems, tmp_path):
    exe = _caller()
    U0 = problems.initialize_4shocks(48)
    ref, n_ref, _, frames, idx = orc.solve_record(U0, gap=3, nthreads=2, t_end=0.02)
    txt = tmp_path / "txt"
    txt.mkdir()
    out = subprocess.run([exe, "48", "0.02", "1", "-", "0", str(txt)], capture_output=True, text=True,
                         check=True)
    steps, _, rho_sum = out.stdout.split()
    assert int(steps) == n_ref
    s = 0.0
    for row in ref[..., 0]:
        for c in row:
            s += float(c)  # the caller's sequential sum
    assert float(rho_sum) == s
    # writeFinalResultTo_txt: rho_/p_/u_/v_<std::to_string(t_end)>s.txt
    exp = tmp_path / "exp"
    exp.mkdir()
    io.write_matrix_txt(str(exp), ref, 0.02)
    names = sorted(os.listdir(txt))
    assert names == ["p_0.020000s.txt", "rho_0.020000s.txt", "u_0.020000s.txt", "v_0.020000s.txt"]
    for nm in names:
        assert (txt / nm).read_bytes() == (exp / nm).read_bytes(), nm
    rec = str(tmp_path / "frames")
    subprocess.run([exe, "48", "0.02", "1", rec, "3"], capture_output=True, text=True, check=True)
    assert _frames(rec) == idx == list(range(0, n_ref, 3))
    for k, f in zip(idx, frames):
        assert np.array_equal(io.read_png(os.path.join(rec, f"{k}.png")),
                              io.colour_viridis(io.normalize_u8(f[..., 0]))), k
