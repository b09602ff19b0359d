"""Host-side checks of the record-mode frame writers and text output
(2d_fvm_cuda_amd/io.py, include/fvm2d_frame.hpp): normalisation, the
viridis table, PNG layout and round trip."""
import importlib
import os
import re
import struct
import subprocess

import numpy as np
import pytest

io = importlib.import_module("2d_fvm_cuda_amd.io")
VIRIDIS_U8 = importlib.import_module("2d_fvm_cuda_amd.viridis").VIRIDIS_U8
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_normalize_u8_minmax_rounding():
    a = np.array([[1.0, 1.5, 2.0], [1.25, 1.75, 1.0 + 1.0 / 510]])
    img = io.normalize_u8(a)
    # (x - min) * 255 / (max - min), rounded to nearest (ties to even)
    assert img.dtype == np.uint8
    assert img.tolist() == [[0, 128, 255], [64, 191, 0]]
    assert io.normalize_u8(np.full((3, 4), 7.0)).max() == 0


def test_frame_png_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.1, 2.0, size=(7, 11))
    path = io.write_frame_png(str(tmp_path / "frames"), rho, 36)
    assert os.path.basename(path) == "36.png"
    head = open(path, "rb").read(24)
    assert head[:8] == b"\x89PNG\r\n\x1a\n" and head[12:16] == b"IHDR"
    assert struct.unpack(">II", head[16:24]) == (11, 7)
    back = io.read_png(path)
    assert back.shape == (7, 11, 3)
    assert np.array_equal(back, io.colour_viridis(io.normalize_u8(rho)))
    lut = np.frombuffer(VIRIDIS_U8, dtype=np.uint8).reshape(256, 3)
    assert np.array_equal(back[io.normalize_u8(rho) == 0][0], lut[0])
    assert np.array_equal(back[io.normalize_u8(rho) == 255][0], lut[255])


def test_viridis_table():
    """The Python and C++ copies of the table are the same, run from
    (68, 1, 84) to (253, 231, 37), and are round(255 v) of the viridis data
    where matplotlib (which ships that data) is importable."""
    lut = np.frombuffer(VIRIDIS_U8, dtype=np.uint8).reshape(256, 3)
    hpp = open(os.path.join(ROOT, "include", "fvm2d_viridis.hpp")).read()
    body = hpp[hpp.index("{", hpp.index("kViridis")) + 1:hpp.index("};")]
    assert [int(x) for x in re.findall(r"\d+", body)] == list(VIRIDIS_U8)
    assert lut[0].tolist() == [68, 1, 84] and lut[255].tolist() == [253, 231, 37]
    try:
        from matplotlib import _cm_listed
    except ImportError:
        pytest.skip("matplotlib not importable")
    ref = np.rint(255.0 * np.asarray(_cm_listed._viridis_data)).astype(np.uint8)
    assert np.array_equal(lut, ref)


@pytest.mark.parametrize("w,h", [(11, 7), (300, 100), (1, 1)])
@pytest.mark.parametrize("stored", [False, True])
def test_cpp_frame_writer(tmp_path, w, h, stored):
    """include/fvm2d_frame.hpp (the drop-in header's frame writer: zlib's
    compress2 loaded at run time, or stored deflate blocks, of which 300 x
    100 needs two) decodes with zlib, and with PIL, to the viridis colouring
    of its input."""
    exe = str(tmp_path / "frame_writer")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "frame_writer.cpp"), "-o", exe], check=True)
    out = str(tmp_path / "f.png")
    env = dict(os.environ, FVM2D_PNG_STORED="1" if stored else "0")
    subprocess.run([exe, out, str(w), str(h)], check=True, env=env)
    i, j = np.meshgrid(np.arange(w), np.arange(h))
    gray = ((7 * i + 13 * j) % 256).astype(np.uint8)
    assert np.array_equal(io.read_png(out), io.colour_viridis(gray))
    raw_bytes = h * (3 * w + 1)
    if stored:
        assert os.path.getsize(out) > raw_bytes
    elif w * h > 1000:
        assert os.path.getsize(out) < raw_bytes // 2  # compressed (libz loads on this host)
    try:
        from PIL import Image
    except ImportError:
        return
    # an independent PNG decoder reads the same pixels
    assert np.array_equal(np.asarray(Image.open(out).convert("RGB")), io.colour_viridis(gray))


def test_matrix_txt_names_and_format(tmp_path):
    U = np.zeros((2, 3, 5))
    U[..., 0] = [[1.0, 2.5, 1.0 / 3.0], [4.0, 5.0, 6.0]]
    io.write_matrix_txt(str(tmp_path), U, 0.3)
    txt = open(tmp_path / "rho_0.300000s.txt").read()
    assert txt == "1 2.5 0.333333 \n4 5 6 \n"
    for name in ("p", "u", "v"):
        assert os.path.exists(tmp_path / f"{name}_0.300000s.txt")


def _dt(orc, U, t, nx, ny, p):
    """dt rule of src/wrappers.cuh:198-203 on state U at time t."""
    smax = max(orc.speed(U[j, i, :4], p["gamma"]) for j in range(ny) for i in range(nx))
    h = min(p["width"] / nx, p["height"] / ny)
    return min(p["cfl"] * h / smax, p["t_end"] - t)


import pytest  # noqa: E402


@pytest.mark.parametrize("strategy", [0, 4])
def test_oracle_record_frames_are_the_prefill_states(orc, problems, strategy):
    """The oracle's record mode (src/wrappers.cuh:253-261) copies the state
    between the branch average and the ghost fill: filling a frame gives the
    state after its step; its ghost rows are the X-sweep of the step's input
    (the Y-sweep leaves them alone), its ghost columns the Y-sweep (the
    X-sweep leaves them alone) and its corners the input itself.  Frames
    come after the first step and every gap steps, indexed 0, gap, 2 gap..."""
    rng = problems.SplitMix64(19 + strategy)
    ny, nx = 28, 36
    U0 = problems.random_four_quadrant(rng, nx, ny)
    p = dict(width=1.0, height=1.0, t_end=0.05, cfl=0.5, alpha=1.9, gamma=1.4, strategy=strategy)
    U, n, t, frames, idx = orc.solve_record(U0, gap=2, nthreads=1, **p)
    assert n >= 5 and idx == list(range(0, n, 2))
    for frame, k in zip(frames, idx):
        s = k + 1  # steps done when the frame is taken
        after, ns, ts, _ = orc.solve(U0, nthreads=1, max_steps=s, **p)
        prev, _, tp, _ = orc.solve(U0, nthreads=1, max_steps=s - 1, **p)
        assert ns == s
        filled = orc.fill(frame, strategy, ts, p["width"], p["height"])
        assert np.array_equal(filled[..., :4], after[..., :4])
        dt = _dt(orc, prev, tp, nx, ny, p)
        X = orc.sweep_x(prev, p["width"], dt, p["alpha"], p["gamma"])
        Y = orc.rotate(orc.sweep_x(orc.rotate(prev), p["height"], dt, p["alpha"], p["gamma"]))
        g = [0, 1, ny - 2, ny - 1]
        c = [0, 1, nx - 2, nx - 1]
        assert np.array_equal(frame[g, 2:nx - 2, :4], X[g, 2:nx - 2, :4])
        assert np.array_equal(frame[2:ny - 2][:, c, :4], Y[2:ny - 2][:, c, :4])
        assert np.array_equal(frame[np.ix_(g, c)][..., :4], prev[np.ix_(g, c)][..., :4])
        # the fill changes the ghost cells: the frame is not the filled state
        assert not np.array_equal(frame[..., :4], after[..., :4])


def test_cpp_header_text_files_match_ostream_format(tmp_path):
    """The C++ header's writeFinalResultTo_txt (values formatted by
    std::to_chars, four files on four threads) writes the bytes of the
    reference's ostream format (src/wrappers.cuh:370-397: default precision,
    "%.6g", a space after each value), which io.write_matrix_txt restates:
    on seeded values over many magnitudes and the printf edge cases."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "txt_writer")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"),
                        os.path.join(root, "tests", "cpp", "txt_writer.cpp"), "-pthread", "-ldl", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    rng = np.random.default_rng(4)
    ny, nx = 37, 53
    U = rng.standard_normal((ny, nx, 5)) * 10.0 ** rng.integers(-12, 12, size=(ny, nx, 5))
    special = [0.0, -0.0, 1.0, 0.5, 123456.5, 1234565.0, 999999.5, 9999995.0, 1e-5, 1e-4, 1e21,
               5e-324, 2.2250738585072014e-308, 0.30000000000000004, 1.0 / 3.0, 1e16]
    U.reshape(-1)[:len(special)] = special
    U[..., 4] = 0.0
    src = tmp_path / "U.bin"
    U.astype("<f8").tofile(src)
    got, exp = tmp_path / "got", tmp_path / "exp"
    got.mkdir()
    exp.mkdir()
    r = subprocess.run([exe, str(src), str(ny), str(nx), "0.3", str(got)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stderr
    io.write_matrix_txt(str(exp), U, 0.3)
    names = sorted(os.listdir(got))
    assert names == ["p_0.300000s.txt", "rho_0.300000s.txt", "u_0.300000s.txt", "v_0.300000s.txt"]
    for nm in names:
        assert (got / nm).read_bytes() == (exp / nm).read_bytes(), nm
