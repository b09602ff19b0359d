"""Host-side checks of the record-mode frame writer and text output
(2d_fvm_cuda_amd/io.py): normalisation, PGM layout, round trip."""
import importlib
import os

import numpy as np

io = importlib.import_module("2d_fvm_cuda_amd.io")


def test_normalize_u8_minmax_rounding():
    a = np.array([[1.0, 1.5, 2.0], [1.25, 1.75, 1.0 + 1.0 / 510]])
    img = io.normalize_u8(a)
    # (x - min) * 255 / (max - min), rounded to nearest (ties to even)
    assert img.dtype == np.uint8
    assert img.tolist() == [[0, 128, 255], [64, 191, 0]]
    assert io.normalize_u8(np.full((3, 4), 7.0)).max() == 0


def test_frame_pgm_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.1, 2.0, size=(7, 11))
    path = io.write_frame_pgm(str(tmp_path / "frames"), rho, 36)
    assert os.path.basename(path) == "36.pgm"
    head = open(path, "rb").read(12)
    assert head.startswith(b"P5\n11 7\n255\n")
    back = io.read_pgm(path)
    assert back.shape == (7, 11)
    assert np.array_equal(back, io.normalize_u8(rho))
    assert back.min() == 0 and back.max() == 255


def test_matrix_txt_names_and_format(tmp_path):
    U = np.zeros((2, 3, 5))
    U[..., 0] = [[1.0, 2.5, 1.0 / 3.0], [4.0, 5.0, 6.0]]
    io.write_matrix_txt(str(tmp_path), U, 0.3)
    txt = open(tmp_path / "rho_0.300000s.txt").read()
    assert txt == "1 2.5 0.333333 \n4 5 6 \n"
    for name in ("p", "u", "v"):
        assert os.path.exists(tmp_path / f"{name}_0.300000s.txt")
