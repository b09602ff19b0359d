"""The C ABI library loads and exports every symbol its headers declare; host
logic that needs no device (parameter validation, error reporting)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header=None):
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h") or (header is not None and h != header):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(fvm2d_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(fvm):
    lib = fvm.lib()
    missing = [n for n in declared_symbols("fvm2d.h") if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols("fvm2d.h")) >= 25


def test_probes_library_exports_every_declared_symbol(fvm):
    lib = fvm.probes_lib()
    declared = declared_symbols("fvm2d_probes.h")
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(fvm.PROBE_EXPORTS) == set(declared)


def test_solver_library_holds_no_probes(fvm):
    """The measurement probes live in libfvm2d_probes.so only."""
    lib = fvm.lib()
    for n in declared_symbols("fvm2d_probes.h"):
        assert not hasattr(lib, n), n


def test_every_header_is_covered():
    """Each public header's declarations belong to one of the two libraries."""
    hs = sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h"))
    assert hs == ["fvm2d.h", "fvm2d_probes.h"]


def test_python_binding_covers_exports(fvm):
    assert set(fvm.EXPORTS) <= set(declared_symbols("fvm2d.h"))


def test_params_default_matches_reference(fvm):
    p = fvm.Params()
    fvm.lib().fvm2d_params_default(ctypes.byref(p))
    # src/wrappers.cuh:49-52
    assert (p.width, p.height, p.t_end, p.cfl, p.alpha, p.gamma) == (1.0, 1.0, 0.0, 0.5, 1.9, 1.4)
    assert p.strategy == 0 and p.max_steps == -1 and p.mode == 0


def test_ctypes_struct_layout_matches_header(fvm):
    assert ctypes.sizeof(fvm.Params) == 6 * 8 + 4 * 4


def test_bad_arguments_are_rejected_without_device(fvm):
    p = fvm.make_params(t_end=0.1)
    h = ctypes.c_void_p()
    rc = fvm.lib().fvm2d_create(4, 100, ctypes.byref(p), ctypes.byref(h))
    assert rc == -1
    assert b"6x6" in fvm.lib().fvm2d_last_error()
    # cell indices are 32-bit signed, as the reference's idx = j * Nx + i
    rc = fvm.lib().fvm2d_create(46341, 46341, ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"2^31-1" in fvm.lib().fvm2d_last_error()
    for field, bad, msg in (("strategy", 5, b"strategy"), ("mode", 2, b"mode"), ("cfl", 0.0, b"CFL"),
                            ("width", -1.0, b"width"), ("alpha", float("nan"), b"NaN")):
        q = fvm.make_params(t_end=0.1)
        setattr(q, field, bad)
        rc = fvm.lib().fvm2d_create(64, 64, ctypes.byref(q), ctypes.byref(h))
        assert rc == -1 and msg in fvm.lib().fvm2d_last_error(), field
    p.gamma = 1.0
    rc = fvm.lib().fvm2d_create(64, 64, ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"gamma" in fvm.lib().fvm2d_last_error()


def test_no_device_fails_loudly(fvm):
    if fvm.device_available():
        pytest.skip("a device is present")
    with pytest.raises(fvm.FVM2DError, match="no HIP device"):
        fvm.solve(np.ones((16, 16, 5)), t_end=0.1)


def test_drop_in_header_compiles_against_abi():
    """A test2D-style caller of the C++ drop-in header compiles and links
    against libfvm2d.so with the host compiler alone."""
    src = os.path.join(ROOT, "tests", "cpp", "drop_in_caller.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "drop_in_caller")
    lib_dir = os.path.join(ROOT, "2d_fvm_cuda_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src,
                        "-L", lib_dir, "-lfvm2d", "-Wl,-rpath," + lib_dir, "-o", out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_drop_in_header_argument_errors(tmp_path):
    """The C++ drop-in header refuses U0 rows of unequal length
    (std::invalid_argument) and an empty mesh (solve() throws
    std::runtime_error from fvm2d_solve's argument check, before any HIP
    call)."""
    src = os.path.join(ROOT, "tests", "cpp", "header_errors.cpp")
    out = str(tmp_path / "header_errors")
    lib_dir = os.path.join(ROOT, "2d_fvm_cuda_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src,
                        "-L", lib_dir, "-lfvm2d", "-Wl,-rpath," + lib_dir, "-o", out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
