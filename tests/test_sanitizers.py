"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle (every entry point, every preset, sizes that are not multiples of
16, the host reference adjoint), the engine's launch-layout logic and
parameter checks, and the drop-in header's PNG frame writer, each built
with g++ -fsanitize=address,undefined from a driver in tests/cpp/ and run;
any report aborts the driver.  The header's threaded text writer also runs
under ThreadSanitizer."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _build_and_run(tmp_path, name, srcs, extra):
    exe = str(tmp_path / name)
    r = subprocess.run(["g++", "-std=c++17", *SAN, *extra, *srcs, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1", OMP_NUM_THREADS="2",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0 and r.stdout.startswith("OK")


def test_oracle_under_sanitizers(tmp_path):
    srcs = [os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.cpp"),
            os.path.join(ROOT, "oracle", "fvm2d_oracle.cpp"),
            os.path.join(ROOT, "oracle", "fvm2d_oracle_adjoint.cpp")]
    _build_and_run(tmp_path, "oracle_san", srcs, ["-ffp-contract=off", "-fopenmp", "-I", os.path.join(ROOT, "oracle")])


def test_engine_layout_under_sanitizers(tmp_path):
    if not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
        pytest.skip("HIP headers not installed")
    srcs = [os.path.join(ROOT, "tests", "cpp", "layout_sanitize.cpp")]
    _build_and_run(tmp_path, "layout_san", srcs, ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                                                  "-Wno-unused-result"])


@pytest.mark.parametrize("stored", ["0", "1"])
def test_frame_writer_under_sanitizers(tmp_path, stored):
    """The drop-in header's PNG frame writer (include/fvm2d_frame.hpp), with
    zlib loaded at run time and with stored blocks, on an image that needs
    several stored blocks."""
    exe = str(tmp_path / "frame_san")
    r = subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "frame_writer.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", FVM2D_PNG_STORED=stored)
    out = str(tmp_path / "f.png")
    r = subprocess.run([exe, out, "513", "301"], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.path.getsize(out) > 0


def test_text_writer_under_thread_sanitizer(tmp_path):
    """The drop-in header's writeFinalResultTo_txt formats its four files on
    four threads: built with g++ -fsanitize=thread and run on a small state,
    no data race is reported."""
    import numpy as np
    exe = str(tmp_path / "txt_tsan")
    r = subprocess.run(["g++", "-std=c++17", "-fsanitize=thread", "-g", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "txt_writer.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    U = np.random.default_rng(2).standard_normal((30, 41, 5))
    src = tmp_path / "U.bin"
    U.astype("<f8").tofile(src)
    out = tmp_path / "out"
    out.mkdir()
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    r = subprocess.run([exe, str(src), "30", "41", "0.1", str(out)], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stderr[-4000:]
    assert len(os.listdir(out)) == 4
