"""fvm2d_solve_device_noalloc (the allocation-free variant of the solve)
against fvm2d_solve on the same input: bit for bit in both arithmetic modes,
two calls in a row with the second on a smaller mesh, and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime libfvm2d is bound to (already loaded under this soname)
    L = ctypes.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


def _run_noalloc(gpu, hip, U0, **prm):
    ny, nx = U0.shape[:2]
    planes = np.ascontiguousarray(np.moveaxis(U0[..., :4], 2, 0))
    nbytes = planes.nbytes
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), nbytes) == 0
    try:
        assert hip.hipMemcpy(d, planes.ctypes.data, nbytes, 1) == 0  # host to device
        p = gpu.make_params(**prm)
        steps, t = ctypes.c_int(0), ctypes.c_double(0.0)
        rc = gpu.lib().fvm2d_solve_device_noalloc(d, nx, ny, ctypes.byref(p), ctypes.byref(steps),
                                                   ctypes.byref(t))
        gpu.check(rc, "fvm2d_solve_device_noalloc")
        out = np.empty_like(planes)
        assert hip.hipMemcpy(out.ctypes.data, d, nbytes, 2) == 0  # device to host
    finally:
        hip.hipFree(d)
    return np.moveaxis(out, 0, 2), steps.value, t.value


@pytest.mark.parametrize("mode", [1, 0])
def test_noalloc_solve_equals_solve_on_two_calls(gpu, problems, mode):
    hip = _hip()
    rng = problems.SplitMix64(91)
    cases = [(problems.random_four_quadrant(rng, 300, 211), dict(t_end=0.03)),
             (problems.random_four_quadrant(rng, 97, 64), dict(t_end=1.0, max_steps=9, strategy=4,
                                                                  width=1.5))]
    for U0, prm in cases:  # the second call on a smaller mesh
        print(f"noalloc mode {mode} shape {U0.shape} {prm}", flush=True)
        ref, n_ref, _ = gpu.solve(U0, mode=mode, **prm)
        got, n, t = _run_noalloc(gpu, hip, U0, mode=mode, **prm)
        assert n == n_ref and n > 0
        assert np.array_equal(got, ref[..., :4])


def test_noalloc_rejects_meshes_over_the_limit(gpu):
    p = gpu.make_params(t_end=0.1)
    steps, t = ctypes.c_int(0), ctypes.c_double(0.0)
    rc = gpu.lib().fvm2d_solve_device_noalloc(ctypes.c_void_p(16), 1025, 1024, ctypes.byref(p),
                                               ctypes.byref(steps), ctypes.byref(t))
    assert rc == -1 and b"no-allocation limit" in gpu.lib().fvm2d_last_error()
