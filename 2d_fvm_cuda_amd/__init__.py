"""fvm2d on MI355X: Python mirror of the reference's solver API over libfvm2d.so.

The package directory name starts with a digit, so import it with
``importlib.import_module("2d_fvm_cuda_amd")``.

``RGRP_2D_CUDA`` keeps the method names, argument meaning and defaults of the
reference's class (src/wrappers.cuh:26-65): a caller builds it from a
[ny][nx][5] vec5d-layout array (row 0 = north, primitive (rho, p, u, v, 0),
two ghost layers included), sets parameters, and ``solve()`` returns the final
state with ghost cells filled.  Everything runs through the C ABI declared in
include/fvm2d.h; there is no CPU fallback: without a GPU (or without the built
library) every compute call raises.
"""
from __future__ import annotations

import ctypes
import os
from enum import IntEnum

import numpy as np

from . import problems  # noqa: F401  (re-exported problem library)

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfvm2d.so")
PROBES_LIB_PATH = os.path.join(PKG, "libfvm2d_probes.so")

FVM2D_OK = 0


class FVM2DError(RuntimeError):
    pass


class ghostCellStrategy(IntEnum):
    """enum class ghostCellStrategy, src/wrappers.cuh:17-24 (same order)."""
    outflow = 0
    reflective = 1
    jet = 2
    jet800 = 3
    doubleMach = 4


MODE_FAST = 0
MODE_EXACT = 1


class Params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_double),
        ("height", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("strategy", ctypes.c_int),
        ("max_steps", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]


EXPORTS = [
    "fvm2d_params_default", "fvm2d_last_error", "fvm2d_device_available", "fvm2d_solve",
    "fvm2d_solve_device_noalloc",
    "fvm2d_create", "fvm2d_set_params", "fvm2d_upload", "fvm2d_upload_planes_device",
    "fvm2d_download_planes_device", "fvm2d_run", "fvm2d_download", "fvm2d_download_frame",
    "fvm2d_download_frame_gray",
    "fvm2d_steps_total",
    "fvm2d_destroy", "fvm2d_set_launch", "fvm2d_get_launch", "fvm2d_autotune", "fvm2d_begin",
    "fvm2d_advance", "fvm2d_sync", "fvm2d_time_steps", "fvm2d_time_kernel",
    "fvm2d_batch_create", "fvm2d_batch_upload", "fvm2d_batch_begin", "fvm2d_batch_run",
    "fvm2d_batch_status", "fvm2d_batch_download", "fvm2d_batch_set_launch",
    "fvm2d_batch_autotune", "fvm2d_batch_destroy", "fvm2d_strip_layout",
    "fvm2d_adj_create", "fvm2d_adj_forward", "fvm2d_adj_backward", "fvm2d_adj_get_stats",
    "fvm2d_adj_set_launch", "fvm2d_adj_destroy",
]

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def hip_runtimes() -> list:
    """Paths of the HIP runtime libraries mapped into this process.

    PyTorch's ROCm wheel bundles its own HIP runtime under the soname of the
    system's (libamdhip64.so.7) while its libraries ask for it by another
    name.  Imported before libfvm2d is loaded, torch's copy is the only HIP
    runtime in the process: libfvm2d binds to it by soname and device
    pointers of torch tensors are valid for the library.  Imported after,
    torch maps a second runtime beside the system's, and only host memory
    can cross between the two."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln})
    except OSError:
        return []


def lib():
    """Load libfvm2d.so (built in-tree by build.py); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FVM2DError(f"{LIB_PATH} is not built (run 2d_fvm_cuda_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(Params)
    sig = {
        "fvm2d_params_default": (None, [P]),
        "fvm2d_last_error": (ctypes.c_char_p, []),
        "fvm2d_device_available": (ctypes.c_int, []),
        "fvm2d_solve": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, P, _ip, _dp]),
        "fvm2d_solve_device_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, P, _ip, _dp]),
        "fvm2d_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(_vp)]),
        "fvm2d_set_params": (ctypes.c_int, [_vp, P]),
        "fvm2d_upload": (ctypes.c_int, [_vp, _dp]),
        "fvm2d_upload_planes_device": (ctypes.c_int, [_vp, _vp]),
        "fvm2d_download_planes_device": (ctypes.c_int, [_vp, _vp]),
        "fvm2d_run": (ctypes.c_int, [_vp, _ip, _dp, _dp]),
        "fvm2d_download": (ctypes.c_int, [_vp, _dp]),
        "fvm2d_download_frame": (ctypes.c_int, [_vp, _dp]),
        "fvm2d_download_frame_gray": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_ubyte)]),
        "fvm2d_steps_total": (ctypes.c_int, [_vp]),
        "fvm2d_destroy": (ctypes.c_int, [_vp]),
        "fvm2d_set_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "fvm2d_get_launch": (ctypes.c_int, [_vp, _ip, _ip, _ip]),
        "fvm2d_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "fvm2d_begin": (ctypes.c_int, [_vp]),
        "fvm2d_advance": (ctypes.c_int, [_vp, ctypes.c_int]),
        "fvm2d_sync": (ctypes.c_int, [_vp, _ip, _dp, _dp]),
        "fvm2d_time_steps": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "fvm2d_time_kernel": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
        "fvm2d_batch_create": (ctypes.c_int, [ctypes.c_int, _ip, _ip, P, ctypes.POINTER(_vp)]),
        "fvm2d_batch_upload": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "fvm2d_batch_begin": (ctypes.c_int, [_vp]),
        "fvm2d_batch_run": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "fvm2d_batch_status": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp]),
        "fvm2d_batch_download": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "fvm2d_batch_set_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "fvm2d_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "fvm2d_batch_destroy": (ctypes.c_int, [_vp]),
        "fvm2d_strip_layout": (ctypes.c_int, [ctypes.c_int] * 7 + [ctypes.POINTER(ctypes.c_int),
                                                                 ctypes.c_int,
                                                                 ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    _bind_adjoint(L)
    return L


def _bind_adjoint(L):
    if hasattr(L, "fvm2d_adj_create"):
        from . import adjoint as _adj
        _adj.bind(L)


def check(rc: int, what: str = "") -> None:
    if rc != FVM2D_OK:
        msg = lib().fvm2d_last_error().decode(errors="replace")
        raise FVM2DError(f"{what} failed ({rc}): {msg}")


PROBE_EXPORTS = ["fvm2d_probe_last_error", "fvm2d_hbm_copy_probe", "fvm2d_valu_probe",
                 "fvm2d_recip_probe"]
_probes = None


def probes_lib():
    """Load libfvm2d_probes.so, the measurement probes (include/fvm2d_probes.h)."""
    global _probes
    if _probes is not None:
        return _probes
    if not os.path.exists(PROBES_LIB_PATH):
        raise FVM2DError(f"{PROBES_LIB_PATH} is not built (run 2d_fvm_cuda_amd/build.py)")
    L = ctypes.CDLL(PROBES_LIB_PATH)
    sig = {
        "fvm2d_probe_last_error": (ctypes.c_char_p, []),
        "fvm2d_hbm_copy_probe": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_int, _dp]),
        "fvm2d_valu_probe": (ctypes.c_int, [ctypes.c_int, _dp]),
        "fvm2d_recip_probe": (ctypes.c_int, [_dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _probes = L
    return L


def _probe_check(rc: int, what: str) -> None:
    if rc != 0:
        msg = probes_lib().fvm2d_probe_last_error().decode(errors="replace")
        raise FVM2DError(f"{what} failed ({rc}): {msg}")


def valu_probe(kind: int) -> float:
    """Wave-level instructions per SIMD per second of one VALU kind (0 fma_f64,
    1 mul_f64, 2 add_f64, 3 rcp_f64, 4 32-bit integer ops)."""
    r = ctypes.c_double(0.0)
    _probe_check(probes_lib().fvm2d_valu_probe(int(kind), ctypes.byref(r)), "fvm2d_valu_probe")
    return r.value


def recip_probe():
    """Max relative errors (raw rcp, refined rcp, raw rsq, refined rsq) of
    the device reciprocals the FAST path uses."""
    out = (ctypes.c_double * 4)()
    _probe_check(probes_lib().fvm2d_recip_probe(out), "fvm2d_recip_probe")
    return tuple(out)


def device_available() -> bool:
    return bool(lib().fvm2d_device_available())


def strip_layout(nx: int, ny: int, strip_rows: int, waves_per_block: int = 4, persistent: int = 0,
                 cu_count: int = 256, blocks_per_cu: int = 2) -> np.ndarray:
    """The strips one launch assigns to its waves, in launch position order:
    an (ntasks, 3) array of (column window, first row, end row).  Needs no
    device."""
    n = ctypes.c_int(0)
    check(lib().fvm2d_strip_layout(nx, ny, strip_rows, waves_per_block, persistent, cu_count,
                                   blocks_per_cu, None, 0, ctypes.byref(n)), "fvm2d_strip_layout")
    out = np.zeros((n.value, 3), dtype=np.int32)
    check(lib().fvm2d_strip_layout(nx, ny, strip_rows, waves_per_block, persistent, cu_count,
                                   blocks_per_cu, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   n.value, ctypes.byref(n)), "fvm2d_strip_layout")
    return out


def hbm_copy_probe(n_doubles: int, reps: int) -> float:
    """Mean ms of one plain device copy of n_doubles FP64 values (the HBM
    counter calibration kernel)."""
    ms = ctypes.c_double(0.0)
    _probe_check(probes_lib().fvm2d_hbm_copy_probe(int(n_doubles), int(reps), ctypes.byref(ms)),
                 "fvm2d_hbm_copy_probe")
    return ms.value


def make_params(width=1.0, height=1.0, t_end=0.0, cfl=0.5, alpha=1.9, gamma=1.4,
                strategy=0, max_steps=-1, mode=MODE_FAST) -> Params:
    return Params(float(width), float(height), float(t_end), float(cfl), float(alpha),
                  float(gamma), int(strategy), int(max_steps), int(mode), 0)


def _as_u5(U5) -> np.ndarray:
    a = np.ascontiguousarray(U5, dtype=np.float64)
    if a.ndim != 3 or a.shape[2] != 5:
        raise ValueError("state must have shape [ny][nx][5]")
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def solve(U5, **params):
    """fvm2d_solve on a copy of U5; returns (U_final, steps, ms)."""
    U = _as_u5(U5).copy()
    p = make_params(**params)
    steps = ctypes.c_int(0)
    ms = ctypes.c_double(0.0)
    check(lib().fvm2d_solve(_ptr(U), U.shape[1], U.shape[0], ctypes.byref(p), ctypes.byref(steps),
                            ctypes.byref(ms)), "fvm2d_solve")
    return U, steps.value, ms.value


class Solver:
    """A resident solver (fvm2d_create / upload / run / download)."""

    def __init__(self, nx: int, ny: int, **params):
        self.nx, self.ny = int(nx), int(ny)
        self.p = make_params(**params)
        h = _vp()
        check(lib().fvm2d_create(self.nx, self.ny, ctypes.byref(self.p), ctypes.byref(h)),
              "fvm2d_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fvm2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **params):
        self.p = make_params(**params)
        check(lib().fvm2d_set_params(self.h, ctypes.byref(self.p)), "fvm2d_set_params")

    def upload(self, U5):
        U = _as_u5(U5)
        if U.shape[:2] != (self.ny, self.nx):
            raise ValueError("shape mismatch")
        check(lib().fvm2d_upload(self.h, _ptr(U)), "fvm2d_upload")

    def download(self) -> np.ndarray:
        U = np.empty((self.ny, self.nx, 5))
        check(lib().fvm2d_download(self.h, _ptr(U)), "fvm2d_download")
        return U

    def download_frame(self) -> np.ndarray:
        """Inside a time loop: the state after the last step as the
        reference's record mode copies it (before that step's ghost fill)."""
        U = np.empty((self.ny, self.nx, 5))
        check(lib().fvm2d_download_frame(self.h, _ptr(U)), "fvm2d_download_frame")
        return U

    def download_frame_gray(self) -> np.ndarray:
        """The density of download_frame() as the reference's 8-bit frame
        image (min-max normalised and rounded on the device): (ny, nx) uint8."""
        g = np.empty((self.ny, self.nx), dtype=np.uint8)
        check(lib().fvm2d_download_frame_gray(self.h, g.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))),
              "fvm2d_download_frame_gray")
        return g

    def upload_planes_device(self, ptr: int):
        check(lib().fvm2d_upload_planes_device(self.h, _vp(ptr)), "fvm2d_upload_planes_device")

    def download_planes_device(self, ptr: int):
        check(lib().fvm2d_download_planes_device(self.h, _vp(ptr)), "fvm2d_download_planes_device")

    def run(self):
        steps = ctypes.c_int(0)
        t = ctypes.c_double(0.0)
        ms = ctypes.c_double(0.0)
        check(lib().fvm2d_run(self.h, ctypes.byref(steps), ctypes.byref(t), ctypes.byref(ms)),
              "fvm2d_run")
        return steps.value, t.value, ms.value

    def steps_total(self) -> int:
        return lib().fvm2d_steps_total(self.h)

    def set_launch(self, strip_rows=0, waves_per_block=0, persistent=-1):
        check(lib().fvm2d_set_launch(self.h, strip_rows, waves_per_block, persistent),
              "fvm2d_set_launch")

    def get_launch(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().fvm2d_get_launch(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              "fvm2d_get_launch")
        return a.value, b.value, c.value

    def autotune(self, steps_per_trial=4) -> float:
        ms = ctypes.c_double(0.0)
        check(lib().fvm2d_autotune(self.h, steps_per_trial, ctypes.byref(ms)), "fvm2d_autotune")
        return ms.value

    def begin(self):
        check(lib().fvm2d_begin(self.h), "fvm2d_begin")

    def advance(self, nsteps: int):
        check(lib().fvm2d_advance(self.h, int(nsteps)), "fvm2d_advance")

    def sync(self):
        s, t, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().fvm2d_sync(self.h, ctypes.byref(s), ctypes.byref(t), ctypes.byref(dt)),
              "fvm2d_sync")
        return s.value, t.value, dt.value

    def time_kernel(self, nsteps: int):
        """(avg_ms, min_ms) of single step-kernel launches (HIP events)."""
        a, m = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(lib().fvm2d_time_kernel(self.h, int(nsteps), ctypes.byref(a), ctypes.byref(m)),
              "fvm2d_time_kernel")
        return a.value, m.value

    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double(0.0)
        check(lib().fvm2d_time_steps(self.h, int(nsteps), ctypes.byref(ms)), "fvm2d_time_steps")
        return ms.value


class Batch:
    """Batched forward: many meshes advanced by one launch per step."""

    def __init__(self, shapes, params_list):
        n = len(shapes)
        self.n = n
        self.shapes = [(int(ny), int(nx)) for (ny, nx) in shapes]
        nx = (ctypes.c_int * n)(*[s[1] for s in self.shapes])
        ny = (ctypes.c_int * n)(*[s[0] for s in self.shapes])
        P = (Params * n)(*[make_params(**pp) for pp in params_list])
        h = _vp()
        check(lib().fvm2d_batch_create(n, nx, ny, P, ctypes.byref(h)), "fvm2d_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fvm2d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, i: int, U5):
        U = _as_u5(U5)
        check(lib().fvm2d_batch_upload(self.h, i, _ptr(U)), "fvm2d_batch_upload")

    def begin(self):
        check(lib().fvm2d_batch_begin(self.h), "fvm2d_batch_begin")

    def run(self, nsteps: int) -> float:
        ms = ctypes.c_double(0.0)
        check(lib().fvm2d_batch_run(self.h, int(nsteps), ctypes.byref(ms)), "fvm2d_batch_run")
        return ms.value

    def status(self, i: int):
        s, t = ctypes.c_int(), ctypes.c_double()
        check(lib().fvm2d_batch_status(self.h, i, ctypes.byref(s), ctypes.byref(t)),
              "fvm2d_batch_status")
        return s.value, t.value

    def download(self, i: int) -> np.ndarray:
        ny, nx = self.shapes[i]
        U = np.empty((ny, nx, 5))
        check(lib().fvm2d_batch_download(self.h, i, _ptr(U)), "fvm2d_batch_download")
        return U

    def set_launch(self, strip_rows=0, waves_per_block=0, persistent=-1):
        check(lib().fvm2d_batch_set_launch(self.h, strip_rows, waves_per_block, persistent),
              "fvm2d_batch_set_launch")

    def autotune(self, steps_per_trial=4) -> float:
        ms = ctypes.c_double(0.0)
        check(lib().fvm2d_batch_autotune(self.h, steps_per_trial, ctypes.byref(ms)),
              "fvm2d_batch_autotune")
        return ms.value


class RGRP_2D_CUDA:
    """The reference's solver class (src/wrappers.cuh:26-65) over libfvm2d.

    Same methods, same defaults (alpha 1.9, gamma 1.4, CFL 0.5, width =
    height = 1, outflow ghosts; src/wrappers.cuh:49-52).  ``solve()``
    continues from the previous result and accumulates the step count, as the
    reference does (src/wrappers.cuh:290-291, 252)."""

    def __init__(self, U0=None):
        self._width, self._height, self._endTime, self._CFL = 1.0, 1.0, 0.0, 0.5
        self._alpha, self._gamma = 1.9, 1.4
        self._Nx, self._Ny = 100, 100
        self._strategy = ghostCellStrategy.outflow
        self._record = False
        self._gap = 1
        self._recordDir = "."
        self._ms = 0.0
        self._steps = 0
        self._mode = MODE_FAST
        self._U0 = None
        self._U = None
        if U0 is not None:
            U = _as_u5(U0)
            self._U0 = U.copy()
            self._Ny, self._Nx = U.shape[0], U.shape[1]

    def setRPInitialData(self, U1, U2, U3, U4, Nx, Ny):
        self._U0 = problems.set_rp_initial_data(U1, U2, U3, U4, Nx, Ny)
        self._Nx, self._Ny = Nx, Ny

    def setSpatialLayout(self, width, height):
        self._width, self._height = float(width), float(height)

    def setTimeLayout(self, endTime, CFL=None):
        self._endTime = float(endTime)
        if CFL is not None:
            self._CFL = float(CFL)

    def setCFL(self, CFL):
        self._CFL = float(CFL)

    def setAlpha(self, alpha):
        self._alpha = float(alpha)

    def setGamma(self, gamma):
        self._gamma = float(gamma)

    def setGhostCellStrategy(self, strategy):
        self._strategy = ghostCellStrategy(int(strategy))

    def setRecordModel(self, on_off, gapPerFrame=10, recordDir="."):
        self._record, self._gap, self._recordDir = bool(on_off), int(gapPerFrame), str(recordDir)

    def setArithmeticMode(self, mode):
        """Extension: MODE_FAST (default) or MODE_EXACT (reference op order)."""
        self._mode = int(mode)

    def solve(self):
        if self._U0 is None:
            raise FVM2DError("no initial data")
        if self._record:
            return self._solve_recording()
        U, steps, ms = solve(self._U0, width=self._width, height=self._height,
                             t_end=self._endTime, cfl=self._CFL, alpha=self._alpha,
                             gamma=self._gamma, strategy=int(self._strategy), mode=self._mode)
        self._U0 = U
        self._U = U
        self._steps += steps
        self._ms = ms
        return U

    def _solve_recording(self):
        """Record mode (src/wrappers.cuh:253-261): a frame of the density
        after the first step and after every gapPerFrame further steps,
        named by the index of the step (0, gap, 2 gap, ...).  The loop runs
        on the device in chunks of gapPerFrame steps; each frame is the
        state as the reference copies it out, after the branch average and
        before that step's ghost fill; its density is normalised to the
        8-bit image on the device (fvm2d_download_frame_gray).  Frames are
        coloured, encoded and written by a pool of host threads while the
        device steps on; like the reference's, the computation time
        includes writing them (the call returns after the last one is
        written)."""
        import time
        from concurrent.futures import ThreadPoolExecutor
        from .io import write_gray_png
        gap = max(1, self._gap)
        workers = max(1, min(16, os.cpu_count() or 1))
        s = Solver(self._Nx, self._Ny, width=self._width, height=self._height,
                   t_end=self._endTime, cfl=self._CFL, alpha=self._alpha, gamma=self._gamma,
                   strategy=int(self._strategy), mode=self._mode)
        try:
            s.upload(self._U0)
            with ThreadPoolExecutor(max_workers=workers) as pool:
                pending = []
                t0 = time.perf_counter()
                s.begin()
                done_steps, due = 0, 1
                while True:
                    s.advance(due - done_steps)
                    steps, _, _ = s.sync()
                    if steps < due:
                        break
                    done_steps = steps
                    gray = s.download_frame_gray()
                    pending.append(pool.submit(write_gray_png, self._recordDir, gray, due - 1))
                    if len(pending) >= 2 * workers:  # bound the frames held in memory
                        pending.pop(0).result()
                    due += gap
                for f in pending:
                    f.result()
                ms = (time.perf_counter() - t0) * 1e3
            U = s.download()
        finally:
            s.close()
        self._U0 = U
        self._U = U
        self._steps += steps
        self._ms = ms
        return U

    def getComputationTime_ms(self):
        return self._ms

    def getTimeStagesCount(self):
        return self._steps

    def writeFinalResultTo_txt(self, dir):
        """src/wrappers.cuh:318-321,370-397: rho_/p_/u_/v_<t_end>s.txt with
        one row per mesh row, values separated by a space, default ostream
        precision (6 significant digits)."""
        from .io import write_matrix_txt
        write_matrix_txt(dir, self._U, self._endTime)

    def clear(self):
        self._U0 = None
        self._U = None
