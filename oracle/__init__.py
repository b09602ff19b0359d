"""ctypes loader for the FP64 host oracle (liboracle_fvm2d.so).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it as the checker.  The product
library never imports or links it.

The C++ restatement follows the reference's ``src/basicFuns.cuh``,
``src/ghostCellPresets.cuh`` and ``src/wrappers.cuh:154-306`` statement for
statement; see ``fvm2d_oracle.cpp`` for the line citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_fvm2d.so")


class OrcParams(ctypes.Structure):
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


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -O2 -ffp-contract=off -fopenmp)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE],
                       check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_solve.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcParams),
                                ctypes.c_int, ip, dp, dp]
        L.orc_solve.restype = ctypes.c_int
        L.orc_solve_record.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcParams),
                                       ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, ip, ip, dp]
        L.orc_solve_record.restype = ctypes.c_int
        L.orc_slope.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.orc_grp.argtypes = [dp, dp, dp, dp, ctypes.c_double, dp, dp, ip]
        L.orc_flux.argtypes = [dp, dp, ctypes.c_double, dp]
        L.orc_cons.argtypes = [dp, ctypes.c_double, dp]
        L.orc_prim.argtypes = [dp, ctypes.c_double, dp]
        L.orc_speed.argtypes = [dp, ctypes.c_double]
        L.orc_speed.restype = ctypes.c_double
        L.orc_fill.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double]
        L.orc_sweep_x.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_rotate.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int]
        L.orc_adj_forward.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcParams),
                                      ctypes.c_int, dp]
        L.orc_adj_backward.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcParams),
                                       ctypes.c_int, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params(width=1.0, height=1.0, t_end=0.0, cfl=0.5, alpha=1.9, gamma=1.4, strategy=0,
           max_steps=-1) -> OrcParams:
    return OrcParams(width, height, t_end, cfl, alpha, gamma, strategy, max_steps, 0, 0)


def solve(U5: np.ndarray, nthreads: int = 0, **kw):
    """Run the oracle solve on a copy of ``U5`` ([ny][nx][5]).

    Returns (U_final, steps, t_final, ms)."""
    U = np.ascontiguousarray(U5, dtype=np.float64).copy()
    ny, nx = U.shape[0], U.shape[1]
    p = params(**kw)
    steps = ctypes.c_int(0)
    t = ctypes.c_double(0.0)
    ms = ctypes.c_double(0.0)
    rc = lib().orc_solve(_dp(U), nx, ny, ctypes.byref(p), int(nthreads), ctypes.byref(steps),
                         ctypes.byref(t), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"orc_solve failed ({rc})")
    return U, steps.value, t.value, ms.value


def solve_record(U5: np.ndarray, gap: int, max_frames: int = 64, nthreads: int = 0, **kw):
    """Record-mode solve (src/wrappers.cuh:253-261) on a copy of ``U5``.

    Returns (U_final, steps, t_final, frames, indices): frames[k] is the
    [ny][nx][5] state copied out for the frame with file index indices[k]
    (after the average, before the ghost fill of that step)."""
    U = np.ascontiguousarray(U5, dtype=np.float64).copy()
    ny, nx = U.shape[0], U.shape[1]
    frames = np.zeros((max_frames, ny, nx, 5))
    idx = np.zeros(max_frames, dtype=np.int32)
    nf, steps = ctypes.c_int(0), ctypes.c_int(0)
    t = ctypes.c_double(0.0)
    p = params(**kw)
    rc = lib().orc_solve_record(_dp(U), nx, ny, ctypes.byref(p), int(nthreads), int(gap), _dp(frames),
                                int(max_frames), ctypes.byref(nf),
                                idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(steps),
                                ctypes.byref(t))
    if rc != 0:
        raise RuntimeError(f"orc_solve_record failed ({rc})")
    k = min(nf.value, max_frames)
    return U, steps.value, t.value, frames[:k], idx[:k].tolist()


def slope(qL, q, qR, h, alpha):
    out = np.zeros(4)
    lib().orc_slope(_dp(np.ascontiguousarray(qL, np.float64)), _dp(np.ascontiguousarray(q, np.float64)),
                    _dp(np.ascontiguousarray(qR, np.float64)), h, alpha, _dp(out))
    return out


def grp(fl, fr, sl, sr, gamma):
    W = np.zeros(5)
    Wt = np.zeros(5)
    br = ctypes.c_int(-1)
    a = [np.ascontiguousarray(x, np.float64) for x in (fl, fr, sl, sr)]
    lib().orc_grp(_dp(a[0]), _dp(a[1]), _dp(a[2]), _dp(a[3]), gamma, _dp(W), _dp(Wt),
                  ctypes.byref(br))
    return W, Wt, br.value


def flux(W, Wt, k):
    F = np.zeros(4)
    lib().orc_flux(_dp(np.ascontiguousarray(W, np.float64)), _dp(np.ascontiguousarray(Wt, np.float64)),
                   k, _dp(F))
    return F


def cons(q, gamma):
    out = np.zeros(4)
    lib().orc_cons(_dp(np.ascontiguousarray(q, np.float64)), gamma, _dp(out))
    return out


def prim(c, gamma):
    out = np.zeros(4)
    lib().orc_prim(_dp(np.ascontiguousarray(c, np.float64)), gamma, _dp(out))
    return out


def speed(q, gamma):
    return lib().orc_speed(_dp(np.ascontiguousarray(q, np.float64)), gamma)


def fill(U5: np.ndarray, strategy: int, t: float, width=1.0, height=1.0) -> np.ndarray:
    U = np.ascontiguousarray(U5, dtype=np.float64).copy()
    lib().orc_fill(_dp(U), U.shape[1], U.shape[0], strategy, width, height, t)
    return U


def sweep_x(U5: np.ndarray, width, dt, alpha, gamma, nthreads=1) -> np.ndarray:
    U = np.ascontiguousarray(U5, dtype=np.float64).copy()
    lib().orc_sweep_x(_dp(U), U.shape[1], U.shape[0], width, dt, alpha, gamma, nthreads)
    return U


def rotate(U5: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U5, dtype=np.float64)
    ny, nx = U.shape[0], U.shape[1]
    out = np.zeros((nx, ny, 5))
    lib().orc_rotate(_dp(U), _dp(out), nx, ny)
    return out


def adj_forward(U0cons: np.ndarray, nsteps: int, **kw) -> np.ndarray:
    """Host forward of the differentiated solve: [4][ny][nx] conserved planes."""
    U = np.ascontiguousarray(U0cons, dtype=np.float64)
    out = np.zeros_like(U)
    p = params(**kw)
    lib().orc_adj_forward(_dp(U), U.shape[2], U.shape[1], ctypes.byref(p), int(nsteps), _dp(out))
    return out


def adj_backward(U0cons: np.ndarray, nsteps: int, gN: np.ndarray, **kw) -> np.ndarray:
    """Host reference gradient dL/dU0 given dL/dU_N (conserved planes)."""
    U = np.ascontiguousarray(U0cons, dtype=np.float64)
    g = np.ascontiguousarray(gN, dtype=np.float64)
    out = np.zeros_like(U)
    p = params(**kw)
    lib().orc_adj_backward(_dp(U), U.shape[2], U.shape[1], ctypes.byref(p), int(nsteps), _dp(g), _dp(out))
    return out
