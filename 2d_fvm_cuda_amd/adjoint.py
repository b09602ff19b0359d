"""Differentiable solve: the hand-written adjoint of libfvm2d behind
``torch.autograd.Function``.

    adj = Adjoint([(ny, nx)], [params], nsteps=400, checkpoint_slots=32)
    U_N = FVMSolve.apply(adj, U0)          # U0: float64 tensor [4, ny, nx]
    loss = 0.5 * ((U_N[0] - target) ** 2).sum()
    loss.backward()                        # U0.grad = dL/dU0

States are conserved planes (rho, rho u, rho v, E) in float64, row 0 =
north, ghost layers included.  CPU tensors (and numpy arrays) cross the
boundary in host memory and the library stages them into HBM; CUDA tensors
cross as device pointers (fvm2d_adj_forward / fvm2d_adj_backward) with no
host copy.  That needs one HIP runtime in the process: import torch before
this package loads libfvm2d, so that libfvm2d binds to torch's runtime
(see ``hip_runtimes``).  The forward runs ``nsteps`` steps with
dt = CFL h / s_max each (no t_end clamp) and stores binomial checkpoints; the
backward recomputes from them and runs the reverse of every step.  Several
instances (a batch) advance together: ``FVMSolve.apply(adj, U0_a, U0_b, ...)``
returns one tensor per instance.
"""
from __future__ import annotations

import ctypes

_vp = ctypes.c_void_p


class AdjStats(ctypes.Structure):
    _fields_ = [
        ("forward_ms", ctypes.c_double),
        ("backward_ms", ctypes.c_double),
        ("forward_step_ms", ctypes.c_double),
        ("reverse_step_ms", ctypes.c_double),
        ("reverse_steps", ctypes.c_int),
        ("recompute_steps", ctypes.c_int),
        ("checkpoint_slots", ctypes.c_int),
        ("schedule_cost", ctypes.c_int),
    ]


EXPORTS = ["fvm2d_adj_create", "fvm2d_adj_forward", "fvm2d_adj_backward", "fvm2d_adj_forward_host",
           "fvm2d_adj_backward_host", "fvm2d_adj_get_stats", "fvm2d_adj_set_launch", "fvm2d_adj_destroy"]


def bind(L):
    from . import Params
    P = ctypes.POINTER(Params)
    ip = ctypes.POINTER(ctypes.c_int)
    pp = ctypes.POINTER(_vp)
    L.fvm2d_adj_create.argtypes = [ctypes.c_int, ip, ip, P, ctypes.c_int, ctypes.c_int, pp]
    L.fvm2d_adj_create.restype = ctypes.c_int
    L.fvm2d_adj_forward.argtypes = [_vp, pp, pp]
    L.fvm2d_adj_forward.restype = ctypes.c_int
    L.fvm2d_adj_backward.argtypes = [_vp, pp, pp]
    L.fvm2d_adj_backward.restype = ctypes.c_int
    L.fvm2d_adj_forward_host.argtypes = [_vp, pp, pp]
    L.fvm2d_adj_forward_host.restype = ctypes.c_int
    L.fvm2d_adj_backward_host.argtypes = [_vp, pp, pp]
    L.fvm2d_adj_backward_host.restype = ctypes.c_int
    L.fvm2d_adj_get_stats.argtypes = [_vp, ctypes.POINTER(AdjStats)]
    L.fvm2d_adj_get_stats.restype = ctypes.c_int
    L.fvm2d_adj_set_launch.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.fvm2d_adj_set_launch.restype = ctypes.c_int
    L.fvm2d_adj_destroy.argtypes = [_vp]
    L.fvm2d_adj_destroy.restype = ctypes.c_int


def _ptrs(arrays):
    arr = (_vp * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = a.ctypes.data
    return arr


def _as_host(x):
    """numpy float64 C-contiguous view/copy of a numpy array or CPU tensor."""
    import numpy as np
    if hasattr(x, "detach"):
        if x.device.type != "cpu":
            raise ValueError("pass every state of a call on the same side: all CUDA tensors "
                             "(device pointers) or all host arrays / CPU tensors")
        x = x.detach().numpy()
    return np.ascontiguousarray(x, dtype=np.float64)


class Adjoint:
    """A forward/adjoint pair for one or more meshes (fvm2d_adj_*)."""

    def __init__(self, shapes, params_list, nsteps: int, checkpoint_slots: int = 0):
        from . import check, lib, make_params, Params
        self._lib = lib()
        n = len(shapes)
        self.n = n
        self.shapes = [(int(a), int(b)) for a, b in shapes]
        self.nsteps = int(nsteps)
        nx = (ctypes.c_int * n)(*[s[1] for s in self.shapes])
        ny = (ctypes.c_int * n)(*[s[0] for s in self.shapes])
        P = (Params * n)(*[make_params(**pp) for pp in params_list])
        h = _vp()
        check(self._lib.fvm2d_adj_create(n, nx, ny, P, self.nsteps, int(checkpoint_slots),
                                         ctypes.byref(h)), "fvm2d_adj_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._lib.fvm2d_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arrays(self, xs):
        assert len(xs) == self.n, "one state per instance"
        out = []
        for x, (ny, nx) in zip(xs, self.shapes):
            a = _as_host(x)
            if a.shape != (4, ny, nx):
                raise ValueError(f"expected shape (4, {ny}, {nx}), got {a.shape}")
            out.append(a)
        return out

    def _call(self, fn, xs, what):
        import numpy as np
        from . import check
        ins = self._arrays(xs)
        outs = [np.empty_like(a) for a in ins]
        check(fn(self.h, _ptrs(ins), _ptrs(outs)), what)
        return outs

    def _device_call(self, fn, xs, what):
        """Device-pointer call on CUDA tensors; returns CUDA tensors."""
        import torch
        from . import check, hip_runtimes
        if len(hip_runtimes()) != 1:
            raise RuntimeError("CUDA tensors need a single HIP runtime in the process: import "
                               "torch before libfvm2d is loaded")
        assert len(xs) == self.n, "one state per instance"
        ins = []
        for x, (ny, nx) in zip(xs, self.shapes):
            if tuple(x.shape) != (4, ny, nx):
                raise ValueError(f"expected shape (4, {ny}, {nx}), got {tuple(x.shape)}")
            ins.append(x.detach().to(torch.float64).contiguous())
        outs = [torch.empty_like(a) for a in ins]
        arr_in, arr_out = (_vp * self.n)(), (_vp * self.n)()
        for i in range(self.n):
            arr_in[i], arr_out[i] = ins[i].data_ptr(), outs[i].data_ptr()
        torch.cuda.synchronize(ins[0].device)  # inputs written by torch's streams
        check(fn(self.h, arr_in, arr_out), what)  # returns after its own stream is done
        return outs

    @staticmethod
    def _on_device(xs):
        return all(hasattr(x, "is_cuda") and x.is_cuda for x in xs)

    def forward(self, U0s):
        """Conserved planes U_N of every instance after nsteps steps: CUDA
        tensors for CUDA inputs, numpy arrays otherwise."""
        if self._on_device(U0s):
            return self._device_call(self._lib.fvm2d_adj_forward, U0s, "fvm2d_adj_forward")
        return self._call(self._lib.fvm2d_adj_forward_host, U0s, "fvm2d_adj_forward_host")

    def backward(self, gNs):
        """dL/dU_0 of every instance given dL/dU_N, for the last forward
        (CUDA tensors for CUDA inputs, numpy arrays otherwise)."""
        if self._on_device(gNs):
            return self._device_call(self._lib.fvm2d_adj_backward, gNs, "fvm2d_adj_backward")
        return self._call(self._lib.fvm2d_adj_backward_host, gNs, "fvm2d_adj_backward_host")

    def stats(self) -> dict:
        from . import check
        st = AdjStats()
        check(self._lib.fvm2d_adj_get_stats(self.h, ctypes.byref(st)), "fvm2d_adj_get_stats")
        return {k: getattr(st, k) for k, _ in AdjStats._fields_}

    def set_launch(self, strip_rows=0, waves_per_block=0, persistent=-1):
        from . import check
        check(self._lib.fvm2d_adj_set_launch(self.h, strip_rows, waves_per_block, persistent),
              "fvm2d_adj_set_launch")


def _function():
    import torch

    class FVMSolve(torch.autograd.Function):
        """U_N = solve(U_0) for every instance of ``adj``; backward is the
        hand-written adjoint with binomial checkpointing."""

        @staticmethod
        def forward(ctx, adj, *U0s):
            ctx.adj = adj
            outs = adj.forward(list(U0s))
            outs = [o if torch.is_tensor(o) else torch.from_numpy(o) for o in outs]
            return outs[0] if len(outs) == 1 else tuple(outs)

        @staticmethod
        def backward(ctx, *gNs):
            gs = ctx.adj.backward([g.contiguous() for g in gNs])
            return (None, *[g if torch.is_tensor(g) else torch.from_numpy(g) for g in gs])

    return FVMSolve


_cache = {}


def __getattr__(name):
    """``FVMSolve`` is created on first access so that importing this module
    (which the library loader does) does not import torch."""
    if name == "FVMSolve":
        if name not in _cache:
            _cache[name] = _function()
        return _cache[name]
    raise AttributeError(name)
