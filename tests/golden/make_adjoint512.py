"""Writes adjoint512.json (+ adjoint512_sub8.npz): the host reference
gradient (oracle/fvm2d_oracle_adjoint.cpp) of L = <gN, U_400> on the 512^2
adjoint workload made tie-free -- problems.fourier_perturbed(512, 512,
seed=1) plus 8 seeded random-phase modes of amplitude <= 0.02 on each of u
and v, so that no limiter starts on its kink -- with gN ~ N(0, 1) (seed 3):
400 steps, dL/dU0 in conserved planes.  Stored: every 8th cell of the
gradient and of the final state U_400 in both directions
(adjoint512_sub8.npz), the gradient's maximum magnitude and the final
state's SHA-256.  About 8 minutes on one core.
    python tests/golden/make_adjoint512.py
"""
import hashlib
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

problems = importlib.import_module("2d_fvm_cuda_amd.problems")
HERE = os.path.dirname(os.path.abspath(__file__))
N, STEPS, STRIDE = 512, 400, 8


def workload():
    """The tie-free 512^2 state (conserved planes) and the output weights."""
    U5 = problems.fourier_perturbed(N, N, seed=1)
    rng = np.random.default_rng(123)
    x = (np.arange(N) + 0.5) / N
    X, Y = np.meshgrid(x, x)
    for c in (2, 3):
        for _ in range(8):
            kx, ky = rng.integers(1, 9, size=2)
            U5[..., c] += rng.uniform(0, 0.02) * np.sin(2 * np.pi * (kx * X + ky * Y) + rng.uniform(0, 2 * np.pi))
    r, p, u, v = U5[..., 0], U5[..., 1], U5[..., 2], U5[..., 3]
    C0 = np.ascontiguousarray(np.stack([r, r * u, r * v, p / 0.4 + 0.5 * r * (u * u + v * v)]))
    gN = np.random.default_rng(3).normal(size=C0.shape)
    return C0, gN


if __name__ == "__main__":
    import oracle

    C0, gN = workload()
    t0 = time.time()
    UN = oracle.adj_forward(C0, STEPS)
    g0 = oracle.adj_backward(C0, STEPS, gN)
    out = {"workload": f"tie-free fourier_perturbed {N}x{N}, {STEPS} steps, CFL 0.5, alpha 1.9, gamma 1.4, "
                       "outflow; L = <gN, U_N>, gN ~ N(0,1) seed 3",
           "steps": STEPS, "stride": STRIDE, "grad_max_abs": float(np.abs(g0).max()),
           "grad_l1": float(np.abs(g0).sum()),
           "sha256_UN": hashlib.sha256(np.ascontiguousarray(UN).tobytes()).hexdigest(),
           "host_s": time.time() - t0}
    np.savez(os.path.join(HERE, "adjoint512_sub8.npz"), grad=np.ascontiguousarray(g0[:, ::STRIDE, ::STRIDE]),
             UN=np.ascontiguousarray(UN[:, ::STRIDE, ::STRIDE]))
    json.dump(out, open(os.path.join(HERE, "adjoint512.json"), "w"), indent=1)
    print(json.dumps(out, indent=1))
