"""Writes fourquad_8192.json (+ fourquad_8192_sub64.npy): the oracle's
8192x8192 four-quadrant solve (initialize_4shocks(8192): the 2000^2 data
scaled; CFL 0.5, alpha 1.9, gamma 1.4, outflow) for 3 steps: step count,
final time, SHA-256 of the final (rho, p, u, v) planes in [ny][nx][4]
float64 order and every 64th cell.  The oracle keeps the reference's
buffers (about 36 GB at this size); ~1 minute on 8 cores.
    python tests/golden/make_fourquad_8192.py
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
import oracle  # noqa: E402

problems = importlib.import_module("2d_fvm_cuda_amd.problems")
HERE = os.path.dirname(os.path.abspath(__file__))
N, STEPS, STRIDE = 8192, 3, 64

t0 = time.time()
U0 = problems.initialize_4shocks(N)
U, steps, t, ms = oracle.solve(U0, nthreads=0, t_end=1e30, max_steps=STEPS)
del U0
P4 = np.ascontiguousarray(U[..., :4])
del U
out = {"workload": f"four-quadrant {N}x{N}, {STEPS} steps, CFL 0.5, alpha 1.9, gamma 1.4, outflow",
       "steps": steps, "t_final": t, "stride": STRIDE,
       "sha256_prim4": hashlib.sha256(P4.tobytes()).hexdigest(), "oracle_ms": ms}
np.save(os.path.join(HERE, "fourquad_8192_sub64.npy"), np.ascontiguousarray(P4[::STRIDE, ::STRIDE]))
json.dump(out, open(os.path.join(HERE, "fourquad_8192.json"), "w"), indent=1)
print(json.dumps(out, indent=1), "wall", time.time() - t0)
