"""Writes fourquad_256.json: the oracle's 256x256 four-quadrant solve to
t_end = 0.3 (alpha 1.9, CFL 0.5, gamma 1.4, outflow) as a regression digest
(step count, SHA-256 of the final (rho, p, u, v) planes in [ny][nx][4]
float64 order, and per-variable sums).  Run from the repository root:
    python tests/golden/make_fourquad_256.py
"""
import hashlib
import importlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

problems = importlib.import_module("2d_fvm_cuda_amd.problems")

U0 = problems.initialize_4shocks(256)
U, steps, t, ms = oracle.solve(U0, nthreads=0, t_end=0.3)
P4 = np.ascontiguousarray(U[..., :4])
out = {
    "workload": "four-quadrant 256x256, t_end 0.3, CFL 0.5, alpha 1.9, gamma 1.4, outflow",
    "steps": steps,
    "t_final": t,
    "sha256_prim4": hashlib.sha256(P4.tobytes()).hexdigest(),
    "sums_prim4": [float(x) for x in P4.reshape(-1, 4).sum(axis=0)],
}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fourquad_256.json")
json.dump(out, open(path, "w"), indent=1)
print(json.dumps(out, indent=1))
