"""Writes fourquad_2000.json (+ fourquad_2000_sub16.npy): the oracle's
2000x2000 four-quadrant solve to t_end = 0.3 (alpha 1.9, CFL 0.5, gamma 1.4,
outflow): step count, final time, SHA-256 of the final (rho, p, u, v) planes
in [ny][nx][4] float64 order, per-variable sums, and every 16th cell of the
final state.  About 1.2e10 cell-updates: ~20 minutes on 8 cores.
    python tests/golden/make_fourquad_2000.py
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

U0 = problems.initialize_4shocks(2000)
t0 = time.time()
U, steps, t, ms = oracle.solve(U0, nthreads=0, t_end=0.3)
P4 = np.ascontiguousarray(U[..., :4])
out = {
    "workload": "four-quadrant 2000x2000, t_end 0.3, CFL 0.5, alpha 1.9, gamma 1.4, outflow",
    "steps": steps,
    "t_final": t,
    "sha256_prim4": hashlib.sha256(P4.tobytes()).hexdigest(),
    "sums_prim4": [float(x) for x in P4.reshape(-1, 4).sum(axis=0)],
    "oracle_ms": ms,
    "oracle_threads": os.cpu_count(),
}
np.save(os.path.join(HERE, "fourquad_2000_sub16.npy"), np.ascontiguousarray(P4[::16, ::16]))
json.dump(out, open(os.path.join(HERE, "fourquad_2000.json"), "w"), indent=1)
print(json.dumps(out, indent=1), "wall", time.time() - t0)
