"""Benchmark legs beyond the plain 2000^2 solve (used by bench.py).

dropin()           the drop-in entry point fvm2d_solve (src/wrappers.cuh:154-306)
                   on the headline workload with its default launch geometry,
                   against the resident, autotuned solver's time per step.
batched()          64 seeded four-quadrant meshes, sides ~U{128..1024}, 200 steps
                   all in one launch per step (hipGraph-captured),
                   against the same meshes solved one after another.
adjoint()          512^2 Fourier-perturbed state, 400 steps forward with
                   binomial checkpointing, then the reverse sweep of the misfit
                   gradient; reports the reverse-step / forward-step time ratio
                   and central differences (h = 1e-6) on 50 seeded cells.
batched_adjoint()  32 meshes with sides ~U{256..768} (the adjoint recipe under
                   seeds 1..32), 300 steps forward + backward in one batch;
                   recipe="riemann": the batched-forward recipe instead (seeded
                   four-quadrant Riemann data per mesh, sides ~U{256..768}).
peak()             8192^2 plain forward (2.1 GB of state), 20 timed steps.
published()        the runs whose wall-clock times the reference publishes
                   (README.md:2, test2D.cu:137-144, its figure's panels (a)
                   and (c)), solved by the drop-in entry point.
published_record() the demo main as published, record mode included
                   (test2D.cu:134-152: a viridis PNG frame every 18 steps,
                   then writeFinalResultTo_txt), through the Python mirror
                   of the reference's class; files in a temporary directory.
Cell-update counts use the full meshes (ghost layers included).
"""
from __future__ import annotations

import os
import shutil
import tempfile
import time

import numpy as np

BYTES_PER_CELL_UPDATE = 288
HBM_PEAK_GBS = 8000.0
G = 1.4


def _cons(U5):
    r, p, u, v = U5[..., 0], U5[..., 1], U5[..., 2], U5[..., 3]
    return np.ascontiguousarray(np.stack([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)]))


def batched(fvm, plain_rate=None, count=64, steps=200, seed=2024):
    Us = fvm.problems.batched_instances(count=count, seed=seed)
    shapes = [U.shape[:2] for U in Us]
    cells = sum(ny * nx for ny, nx in shapes)
    prm = [dict(t_end=1e30, max_steps=steps) for _ in Us]
    b = fvm.Batch(shapes, prm)
    for i, U in enumerate(Us):
        b.upload(i, U)
    tuned = b.autotune(steps_per_trial=3)
    for i, U in enumerate(Us):
        b.upload(i, U)
    b.begin()
    ms = b.run(steps)
    done = [b.status(i)[0] for i in range(count)]
    b.close()
    rate = cells * steps / (ms / 1000.0)
    # the same meshes one after another, each on its own resident solver
    seq_ms = 0.0
    for U in Us:
        s = fvm.Solver(U.shape[1], U.shape[0], t_end=1e30, max_steps=steps)
        s.upload(U)
        n, _, m = s.run()
        seq_ms += m
        s.close()
    seq_rate = cells * steps / (seq_ms / 1000.0)
    return {"instances": count, "steps": steps, "cells": cells, "ms": ms,
            "cell_updates_per_s": rate, "sequential_ms": seq_ms,
            "sequential_cell_updates_per_s": seq_rate, "speedup_vs_sequential": rate / seq_rate,
            "all_instances_stepped": all(d == steps for d in done), "autotune_ms_per_step": tuned}


def dropin(fvm, n=2000, steps=-1, mode=0, resident_ms_per_step=None):
    """fvm2d_solve (create, upload, time loop, download, destroy) on the
    four-quadrant n x n workload with the library's default launch geometry;
    ms_per_step is the call's own device time-loop figure."""
    U0 = fvm.problems.initialize_4shocks(n)
    t0 = time.perf_counter()
    _, nsteps, ms = fvm.solve(U0, t_end=0.3, cfl=0.5, alpha=1.9, gamma=1.4, strategy=0, mode=mode,
                              max_steps=steps)
    wall = time.perf_counter() - t0
    rec = {"mesh": [n, n], "steps": nsteps, "loop_ms": ms, "ms_per_step": ms / max(nsteps, 1),
           "cell_updates_per_s": n * n * nsteps / (ms / 1000.0), "call_wall_s": wall}
    if resident_ms_per_step:
        rec["over_resident_autotuned"] = rec["ms_per_step"] / resident_ms_per_step
    return rec


def _velocity_modes(U5s, n, seed=123, modes=8, amp=0.02):
    """Seeded random-phase modes (k ~ U{1..8}^2, amplitude ~ U[0, amp]) added
    to u and v of each state in place: no limiter starts on its kink."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    for U in U5s:
        for c in (2, 3):
            for _ in range(modes):
                kx, ky = rng.integers(1, 9, size=2)
                U[..., c] += rng.uniform(0, amp) * np.sin(2 * np.pi * (kx * X + ky * Y) +
                                                          rng.uniform(0, 2 * np.pi))


def _misfit_gradient(adj, C0, target):
    """L = 1/2 sum (rho_N - target)^2 and dL/dU_0."""
    UN = adj.forward([C0])[0]
    loss = 0.5 * float(((UN[0] - target) ** 2).sum())
    gN = np.zeros_like(UN)
    gN[0] = UN[0] - target
    return loss, adj.backward([gN])[0]


def _fd_devs(adj, C0, target, g, cells, h=1e-6, seed=11):
    """Central differences of the misfit at ``cells`` seeded (component, cell)
    pairs against the gradient, relative to max(|FD|, 1e-6 max|g|)."""
    n = C0.shape[1]
    rng = np.random.default_rng(seed)
    devs = []
    for _ in range(cells):
        k, j, i = rng.integers(0, 4), rng.integers(0, n), rng.integers(0, n)
        Cp, Cm = C0.copy(), C0.copy()
        Cp[k, j, i] += h
        Cm[k, j, i] -= h
        a = adj.forward([Cp])[0][0]
        b = adj.forward([Cm])[0][0]
        fd = float((0.5 * (a - b) * (a + b - 2 * target)).sum()) / (2 * h)
        devs.append(abs(fd - g[k, j, i]) / max(abs(fd), 1e-6 * np.abs(g).max()))
    return np.array(devs) if devs else np.zeros(1)


def adjoint(fvm, n=512, steps=400, slots=32, fd_cells=50):
    """Misfit gradient through the adjoint C ABI (host-pointer entry points;
    the torch.autograd wrapper over the same calls is exercised by the tests),
    timed on the specified data; central differences at ``fd_cells`` seeded
    cells on the specified data (u = v = 0.1 exactly: limiters on their
    kinks) and on the same data with random-phase velocity modes added."""
    import importlib
    adjmod = importlib.import_module("2d_fvm_cuda_amd.adjoint")
    C0 = _cons(fvm.problems.fourier_perturbed(n, n, seed=1))
    Ct = _cons(fvm.problems.fourier_perturbed(n, n, seed=2))
    adj = adjmod.Adjoint([(n, n)], [{}], nsteps=steps, checkpoint_slots=slots)
    target = adj.forward([Ct])[0][0]
    t0 = time.time()
    loss, g = _misfit_gradient(adj, C0, target)
    wall = time.time() - t0
    st = adj.stats()
    devs = _fd_devs(adj, C0, target, g, fd_cells)
    devs_v = np.zeros(1)
    if fd_cells > 0:
        U5s = [fvm.problems.fourier_perturbed(n, n, seed=1), fvm.problems.fourier_perturbed(n, n, seed=2)]
        _velocity_modes(U5s, n)
        C0v, Ctv = _cons(U5s[0]), _cons(U5s[1])
        target_v = adj.forward([Ctv])[0][0]
        _, g_v = _misfit_gradient(adj, C0v, target_v)
        devs_v = _fd_devs(adj, C0v, target_v, g_v, fd_cells)
    adj.close()
    return {"mesh": [n, n], "steps": steps, "checkpoint_slots": slots, "loss": loss,
            "forward_ms": st["forward_ms"], "backward_ms": st["backward_ms"],
            "forward_step_ms": st["forward_step_ms"], "reverse_step_ms": st["reverse_step_ms"],
            "reverse_over_forward_step": st["reverse_step_ms"] / st["forward_step_ms"],
            "gradient_over_forward_total": (st["forward_ms"] + st["backward_ms"]) / st["forward_ms"],
            "recompute_steps": st["recompute_steps"], "schedule_cost": st["schedule_cost"],
            "fd_cells": fd_cells, "fd_h": 1e-6, "fd_rel_dev_median": float(np.median(devs)),
            "fd_rel_dev_p90": float(np.percentile(devs, 90)), "fd_rel_dev_max": float(devs.max()),
            "tie_free_fd_rel_dev_median": float(np.median(devs_v)),
            "tie_free_fd_rel_dev_p90": float(np.percentile(devs_v, 90)),
            "tie_free_fd_rel_dev_max": float(devs_v.max()), "wall_s": wall}


def batched_adjoint(fvm, count=32, steps=300, slots=24, seed=77, recipe="fourier"):
    import importlib
    adjmod = importlib.import_module("2d_fvm_cuda_amd.adjoint")
    if recipe == "riemann":
        Us = fvm.problems.batched_instances(count=count, seed=seed, lo=256, hi=768)
        shapes = [U.shape[:2] for U in Us]
        Cs = [_cons(U) for U in Us]
    else:
        rng = fvm.problems.SplitMix64(seed)
        shapes = [(rng.randint(256, 768), rng.randint(256, 768)) for _ in range(count)]
        Cs = [_cons(fvm.problems.fourier_perturbed(nx, ny, seed=i + 1)) for i, (ny, nx) in enumerate(shapes)]
    adj = adjmod.Adjoint(shapes, [{}] * count, nsteps=steps, checkpoint_slots=slots)
    UNs = adj.forward(Cs)
    adj.backward([np.ones_like(u) for u in UNs])
    st = adj.stats()
    cells = sum(a * b for a, b in shapes)
    adj.close()
    return {"recipe": recipe, "instances": count, "steps": steps, "cells": cells, "checkpoint_slots": slots,
            "forward_ms": st["forward_ms"], "backward_ms": st["backward_ms"],
            "forward_cell_updates_per_s": cells * steps / (st["forward_ms"] / 1000.0),
            "reverse_over_forward_step": st["reverse_step_ms"] / st["forward_step_ms"],
            "recompute_steps": st["recompute_steps"]}


def peak(fvm, n=8192, steps=20, warmup=3):
    U0 = fvm.problems.initialize_4shocks(n)
    s = fvm.Solver(n, n, t_end=1e30)
    s.upload(U0)
    del U0
    launch = s.get_launch()
    s.begin()
    s.advance(warmup)
    s.sync()
    ms = s.time_steps(steps)
    k_avg, k_min = s.time_kernel(5)
    s.close()
    rate = n * n * steps / (ms / 1000.0)
    step_s = ms / steps / 1000.0
    rec = {"mesh": [n, n], "steps": steps, "ms": ms, "ms_per_step": ms / steps, "cell_updates_per_s": rate,
           "launch": list(launch), "kernel_avg_ms": k_avg,
           # the metric's accounting: 288 B per cell-update (four sweeps'
           # reads and writes and the partner read), not bytes moved
           "algorithmic_GB_s": BYTES_PER_CELL_UPDATE * n * n / step_s / 1e9}
    # bytes the kernel moves, from the committed counter profile of this
    # workload (FETCH_SIZE / WRITE_SIZE, copy-calibrated)
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                        "step_kernel_8192_traffic.json")
    if os.path.exists(path):
        b = json.load(open(path))["bytes_per_launch"]
        rec["hbm_measured"] = {"bytes_per_launch": b, "bytes_per_cell_update": b / (n * n),
                               "GB_s": b / step_s / 1e9, "frac": b / step_s / 1e9 / HBM_PEAK_GBS}
    return rec


# The reference's published wall-clock runs (four-quadrant data, CFL 0.5,
# outflow): the demo main of test2D.cu:137-144 (2000^2, alpha 2.0, t_end
# 0.8; README.md:2 "finishing in 15 minutes" on an RTX 4060, with a PNG
# frame every 18 steps) and the figure's panels (a) (1000^2, alpha 1.9,
# t_end 0.35: 1 min 18 s) and (c) (2000^2, alpha 1.7, t_end 0.35: 10 min
# 44 s).  Frames are not written here: the time is the solve call's.
PUBLISHED_RUNS = (("demo 2000^2 t_end 0.8 alpha 2.0", 2000, 0.8, 2.0, 900.0),
                  ("figure (a) 1000^2 t_end 0.35 alpha 1.9", 1000, 0.35, 1.9, 78.0),
                  ("figure (c) 2000^2 t_end 0.35 alpha 1.7", 2000, 0.35, 1.7, 644.0))


def published(fvm):
    out = []
    for name, n, t_end, alpha, ref_s in PUBLISHED_RUNS:
        U0 = fvm.problems.initialize_4shocks(n)
        t0 = time.perf_counter()
        _, nsteps, ms = fvm.solve(U0, t_end=t_end, cfl=0.5, alpha=alpha, gamma=1.4, strategy=0)
        wall = time.perf_counter() - t0
        out.append({"run": name, "mesh": [n, n], "steps": nsteps, "loop_s": ms / 1000.0,
                    "call_wall_s": wall, "cell_updates_per_s": n * n * nsteps / (ms / 1000.0),
                    "published_wall_s": ref_s, "published_over_call_wall": ref_s / wall})
    return out


def published_record(fvm):
    """test2D.cu:134-152 as written: 2000^2 four-quadrant data, [0,1]^2,
    t_end 0.8, CFL 0.5, outflow, alpha 2.0, gamma 1.4, a frame every 18
    steps into ./pngs, then writeFinalResultTo_txt(".") -- here into a
    temporary directory that is removed afterwards.  The published time
    ("finishing in 15 minutes", README.md:2) is this program's."""
    d = tempfile.mkdtemp(prefix="fvm2d_demo_")
    try:
        r = fvm.RGRP_2D_CUDA(fvm.problems.initialize_4shocks(2000))
        r.setSpatialLayout(1.0, 1.0)
        r.setTimeLayout(0.8, 0.5)
        r.setGhostCellStrategy(fvm.ghostCellStrategy.outflow)
        r.setAlpha(2.0)
        r.setGamma(1.4)
        pngs = os.path.join(d, "pngs")
        r.setRecordModel(True, 18, pngs)
        t0 = time.perf_counter()
        r.solve()
        r.writeFinalResultTo_txt(d)
        wall = time.perf_counter() - t0
        frames = sorted(os.listdir(pngs))
        nbytes = sum(os.path.getsize(os.path.join(pngs, f)) for f in frames)
        return {"run": "demo main 2000^2 t_end 0.8 alpha 2.0, PNG frame every 18 steps",
                "steps": r.getTimeStagesCount(), "frames": len(frames), "frame_MB": nbytes / 1e6 / len(frames),
                "computation_time_s": r.getComputationTime_ms() / 1000.0, "program_wall_s": wall,
                "published_wall_s": 900.0, "published_over_program_wall": 900.0 / wall}
    finally:
        shutil.rmtree(d, ignore_errors=True)
