"""GPU adjoint (fvm2d_adj_*) against the FP64 host reference gradient,
central finite differences of the forward, and its own invariants."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 1.4


def to_cons(U5):
    r, p, u, v = U5[..., 0], U5[..., 1], U5[..., 2], U5[..., 3]
    return np.ascontiguousarray(np.stack([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)]))


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def adjmod(gpu):
    import importlib
    return importlib.import_module("2d_fvm_cuda_amd.adjoint")


def smooth_state(problems, nx, ny, seed, vel=True):
    U5 = problems.fourier_perturbed(nx, ny, seed)
    if vel:
        x = (np.arange(nx) + 0.5) / nx
        y = (np.arange(ny) + 0.5) / ny
        X, Y = np.meshgrid(x, y)
        U5[..., 2] += 0.05 * np.sin(2 * np.pi * (X + 2 * Y) + seed)
        U5[..., 3] += 0.05 * np.cos(2 * np.pi * (2 * X - Y) + 0.5 * seed)
    return to_cons(U5)


@pytest.mark.parametrize("shape,strategy,steps", [((40, 36), 0, 10), ((61, 130), 0, 7),
                                                  ((120, 70), 3, 6), ((50, 64), 4, 5),
                                                  ((33, 47), 2, 4)])
def test_gpu_gradient_matches_host_reference(gpu, orc, problems, adjmod, torch_mod, shape, strategy, steps):
    ny, nx = shape
    print(f"case shape {shape} strategy {strategy} steps {steps}", flush=True)
    C0 = smooth_state(problems, nx, ny, seed=nx + ny)
    rng = np.random.default_rng(nx)
    gN = rng.normal(size=C0.shape)
    prm = dict(strategy=strategy, width=nx / ny if strategy in (2, 3, 4) else 1.0)
    ref_UN = orc.adj_forward(C0, steps, **prm)
    ref_g0 = orc.adj_backward(C0, steps, gN, **prm)
    adj = adjmod.Adjoint([shape], [prm], nsteps=steps, checkpoint_slots=3)
    t = torch_mod.from_numpy(C0).cuda()
    UN = adj.forward([t])[0].cpu().numpy()
    np.testing.assert_allclose(UN, ref_UN, rtol=1e-12, atol=1e-13)
    g0 = adj.backward([torch_mod.from_numpy(gN).cuda()])[0].cpu().numpy()
    err = np.abs(g0 - ref_g0).max() / np.abs(ref_g0).max()
    print(f"max rel gradient error {err:.3e}")
    assert err <= 1e-11
    adj.close()


def test_checkpoint_budget_does_not_change_gradient(gpu, problems, adjmod, torch_mod):
    C0 = smooth_state(problems, 64, 48, seed=2)
    rng = np.random.default_rng(0)
    gN = torch_mod.from_numpy(rng.normal(size=C0.shape)).cuda()
    t = torch_mod.from_numpy(C0).cuda()
    res = []
    for slots in (1, 2, 5, 0):
        adj = adjmod.Adjoint([(48, 64)], [{}], nsteps=17, checkpoint_slots=slots)
        adj.forward([t])
        g = adj.backward([gN])[0].cpu().numpy()
        st = adj.stats()
        print(slots, st)
        assert st["reverse_steps"] == 17
        res.append((g, st))
        adj.close()
    for g, _ in res[1:]:
        assert np.array_equal(g, res[0][0])
    # fewer slots, more recomputation
    assert res[0][1]["recompute_steps"] >= res[1][1]["recompute_steps"] >= res[2][1]["recompute_steps"]
    assert res[3][1]["recompute_steps"] == 0


def test_adjoint_workload_central_fd(gpu, problems, adjmod, torch_mod):
    """The adjoint workload's recipe (background rho = p = 1, u = v = 0.1,
    16 seeded modes on rho and p, 2x pressure in a centred disc) on a 128^2
    mesh for 60 steps; loss = 1/2 sum (rho_N - rho_target)^2 with the target
    from a second seed; central differences (h = 1e-6) on 30 random cells."""
    nx = ny = 128
    steps = 60
    C0 = to_cons(problems.fourier_perturbed(nx, ny, seed=1))
    Ct = to_cons(problems.fourier_perturbed(nx, ny, seed=2))
    adj = adjmod.Adjoint([(ny, nx)], [{}], nsteps=steps, checkpoint_slots=8)
    target = torch_mod.from_numpy(adj.forward([torch_mod.from_numpy(Ct).cuda()])[0].cpu().numpy()[0]).cuda()
    x = torch_mod.from_numpy(C0).cuda().requires_grad_(True)
    FVMSolve = adjmod.FVMSolve
    UN = FVMSolve.apply(adj, x)
    loss = 0.5 * ((UN[0] - target) ** 2).sum()
    loss.backward()
    g = x.grad.cpu().numpy()
    rng = np.random.default_rng(7)
    rho_N0 = UN[0].detach()
    worst = 0.0
    for _ in range(30):
        k, j, i = rng.integers(0, 4), rng.integers(0, ny), rng.integers(0, nx)
        h = 1e-6
        Cp, Cm = C0.copy(), C0.copy()
        Cp[k, j, i] += h
        Cm[k, j, i] -= h
        a = adj.forward([torch_mod.from_numpy(Cp).cuda()])[0][0]
        b = adj.forward([torch_mod.from_numpy(Cm).cuda()])[0][0]
        fd = float((0.5 * (a - b) * (a + b - 2 * target)).sum()) / (2 * h)
        scale = max(abs(fd), 1e-6 * np.abs(g).max())
        worst = max(worst, abs(fd - g[k, j, i]) / scale)
    print(f"worst relative FD deviation {worst:.3e}")
    assert worst < 1e-4
    del rho_N0
    adj.close()


def test_autograd_matches_direct_backward(gpu, problems, adjmod, torch_mod):
    C0 = smooth_state(problems, 40, 40, seed=4)
    adj = adjmod.Adjoint([(40, 40)], [{}], nsteps=8, checkpoint_slots=2)
    x = torch_mod.from_numpy(C0).cuda().requires_grad_(True)
    UN = adjmod.FVMSolve.apply(adj, x)
    w = torch_mod.randn_like(UN)
    (UN * w).sum().backward()
    adj.forward([x.detach()])
    g = adj.backward([w])[0]
    assert torch_mod.equal(x.grad, g)
    adj.close()


def test_batched_adjoint_matches_single(gpu, problems, adjmod, torch_mod):
    shapes = [(40, 52), (70, 33), (64, 64)]
    Cs = [smooth_state(problems, nx, ny, seed=ny) for (ny, nx) in shapes]
    rng = np.random.default_rng(3)
    gNs = [torch_mod.from_numpy(rng.normal(size=C.shape)).cuda() for C in Cs]
    adj = adjmod.Adjoint(shapes, [{}] * 3, nsteps=9, checkpoint_slots=3)
    UNs = adj.forward([torch_mod.from_numpy(C).cuda() for C in Cs])
    g0s = adj.backward(gNs)
    for i, (shape, C) in enumerate(zip(shapes, Cs)):
        one = adjmod.Adjoint([shape], [{}], nsteps=9, checkpoint_slots=3)
        UN = one.forward([torch_mod.from_numpy(C).cuda()])[0]
        g0 = one.backward([gNs[i]])[0]
        assert torch_mod.equal(UN, UNs[i])
        assert torch_mod.equal(g0, g0s[i])
        one.close()
    adj.close()
