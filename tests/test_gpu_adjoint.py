"""GPU adjoint (fvm2d_adj_*) against the FP64 host reference gradient,
central finite differences of the forward, and its own invariants.
States and gradients cross the C ABI in host memory (numpy / CPU tensors)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 1.4


def to_cons(U5):
    r, p, u, v = U5[..., 0], U5[..., 1], U5[..., 2], U5[..., 3]
    return np.ascontiguousarray(np.stack([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)]))


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
def test_gpu_gradient_matches_host_reference(gpu, orc, problems, adjmod, shape, strategy, steps):
    ny, nx = shape
    print(f"case shape {shape} strategy {strategy} steps {steps}", flush=True)
    C0 = smooth_state(problems, nx, ny, seed=nx + ny)
    rng = np.random.default_rng(nx)
    gN = rng.normal(size=C0.shape)
    prm = dict(strategy=strategy, width=nx / ny if strategy in (2, 3, 4) else 1.0)
    ref_UN = orc.adj_forward(C0, steps, **prm)
    ref_g0 = orc.adj_backward(C0, steps, gN, **prm)
    adj = adjmod.Adjoint([shape], [prm], nsteps=steps, checkpoint_slots=3)
    UN = adj.forward([C0])[0]
    np.testing.assert_allclose(UN, ref_UN, rtol=1e-12, atol=1e-13)
    g0 = adj.backward([gN])[0]
    err = np.abs(g0 - ref_g0).max() / np.abs(ref_g0).max()
    print(f"max rel gradient error {err:.3e}")
    assert err <= 1e-11
    adj.close()


def test_checkpoint_budget_does_not_change_gradient(gpu, problems, adjmod):
    C0 = smooth_state(problems, 64, 48, seed=2)
    gN = np.random.default_rng(0).normal(size=C0.shape)
    res = []
    for slots in (1, 2, 5, 0):
        adj = adjmod.Adjoint([(48, 64)], [{}], nsteps=17, checkpoint_slots=slots)
        adj.forward([C0])
        g = adj.backward([gN])[0]
        st = adj.stats()
        print(slots, st)
        assert st["reverse_steps"] == 17
        res.append((g, st))
        adj.close()
    for g, _ in res[1:]:
        assert np.array_equal(g, res[0][0])
    # fewer slots, more recomputation; storing every step recomputes nothing
    assert res[0][1]["recompute_steps"] >= res[1][1]["recompute_steps"] >= res[2][1]["recompute_steps"]
    assert res[3][1]["recompute_steps"] == 0


def test_adjoint_workload_central_fd(gpu, problems, adjmod):
    """The adjoint workload's recipe (background rho = p = 1, u = v = 0.1,
    16 seeded modes on rho and p, 2x pressure in a centred disc) on 512^2
    for 400 steps with binomial checkpointing; loss = 1/2 sum (rho_N -
    rho_target)^2, the target from a second seed; central differences with
    h = 1e-6 on 50 random cells."""
    import torch
    n, steps = 512, 400
    C0 = to_cons(problems.fourier_perturbed(n, n, seed=1))
    Ct = to_cons(problems.fourier_perturbed(n, n, seed=2))
    adj = adjmod.Adjoint([(n, n)], [{}], nsteps=steps, checkpoint_slots=32)
    target = adj.forward([Ct])[0][0]
    x = torch.from_numpy(C0.copy()).requires_grad_(True)
    UN = adjmod.FVMSolve.apply(adj, x)
    loss = 0.5 * ((UN[0] - torch.from_numpy(target)) ** 2).sum()
    loss.backward()
    g = x.grad.numpy()
    st = adj.stats()
    print(st)
    assert st["recompute_steps"] <= st["schedule_cost"]
    gmax = np.abs(g).max()

    def fd(k, j, i, h):
        Cp, Cm = C0.copy(), C0.copy()
        Cp[k, j, i] += h
        Cm[k, j, i] -= h
        a = adj.forward([Cp])[0][0]
        b = adj.forward([Cm])[0][0]
        return float((0.5 * (a - b) * (a + b - 2 * target)).sum()) / (2 * h)

    # The limiter and the GRP branch tests make the scheme piecewise smooth:
    # a central difference whose stencil straddles a branch switch differs
    # from the derivative along the taken branch.  A pair is accepted when
    # h = 1e-6 agrees, or when a second step size (1e-7) agrees and the two
    # differences disagree with each other (a switch inside [x-h, x+h]).
    rng = np.random.default_rng(7)
    devs, worst = [], 0.0
    for _ in range(50):
        k, j, i = rng.integers(0, 4), rng.integers(0, n), rng.integers(0, n)
        f1 = fd(k, j, i, 1e-6)
        scale = max(abs(f1), 1e-6 * gmax)
        d1 = abs(f1 - g[k, j, i]) / scale
        devs.append(d1)
        if d1 > 1e-5:
            f2 = fd(k, j, i, 1e-7)
            d2 = abs(f2 - g[k, j, i]) / scale
            d12 = abs(f1 - f2) / scale
            print(f"pair ({k},{j},{i}): adj {g[k, j, i]:.9e} fd(1e-6) {f1:.9e} fd(1e-7) {f2:.9e} "
                  f"dev {d1:.2e}/{d2:.2e} fd spread {d12:.2e}")
            d1 = min(d1, d2) if d12 > min(d1, d2) else d1
        worst = max(worst, d1)
    print(f"FD deviations at h = 1e-6: max {max(devs):.3e} median {np.median(devs):.3e}; "
          f"accepted worst {worst:.3e}")
    assert np.median(devs) < 1e-5
    assert worst < 1e-4
    adj.close()


def test_autograd_matches_direct_backward(gpu, problems, adjmod):
    import torch
    C0 = smooth_state(problems, 40, 40, seed=4)
    adj = adjmod.Adjoint([(40, 40)], [{}], nsteps=8, checkpoint_slots=2)
    x = torch.from_numpy(C0.copy()).requires_grad_(True)
    UN = adjmod.FVMSolve.apply(adj, x)
    w = torch.from_numpy(np.random.default_rng(1).normal(size=C0.shape))
    (UN * w).sum().backward()
    adj.forward([C0])
    g = adj.backward([w.numpy()])[0]
    assert np.array_equal(x.grad.numpy(), g)
    adj.close()


def test_batched_adjoint_matches_single(gpu, problems, adjmod):
    shapes = [(40, 52), (70, 33), (64, 64)]
    Cs = [smooth_state(problems, nx, ny, seed=ny) for (ny, nx) in shapes]
    rng = np.random.default_rng(3)
    gNs = [rng.normal(size=C.shape) for C in Cs]
    adj = adjmod.Adjoint(shapes, [{}] * 3, nsteps=9, checkpoint_slots=3)
    UNs = adj.forward(Cs)
    g0s = adj.backward(gNs)
    for i, (shape, C) in enumerate(zip(shapes, Cs)):
        one = adjmod.Adjoint([shape], [{}], nsteps=9, checkpoint_slots=3)
        UN = one.forward([C])[0]
        g0 = one.backward([gNs[i]])[0]
        assert np.array_equal(UN, UNs[i])
        assert np.array_equal(g0, g0s[i])
        one.close()
    adj.close()
