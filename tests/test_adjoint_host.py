"""CPU checks of the host reference gradient (oracle/fvm2d_oracle_adjoint.cpp):
its forward equals the oracle's literal restatement to rounding, and its
gradient matches central finite differences of its forward."""
import numpy as np
import pytest

G = 1.4


def to_cons(U5):
    r, p, u, v = U5[..., 0], U5[..., 1], U5[..., 2], U5[..., 3]
    return np.ascontiguousarray(np.stack([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)]))


def cons_to_prim5(C):
    r, mu, mv, E = C
    u, v = mu / r, mv / r
    out = np.zeros(C.shape[1:] + (5,))
    out[..., 0], out[..., 1], out[..., 2], out[..., 3] = r, (G - 1) * (E - 0.5 * (mu * u + mv * v)), u, v
    return out


def small_problem(problems, nx=22, ny=18, seed=3):
    """Smooth seeded data: the adjoint workload's recipe plus random-phase
    velocity modes, so that no component is flat and no two neighbouring
    values are equal by symmetry (the limiter is not differentiable where a
    one-sided difference is zero; there a perturbation picks a
    direction-dependent branch)."""
    U5 = problems.fourier_perturbed(nx, ny, seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rng = np.random.default_rng(1000 + seed)
    for c in (2, 3):
        for _ in range(4):
            kx, ky = rng.integers(1, 4, size=2)
            U5[..., c] += rng.uniform(0.01, 0.03) * np.sin(2 * np.pi * (kx * X + ky * Y) +
                                                           rng.uniform(0, 2 * np.pi))
    return to_cons(U5)


@pytest.mark.parametrize("strategy", [0, 2, 3, 4])
def test_host_forward_matches_oracle(orc, problems, strategy):
    """The FAST-numerics forward agrees with the literal oracle to rounding
    (both run nsteps fixed CFL steps)."""
    C0 = small_problem(problems, 26, 20, seed=strategy + 1)
    n = 8
    got = orc.adj_forward(C0, n, strategy=strategy)
    ref, steps, _, _ = orc.solve(cons_to_prim5(C0), nthreads=1, t_end=1e9, max_steps=n, strategy=strategy)
    assert steps == n
    np.testing.assert_allclose(got, to_cons(ref), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("strategy,gas", [(0, {}), (3, {}), (4, {}),
                                          (0, dict(gamma=5.0 / 3.0, alpha=2.0, cfl=0.4)),
                                          (4, dict(gamma=1.1, alpha=1.5))])
def test_host_gradient_directional_fd(orc, problems, strategy, gas):
    """<dL/dU0, v> = (L(U0 + eps v) - L(U0 - eps v)) / (2 eps) for
    L = <w, U_N> with random w and a random smooth direction v; also at
    gamma 5/3 (alpha 2, the demo's value; CFL 0.4) and gamma 1.1."""
    C0 = small_problem(problems, 22, 18, seed=11 + strategy)
    n = 6
    rng = np.random.default_rng(strategy)
    w = rng.normal(size=C0.shape)
    kw = dict(strategy=strategy, **gas)
    g0 = orc.adj_backward(C0, n, w, **kw)
    for trial in range(3):
        v = rng.normal(size=C0.shape) * np.abs(C0) * 1e-2
        ad = np.sum(g0 * v)
        scale = np.sum(np.abs(g0 * v))
        errs = []
        # a step may straddle a limiter or GRP branch switch of one cell, so
        # the better of two step sizes is taken
        for eps in (1e-5, 1e-6):
            d = orc.adj_forward(C0 + eps * v, n, **kw) - orc.adj_forward(C0 - eps * v, n, **kw)
            fd = np.sum(w * d) / (2 * eps)  # difference per cell first
            errs.append(abs(fd - ad))
        assert min(errs) <= 1e-6 * scale, (trial, errs, scale)


def test_host_gradient_single_cells_fd(orc, problems):
    """Every (component, cell) of a 20 x 20 mesh over 5 steps: the gradient
    of L = 1/2 sum (rho_N - target)^2 against the central and one-sided
    difference quotients (h = 1e-6 scaled), the best of the three counting
    (a quotient whose interval holds a branch switch is not the derivative
    along the taken branch)."""
    C0 = small_problem(problems, 20, 20, seed=5)
    n = 5
    rng = np.random.default_rng(1)
    target = rng.normal(size=C0.shape[1:]) * 0.1 + C0[0]

    UN = orc.adj_forward(C0, n)
    r0 = UN[0]

    def dloss(Cp):
        """L(Cp) - L(C0) evaluated cell by cell before summing."""
        a = orc.adj_forward(Cp, n)[0]
        return np.sum(0.5 * (a - r0) * (a + r0 - 2 * target))

    gN = np.zeros_like(C0)
    gN[0] = UN[0] - target
    g0 = orc.adj_backward(C0, n, gN)
    gmax = np.abs(g0).max()
    worst, bad = 0.0, 0
    for k in range(4):
        for j in range(20):
            for i in range(20):
                h = 1e-6 * max(1.0, abs(C0[k, j, i]))
                Cp, Cm = C0.copy(), C0.copy()
                Cp[k, j, i] += h
                Cm[k, j, i] -= h
                fp, fm = dloss(Cp) / h, -dloss(Cm) / h
                a = g0[k, j, i]
                d = min(abs(a - 0.5 * (fp + fm)), abs(a - fp), abs(a - fm)) / max(abs(a), 1e-3 * gmax)
                worst = max(worst, d)
                bad += d > 1e-6
    print(f"worst best-of-three deviation {worst:.2e}, pairs above 1e-6: {bad} of 1600")
    assert bad <= 8
    assert worst < 1e-5
