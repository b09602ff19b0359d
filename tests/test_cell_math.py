"""CPU checks of the product's per-cell numerics (host build of
fvm2d_cell.h / fvm2d_adjoint_cell.h):

* EXACT mode reproduces the oracle bit for bit (same op order, no FMA);
* FAST mode (shared reciprocals + FMA) agrees with EXACT to rounding level
  on every GRP branch;
* every hand-written adjoint equals the transpose of the forward Jacobian,
  checked against fourth-order central differences of the forward.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "cell_host.cpp")
SO = os.path.join(HERE, "cpp", "libcell_host.so")
G = 1.4


@pytest.fixture(scope="module")
def cm():
    deps = [SRC] + [os.path.join(HERE, "..", "2d_fvm_cuda_amd", "csrc", f)
                    for f in ("fvm2d_cell.h", "fvm2d_adjoint_cell.h")]
    if not os.path.exists(SO) or any(os.path.getmtime(d) > os.path.getmtime(SO) for d in deps):
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-shared",
                        "-fPIC", SRC, "-o", SO], check=True)
    L = ctypes.CDLL(SO)
    dp = ctypes.POINTER(ctypes.c_double)
    L.cm_grp.argtypes = [ctypes.c_int, dp, dp, dp, dp, ctypes.c_double, dp, dp]
    L.cm_grp.restype = ctypes.c_int
    L.cm_flux.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, dp]
    L.cm_slope.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_double, dp]
    L.cm_update.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_double, dp]
    L.cm_speed.argtypes = [ctypes.c_int, dp, ctypes.c_double]
    L.cm_speed.restype = ctypes.c_double
    L.cm_grp_adj.argtypes = [dp, dp, dp, dp, ctypes.c_double, dp, dp, dp, dp, dp, dp]
    L.cm_flux_adj.argtypes = [dp, dp, ctypes.c_double, dp, dp, dp]
    L.cm_flux_adj.restype = ctypes.c_double
    L.cm_slope_adj.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
    L.cm_prim_adj.argtypes = [dp, dp, ctypes.c_double, dp]
    L.cm_cons_adj.argtypes = [dp, dp, ctypes.c_double, dp]
    L.cm_speed_grad.argtypes = [dp, ctypes.c_double, dp]
    return L


def P(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def grp(cm, exact, x16):
    W, pW = P(np.zeros(5))
    Wt, pWt = P(np.zeros(5))
    parts = [P(x16[4 * i:4 * i + 4]) for i in range(4)]
    br = cm.cm_grp(exact, parts[0][1], parts[1][1], parts[2][1], parts[3][1], G, pW, pWt)
    return np.concatenate([W, Wt]), br


def random_interface(rng, branch):
    """Face states / slopes that land in the requested GRP branch."""
    while True:
        WL = np.array([rng.uniform(0.3, 2.5), rng.uniform(0.3, 2.5), rng.uniform(-1, 1), rng.normal()])
        WR = np.array([rng.uniform(0.3, 2.5), rng.uniform(0.3, 2.5), rng.uniform(-1, 1), rng.normal()])
        if branch == 0:
            WL[2] += 4.0
            WR[2] += 4.0
        elif branch == 1:
            WL[2] -= 4.0
            WR[2] -= 4.0
        x = np.concatenate([WL, WR, rng.normal(size=4), rng.normal(size=4)])
        return x


def branch_of(cm, x):
    return grp(cm, 0, x)[1]


# ------------------------------------------------------------------ forward
def test_exact_mode_is_bitwise_oracle(cm, orc):
    rng = np.random.default_rng(0)
    for _ in range(2000):
        x = random_interface(rng, rng.integers(0, 3))
        got, br = grp(cm, 1, x)
        W, Wt, obr = orc.grp(x[0:4], x[4:8], x[8:12], x[12:16], G)
        assert br == obr
        assert np.array_equal(got, np.concatenate([W, Wt]))
        k = rng.uniform(1e-4, 1e-2)
        F, pF = P(np.zeros(4))
        a, pa = P(W)
        b, pb = P(Wt)
        cm.cm_flux(1, pa, pb, k, pF)
        assert np.array_equal(F, orc.flux(W, Wt, k))
        q3 = rng.normal(size=(3, 4))
        s, ps = P(np.zeros(4))
        cm.cm_slope(1, P(q3[0])[1], P(q3[1])[1], P(q3[2])[1], 0.01, 1.9, ps)
        assert np.array_equal(s, orc.slope(q3[0], q3[1], q3[2], 0.01, 1.9))


def test_exact_update_matches_oracle_update(cm, orc):
    rng = np.random.default_rng(1)
    for _ in range(500):
        q = np.array([rng.uniform(0.2, 2), rng.uniform(0.2, 2), rng.normal(), rng.normal()])
        FL, FR = rng.normal(size=4) * 0.1, rng.normal(size=4) * 0.1
        lam = rng.uniform(0.01, 0.3)
        out, po = P(np.zeros(4))
        cm.cm_update(1, P(q)[1], P(FL)[1], P(FR)[1], lam, G, po)
        c = orc.cons(q, G) - (FR - FL) * lam
        assert np.array_equal(out, orc.prim(c, G))


@pytest.mark.parametrize("branch", [0, 1, 2])
def test_fast_grp_matches_exact(cm, branch):
    rng = np.random.default_rng(10 + branch)
    seen = set()
    for _ in range(3000):
        x = random_interface(rng, branch)
        a, br = grp(cm, 1, x)
        b, brf = grp(cm, 0, x)
        assert br == brf
        seen.add(br)
        np.testing.assert_allclose(b, a, rtol=1e-11, atol=1e-11 * np.abs(a).max())
    if branch == 2:
        assert {2, 3} <= seen
    else:
        assert branch in seen


def test_fast_flux_update_slope_speed_match_exact(cm):
    rng = np.random.default_rng(2)
    for _ in range(1000):
        W = np.array([rng.uniform(0.2, 2), rng.uniform(0.2, 2), rng.normal(), rng.normal(), rng.uniform(0.5, 5)])
        Wt = rng.normal(size=5)
        k = rng.uniform(1e-4, 1e-2)
        F1, p1 = P(np.zeros(4))
        F2, p2 = P(np.zeros(4))
        cm.cm_flux(1, P(W)[1], P(Wt)[1], k, p1)
        cm.cm_flux(0, P(W)[1], P(Wt)[1], k, p2)
        np.testing.assert_allclose(F2, F1, rtol=1e-13, atol=1e-14)
        q = W[:4]
        FL, FR = rng.normal(size=4) * 0.1, rng.normal(size=4) * 0.1
        o1, q1 = P(np.zeros(4))
        o2, q2 = P(np.zeros(4))
        cm.cm_update(1, P(q)[1], P(FL)[1], P(FR)[1], 0.1, G, q1)
        cm.cm_update(0, P(q)[1], P(FL)[1], P(FR)[1], 0.1, G, q2)
        np.testing.assert_allclose(o2, o1, rtol=1e-12, atol=1e-14)
        q3 = rng.normal(size=(3, 4))
        s1, ps1 = P(np.zeros(4))
        s2, ps2 = P(np.zeros(4))
        cm.cm_slope(1, P(q3[0])[1], P(q3[1])[1], P(q3[2])[1], 0.01, 1.9, ps1)
        cm.cm_slope(0, P(q3[0])[1], P(q3[1])[1], P(q3[2])[1], 0.01, 1.9, ps2)
        np.testing.assert_allclose(s2, s1, rtol=1e-14)
        assert cm.cm_speed(0, P(q)[1], G) == pytest.approx(cm.cm_speed(1, P(q)[1], G), rel=1e-15)


# ------------------------------------------------------------------ adjoints
def fd_jac(f, x, h):
    """Fourth-order central differences of f at x, step h * max(1, |x_i|)."""
    cols = []
    for i in range(len(x)):
        d = h * max(1.0, abs(x[i]))
        e = np.zeros_like(x)
        e[i] = d
        cols.append((-f(x + 2 * e) + 8 * f(x + e) - 8 * f(x - e) + f(x - 2 * e)) / (12 * d))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("branch", [0, 1, 2])
def test_grp_adjoint_is_jacobian_transpose(cm, branch):
    rng = np.random.default_rng(20 + branch)
    checked = set()
    for _ in range(60):
        x = random_interface(rng, branch)
        br = branch_of(cm, x)

        def f(z):
            out, b = grp(cm, 0, z)
            return out if b == br else np.full(10, np.nan)

        J = fd_jac(f, x, 1e-5)
        if not np.all(np.isfinite(J)):
            continue  # too close to a branch switch for the difference stencil
        checked.add(br)
        wb = rng.normal(size=10)
        outs = [P(np.zeros(4)) for _ in range(4)]
        parts = [P(x[4 * i:4 * i + 4]) for i in range(4)]
        cm.cm_grp_adj(parts[0][1], parts[1][1], parts[2][1], parts[3][1], G, P(wb[:5])[1],
                      P(wb[5:])[1], outs[0][1], outs[1][1], outs[2][1], outs[3][1])
        adj = np.concatenate([o[0] for o in outs])
        exp = J.T @ wb
        np.testing.assert_allclose(adj, exp, rtol=1e-6, atol=1e-7 * (1 + np.abs(exp).max()))
    assert checked


def test_flux_adjoint_is_jacobian_transpose(cm):
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = np.concatenate([[rng.uniform(0.2, 2), rng.uniform(0.2, 2), rng.normal(), rng.normal(),
                             rng.uniform(0.5, 5)], rng.normal(size=5), [rng.uniform(1e-3, 1e-1)]])

        def f(z):
            F, pF = P(np.zeros(4))
            cm.cm_flux(0, P(z[:5])[1], P(z[5:10])[1], z[10], pF)
            return F

        J = fd_jac(f, x, 1e-5)
        fb = rng.normal(size=4)
        Wb, pWb = P(np.zeros(5))
        Wtb, pWtb = P(np.zeros(5))
        kb = cm.cm_flux_adj(P(x[:5])[1], P(x[5:10])[1], x[10], P(fb)[1], pWb, pWtb)
        np.testing.assert_allclose(np.concatenate([Wb, Wtb, [kb]]), J.T @ fb, rtol=1e-7, atol=1e-9)


def test_slope_adjoint_is_jacobian_transpose(cm):
    rng = np.random.default_rng(4)
    for _ in range(200):
        x = rng.normal(size=12)

        def f(z):
            s, ps = P(np.zeros(4))
            cm.cm_slope(0, P(z[:4])[1], P(z[4:8])[1], P(z[8:])[1], 0.01, 1.9, ps)
            return s

        J = fd_jac(f, x, 1e-7)
        sb = rng.normal(size=4)
        outs = [P(np.zeros(4)) for _ in range(3)]
        cm.cm_slope_adj(P(x[:4])[1], P(x[4:8])[1], P(x[8:])[1], 0.01, 1.9, P(sb)[1], outs[0][1],
                        outs[1][1], outs[2][1])
        adj = np.concatenate([o[0] for o in outs])
        np.testing.assert_allclose(adj, J.T @ sb, rtol=1e-6, atol=1e-6)


def test_update_adjoint_is_jacobian_transpose(cm):
    """prim_adj and cons_adj compose to the transpose of d update / d q and
    d update / d F."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        q = np.array([rng.uniform(0.3, 2), rng.uniform(0.3, 2), rng.normal(), rng.normal()])
        FL, FR = rng.normal(size=4) * 0.05, rng.normal(size=4) * 0.05
        lam = 0.2

        def f(z):
            o, po = P(np.zeros(4))
            cm.cm_update(0, P(z[:4])[1], P(z[4:8])[1], P(z[8:])[1], lam, G, po)
            return o

        x = np.concatenate([q, FL, FR])
        J = fd_jac(f, x, 1e-6)
        qb = rng.normal(size=4)
        out = f(x)
        Ub, pUb = P(np.zeros(4))
        cm.cm_prim_adj(P(out)[1], P(qb)[1], G, pUb)
        qadj, pq = P(np.zeros(4))
        cm.cm_cons_adj(P(q)[1], P(Ub)[1], G, pq)
        exp = J.T @ qb
        np.testing.assert_allclose(qadj, exp[:4], rtol=1e-7, atol=1e-8)
        np.testing.assert_allclose(lam * Ub, exp[4:8], rtol=1e-7, atol=1e-8)   # d/dFL = +lam Ub
        np.testing.assert_allclose(-lam * Ub, exp[8:], rtol=1e-7, atol=1e-8)  # d/dFR = -lam Ub


def test_speed_gradient(cm):
    rng = np.random.default_rng(6)
    for _ in range(100):
        q = np.array([rng.uniform(0.3, 2), rng.uniform(0.3, 2), rng.normal(), rng.normal()])
        J = fd_jac(lambda z: np.array([cm.cm_speed(0, P(z)[1], G)]), q, 1e-6)
        g, pg = P(np.zeros(4))
        cm.cm_speed_grad(P(q)[1], G, pg)
        np.testing.assert_allclose(g, J[0], rtol=1e-7, atol=1e-9)
