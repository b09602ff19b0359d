"""Pins of the FP64 host oracle against mathematics that fixes its results.

The reference ships no tests, fixtures or golden vectors, and its CUDA
sources cannot be compiled here, so each routine of the oracle is checked
against closed forms, textbook special cases, an emulation of the
reference's GPU execution model, and whole-solve invariants.  Each check is
chosen so that a plausible restatement mistake (dropped term, wrong sign,
swapped left/right operand) fails it.
"""
import math

import numpy as np
import pytest
import sympy as sp

G = 1.4


# ----------------------------------------------------------------- flux
def test_flux_is_state_flux_plus_half_step_time_derivative(orc):
    """flux(W, W_t, k) = f(W) + (k/2) d/dtau f(W + tau W_t) at tau = 0
    (src/basicFuns.cuh:14-34), with f the Euler flux in the (mass, normal
    momentum, energy, tangential momentum) order using the interface e."""
    r, p, u, v, e, tau = sp.symbols("r p u v e tau")
    f = [r * u, r * u * u + p, r * u * (e + u * u / 2 + v * v / 2) + p * u, r * u * v]
    rng = np.random.default_rng(1)
    for _ in range(50):
        W = np.array([rng.uniform(0.1, 3), rng.uniform(0.1, 3), rng.uniform(-2, 2),
                      rng.uniform(-2, 2), rng.uniform(0.1, 5)])
        Wt = rng.normal(size=5) * 3
        k = rng.uniform(1e-4, 1e-1)
        sub = {r: W[0] + tau * Wt[0], p: W[1] + tau * Wt[1], u: W[2] + tau * Wt[2],
               v: W[3] + tau * Wt[3], e: W[4] + tau * Wt[4]}
        exp = []
        for fk in f:
            g = fk.subs(sub)
            val = float(g.subs(tau, 0)) + k / 2 * float(sp.diff(g, tau).subs(tau, 0))
            exp.append(val)
        got = orc.flux(W, Wt, k)
        np.testing.assert_allclose(got, exp, rtol=1e-13, atol=1e-13)


# ----------------------------------------------------------------- limiter
def textbook_minmod_slope(qL, q, qR, h, alpha):
    """Generalised minmod of (alpha dq-/h, dq_c/(2h), alpha dq+/h)."""
    a = np.array([(qR - qL) * 0.5 / h, alpha * (q - qL) / h, alpha * (qR - q) / h])
    if np.all(a > 0):
        return a.min()
    if np.all(a < 0):
        return a.max()
    return 0.0


@pytest.mark.parametrize("alpha", [1.0, 1.7, 1.9, 2.0])
def test_limiter_matches_textbook_minmod(orc, alpha):
    rng = np.random.default_rng(int(alpha * 10))
    h = 0.01
    for _ in range(400):
        q3 = rng.normal(size=(3, 4))
        # include exact ties and zeros
        if rng.random() < 0.2:
            q3[1] = q3[0]
        if rng.random() < 0.1:
            q3[2] = 2 * q3[1] - q3[0]
        got = orc.slope(q3[0], q3[1], q3[2], h, alpha)
        for c in range(4):
            exp = textbook_minmod_slope(q3[0, c], q3[1, c], q3[2, c], h, alpha)
            assert got[c] == pytest.approx(exp, rel=1e-14, abs=1e-300)


def test_limiter_hand_cases(orc):
    h, al = 1.0, 1.9
    s = orc.slope([1, 1, 1, 1], [2, 2, 2, 2], [3, 3, 3, 3], h, al)  # linear -> central
    assert np.all(s == 1.0)
    s = orc.slope([1, 1, 1, 1], [2, 2, 2, 2], [1, 1, 1, 1], h, al)  # extremum -> 0
    assert np.all(s == 0.0)
    s = orc.slope([5, 5, 5, 5], [5, 5, 5, 5], [5, 5, 5, 5], h, al)  # constant -> 0
    assert np.all(s == 0.0)
    s = orc.slope([3, 3, 3, 3], [2, 2, 2, 2], [1, 1, 1, 1], h, al)  # decreasing
    assert np.all(s == -1.0)
    s = orc.slope([0, 0, 0, 0], [0.1, 0.1, 0.1, 0.1], [2, 2, 2, 2], h, al)  # one-sided limits
    np.testing.assert_allclose(s, 0.1 * 1.9, rtol=1e-15)


# ----------------------------------------------------------------- conversions
def test_cons_prim_closed_form_and_roundtrip(orc):
    rng = np.random.default_rng(3)
    for _ in range(200):
        q = np.array([rng.uniform(0.05, 5), rng.uniform(0.01, 10), rng.normal(), rng.normal()])
        c = orc.cons(q, G)
        np.testing.assert_allclose(
            c, [q[0], q[0] * q[2], q[1] / (G - 1) + 0.5 * q[0] * (q[2] ** 2 + q[3] ** 2), q[0] * q[3]],
            rtol=1e-15)
        np.testing.assert_allclose(orc.prim(c, G), q, rtol=1e-12, atol=1e-14)


def test_speed_closed_form(orc):
    q = np.array([0.138, 0.029, 1.206, 1.206])
    assert orc.speed(q, G) == pytest.approx(math.hypot(1.206, 1.206) + math.sqrt(G * 0.029 / 0.138),
                                            rel=1e-15)


# ----------------------------------------------------------------- GRP
def euler_prim_dt(W, S):
    """Primitive Euler time derivatives -A(W) W_x (rho, p, un, ut, e)."""
    r, p, u, v = W
    rx, px, ux, vx = S
    e_x = (px / r - p * rx / r ** 2) / (G - 1)
    return np.array([-(r * ux + u * rx), -(u * px + G * p * ux), -(u * ux + px / r), -u * vx,
                     -(u * e_x + p / r * ux)])


def test_grp_identical_states_zero_slopes(orc):
    W = np.array([1.3, 2.1, 0.4, -0.7])
    Wout, Wt, _ = orc.grp(W, W, np.zeros(4), np.zeros(4), G)
    np.testing.assert_allclose(Wout[:4], W, rtol=1e-14)
    assert Wout[4] == pytest.approx(W[1] / W[0] / (G - 1), rel=1e-14)
    np.testing.assert_allclose(Wt, 0.0, atol=1e-14)


@pytest.mark.parametrize("side", ["left", "right"])
def test_grp_supersonic_is_upwind_euler(orc, side):
    """Supersonic interfaces return the upwind face state and the primitive
    Euler derivatives of the upwind slopes (src/basicFuns.cuh:129-146)."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        WL = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), rng.uniform(3, 5), rng.normal()])
        WR = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), rng.uniform(3, 5), rng.normal()])
        SL, SR = rng.normal(size=4), rng.normal(size=4)
        if side == "right":
            WL[2], WR[2] = -WL[2], -WR[2]
        W, Wt, br = orc.grp(WL, WR, SL, SR, G)
        up, S = (WL, SL) if side == "left" else (WR, SR)
        assert br == (0 if side == "left" else 1)
        np.testing.assert_allclose(W[:4], up, rtol=1e-15)
        exp = euler_prim_dt(up, S)
        # pressure: -(u p_x + c^2/rho u_x) with the relaxation c; equals
        # gamma p u_x only when the relaxation max-term vanishes, so compare
        # the other components exactly and p_t via its formula bounds
        np.testing.assert_allclose(Wt[[0, 2, 3, 4]], exp[[0, 2, 3, 4]], rtol=1e-12, atol=1e-12)
        # continuous data: the relaxation speed is rho a and p_t is Euler's
        W, Wt, _ = orc.grp(up, up, S, S, G)
        np.testing.assert_allclose(Wt, euler_prim_dt(up, S), rtol=1e-12, atol=1e-12)


def continuous_star_dt(W, S):
    """Star-region GRP time derivatives for continuous data (W_L = W_R,
    equal slopes) derived from src/basicFuns.cuh:149-185, including the
    reference's use of u = 0 inside k_l, k_r, d_l, d_r: the Euler
    derivatives plus the terms proportional to u c_x that this produces."""
    r, p, u, v = W
    rx, px, ux, vx = S
    c = math.sqrt(G * p * r)                        # Lagrangian sound speed
    cx = 0.5 * math.sqrt(G / p / r) * (p * rx + r * px)
    base = euler_prim_dt(W, S)
    a = u * r / c
    d_r = -r * u * cx / (2 * c) * ((1 + a) if u <= 0 else (1 - a))
    d_p = -c * u * cx / (2 * r)
    d_u = -u * u * cx / (2 * c)
    d_e = -p * u * cx / (2 * r * c)
    return base + np.array([d_r, d_p, d_u, 0.0, d_e])


def test_grp_star_continuous_limit(orc):
    """W_L = W_R, subsonic: the GRP returns the state itself and the time
    derivatives of the closed form above for both signs of u*."""
    rng = np.random.default_rng(9)
    for _ in range(300):
        W = np.array([rng.uniform(0.3, 3), rng.uniform(0.3, 3), 0.0, rng.normal()])
        W[2] = rng.uniform(-0.9, 0.9) * math.sqrt(G * W[1] / W[0])  # subsonic
        S = rng.normal(size=4)
        Wout, Wt, br = orc.grp(W, W, S, S, G)
        assert br in (2, 3) and br == (2 if W[2] <= 0 else 3)
        np.testing.assert_allclose(Wout[:4], W, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(Wt, continuous_star_dt(W, S), rtol=1e-11, atol=1e-11)


def test_grp_star_values_are_relaxation_solution(orc):
    """Star-region values for discontinuous data: u*, p* of the Suliciu
    relaxation solver with Bouchut's wave-speed choice, 1/rho* = 1/rho_K +
    (u* - u_K)/c_K on the upwind side, and the outer-wave invariant
    e - p^2/(2 c^2)."""
    rng = np.random.default_rng(13)
    kap = (G + 1) / 2
    n = 0
    while n < 300:
        WL = np.array([rng.uniform(0.2, 3), rng.uniform(0.2, 3), rng.uniform(-1, 1), rng.normal()])
        WR = np.array([rng.uniform(0.2, 3), rng.uniform(0.2, 3), rng.uniform(-1, 1), rng.normal()])
        W, _, br = orc.grp(WL, WR, rng.normal(size=4), rng.normal(size=4), G)
        if br < 2:
            continue
        n += 1
        rl, pl, ul, vl = WL
        rr, pr, ur, vr = WR
        al, ar = math.sqrt(G * pl / rl), math.sqrt(G * pr / rr)
        if pr >= pl:
            cl = rl * (al + kap * max(0.0, (pr - pl) / (rr * ar) + ul - ur))
            cr = rr * (ar + kap * max(0.0, (pl - pr) / cl + ul - ur))
        else:
            cr = rr * (ar + kap * max(0.0, (pl - pr) / (rl * al) + ul - ur))
            cl = rl * (al + kap * max(0.0, (pr - pl) / cr + ul - ur))
        us = (cl * ul + cr * ur + pl - pr) / (cl + cr)
        ps = (cr * pl + cl * pr + cl * cr * (ul - ur)) / (cl + cr)
        assert W[2] == pytest.approx(us, rel=1e-12, abs=1e-13)
        assert W[1] == pytest.approx(ps, rel=1e-12)
        if us <= 0:
            rK, pK, uK, vK, cK, sgn = rr, pr, ur, vr, cr, -1.0
        else:
            rK, pK, uK, vK, cK, sgn = rl, pl, ul, vl, cl, 1.0
        assert 1.0 / W[0] == pytest.approx(1.0 / rK + sgn * (us - uK) / cK, rel=1e-12)
        eK = pK / rK / (G - 1)
        assert W[4] - ps ** 2 / (2 * cK ** 2) == pytest.approx(eK - pK ** 2 / (2 * cK ** 2), rel=1e-12)
        assert W[3] == vK


def test_grp_mirror_symmetry(orc):
    """Reflecting x -> -x (swap sides, negate normal velocity and every
    slope except that of un) mirrors the interface solution."""
    rng = np.random.default_rng(21)
    flip = np.array([1, 1, -1, 1])
    for _ in range(300):
        WL = np.array([rng.uniform(0.2, 3), rng.uniform(0.2, 3), rng.uniform(-2, 2), rng.normal()])
        WR = np.array([rng.uniform(0.2, 3), rng.uniform(0.2, 3), rng.uniform(-2, 2), rng.normal()])
        SL, SR = rng.normal(size=4), rng.normal(size=4)
        W, Wt, br = orc.grp(WL, WR, SL, SR, G)
        if abs(W[2]) < 1e-9:
            continue
        Wm, Wtm, brm = orc.grp(WR * flip, WL * flip, -SR * flip, -SL * flip, G)
        mirror = np.array([1, 1, -1, 1, 1])
        np.testing.assert_allclose(Wm, W * mirror, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(Wtm[1:], (Wt * mirror)[1:], rtol=1e-9, atol=1e-9)
        # rho_t: the 1.5 (u* - u_K) c_Kx term of src/basicFuns.cuh:165,179 has
        # the same sign in both star branches, so mirroring the problem moves
        # rho_t by 3 u* rho*^3 (u* - u_K) c_Kx / (rho_K c_K^2), K = upwind side
        asym = 0.0
        if br >= 2:
            cl, cr = relax_speeds(WL, WR)
            K, SK, cK = (WR, SR, cr) if br == 2 else (WL, SL, cl)
            cKx = 0.5 * math.sqrt(G / K[1] / K[0]) * (K[1] * SK[0] + K[0] * SK[1])
            asym = 3 * W[2] * W[0] ** 3 * (W[2] - K[2]) * cKx / (K[0] * cK ** 2)
        assert Wtm[0] - Wt[0] == pytest.approx(asym, rel=1e-8, abs=1e-9)


def relax_speeds(WL, WR):
    """Bouchut's relaxation speeds (Lagrangian) used by the GRP."""
    kap = (G + 1) / 2
    rl, pl, ul, _ = WL
    rr, pr, ur, _ = WR
    al, ar = math.sqrt(G * pl / rl), math.sqrt(G * pr / rr)
    if pr >= pl:
        cl = rl * (al + kap * max(0.0, (pr - pl) / (rr * ar) + ul - ur))
        cr = rr * (ar + kap * max(0.0, (pl - pr) / cl + ul - ur))
    else:
        cr = rr * (ar + kap * max(0.0, (pl - pr) / (rl * al) + ul - ur))
        cl = rl * (al + kap * max(0.0, (pr - pl) / cr + ul - ur))
    return cl, cr


# ----------------------------------------------------------------- ghost fill
def emulate_reference_fill(U, strategy, t, order_rng):
    """Runs a reference ghost kernel (src/ghostCellPresets.cuh:47-241) under
    the CUDA execution model it was written for: 16x16 blocks, warps of 16
    columns x 2 rows executing each statement in lockstep (every active
    thread loads, then every active thread stores), warps in random order."""
    U = U.copy()
    Ny, Nx = U.shape[:2]
    post = np.array([8, 116.5, 4.125 * math.sqrt(3.0), -4.125, 0.0])
    pre = np.array([1.4, 1.0, 0.0, 0.0, 0.0])
    warps = [(bx, by, w) for by in range((Ny + 15) // 16) for bx in range((Nx + 15) // 16)
             for w in range(8)]
    order_rng.shuffle(warps)
    flat = U.reshape(-1, 5)

    def lockstep(threads, stmt):
        vals = []
        for (i, j) in threads:
            r = stmt(i, j)
            if r is not None:
                vals.append(r)
        for idx, val in vals:
            flat[idx] = val

    for bx, by, w in warps:
        threads = [(bx * 16 + x, by * 16 + 2 * w + y) for y in range(2) for x in range(16)]
        threads = [(i, j) for (i, j) in threads if i < Nx and j < Ny]

        def row_phase(i, j):
            idx = j * Nx + i
            if strategy in (0, 2, 3):
                if j == 0:
                    return idx, flat[idx + 2 * Nx].copy()
                if j == 1:
                    return idx, flat[idx + Nx].copy()
                if j == Ny - 2:
                    return idx, flat[idx - Nx].copy()
                if j == Ny - 1:
                    return idx, flat[idx - (3 if strategy == 3 else 2) * Nx].copy()
                return None
            if j in (0, 1):
                temp = float(Ny) / math.sqrt(3.0)
                dist = float(Ny) * 20 * t / math.sqrt(3.0)
                return idx, (post if i <= Nx // 24 + int(temp) + int(dist) else pre).copy()
            if j == Ny - 1:
                src = flat[idx - (2 if i <= Nx // 24 else 3) * Nx].copy()
                if i > Nx // 24:
                    src[3] = -src[3]
                return idx, src
            if j == Ny - 2:
                src = flat[idx - Nx].copy()
                if i > Nx // 24:
                    src[3] = -src[3]
                return idx, src
            return None

        def col_phase(i, j):
            idx = j * Nx + i
            if strategy == 0:
                off = {0: 2, 1: 1, Nx - 2: -1, Nx - 1: -2}.get(i)
                return None if off is None else (idx, flat[idx + off].copy())
            if strategy in (2, 3):
                if i in (0, 1):
                    if strategy == 2:
                        on = 2 * Ny // 5 + 1 <= j <= 3 * Ny // 5
                        return idx, np.array([5.0, 0.4127, 800.0, 0, 0] if on else [0.5, 0.4127, 0, 0, 0])
                    on = j >= 9 * Ny // 10
                    return idx, np.array([1.4, 1, 800, 0, 0] if on else [0.14, 1, 0, 0, 0], dtype=float)
                off = {Nx - 2: -1, Nx - 1: -2}.get(i)
                return None if off is None else (idx, flat[idx + off].copy())
            if i in (0, 1):
                return idx, post.copy()
            if i in (Nx - 2, Nx - 1):
                return idx, pre.copy()
            return None

        lockstep(threads, row_phase)
        lockstep(threads, col_phase)
    return U


@pytest.mark.parametrize("strategy", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("shape", [(32, 32), (48, 32), (32, 64)])
def test_fill_matches_warp_lockstep_emulation(orc, strategy, shape):
    rng = np.random.default_rng(strategy * 7 + shape[1])
    U = rng.uniform(0.5, 2.0, size=shape + (5,))
    U[..., 4] = 0.0
    t = 0.0123
    exp = emulate_reference_fill(U, strategy, t, np.random.default_rng(1))
    exp2 = emulate_reference_fill(U, strategy, t, np.random.default_rng(2))
    np.testing.assert_array_equal(exp, exp2)  # warp order does not matter
    got = orc.fill(U, strategy, t)
    np.testing.assert_array_equal(got[..., :4], exp[..., :4])


# ----------------------------------------------------------------- rotation
def test_rotation_is_involution_with_uv_swap(orc):
    rng = np.random.default_rng(4)
    U = rng.normal(size=(7, 11, 5))
    R = orc.rotate(U)
    assert R.shape == (11, 7, 5)
    np.testing.assert_array_equal(orc.rotate(R)[..., :4], U[..., :4])
    # out[(Nx-1-x)][(Ny-1-y)] = in[y][x] with u, v swapped
    np.testing.assert_array_equal(R[11 - 1 - 3, 7 - 1 - 2, [0, 1, 3, 2]], U[2, 3, :4])


def test_solve_is_rotation_equivariant(orc, problems):
    """solve(R U0) = R solve(U0) bit for bit on a square mesh with equal
    width and height: the Y-sweep is the X-sweep of the rotated array."""
    rng = problems.SplitMix64(3)
    U0 = problems.random_four_quadrant(rng, 40, 40)
    a, na, _, _ = orc.solve(U0, nthreads=2, t_end=0.05)
    b, nb, _, _ = orc.solve(orc.rotate(U0), nthreads=2, t_end=0.05)
    assert na == nb
    np.testing.assert_array_equal(orc.rotate(a)[..., :4], b[..., :4])


def test_4shocks_2000_style_symmetry(orc, problems):
    """At N = 1000 (N/5 + 4N/5 = N) the four-quadrant data is symmetric under
    the rotation, hence so is every step of the solve (bitwise)."""
    U0 = problems.initialize_4shocks(40)
    np.testing.assert_array_equal(orc.rotate(U0)[..., :4], U0[..., :4])
    U, n, _, _ = orc.solve(U0, nthreads=2, t_end=1.0, max_steps=6)
    np.testing.assert_array_equal(orc.rotate(U)[..., :4], U[..., :4])


# ----------------------------------------------------------------- invariants
def test_uniform_state_stays_uniform(orc):
    U0 = np.zeros((24, 20, 5))
    U0[...] = [1.2, 0.8, 0.3, -0.4, 0.0]
    U, n, _, _ = orc.solve(U0, nthreads=1, t_end=1.0, max_steps=10)
    assert n == 10
    np.testing.assert_allclose(U[..., :4], U0[..., :4], rtol=1e-14)


def test_y_independent_data_keeps_rows_identical(orc, problems):
    row = np.zeros((40, 5))
    row[:20] = [1.0, 1.0, 0.0, 0.0, 0.0]
    row[20:] = [0.125, 0.1, 0.0, 0.0, 0.0]
    U0 = np.ascontiguousarray(np.broadcast_to(row, (16, 40, 5)))
    U, n, _, _ = orc.solve(U0, nthreads=1, t_end=0.1)
    assert n > 5
    for j in range(1, 16):
        np.testing.assert_array_equal(U[j], U[0])
    assert np.all(U[..., 3] == 0.0)


def test_interior_mass_conserved_before_waves_reach_ghosts(orc, problems):
    U0 = np.zeros((80, 80, 5))
    U0[...] = [1.0, 1.0, 0.2, -0.1, 0.0]
    U0[35:45, 30:50] = [2.0, 3.0, -0.3, 0.4, 0.0]  # blob away from the ghosts
    U, n, _, _ = orc.solve(U0, nthreads=2, t_end=1.0, max_steps=5)
    m0 = U0[2:-2, 2:-2, 0].sum()
    m1 = U[2:-2, 2:-2, 0].sum()
    assert abs(m1 - m0) / m0 < 1e-13


def test_step_count_fourquad_256(orc, problems):
    """256^2 four-quadrant to t_end = 0.3 takes 403 steps and stays
    positive; the final-state digest is the committed regression fixture."""
    import hashlib
    import json
    import os
    U0 = problems.initialize_4shocks(256)
    U, n, t, _ = orc.solve(U0, nthreads=0, t_end=0.3)
    assert n == 403 and t == pytest.approx(0.3, abs=0)
    assert np.all(U[..., 0] > 0) and np.all(U[..., 1] > 0)
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fourquad_256.json")))
    assert fx["steps"] == n
    assert hashlib.sha256(np.ascontiguousarray(U[..., :4]).tobytes()).hexdigest() == fx["sha256_prim4"]


# ----------------------------------------------------------------- GRP time derivatives
# The pins below fix the GRP's interface time derivatives for discontinuous
# data (src/basicFuns.cuh:129-185).  Branch labels: 0 left supersonic,
# 1 right supersonic, 2 star region with u* <= 0, 3 star region with u* > 0.

def _rand_state(rng, umax):
    return np.array([rng.uniform(0.2, 3), rng.uniform(0.2, 3), rng.uniform(-umax, umax),
                     rng.normal()])


def _grp_bins(orc, rng, per_bin, slope_fn=None):
    """Random interface problems, per_bin of each (branch, p_R >= p_L) pair:
    all four branches under both pressure orderings."""
    bins = {(b, o): [] for b in range(4) for o in (False, True)}
    tries = 0
    while any(len(v) < per_bin for v in bins.values()):
        tries += 1
        assert tries < 200000, "sampler cannot reach every branch"
        umax = rng.choice([1.5, 6.0])
        WL, WR = _rand_state(rng, umax), _rand_state(rng, umax)
        if slope_fn is None:
            SL, SR = rng.normal(size=4), rng.normal(size=4)
        else:
            SL, SR = slope_fn(rng, WL), slope_fn(rng, WR)
        _, _, br = orc.grp(WL, WR, SL, SR, G)
        key = (br, bool(WR[1] >= WL[1]))
        if len(bins[key]) < per_bin:
            bins[key].append((WL, WR, SL, SR))
    return bins


def test_grp_zero_slopes_give_zero_time_derivatives(orc):
    """(a) Piecewise-constant data: every branch returns W_t = 0 for any
    jump, under both pressure orderings (the relaxation Riemann solution is
    self-similar, so the interface state does not change in time)."""
    rng = np.random.default_rng(101)
    z = np.zeros(4)
    for key, cases in _grp_bins(orc, rng, 40).items():
        for WL, WR, _, _ in cases:
            W, Wt, br = orc.grp(WL, WR, z, z, G)
            assert br == key[0]
            np.testing.assert_array_equal(Wt, 0.0)


def test_grp_time_derivatives_linear_in_slopes(orc):
    """(b) For fixed face states the branch, the relaxation speeds and the
    star values do not depend on the slopes, and W_t is linear in the 16
    slope inputs: W_t(a S1 + b S2) = a W_t(S1) + b W_t(S2), and the 5 x 16
    matrix assembled from unit slopes reproduces W_t for random slopes."""
    rng = np.random.default_rng(102)
    for key, cases in _grp_bins(orc, rng, 25).items():
        for WL, WR, SL, SR in cases:
            W0, _, _ = orc.grp(WL, WR, np.zeros(4), np.zeros(4), G)
            J = np.zeros((5, 8))
            for k in range(8):
                e = np.zeros(8)
                e[k] = 1.0
                Wk, J[:, k], brk = orc.grp(WL, WR, e[:4], e[4:], G)
                assert brk == key[0]
                np.testing.assert_array_equal(Wk, W0)
            S = np.concatenate([SL, SR])
            a, b = rng.uniform(-3, 3, size=2)
            S2 = rng.normal(size=8)
            _, Wt1, _ = orc.grp(WL, WR, SL, SR, G)
            _, Wt2, _ = orc.grp(WL, WR, S2[:4], S2[4:], G)
            _, Wt12, _ = orc.grp(WL, WR, a * SL + b * S2[:4], a * SR + b * S2[4:], G)
            scale = np.abs(J).sum(axis=1) * (abs(a) + abs(b)) * 4 + 1e-300
            assert np.all(np.abs(Wt12 - (a * Wt1 + b * Wt2)) <= 1e-13 * scale)
            assert np.all(np.abs(Wt1 - J @ S) <= 1e-13 * (np.abs(J) @ np.abs(S) + 1e-300))


@pytest.mark.parametrize("exact_pow2", [True, False])
def test_grp_dimensional_scaling(orc, exact_pow2):
    """(c) Scaling laws of the Euler equations that the relaxation GRP keeps:
    (rho, p) -> lam (rho, p) with (rho_x, p_x) scaled alike leaves u, v, e
    and their derivatives unchanged and scales rho_t, p_t by lam;
    (u, v, p) -> (kap u, kap v, kap^2 p) with slopes scaled alike scales
    (rho, p, u, v, e)_t by (kap, kap^3, kap^2, kap^2, kap^3).  With powers of
    two every operation commutes with the scaling, so the results are equal
    bit for bit."""
    rng = np.random.default_rng(103 + exact_pow2)
    for key, cases in _grp_bins(orc, rng, 20).items():
        for WL, WR, SL, SR in cases:
            W, Wt, br = orc.grp(WL, WR, SL, SR, G)
            lam = 2.0 ** rng.integers(-3, 4) if exact_pow2 else rng.uniform(0.3, 3.0)
            kap = 2.0 ** rng.integers(-3, 4) if exact_pow2 else rng.uniform(0.3, 3.0)
            m = np.array([lam, lam, 1, 1])
            Wl, Wtl, brl = orc.grp(WL * m, WR * m, SL * m, SR * m, G)
            mv = np.array([1, kap * kap, kap, kap])
            Wk, Wtk, brk = orc.grp(WL * mv, WR * mv, SL * mv, SR * mv, G)
            assert brl == br and brk == br
            expW_l, expWt_l = W * np.array([lam, lam, 1, 1, 1]), Wt * np.array([lam, lam, 1, 1, 1])
            expW_k = W * np.array([1, kap ** 2, kap, kap, kap ** 2])
            expWt_k = Wt * np.array([kap, kap ** 3, kap ** 2, kap ** 2, kap ** 3])
            if exact_pow2:
                np.testing.assert_array_equal(Wl, expW_l)
                np.testing.assert_array_equal(Wtl, expWt_l)
                np.testing.assert_array_equal(Wk, expW_k)
                np.testing.assert_array_equal(Wtk, expWt_k)
            else:
                for got, exp in ((Wl, expW_l), (Wtl, expWt_l), (Wk, expW_k), (Wtk, expWt_k)):
                    np.testing.assert_allclose(got, exp, rtol=1e-11,
                                               atol=1e-12 * np.abs(exp).max())


def _acoustic_interface(WL, WR, SL, SR):
    """Linear acoustics at a material interface at rest (u = 0 on both sides,
    equal pressure): the right-going invariant p + Z_l u arrives from the
    left and the left-going p - Z_r u from the right (Z = rho a), so with
    k_l = -(p_x + Z u_x)_l and k_r = -(p_x - Z u_x)_r
    p_t = (Z_r a_l k_l - Z_l a_r k_r)/(Z_l + Z_r),
    u_t = (a_l k_l + a_r k_r)/(Z_l + Z_r); the interface is a stationary
    particle path on the right side (u* = 0 selects it), where entropy is
    constant in time: rho_t = p_t / a_r^2 and e_t = p_t / (gamma rho_r)."""
    rl, p, _, _ = WL
    rr, _, _, vr = WR
    al, ar = math.sqrt(G * p / rl), math.sqrt(G * p / rr)
    Zl, Zr = rl * al, rr * ar
    kl = -(SL[1] + Zl * SL[2])
    kr = -(SR[1] - Zr * SR[2])
    pt = (Zr * al * kl - Zl * ar * kr) / (Zl + Zr)
    ut = (al * kl + ar * kr) / (Zl + Zr)
    return (np.array([rr, p, 0.0, vr, p / rr / (G - 1)]),
            np.array([pt / ar ** 2, pt, ut, 0.0, pt / (G * rr)]))


@pytest.mark.parametrize("mismatch", [False, True])
def test_grp_acoustic_limit(orc, mismatch):
    """(d) u_L = u_R = 0, equal pressures, slopes S_L != S_R: the GRP gives
    the characteristic-upwind solution of linear acoustics.  Without
    mismatch W_L = W_R (p_t = a (k_l - k_r)/2, u_t = (k_l + k_r)/(2 rho));
    with mismatch the densities differ, so the two sides have different
    impedances and every left/right operand of the star formulas is
    distinguishable."""
    rng = np.random.default_rng(104 + mismatch)
    for _ in range(300):
        rl, p = rng.uniform(0.2, 3), rng.uniform(0.2, 3)
        rr = rng.uniform(0.2, 3) if mismatch else rl
        WL = np.array([rl, p, 0.0, rng.normal()])
        WR = np.array([rr, p, 0.0, rng.normal() if mismatch else WL[3]])
        SL, SR = rng.normal(size=4), rng.normal(size=4)
        W, Wt, br = orc.grp(WL, WR, SL, SR, G)
        assert br == 2  # u* = 0 exactly: the right side of the contact
        expW, expWt = _acoustic_interface(WL, WR, SL, SR)
        np.testing.assert_allclose(W, expW, rtol=1e-14)
        np.testing.assert_allclose(Wt, expWt, rtol=1e-11, atol=1e-12 * np.abs(Wt).max())


def _relaxation_interface(WL, WR, SL, SR, cl, cr):
    """Interface values and Eulerian time derivatives at x = 0 of the
    Lagrangian relaxation system (Suliciu; Bouchut 2004) with the impedance
    c frozen on each side:  tau_t - u_m = 0, u_t + pi_m = 0,
    pi_t + c^2 u_m = 0 (m: mass coordinate, dm = rho dx), v_t = 0 and
    (e - pi^2 / 2c^2)_t = 0, for piecewise-linear data.  Solved by
    characteristics from the numerical eigen-decomposition of each side's
    3 x 3 system:
      * supersonic: the upwind data itself, q_t = -A q_m - rho u q_m;
      * star region: the invariants arriving from the data fix (u*, pi*)
        and their material derivatives at the contact, the zero-speed field
        fixes tau* on the contact's side K, the field reflected into side K
        is carried from the contact, and the Eulerian derivative at the
        fixed interface is D/Dt - rho* u* d/dm on side K (the contact moves
        with u*, K = left if u* > 0)."""
    def side(W, S, c):
        r, p, u, v = W
        rx, px, ux, vx = S
        z = np.array([1.0 / r, u, p])
        zm = np.array([-rx / r ** 2, ux, px]) / r  # d/dm = (1/rho) d/dx
        A = np.array([[0.0, -1.0, 0.0], [0.0, 0.0, 1.0], [0.0, c * c, 0.0]])
        lam, Rv = np.linalg.eig(A)
        order = np.argsort(lam.real)  # -c, 0, +c
        lam, Rv = lam.real[order], Rv.real[:, order]
        L = np.linalg.inv(Rv)
        ex = (px / r - p * rx / r ** 2) / (G - 1)
        return dict(r=r, p=p, u=u, v=v, c=c, z=z, zm=zm, A=A, lam=lam, R=Rv, L=L,
                    e=p / r / (G - 1), em=ex / r, vm=vx / r)

    sl, sr = side(WL, SL, cl), side(WR, SR, cr)
    if WL[2] - cl / WL[0] >= 0.0 or WR[2] + cr / WR[0] <= 0.0:
        K = sl if WL[2] - cl / WL[0] >= 0.0 else sr
        r, u, p = K["r"], K["u"], K["p"]
        Dz = -K["A"] @ K["zm"]
        zt = Dz - r * u * K["zm"]
        De = -p * K["zm"][1]           # e_t = -pi u_m
        et = De - r * u * K["em"]
        W = np.array([r, p, u, K["v"], K["e"]])
        return W, np.array([-r * r * zt[0], zt[2], zt[1], -r * u * K["vm"], et])
    # incoming invariants: +c from the left data, -c from the right data;
    # neither involves tau (checked), so they fix (u*, pi*) and D/Dt of both
    lp, lm = sl["L"][2], sr["L"][0]
    assert abs(lp[0]) < 1e-12 * np.abs(lp).max() and abs(lm[0]) < 1e-12 * np.abs(lm).max()
    M = np.array([lp[1:], lm[1:]])
    us, ps = np.linalg.solve(M, [lp @ sl["z"], lm @ sr["z"]])
    Du, Dp = np.linalg.solve(M, [-sl["lam"][2] * (lp @ sl["zm"]), -sr["lam"][0] * (lm @ sr["zm"])])
    K = sl if us > 0.0 else sr
    l0 = K["L"][1]                    # zero-speed field: stationary on each side
    taus = (l0 @ K["z"] - l0[1] * us - l0[2] * ps) / l0[0]
    Dtau = -(l0[1] * Du + l0[2] * Dp) / l0[0]
    Dz = np.array([Dtau, Du, Dp])
    # d/dm of the characteristic fields in the star region on side K
    wm = K["L"] @ K["zm"]             # fields arriving from the data
    refl = 2 if K is sr else 0        # field carried from the contact into K
    wm[refl] = -(K["L"][refl] @ Dz) / K["lam"][refl]
    zm = K["R"] @ wm
    rs = 1.0 / taus
    zt = Dz - rs * us * zm
    c2 = K["c"] ** 2
    es = K["e"] + (ps * ps - K["p"] ** 2) / (2 * c2)
    em = (K["em"] - K["p"] * K["zm"][2] / c2) + ps * zm[2] / c2
    et = ps / c2 * Dp - rs * us * em
    W = np.array([rs, ps, us, K["v"], es])
    return W, np.array([-rs * rs * zt[0], zt[2], zt[1], -rs * us * K["vm"], et])


def test_grp_matches_relaxation_system_with_frozen_impedance(orc):
    """Every branch, discontinuous data, u* of either sign, both pressure
    orderings, relaxation max-terms on and off: where the slopes leave the
    Lagrangian sound speed unchanged (p rho_x + rho p_x = 0, i.e. c_x = 0 in
    src/basicFuns.cuh:104-105) the GRP's values and time derivatives are
    those of the relaxation system solved by characteristics above.  This
    fixes every star-region term of src/basicFuns.cuh:147-185 except the
    ones proportional to c_x; left/right operand swaps, dropped terms and
    wrong powers of c or rho all fail it."""
    rng = np.random.default_rng(105)

    def frozen(rng, W):
        s = rng.normal(size=4)
        s[0] = -W[0] * s[1] / W[1]
        return s

    nmax = 0
    for key, cases in _grp_bins(orc, rng, 60, slope_fn=frozen).items():
        for WL, WR, SL, SR in cases:
            W, Wt, br = orc.grp(WL, WR, SL, SR, G)
            cl, cr = relax_speeds(WL, WR)
            kap = (G + 1) / 2
            active = (WR[1] - WL[1]) / (WR[0] * math.sqrt(G * WR[1] / WR[0])) + WL[2] - WR[2] > 0
            nmax += bool(active)
            expW, expWt = _relaxation_interface(WL, WR, SL, SR, cl, cr)
            np.testing.assert_allclose(W, expW, rtol=1e-11, atol=1e-13)
            np.testing.assert_allclose(Wt, expWt, rtol=1e-9, atol=1e-10 * np.abs(expWt).max())
    assert nmax > 50  # the relaxation max-term was active in many cases
