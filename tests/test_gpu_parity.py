"""GPU parity of the fused step kernel against the FP64 host oracle.

EXACT mode (reference op order, no FMA contraction) must reproduce the oracle
bit for bit; FAST mode (shared reciprocals, FMA) must stay within relative
L1 <= 1e-10 on (rho, rho u, rho v, E) -- the tolerance the plain solve is
held to.  All calls go through the C ABI of libfvm2d.so.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import rel_errors

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-10


def _cases(problems):
    rng = problems.SplitMix64(7)
    cases = []
    # (name, U0, params)
    cases.append(("4shocks64", problems.initialize_4shocks(64), dict(t_end=0.05)))
    cases.append(("rand37x23", problems.random_four_quadrant(rng, 37, 23), dict(t_end=1.0, max_steps=6)))
    cases.append(("rand6x6", problems.random_four_quadrant(rng, 6, 6), dict(t_end=1.0, max_steps=4)))
    cases.append(("rand65x131", problems.random_four_quadrant(rng, 65, 131), dict(t_end=1.0, max_steps=5)))
    cases.append(("rand125x40", problems.random_four_quadrant(rng, 125, 40), dict(t_end=1.0, max_steps=5)))
    cases.append(("2mach24", problems.initialize_2Mach(24), dict(t_end=0.01, max_steps=12, strategy=4,
                                                                   width=4.0, height=1.0)))
    cases.append(("reflective4cds", problems.initialize_4cds(40), dict(t_end=0.05, max_steps=10, strategy=1)))
    cases.append(("jet30", problems.initialize_jet(30), dict(t_end=1e-3, max_steps=8, strategy=2,
                                                               width=2.0, height=1.0)))
    cases.append(("jet800_20", problems.initialize_jet800(20), dict(t_end=1e-3, max_steps=8, strategy=3,
                                                                      width=3.0, height=1.0)))
    cases.append(("rp70x130", problems.set_rp_initial_data(
        [1.0, 1.0, 0.75, -0.5, 0], [2.0, 1.0, 0.75, 0.5, 0], [1.0, 1.0, -0.75, 0.5, 0],
        [3.0, 1.0, -0.75, -0.5, 0], 70, 130), dict(t_end=0.02, alpha=1.7, cfl=0.4)))
    # pressure ratio 1e18 (test2D.cu:118-133 on [0,8]^2): supersonic
    # interfaces on both sides of the blast
    cases.append(("sedov34", problems.initialize_Sedov1D(34), dict(t_end=1.0, max_steps=10, width=8.0,
                                                                     height=8.0)))
    # setGamma / setAlpha away from the defaults (src/wrappers.cuh:132-140):
    # a monatomic gas, gamma = 3 with the minmod limiter, the demo's MC
    # limiter alpha = 2 (test2D.cu:141), and double Mach at gamma = 1.1
    cases.append(("gamma53", problems.random_four_quadrant(rng, 70, 52),
                  dict(t_end=1.0, max_steps=8, gamma=5.0 / 3.0)))
    cases.append(("gamma3_alpha1", problems.random_four_quadrant(rng, 45, 61),
                  dict(t_end=1.0, max_steps=6, gamma=3.0, alpha=1.0)))
    cases.append(("alpha2_4shocks96", problems.initialize_4shocks(96), dict(t_end=0.05, alpha=2.0)))
    cases.append(("gamma11_2mach24", problems.initialize_2Mach(24),
                  dict(t_end=0.01, max_steps=10, strategy=4, width=4.0, height=1.0, gamma=1.1)))
    # extreme aspect ratios: two interior columns over many strips, and three
    # interior rows over many windows
    cases.append(("rand6x3001", problems.random_four_quadrant(rng, 6, 3001),
                  dict(t_end=1.0, max_steps=7, width=0.01, height=5.0)))
    cases.append(("rand2999x7", problems.random_four_quadrant(rng, 2999, 7),
                  dict(t_end=1.0, max_steps=7, width=5.0, height=0.01)))
    return cases


NCASES = 17


@pytest.fixture(scope="module")
def cases(problems):
    return _cases(problems)


@pytest.mark.parametrize("idx", range(NCASES))
def test_exact_mode_bitwise(gpu, orc, cases, idx):
    name, U0, prm = cases[idx]
    print(f"case {name} shape {U0.shape} {prm}", flush=True)
    ref, steps_ref, _, _ = orc.solve(U0, nthreads=4, **prm)
    got, steps, _ = gpu.solve(U0, mode=gpu.MODE_EXACT, **prm)
    assert steps == steps_ref
    diff = np.argwhere(got[..., :4] != ref[..., :4])
    assert diff.size == 0, f"{name}: {len(diff)} differing values, first {diff[:5].tolist()}"


@pytest.mark.parametrize("idx", range(NCASES))
def test_fast_mode_parity(gpu, orc, cases, idx):
    name, U0, prm = cases[idx]
    print(f"case {name} shape {U0.shape} {prm}", flush=True)
    ref, steps_ref, _, _ = orc.solve(U0, nthreads=4, **prm)
    got, steps, _ = gpu.solve(U0, mode=gpu.MODE_FAST, **prm)
    l1, linf = rel_errors(got, ref, prm.get("gamma", 1.4))
    print(f"{name}: steps {steps}/{steps_ref} relL1 {l1} relLinf {linf}")
    assert abs(steps - steps_ref) <= 1
    assert np.all(l1 <= TOL_L1), (name, l1, linf)


def test_fourquad_256(gpu, orc, problems):
    """256^2 four-quadrant to t_end = 0.3 (the reference's host-runnable case)."""
    U0 = problems.initialize_4shocks(256)
    prm = dict(t_end=0.3)
    ref, steps_ref, t_ref, _ = orc.solve(U0, nthreads=0, **prm)
    ex, steps_ex, _ = gpu.solve(U0, mode=gpu.MODE_EXACT, **prm)
    assert steps_ex == steps_ref == 403
    assert np.array_equal(ex[..., :4], ref[..., :4])
    fa, steps_fa, _ = gpu.solve(U0, mode=gpu.MODE_FAST, **prm)
    l1, linf = rel_errors(fa, ref)
    print(f"256^2 fast: steps {steps_fa} relL1 {l1} relLinf {linf}")
    assert steps_fa == steps_ref
    assert np.all(l1 <= TOL_L1)


def test_fourquad_2000(gpu, problems):
    """2000^2 four-quadrant to t_end = 0.3, the headline workload, against
    the oracle's committed record (tests/golden/make_fourquad_2000.py): EXACT mode
    bit for bit (step count, SHA-256 of the final (rho, p, u, v) planes and
    every 16th cell), FAST mode within relative L1 1e-10 of it."""
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    golden = json.load(open(os.path.join(gdir, "fourquad_2000.json")))
    sub16 = np.load(os.path.join(gdir, "fourquad_2000_sub16.npy"))
    U0 = problems.initialize_4shocks(2000)
    ex, steps_ex, _ = gpu.solve(U0, mode=gpu.MODE_EXACT, t_end=0.3)
    assert steps_ex == golden["steps"] == 3716
    P4 = np.ascontiguousarray(ex[..., :4])
    assert np.array_equal(P4[::16, ::16], sub16)
    assert hashlib.sha256(P4.tobytes()).hexdigest() == golden["sha256_prim4"]
    fa, steps_fa, _ = gpu.solve(U0, mode=gpu.MODE_FAST, t_end=0.3)
    l1, linf = rel_errors(fa, ex)
    print(f"2000^2 fast vs exact: steps {steps_fa} relL1 {l1} relLinf {linf}")
    assert steps_fa == golden["steps"]
    assert np.all(l1 <= TOL_L1)


@pytest.mark.parametrize("mode", [0, 1])
def test_launch_geometry_invariance(gpu, problems, mode):
    """Results do not depend on strip height, waves per block or the
    persistent ticket scheduler (bitwise)."""
    rng = problems.SplitMix64(11)
    U0 = problems.random_four_quadrant(rng, 133, 97)
    base = None
    for (R, W, P) in [(16, 4, 0), (4, 1, 1), (7, 2, 0), (64, 4, 1), (33, 1, 0)]:
        s = gpu.Solver(133, 97, t_end=1.0, max_steps=7, mode=mode)
        s.set_launch(R, W, P)
        s.upload(U0)
        steps, _, _ = s.run()
        U = s.download()
        s.close()
        assert steps == 7
        if base is None:
            base = U
        else:
            assert np.array_equal(U, base), (R, W, P)


@pytest.mark.parametrize("mode", [0, 1])
def test_round_filling_strips_invariance(gpu, problems, mode):
    """At 2000 x 2000 and 34-row strips the 2006 strips are re-cut into the
    2048 strips of 33-34 rows that fill one round of wave slots; the result
    equals that of plain 16-row strips and of the persistent scheduler
    (bitwise)."""
    U0 = problems.initialize_4shocks(2000)
    base = None
    for (R, W, P) in [(16, 4, 0), (34, 4, 0), (34, 4, 1)]:
        s = gpu.Solver(2000, 2000, t_end=1.0, max_steps=4, mode=mode)
        s.set_launch(R, W, P)
        s.upload(U0)
        steps, _, _ = s.run()
        U = s.download()
        s.close()
        assert steps == 4
        if base is None:
            base = U
        else:
            assert np.array_equal(U, base), (R, W, P)


def test_resume_like_reference(gpu, orc, problems):
    """A second solve() continues from the previous result, t restarts at 0
    and the step count accumulates (src/wrappers.cuh:191,252,290-291): the
    resident solver run twice equals two oracle solves chained the same way."""
    U0 = problems.initialize_4shocks(48)
    s = gpu.Solver(48, 48, t_end=0.02, mode=gpu.MODE_EXACT)
    s.upload(U0)
    n1, t1, _ = s.run()
    U1 = s.download()
    n2, t2, _ = s.run()
    U2 = s.download()
    total = s.steps_total()
    s.close()
    r1, m1, rt1, _ = orc.solve(U0, nthreads=2, t_end=0.02)
    r2, m2, rt2, _ = orc.solve(r1, nthreads=2, t_end=0.02)
    assert (n1, n2) == (m1, m2) and total == m1 + m2
    assert np.array_equal(U1[..., :4], r1[..., :4]) and np.array_equal(U2[..., :4], r2[..., :4])
    assert (t1, t2) == (rt1, rt2)


def test_drop_in_class_resume_and_counters(gpu, orc, problems):
    """RGRP_2D_CUDA (the Python mirror of the reference class): solve()
    twice continues from the previous result, getTimeStagesCount()
    accumulates, clear() drops the state (src/wrappers.cuh:308-316,347-353)."""
    U0 = problems.initialize_4shocks(40)
    c = gpu.RGRP_2D_CUDA(U0)
    c.setTimeLayout(0.01, 0.5)
    c.setArithmeticMode(gpu.MODE_EXACT)
    a = c.solve()
    b = c.solve()
    r1, m1, _, _ = orc.solve(U0, nthreads=2, t_end=0.01)
    r2, m2, _, _ = orc.solve(r1, nthreads=2, t_end=0.01)
    assert np.array_equal(a[..., :4], r1[..., :4]) and np.array_equal(b[..., :4], r2[..., :4])
    assert c.getTimeStagesCount() == m1 + m2
    c.clear()
    with pytest.raises(gpu.FVM2DError):
        c.solve()


@pytest.mark.parametrize("mode", [0, 1])
def test_zero_end_time_takes_no_step(gpu, orc, problems, mode):
    """t_end = 0: the loop test t < t_end fails at once (src/wrappers.cuh:196),
    so no step and no ghost fill runs and the state comes back unchanged."""
    rng = problems.SplitMix64(41)
    U0 = problems.random_four_quadrant(rng, 30, 26)
    ref, n_ref, _, _ = orc.solve(U0, nthreads=1, t_end=0.0)
    got, n, _ = gpu.solve(U0, t_end=0.0, mode=mode)
    assert n == n_ref == 0
    assert np.array_equal(got[..., :4], U0[..., :4]) and np.array_equal(ref[..., :4], U0[..., :4])


@pytest.mark.parametrize("mode", [0, 1])
def test_tiny_step_ratio_stops_the_loop(gpu, orc, problems, mode):
    """h / s_max < 1e-20 ends the loop before the step (src/wrappers.cuh:204-207):
    one cell with a pressure of 1e44 gives s_max ~ 1e22."""
    U0 = problems.initialize_4shocks(40)
    U0[20, 20, 1] = 1e44
    ref, n_ref, t_ref, _ = orc.solve(U0, nthreads=1, t_end=0.3)
    s = gpu.Solver(40, 40, t_end=0.3, mode=mode)
    s.upload(U0)
    n, t, _ = s.run()
    got = s.download()
    s.close()
    assert n == n_ref == 0 and t == t_ref == 0.0
    assert np.array_equal(got[..., :4], U0[..., :4])


def test_batch_matches_oracle(gpu, orc, problems):
    """Each instance of a batched run (own size, own step count, own dt)
    equals the oracle's solve of that mesh: bitwise in EXACT mode, within
    relative L1 1e-10 in FAST mode."""
    rng = problems.SplitMix64(5)
    Us = [problems.random_four_quadrant(rng, nx, ny)
          for nx, ny in [(40, 33), (128, 64), (17, 90), (200, 150), (64, 64)]]
    steps = [6, 9, 3, 7, 5]
    refs = [orc.solve(U, nthreads=4, t_end=10.0, max_steps=k) for U, k in zip(Us, steps)]
    for mode in (0, 1):
        prms = [dict(t_end=10.0, max_steps=k, mode=mode) for k in steps]
        b = gpu.Batch([U.shape[:2] for U in Us], prms)
        for i, U in enumerate(Us):
            b.upload(i, U)
        b.begin()
        b.run(12)
        for i in range(len(Us)):
            n, _ = b.status(i)
            ref, n_ref, _, _ = refs[i]
            assert n == n_ref == steps[i]
            got = b.download(i)
            if mode == 1:
                assert np.array_equal(got[..., :4], ref[..., :4]), i
            else:
                l1, linf = rel_errors(got, ref)
                assert np.all(l1 <= TOL_L1), (i, l1, linf)
        b.close()


def test_batch_of_mixed_cases_matches_oracle(gpu, orc, cases):
    """Every parity case in one batch: per-instance presets (outflow,
    reflective, jet, jet800, double Mach), gamma, alpha, CFL, domain size,
    t_end and step limit, meshes from 6x6 to 6x3001 and 2999x7 side by side
    in each launch.  Each instance equals its own oracle solve: bitwise in
    EXACT mode, within relative L1 1e-10 in FAST mode."""
    refs = [orc.solve(U0, nthreads=4, **prm) for _, U0, prm in cases]
    for mode in (1, 0):
        b = gpu.Batch([U0.shape[:2] for _, U0, _ in cases], [dict(prm, mode=mode) for _, _, prm in cases])
        for i, (_, U0, _) in enumerate(cases):
            b.upload(i, U0)
        b.begin()
        b.run(100000)
        for i, (name, _, prm) in enumerate(cases):
            ref, n_ref, _, _ = refs[i]
            n, _ = b.status(i)
            got = b.download(i)
            print(f"mode {mode} case {name}: steps {n}/{n_ref}", flush=True)
            if mode == 1:
                assert n == n_ref, name
                assert np.array_equal(got[..., :4], ref[..., :4]), name
            else:
                assert abs(n - n_ref) <= 1, name
                l1, linf = rel_errors(got, ref, prm.get("gamma", 1.4))
                assert np.all(l1 <= TOL_L1), (name, l1, linf)
        b.close()


def test_autotune_keeps_state_and_results(gpu, problems):
    """fvm2d_autotune / fvm2d_batch_autotune time candidate launch shapes on
    scratch copies: the resident state is unchanged afterwards, the chosen
    shape is a valid one, and the run that follows equals a run with the
    default shape (bitwise)."""
    rng = problems.SplitMix64(23)
    U0 = problems.random_four_quadrant(rng, 300, 211)
    outs = []
    for tune in (False, True):
        s = gpu.Solver(300, 211, t_end=10.0, max_steps=5, mode=1)
        s.upload(U0)
        if tune:
            best = s.autotune(steps_per_trial=2)
            assert best > 0.0
            R, W, P = s.get_launch()
            assert R >= 1 and 1 <= W <= 4 and P in (0, 1)
            assert np.array_equal(s.download(), U0)
        steps, _, _ = s.run()
        assert steps == 5
        outs.append(s.download())
        s.close()
    assert np.array_equal(outs[0], outs[1])

    Us = [problems.random_four_quadrant(rng, nx, ny) for nx, ny in [(90, 70), (33, 120), (160, 48)]]
    prms = [dict(t_end=10.0, max_steps=4, mode=1)] * len(Us)
    res = []
    for tune in (False, True):
        b = gpu.Batch([U.shape[:2] for U in Us], prms)
        for i, U in enumerate(Us):
            b.upload(i, U)
        if tune:
            assert b.autotune(steps_per_trial=2) > 0.0
            for i, U in enumerate(Us):
                assert np.array_equal(b.download(i), U)
        b.begin()
        b.run(8)
        res.append([b.download(i) for i in range(len(Us))])
        b.close()
    for a, c in zip(res[0], res[1]):
        assert np.array_equal(a, c)


def test_nan_state_reports_error(gpu, problems):
    U0 = problems.initialize_4shocks(32)
    U0[10, 10, 1] = -1.0  # negative pressure -> NaN sound speed
    with pytest.raises(gpu.FVM2DError, match="non-finite"):
        gpu.solve(U0, t_end=0.1)


def test_fast_reciprocals_accuracy(gpu):
    """The FAST path's one-Newton-step reciprocal and reciprocal square root
    stay within a few ulps (the hardware estimates carry ~24 bits)."""
    raw_rcp, rcp, raw_rsq, rsq = gpu.recip_probe()
    print(f"rcp raw {raw_rcp:.3e} refined {rcp:.3e}; rsq raw {raw_rsq:.3e} refined {rsq:.3e}")
    assert raw_rcp < 1e-6 and raw_rsq < 1e-6
    assert rcp < 1e-14 and rsq < 1e-14
