"""Pins the L-BFGS-B restatement (oracle/lbfgsb_oracle.py) against scipy's L-BFGS-B, an independent
implementation of the same published method (version 3.0), driven step by step through the same
reverse-communication protocol: every trial point, the number of evaluations and the termination
reason must agree.  Also pins the gradient shift search built on it (shsrch_minimize) by its
invariants and by recovering a known shift on noiseless synthetic data.
"""
import numpy as np
import pytest
from scipy.optimize import _lbfgsb

import lbfgsb_oracle as L
import pftcc_oracle as O
from cases import make_case

SCIPY_TASK = {401: "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL",
              402: "CONVERGENCE: REL_REDUCTION_OF_F_<=_FACTR*EPSMCH"}


def scipy_trace(fun, x0, l, u, nbd, m=10, factr=1e7, pgtol=1e-5, maxls=20):
    n = len(x0)
    x = np.array(x0, float)
    f, g = 0.0, np.zeros(n)
    wa = np.zeros(2 * m * n + 5 * n + 11 * m * m + 8 * m)
    iwa = np.zeros(3 * n, np.int32)
    task = np.zeros(2, np.int32)
    ln_task = np.zeros(2, np.int32)
    lsave = np.zeros(4, np.int32)
    isave = np.zeros(44, np.int32)
    dsave = np.zeros(29)
    lo, up, nb = np.array(l, float), np.array(u, float), np.array(nbd, np.int32)
    xs = []
    while True:
        _lbfgsb.setulb(m, x, lo, up, nb, f, g, factr, pgtol, wa, iwa, task, lsave, isave, dsave, maxls, ln_task)
        if task[0] == 3:
            xs.append(x.copy())
            f, g = fun(x)
            g = np.asarray(g, float)
        elif task[0] != 1:
            break
    # task[0] == 4: converged, task[1] the reason (401 projected gradient, 402 relative reduction)
    return xs, (int(task[1]) if task[0] == 4 else -int(task[0])), x.copy()


def oracle_trace(fun, x0, l, u, nbd, m=10):
    lb = L.Lbfgsb(len(x0), m, l, u, nbd)
    gen = lb.minimize(np.array(x0, float))
    req = next(gen)
    xs = []
    try:
        while True:
            if req[0] == "FG":
                xs.append(req[1].copy())
                f, g = fun(req[1])
                req = gen.send((f, np.asarray(g, float)))
            else:
                req = gen.send(None)
    except StopIteration as e:
        task, x, f, nfgv = e.value
    assert nfgv == len(xs)
    return xs, task, x


def rosen(x):
    f = 100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2
    g = np.array([-400 * x[0] * (x[1] - x[0] ** 2) - 2 * (1 - x[0]), 200 * (x[1] - x[0] ** 2)])
    return f, g


def bump(x):
    """Shaped like the shift-search objective: -cc of a Gaussian peak with a ripple."""
    a, b = 1.3, -0.7
    r2 = (x[0] - a) ** 2 + 0.6 * (x[1] - b) ** 2
    e = np.exp(-r2 / 4.0)
    c = np.cos(0.8 * x[0])
    f = -(e * (0.9 + 0.1 * c))
    dr = np.array([2 * (x[0] - a), 1.2 * (x[1] - b)])
    g = -(-e / 4.0 * dr * (0.9 + 0.1 * c) + e * np.array([-0.08 * np.sin(0.8 * x[0]), 0.0]))
    return f, g


def quad5(x):
    A = np.diag([1.0, 3.0, 10.0, 0.5, 2.0]) + 0.3 * np.ones((5, 5))
    b = np.array([1.0, -2.0, 4.0, 0.5, -3.0])
    return 0.5 * x @ A @ x - b @ x, A @ x - b


CASES = [
    ("rosen boxed", rosen, (-1.2, 1.0), (-1.5, -0.5), (2, 2), (2, 2)),
    ("rosen unbounded", rosen, (-1.2, 1.0), (0, 0), (0, 0), (0, 0)),
    ("rosen active upper bound", rosen, (0.5, 0.5), (-1.5, -0.5), (0.8, 2), (2, 2)),
    ("rosen lower/upper only", rosen, (-1.0, 1.5), (-1.1, 0.0), (0.0, 1.2), (1, 3)),
    ("bump in a box", bump, (0.0, 0.0), (-3, -3), (3, 3), (2, 2)),
    ("quadratic 5-D, active bounds", quad5, (0, 0, 0, 0, 0), (-1, -1, -0.2, -1, -1), (1, 0.5, 0.3, 1, 1),
     (2, 2, 2, 2, 2)),
]


@pytest.mark.parametrize("name,fun,x0,l,u,nbd", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("m", [3, 10])
def test_trial_points_match_scipy(name, fun, x0, l, u, nbd, m):
    a_x, a_task, a_fin = scipy_trace(fun, x0, l, u, nbd, m=m)
    b_x, b_task, b_fin = oracle_trace(fun, x0, l, u, nbd, m=m)
    assert len(a_x) == len(b_x), (len(a_x), len(b_x))
    a, b = np.array(a_x), np.array(b_x)
    scale = max(1.0, float(np.abs(a).max()))
    # rounding differences of the two B representations grow slowly along long trajectories
    np.testing.assert_allclose(b[:20], a[:20], rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-5 * scale)
    np.testing.assert_allclose(b_fin, a_fin, rtol=0, atol=1e-5 * scale)
    if a_task in SCIPY_TASK:
        assert b_task == SCIPY_TASK[a_task]
    else:
        assert not b_task.startswith("CONV")


@pytest.mark.parametrize("m", [3, 10])
def test_variable_leaving_the_free_set(m):
    """From the corner (1, -1) of the box the x variable leaves the free set.  The published method
    keeps its two correction pairs there (the reduced middle matrix, updated incrementally in formk,
    stays positive definite: S_A'S_A = s_x s_x'); scipy 1.15.3's L-BFGS-B discards them and takes a
    plain Cauchy step instead.  Both agree up to the corner and converge to the same constrained
    optimum (1, -0.7)."""
    args = (bump, (0.0, 0.0), (-1, -1), (1, 1), (2, 2))
    a_x, a_task, a_fin = scipy_trace(*args, m=m)
    b_x, b_task, b_fin = oracle_trace(*args, m=m)
    np.testing.assert_allclose(np.array(b_x[:3]), np.array(a_x[:3]), rtol=0, atol=1e-12)
    np.testing.assert_allclose(b_x[2], [1.0, -1.0])
    # both stop on the projected-gradient test (pgtol 1e-5, curvature ~0.23 along y)
    np.testing.assert_allclose(b_fin, a_fin, atol=1e-4)
    np.testing.assert_allclose(b_fin, [1.0, -0.7], atol=1e-4)
    assert b_task == SCIPY_TASK[a_task]
    # the oracle's Newton step from the corner is the exact line minimum along y (a quadratic
    # model with the stored pairs): a descent step, so no reset is called for
    assert b_x[3][1] > -0.9


def test_line_search_constants_are_parameters():
    """Moré-Thuente on phi(a) = (a - 3)^2 from a = 0 with the first trial at a = 1: with scipy's
    constants (ftol 1e-3, gtol 0.9) a = 1 already satisfies the strong Wolfe conditions
    (phi'(1) = -4, |phi'(1)| <= 0.9 * 6); with the shift search's (0.1, 1e-3) the search continues
    to |phi'(a)| <= 6e-3, i.e. a within 3e-3 of the minimiser 3."""
    def phi(a):
        return (a - 3.0) ** 2, 2.0 * (a - 3.0)

    for ftol, gtol in ((1e-3, 0.9), (0.1, 1e-3)):
        ls = L.Dcsrch(ftol, gtol, 0.1)
        f0, g0 = phi(0.0)
        assert ls.start(f0, g0, 1.0, 0.0, 100.0) == "FG"
        stp, n = 1.0, 0
        while True:
            f, g = phi(stp)
            task, nxt = ls.step(f, g, stp)
            n += 1
            if task != "FG":
                break
            stp = nxt
        assert task == "CONVERGENCE"
        if gtol == 0.9:
            assert (stp, n) == (1.0, 1)
        else:
            assert n > 1 and abs(phi(stp)[1]) <= gtol * 6.0 and abs(stp - 3.0) <= 3e-3


def _pair_objective(pf, iref, ip):
    def fdf(sh, irot):
        f, g = pf.gencorr_grad_for_rot_8(iref, ip, sh, irot)
        return f, g

    def gencorrs(sh32):
        return pf.gencorrs(iref, ip, tuple(float(v) for v in sh32), method="fft")
    return fdf, gencorrs


@pytest.mark.parametrize("opt_angle", [True, False])
def test_shift_search_recovers_a_known_shift(opt_angle):
    """A noiseless particle made from reference 0 shifted by a sub-pixel vector: the search must
    return that shift (rotated to the reference frame as the reference does) and a cc close to 1."""
    pf, _ = make_case(seed=4, n=48, kfrom=2, kto=16, nrefs=2, nptcls=1, ctf=False, ldim=64)
    true = np.array([1.37, -0.81])
    # particle = reference times the phase of shift -true: gencorrs at shift s applies the phase
    # of s to the reference, so the best shift is -true
    pf.ptcls[0] = (pf.refs_even[0] * O.shmat_8(pf.grid, -true)).astype(np.complex64)
    fdf, gencorrs = _pair_objective(pf, 0, 0)
    trace = []
    cxy, irot = L.shsrch_minimize(fdf, gencorrs, pf.grid.angtab_deg, trs=3.0, irot=1, opt_angle=opt_angle,
                                  trace=trace)
    assert irot == 1
    assert cxy[0] > 0.9999
    np.testing.assert_allclose(cxy[1:], -true, atol=2e-3)     # rotmat2d(0) = I
    # the returned cc is the objective at the returned shift (FP32)
    cc = gencorrs(cxy[1:])[0] if opt_angle else fdf(cxy[1:].astype(np.float64), 1)[0]
    assert abs(cc - cxy[0]) < 1e-5 * abs(cc) + 1e-6


def test_shift_search_invariants_on_random_pairs():
    """On random pairs (rings uncorrelated, so the optimum is often on the box): the result is never
    worse than the start, lies in the box, and at the optimum the projected gradient is small or a
    bound is active."""
    pf, _ = make_case(seed=9, n=32, kfrom=2, kto=10, nrefs=3, nptcls=3, ctf=True, ldim=32)
    trs = 2.0
    for r in range(3):
        for p in range(3):
            fdf, gencorrs = _pair_objective(pf, r, p)
            c0 = gencorrs(np.zeros(2, np.float32))
            cxy, irot = L.shsrch_minimize(fdf, gencorrs, pf.grid.angtab_deg, trs=trs, opt_angle=True)
            if irot == 0:
                continue
            assert cxy[0] > c0.max()
            rot = pf.grid.angtab_deg[irot - 1]
            # undo the rotation to get the search-frame shift back
            sh = L.rotmat2d_apply(cxy[1:], -rot)
            assert np.all(np.abs(sh) <= trs + 1e-5)


def test_shift_search_returns_shift_rotated_to_reference_frame():
    """The returned shift is matmul(shift, rotmat2d(get_rot(irot))) (simple_pftcc_shsrch_grad.f90:
    295-297; rotmat2d simple_math.f90:1549-1559: mat = [[c, s], [-s, c]]).  A noiseless particle
    that is the reference shifted by s* and rotated by 90 degrees (irot = nrots/4 + 1) is found at
    that rotation and shift, and matmul((sx, sy), rotmat2d(90)) = (-sy, sx) is returned."""
    n = 48
    pf, _ = make_case(seed=5, n=n, kfrom=2, kto=16, nrefs=1, nptcls=1, ctf=False, ldim=64)
    star = np.array([0.85, -1.2])
    rho = n // 2                                   # 90 degrees of 2n = 96 rotations
    q = pf.refs_even[0].astype(np.complex128) * O.shmat_8(pf.grid, star)
    pf.ptcls[0] = np.roll(O.full_ring(q), rho, axis=-1)[:, :n].astype(np.complex64)
    fdf, gencorrs = _pair_objective(pf, 0, 0)
    cxy, irot = L.shsrch_minimize(fdf, gencorrs, pf.grid.angtab_deg, trs=3.0, opt_angle=True)
    assert irot == rho + 1
    assert abs(float(pf.grid.angtab_deg[irot - 1]) - 90.0) < 1e-4
    assert cxy[0] > 0.9999
    np.testing.assert_allclose(cxy[1:], [-star[1], star[0]], atol=2e-3)
