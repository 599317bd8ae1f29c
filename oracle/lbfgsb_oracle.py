"""L-BFGS-B for small dimensions, and the reference's gradient shift search built on it
(TEST INFRASTRUCTURE ONLY).

``Lbfgsb.minimize`` restates the bound-constrained limited-memory BFGS method of Byrd, Lu,
Nocedal & Zhu (1995), version 3.0 (Morales & Nocedal 2011), as the reference carries it in
``src/main/simple_opt_lbfgsb.f90`` (mainlb :384-917, cauchy :1093-1491, subsm :2381-2530,
lnsrlb :2125-2230, matupd :2232-2290, projgr :2292-2317, Moré-Thuente dcsrch/dcstep :2532-2830),
with its reverse-communication protocol (``lbfgsbmin`` :103-129).  It works with the BFGS matrix
itself, B = the m most recent BFGS updates applied to theta I (Byrd, Nocedal & Schnabel 1994,
Theorem 2.3), instead of the compact factored form, which for the two shift variables is the
cheaper representation; the quantities the method computes from B (the Cauchy path derivatives,
the reduced Newton step) are the same in exact arithmetic.  The positive-definiteness tests of the
compact form (formt's and formk's Cholesky factorisations), which make the reference discard its
memory, are carried out on the same matrices.

``shsrch_minimize`` restates ``pftcc_shsrch_grad%minimize`` (``src/main/simple_pftcc_shsrch_grad.f90:254-338``)
with the optimizer settings of ``grad_shsrch_new`` (:46-82) and ``opt_spec`` (``simple_opt_spec.f90:22-40,271``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this file.

Parity status: ``Lbfgsb`` is pinned by ``tests/test_oracle_lbfgsb.py`` against scipy's L-BFGS-B
(``scipy.optimize._lbfgsb.setulb``, an independent implementation of version 3.0) driven
step by step through the same reverse-communication protocol: every trial point, the termination
reason and the final iterate agree on bounded and unbounded test functions.  scipy fixes the line
search constants at ftol 1e-3, gtol 0.9; the reference passes the shift search's own (0.1, 1e-3),
which the same code takes as parameters.  ``shsrch_minimize`` is pinned by its invariants
(the returned cc is the objective at the returned shift and rotation, never worse than the
start, stationarity or an active bound at the optimum) and by a recovery test on synthetic
data with a known shift.
"""
from __future__ import annotations

import math

import numpy as np

EPSMCH = float(np.finfo(np.float64).eps)
BIG = 1.0e10
P5, P66 = 0.5, 0.66
XTRAPL, XTRAPU = 1.1, 4.0


# ---------------------------------------------------------------------------------------------
# Moré-Thuente line search (dcsrch / dcstep, simple_opt_lbfgsb.f90:2532-2830)
# ---------------------------------------------------------------------------------------------
def dcstep(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax, zero_tol=0.0):
    """One safeguarded step of the Moré-Thuente interval update; returns the updated
    (stx, fx, dx, sty, fy, dy, stp, brackt).  zero_tol > 0: the reference's is_zero(gamma)."""
    sgnd = dp * (dx / abs(dx))
    if fp > fx:
        # higher function value: the minimum is bracketed; cubic vs quadratic step
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        stpc = stx + (p / q) * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0.0:
        # derivatives of opposite sign: bracketed; cubic vs secant step
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        stpc = stp + (p / q) * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(dp) < abs(dx):
        # derivative decreases in magnitude
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt(max(0.0, (theta / s) ** 2 - (dx / s) * (dp / s)))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        gamma_nonzero = abs(gamma) >= zero_tol if zero_tol > 0 else gamma != 0.0
        if r < 0.0 and gamma_nonzero:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stpc - stp) < abs(stpq - stp) else stpq
            if stp > stx:
                stpf = min(stp + P66 * (sty - stp), stpf)
            else:
                stpf = max(stp + P66 * (sty - stp), stpf)
        else:
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            stpf = max(stpmin, min(stpmax, stpf))
    else:
        # derivative does not decrease: cubic step to the other end, or the step limit
        if brackt:
            theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp
            s = max(abs(theta), abs(dy), abs(dp))
            gamma = s * math.sqrt((theta / s) ** 2 - (dy / s) * (dp / s))
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            stpf = stp + (p / q) * (sty - stp)
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    # interval update
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0.0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    return stx, fx, dx, sty, fy, dy, stpf, brackt


class Dcsrch:
    """State of one Moré-Thuente search (the reference's isave(2)/dsave(13) of dcsrch)."""

    def __init__(self, ftol, gtol, xtol, eq_tol=0.0):
        self.ftol, self.gtol, self.xtol, self.eq_tol = ftol, gtol, xtol, eq_tol

    def start(self, f, g, stp, stpmin, stpmax):
        """task START: returns 'FG' or an 'ERROR' string."""
        if stp < stpmin:
            return "ERROR: STP .LT. STPMIN"
        if stp > stpmax:
            return "ERROR: STP .GT. STPMAX"
        if g >= 0.0:
            return "ERROR: INITIAL G .GE. ZERO"
        self.stpmin, self.stpmax = stpmin, stpmax
        self.brackt = False
        self.stage = 1
        self.finit, self.ginit = f, g
        self.gtest = self.ftol * g
        self.width = stpmax - stpmin
        self.width1 = self.width / P5
        self.stx, self.fx, self.gx = 0.0, f, g
        self.sty, self.fy, self.gy = 0.0, f, g
        self.stmin = 0.0
        self.stmax = stp + XTRAPU * stp
        return "FG"

    def step(self, f, g, stp):
        """Continues with (f, g) at stp: returns (task, next stp)."""
        eq = (lambda a, b: abs(a - b) < self.eq_tol) if self.eq_tol > 0 else (lambda a, b: a == b)
        ftest = self.finit + stp * self.gtest
        if self.stage == 1 and f <= ftest and g >= 0.0:
            self.stage = 2
        task = None
        if self.brackt and (stp <= self.stmin or stp >= self.stmax):
            task = "WARNING: ROUNDING ERRORS PREVENT PROGRESS"
        if self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax:
            task = "WARNING: XTOL TEST SATISFIED"
        if eq(stp, self.stpmax) and f <= ftest and g <= self.gtest:
            task = "WARNING: STP = STPMAX"
        if eq(stp, self.stpmin) and (f > ftest or g >= self.gtest):
            task = "WARNING: STP = STPMIN"
        if f <= ftest and abs(g) <= self.gtol * (-self.ginit):
            task = "CONVERGENCE"
        if task is not None:
            return task, stp
        if self.stage == 1 and f <= self.fx and f > ftest:
            # modified function psi(t) = f(t) - f(0) - ftol t f'(0) in the first stage
            gt = self.gtest
            r = dcstep(self.stx, self.fx - self.stx * gt, self.gx - gt, self.sty, self.fy - self.sty * gt,
                       self.gy - gt, stp, f - stp * gt, g - gt, self.brackt, self.stmin, self.stmax,
                       self.eq_tol)
            stx, fxm, gxm, sty, fym, gym, stp, self.brackt = r
            self.stx, self.sty = stx, sty
            self.fx, self.fy = fxm + stx * gt, fym + sty * gt
            self.gx, self.gy = gxm + gt, gym + gt
        else:
            r = dcstep(self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, stp, f, g, self.brackt,
                       self.stmin, self.stmax, self.eq_tol)
            self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, stp, self.brackt = r
        if self.brackt:
            if abs(self.sty - self.stx) >= P66 * self.width1:
                stp = self.stx + P5 * (self.sty - self.stx)
            self.width1 = self.width
            self.width = abs(self.sty - self.stx)
            self.stmin = min(self.stx, self.sty)
            self.stmax = max(self.stx, self.sty)
        else:
            self.stmin = stp + XTRAPL * (stp - self.stx)
            self.stmax = stp + XTRAPU * (stp - self.stx)
        stp = min(max(stp, self.stpmin), self.stpmax)
        if (self.brackt and (stp <= self.stmin or stp >= self.stmax)) or \
                (self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax):
            stp = self.stx
        return "FG", stp


# ---------------------------------------------------------------------------------------------
# L-BFGS-B
# ---------------------------------------------------------------------------------------------
def _chol_ok(A: np.ndarray) -> bool:
    """LINPACK dpofa succeeds (every pivot > 0) on the symmetric matrix A."""
    k = A.shape[0]
    L = np.zeros_like(A)
    for j in range(k):
        s = A[j, j] - np.dot(L[j, :j], L[j, :j])
        if s <= 0.0:
            return False
        L[j, j] = math.sqrt(s)
        for i in range(j + 1, k):
            L[i, j] = (A[i, j] - np.dot(L[i, :j], L[j, :j])) / L[j, j]
    return True


class Lbfgsb:
    """L-BFGS-B with the reference's reverse communication.

    ``minimize(x0)`` is a generator.  It yields ``('FG', x)`` and expects ``(f, g)`` back; yields
    ``('NEW_X', x, f, nfgv)`` after each accepted iterate and expects ``None`` (continue) or
    ``'STOP'``; and finishes by returning ``(task, x, f, nfgv)`` (StopIteration.value).
    nbd: 0 unbounded, 1 lower, 2 both, 3 upper (setulb's convention).
    eq_tol > 0 reproduces the reference's tolerant equality tests (is_equal, |a - b| < 1e-10, and
    is_gt_zero); 0 gives the exact tests of the published code (scipy)."""

    def __init__(self, n, m, l, u, nbd, factr=1.0e7, pgtol=1.0e-5, ftol=1.0e-3, gtol=0.9, xtol=0.1,
                 maxls=20, eq_tol=0.0):
        self.n, self.m = n, m
        self.l = np.asarray(l, float)
        self.u = np.asarray(u, float)
        self.nbd = np.asarray(nbd, int)
        self.factr, self.pgtol = factr, pgtol
        self.ftol, self.gtol, self.xtol = ftol, gtol, xtol
        self.maxls = maxls
        self.eq_tol = eq_tol

    # -- helpers ----------------------------------------------------------------------------
    def _eq(self, a, b):
        return abs(a - b) < self.eq_tol if self.eq_tol > 0 else a == b

    def _B(self, pairs, theta):
        """BFGS matrix from theta I and the stored (s, y) pairs, oldest first."""
        B = theta * np.eye(self.n)
        for s, y in pairs:
            Bs = B @ s
            B = B - np.outer(Bs, Bs) / float(s @ Bs) + np.outer(y, y) / float(y @ s)
        return B

    def _projgr(self, x, g):
        """projgr (:2292-2317): infinity norm of the projected gradient."""
        m = 0.0
        for i in range(self.n):
            gi = g[i]
            if self.nbd[i] != 0:
                if gi < 0.0:
                    if self.nbd[i] >= 2:
                        gi = max(x[i] - self.u[i], gi)
                elif self.nbd[i] <= 2:
                    gi = min(x[i] - self.l[i], gi)
            m = max(m, abs(gi))
        return m

    def _formt_ok(self, pairs, theta):
        """formt (:2050-2085): T = theta S'S + L D^-1 L' must factor."""
        c = len(pairs)
        if c == 0:
            return True
        S = np.array([p[0] for p in pairs])
        Y = np.array([p[1] for p in pairs])
        SY = S @ Y.T                     # SY[i, k] = s_i . y_k
        T = theta * (S @ S.T)
        for i in range(c):
            for j in range(c):
                k1 = min(i, j)
                T[i, j] += sum(SY[i, k] * SY[j, k] / SY[k, k] for k in range(k1))
        return _chol_ok(T)

    def _formk_ok(self, pairs, theta, free):
        """formk (:1552-1750): the middle matrix of the reduced compact form must factor,
        K11 = Y_Z'Y_Z / theta + D, K12 from S_Z'Y_Z (upper) and S_A'Y_A (strict lower), K22 =
        theta S_A'S_A; chol(K11) then chol(K22 + X'X) with X = chol(K11)'^-1 K12."""
        c = len(pairs)
        S = np.array([p[0] for p in pairs])
        Y = np.array([p[1] for p in pairs])
        fr = np.zeros(self.n, bool)
        fr[list(free)] = True
        YZ, SZ, SA, YA = Y[:, fr], S[:, fr], S[:, ~fr], Y[:, ~fr]
        D = np.diag([float(S[i] @ Y[i]) for i in range(c)])
        K11 = YZ @ YZ.T / theta + D
        szyz = SZ @ YZ.T
        saya = SA @ YA.T
        K12 = np.zeros((c, c))
        for iy in range(c):          # column iy (s index)
            for jy in range(c):      # row jy (y index)
                K12[jy, iy] = szyz[iy, jy] if jy >= iy else -saya[iy, jy]
        if not _chol_ok(K11):
            return False
        U = np.linalg.cholesky(K11).T
        X = np.linalg.solve(U.T, K12)
        K22 = theta * (SA @ SA.T) + X.T @ X
        return _chol_ok(K22)

    def _cauchy(self, x, g, B, theta, sbgnrm, iwhere):
        """Generalized Cauchy point (cauchy, :1093-1491) of the model g'z + z'Bz/2 along the
        projected steepest-descent path; updates iwhere, returns xcp."""
        n, l, u, nbd = self.n, self.l, self.u, self.nbd
        xcp = x.copy()
        if sbgnrm <= 0.0:
            return xcp
        d = np.zeros(n)
        bps = []                      # (t, i) breakpoints
        nfree_nobp = 0
        bnded = True
        for i in range(n):
            neggi = -g[i]
            tl = tu = 0.0
            if iwhere[i] != 3 and iwhere[i] != -1:
                if nbd[i] <= 2:
                    tl = x[i] - l[i]
                if nbd[i] >= 2:
                    tu = u[i] - x[i]
                xlower = nbd[i] <= 2 and tl <= 0.0
                xupper = nbd[i] >= 2 and tu <= 0.0
                iwhere[i] = 0
                if xlower:
                    if neggi <= 0.0:
                        iwhere[i] = 1
                elif xupper:
                    if neggi >= 0.0:
                        iwhere[i] = 2
                elif abs(neggi) <= 0.0:
                    iwhere[i] = -3
            if iwhere[i] != 0 and iwhere[i] != -1:
                continue
            d[i] = neggi
            if nbd[i] <= 2 and nbd[i] != 0 and neggi < 0.0:
                bps.append((tl / (-neggi), i))
            elif nbd[i] >= 2 and neggi > 0.0:
                bps.append((tu / neggi, i))
            else:
                nfree_nobp += 1
                if abs(neggi) > 0.0:
                    bnded = False
        if not bps and nfree_nobp == 0:
            return xcp
        z = np.zeros(n)               # displacement of the path point from x
        f1 = float(g @ d)
        f2 = float(d @ B @ d)
        f2_org = f2
        dtm = -f1 / f2
        tsum = 0.0
        nbreak = len(bps)
        if nbreak:
            # the heap of the reference returns breakpoints in ascending order; the first one is
            # the earliest minimum, ties keep input order
            first = min(range(nbreak), key=lambda k: bps[k][0])
            rest = [bps[k] for k in range(nbreak) if k != first]
            order = [bps[first]] + sorted(rest, key=lambda b: b[0])
            tj = 0.0
            nleft = nbreak
            for tb, ibp in order:
                tj0, tj = tj, tb
                dt = tj - tj0
                if dtm < dt:
                    break
                tsum += dt
                nleft -= 1
                dibp = d[ibp]
                d[ibp] = 0.0
                if dibp > 0.0:
                    zibp = u[ibp] - x[ibp]
                    xcp[ibp] = u[ibp]
                    iwhere[ibp] = 2
                else:
                    zibp = l[ibp] - x[ibp]
                    xcp[ibp] = l[ibp]
                    iwhere[ibp] = 1
                if nleft == 0 and nbreak == n:
                    return xcp
                # path point at tsum: moving variables at tsum d, the fixed one at its bound
                moving = d != 0.0
                z = np.where(moving, tsum * d, z)
                z[ibp] = zibp
                f1 = float(g @ d + z @ B @ d)
                f2 = float(d @ B @ d)
                f2 = max(EPSMCH * f2_org, f2)
                if nleft > 0:
                    dtm = -f1 / f2
                elif bnded:
                    f1 = f2 = 0.0
                    dtm = 0.0
                else:
                    dtm = -f1 / f2
            # fall through to the final segment
        if dtm <= 0.0:
            dtm = 0.0
        tsum += dtm
        xcp = xcp + tsum * d
        return xcp

    def _subsm(self, x, g, xcp, B, free):
        """Subspace minimisation (cmprlb :1493-1550 + subsm :2381-2530): Newton step of the model
        over the free variables from xcp, then the projection / backtracking rule of version 3.0."""
        l, u, nbd = self.l, self.u, self.nbd
        fr = list(free)
        r = -(g + B @ (xcp - x))[fr]
        du = np.linalg.solve(B[np.ix_(fr, fr)], r)
        xp = xcp.copy()
        xn = xcp.copy()
        iword = 0
        for i, k in enumerate(fr):
            dk, xk = du[i], xcp[k]
            if nbd[k] == 0:
                xn[k] = xk + dk
            elif nbd[k] == 1:
                xn[k] = max(l[k], xk + dk)
                if self._eq(xn[k], l[k]):
                    iword = 1
            elif nbd[k] == 2:
                xn[k] = min(u[k], max(l[k], xk + dk))
                if self._eq(xn[k], l[k]) or self._eq(xn[k], u[k]):
                    iword = 1
            else:
                xn[k] = min(u[k], xk + dk)
                if self._eq(xn[k], u[k]):
                    iword = 1
        if iword == 0:
            return xn
        dd_p = float((xn - x) @ g)
        if not (dd_p > (self.eq_tol if self.eq_tol > 0 else 0.0)):
            return xn
        # the projected point is not a descent direction: truncate the Newton step to the box
        xn = xp
        alpha, temp1, ibd = 1.0, 1.0, -1
        for i, k in enumerate(fr):
            dk = du[i]
            if nbd[k] != 0:
                if dk < 0.0 and nbd[k] <= 2:
                    temp2 = l[k] - xn[k]
                    if temp2 >= 0.0:
                        temp1 = 0.0
                    elif dk * alpha < temp2:
                        temp1 = temp2 / dk
                elif dk > 0.0 and nbd[k] >= 2:
                    temp2 = u[k] - xn[k]
                    if temp2 <= 0.0:
                        temp1 = 0.0
                    elif dk * alpha > temp2:
                        temp1 = temp2 / dk
                if temp1 < alpha:
                    alpha, ibd = temp1, i
        du = du.copy()
        if alpha < 1.0:
            k = fr[ibd]
            if du[ibd] > 0.0:
                xn[k] = u[k]
                du[ibd] = 0.0
            elif du[ibd] < 0.0:
                xn[k] = l[k]
                du[ibd] = 0.0
        for i, k in enumerate(fr):
            xn[k] = xn[k] + alpha * du[i]
        return xn

    # -- the method -------------------------------------------------------------------------
    def minimize(self, x0):
        n, m = self.n, self.m
        l, u, nbd = self.l, self.u, self.nbd
        x = np.array(x0, float)
        for i in range(n):               # errclb
            if nbd[i] == 2 and l[i] > u[i]:
                return "ERROR: NO FEASIBLE SOLUTION", x, None, 0
        # active (:919-990): project x0, classify the variables
        cnstnd, boxed = False, True
        iwhere = np.zeros(n, int)
        for i in range(n):
            if nbd[i] > 0:
                if nbd[i] <= 2 and x[i] <= l[i]:
                    x[i] = l[i]
                elif nbd[i] >= 2 and x[i] >= u[i]:
                    x[i] = u[i]
        for i in range(n):
            if nbd[i] != 2:
                boxed = False
            if nbd[i] == 0:
                iwhere[i] = -1
            else:
                cnstnd = True
                iwhere[i] = 3 if (nbd[i] == 2 and u[i] - l[i] <= 0.0) else 0
        tol = self.factr * EPSMCH
        pairs: list = []
        theta = 1.0
        iupdat = 0
        updatd = False
        itr = 0
        nfgv = 0
        free = list(range(n))
        f, g = yield ("FG", x.copy())
        g = np.asarray(g, float)
        nfgv = 1
        sbgnrm = self._projgr(x, g)
        if sbgnrm <= self.pgtol:
            return "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL", x, f, nfgv

        def reset():
            nonlocal pairs, theta, iupdat, updatd
            pairs, theta, iupdat, updatd = [], 1.0, 0, False

        while True:
            # ---- search direction (label 222)
            B = self._B(pairs, theta)
            if not cnstnd and pairs:
                z = x.copy()
                wrk = updatd
            else:
                z = self._cauchy(x, g, B, theta, sbgnrm, iwhere)
                # freev (:1752-1800)
                new_free = [i for i in range(n) if iwhere[i] <= 0]
                enter = leave = False
                if itr > 0 and cnstnd:
                    leave = any(iwhere[k] > 0 for k in free)
                    enter = any(iwhere[k] <= 0 for k in range(n) if k not in free)
                wrk = leave or enter or updatd
                free = new_free
            if free and pairs:
                if wrk and not self._formk_ok(pairs, theta, free):
                    reset()
                    continue
                z = self._subsm(x, g, z, B, free)
            d = z - x
            # ---- line search (lnsrlb, :2125-2230)
            dtd = float(d @ d)
            dnorm = math.sqrt(dtd)
            stpmx = BIG
            if cnstnd:
                if itr == 0:
                    stpmx = 1.0
                else:
                    for i in range(n):
                        a1 = d[i]
                        if nbd[i] != 0:
                            if a1 < 0.0 and nbd[i] <= 2:
                                a2 = l[i] - x[i]
                                if a2 >= 0.0:
                                    stpmx = 0.0
                                elif a1 * stpmx < a2:
                                    stpmx = a2 / a1
                            elif a1 > 0.0 and nbd[i] >= 2:
                                a2 = u[i] - x[i]
                                if a2 <= 0.0:
                                    stpmx = 0.0
                                elif a1 * stpmx > a2:
                                    stpmx = a2 / a1
            stp = min(1.0 / dnorm, stpmx) if (itr == 0 and not boxed) else 1.0
            t, r, fold = x.copy(), g.copy(), f
            ifun = iback = 0
            ls = Dcsrch(self.ftol, self.gtol, self.xtol, self.eq_tol)
            info = 0
            gd = float(g @ d)
            gdold = gd
            if gd >= 0.0:
                info = -4
                task = None
            else:
                task = ls.start(f, gd, stp, 0.0, stpmx)
                if task.startswith("ERROR"):
                    info = -4
            while info == 0:
                # a trial point: dcsrch asked for f and g at stp
                ifun += 1
                nfgv += 1
                iback = ifun - 1
                if iback >= self.maxls:
                    break
                x = z.copy() if self._eq(stp, 1.0) else stp * d + t
                f, g = yield ("FG", x.copy())
                g = np.asarray(g, float)
                gd = float(g @ d)
                task, stp = ls.step(f, gd, stp)
                if task.startswith("CONV") or task.startswith("WARN"):
                    break
                if task.startswith("ERROR"):
                    info = -4
            if info != 0 or iback >= self.maxls:
                # line search failed: back to the last iterate
                x, g, f = t, r, fold
                if not pairs:
                    if info == 0:
                        nfgv -= 1
                    return "ABNORMAL_TERMINATION_IN_LNSRCH", x, f, nfgv
                if info == 0:
                    nfgv -= 1
                reset()
                continue
            itr += 1
            sbgnrm = self._projgr(x, g)
            cmd = yield ("NEW_X", x.copy(), f, nfgv)
            if cmd == "STOP":
                return "STOP", x, f, nfgv
            # ---- tests and update (label 777)
            if sbgnrm <= self.pgtol:
                return "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL", x, f, nfgv
            ddum = max(abs(fold), abs(f), 1.0)
            if fold - f <= tol * ddum:
                return "CONVERGENCE: REL_REDUCTION_OF_F_<=_FACTR*EPSMCH", x, f, nfgv
            y = g - r
            if self._eq(stp, 1.0):
                dr = gd - gdold
                ddum = -gdold
                s = d
            else:
                dr = (gd - gdold) * stp
                s = stp * d
                ddum = -gdold * stp
            if dr <= EPSMCH * ddum:
                updatd = False
                continue
            updatd = True
            iupdat += 1
            pairs = (pairs + [(s.copy(), y.copy())])[-m:]
            theta = float(y @ y) / dr
            if not self._formt_ok(pairs, theta):
                reset()


# ---------------------------------------------------------------------------------------------
# The reference's gradient shift search
# ---------------------------------------------------------------------------------------------
SHSRCH_MAXITS = 60      # simple_strategy2D_srch.f90:55, simple_strategy3D_srch.f90:123
SHSRCH_MAX_EVALS = 5    # pftcc_shsrch_grad%max_evals (simple_pftcc_shsrch_grad.f90:29)


def rotmat2d_apply(v, ang_deg):
    """cxy(2:3) = matmul(v, rotmat2d(ang)) in FP32 (simple_math.f90:1550-1559)."""
    f = np.float32
    a = f(f(ang_deg) * f(np.pi) / f(180.0))
    c, s = f(np.cos(a)), f(np.sin(a))
    v0, v1 = f(v[0]), f(v[1])
    return np.array([f(v0 * c + v1 * (-s)), f(v0 * s + v1 * c)], np.float32)


COARSE_NUM_STEPS = 5    # simple_pftcc_shsrch_grad.f90:15


def _coarse_grid(trs):
    """The 5 x 5 grid of coarse_search / coarse_search_opt_angle (:340-395): step (hi - lo)/5 in FP64
    from the FP32 limits, points lo + step/2 + i step, x outer."""
    f32 = np.float32
    lo, hi = f32(-trs), f32(trs)
    step = float(f32(hi - lo)) / float(COARSE_NUM_STEPS)
    pts = []
    for ix in range(COARSE_NUM_STEPS):
        x = float(lo) + step / 2.0 + float(ix) * step
        for iy in range(COARSE_NUM_STEPS):
            y = float(lo) + step / 2.0 + float(iy) * step
            pts.append((x, y))
    return pts


def shsrch_minimize(fdf, gencorrs, angtab_deg, trs, irot=None, opt_angle=True, maxits=SHSRCH_MAXITS,
                    max_evals=SHSRCH_MAX_EVALS, eq_tol=1e-10, trace=None, coarse_init=False):
    """pftcc_shsrch_grad%minimize (simple_pftcc_shsrch_grad.f90:254-338) for one (reference,
    particle) pair.

    fdf(shift_f64[2], irot) -> (cc, dcc[2])   gencorr_grad_for_rot_8 (FP64)
    gencorrs(shift_f32[2]) -> cc[nrots]       gencorrs with a shift (the FP32 search, :1502-1534)
    Returns (cxy[3] FP32: cc and the shift rotated to the reference frame, irot) with irot = 0 when
    no better solution than the start was found (the reference's convention).
    Optimiser: L-BFGS-B, m = 15, factr 1e7, pgtol 1e-5, line search ftol 0.1, gtol 1e-3, xtol 0.1,
    bounds [-trs, trs]^2 (grad_shsrch_new :46-82; opt_spec :22-40,271)."""
    f32 = np.float32
    lb = Lbfgsb(2, 15, [-trs, -trs], [trs, trs], [2, 2], factr=1.0e7, pgtol=1.0e-5, ftol=float(f32(0.1)),
                gtol=float(f32(1e-3)), xtol=0.1, maxls=20, eq_tol=eq_tol)
    state = {"cur": 0}

    def run_minimize(x32):
        """lbfgsb_minimize (:56-129): returns (lowest_cost FP32, x FP32)."""
        gen = lb.minimize(np.asarray(x32, np.float32).astype(np.float64))
        req = next(gen)
        try:
            while True:
                if req[0] == "FG":
                    cc, dcc = fdf(req[1], state["cur"])
                    req = gen.send((-cc, -np.asarray(dcc, float)))
                else:
                    _, xn, fn, nfgv = req
                    if opt_angle:
                        corrs = gencorrs(xn.astype(np.float32))
                        state["cur"] = int(np.argmax(corrs)) + 1
                    if trace is not None:
                        trace.append(("NEW_X", xn.copy(), fn, state["cur"]))
                    req = gen.send("STOP" if nfgv > maxits else None)
        except StopIteration as e:
            task, xf, ff, _ = e.value
        return f32(ff), xf.astype(np.float32)

    if opt_angle:
        x = np.zeros(2, np.float32)
        corrs = gencorrs(x)
        state["cur"] = int(np.argmax(corrs)) + 1
        lowest_overall = -float(corrs[state["cur"] - 1])
        if coarse_init:
            # coarse_search_opt_angle (:368-395): gencorrs at the FP32 grid points, strict minimum
            best, brot, bxy = np.inf, 0, (0.0, 0.0)
            for (gx, gy) in _coarse_grid(trs):
                c = gencorrs(np.array([gx, gy], np.float32))
                loc = int(np.argmax(c)) + 1
                cost = -float(c[loc - 1])
                if cost < best:
                    best, brot, bxy = cost, loc, (gx, gy)
            if brot != 0:
                x = np.array(bxy, np.float32)
                state["cur"] = brot
        for _ in range(max_evals):
            lowest, x = run_minimize(x)
            corrs = gencorrs(x)
            loc = int(np.argmax(corrs)) + 1
            if loc == state["cur"]:
                break
            state["cur"] = loc
        lowest = -float(f32(corrs[state["cur"] - 1]))
        if lowest < lowest_overall:
            irot_out = state["cur"]
            cxy = np.zeros(3, np.float32)
            cxy[0] = f32(-lowest)
            cxy[1:] = rotmat2d_apply(x, angtab_deg[irot_out - 1])
            return cxy, irot_out
        return None, 0
    state["cur"] = int(irot)
    lowest_overall = -fdf(np.zeros(2), state["cur"])[0]
    x0 = np.zeros(2, np.float32)
    if coarse_init:
        # coarse_search (:340-366): the FP64 objective at the grid points, strict minimum
        best, bxy = np.inf, (0.0, 0.0)
        for (gx, gy) in _coarse_grid(trs):
            cost = -fdf(np.array([gx, gy]), state["cur"])[0]
            if cost < best:
                best, bxy = cost, (gx, gy)
        if best < lowest_overall:
            lowest_overall = best
            x0 = np.array(bxy, np.float32)
    lowest, x = run_minimize(x0)
    if float(lowest) < lowest_overall:
        cxy = np.zeros(3, np.float32)
        cxy[0] = f32(-float(lowest))
        cxy[1:] = rotmat2d_apply(x, angtab_deg[int(irot) - 1])
        return cxy, int(irot)
    return None, 0
