"""L-BFGS-B in reverse communication, for the problems the reference's quench hands it: every
variable unbounded.

The reference quenches with SETULB of L-BFGS-B 3.0 (Byrd, Lu, Nocedal, Zhu, SIAM J. Sci. Comput.
16 (1995) 1190-1208; Zhu, Byrd, Lu, Nocedal, ACM TOMS 23 (1997) 550-560; Morales and Nocedal,
ACM TOMS 38 (2011) 7), vendored as LIB/sor/f/LBFGSB/lbfgsb.f, and its driver
DO_LBFGSB_FORSTEPS_DEV (MDLIB/sor/CommonGPU/MD_LBFGSScheme_GPU.F90:177-388) sets NBD = 0 for
every variable.  With no bounds the method reduces to this iteration (mainlb, lbfgsb.f:287-1000):

- the first iteration, and the first after a restart, moves along d = -g (the generalized Cauchy
  point of the model B = theta I with theta = 1 is x - g), later ones along d = -H g, H the
  limited-memory BFGS inverse of B = theta I - W M W' built from the last m pairs
  (s, y) with theta = y'y / y's of the newest pair (the subspace minimisation over all free
  variables is the unconstrained minimiser of the model);
- the line search is MINPACK-2's dcsrch/dcstep (Moré and Thuente, ACM TOMS 20 (1994) 286-307)
  with ftol = 1e-3, gtol = 0.9, xtol = 0.1, stpmin = 0, stpmax = 1e10, the first trial step
  1/|d| on the first iteration and 1 afterwards (a trial of step 1 is the model minimiser z
  itself); a search that fails (an ascent direction, or 20 backtracks) restores the last iterate
  and restarts from the steepest descent, or ends with ABNORMAL_TERMINATION_IN_LNSRCH when it
  already was a steepest-descent step;
- a new iterate is accepted when the search converges or warns; then the algorithm stops if
  max |g| <= pgtol, or if (f_old - f) <= factr * eps * max(|f_old|, |f|, 1);
- the pair (s, y) = (stp d, g_new - g_old) enters the memory unless y's = stp (g_new.d - g_old.d)
  <= eps * (-stp g_old.d).

H g is formed by the two-loop recursion (Nocedal, Math. Comp. 35 (1980) 773-782), which is the
same matrix as the compact form of mainlb in exact arithmetic.  Tasks and their strings follow
SETULB: 'FG' (evaluate f, g at x and call again), 'NEW_X' (an iterate was accepted; call again),
'CONVERGENCE: ...', 'ABNORMAL_TERMINATION_IN_LNSRCH', 'ERROR: ...'.
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

EPSMCH = float(np.finfo(np.float64).eps)
BIG = 1.0e10
FTOL, GTOL, XTOL = 1.0e-3, 0.9, 0.1


def dcstep(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    """dcstep (MINPACK-2): the safeguarded step of the Moré-Thuente search from the interval
    endpoints (stx, fx, dx), (sty, fy, dy) and the trial (stp, fp, dp).  Returns the updated
    (stx, fx, dx, sty, fy, dy, stp, brackt)."""
    sgnd = dp * (dx / abs(dx))
    if fp > fx:
        # a higher value: the minimum is bracketed; cubic step if closer to stx, else the mean
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        r = p / q
        stpc = stx + r * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0.0:
        # lower value, derivatives of opposite sign: bracketed; the step farther from stp
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        r = p / q
        stpc = stp + r * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(dp) < abs(dx):
        # lower value, same sign, |derivative| decreasing: the cubic may not have a minimiser
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt(max(0.0, (theta / s) ** 2 - (dx / s) * (dp / s)))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        if r < 0.0 and gamma != 0.0:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stpc - stp) < abs(stpq - stp) else stpq
            if stp > stx:
                stpf = min(stp + 0.66 * (sty - stp), stpf)
            else:
                stpf = max(stp + 0.66 * (sty - stp), stpf)
        else:
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            stpf = min(stpmax, stpf)
            stpf = max(stpmin, stpf)
    else:
        # lower value, same sign, |derivative| not decreasing
        if brackt:
            theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp
            s = max(abs(theta), abs(dy), abs(dp))
            gamma = s * math.sqrt((theta / s) ** 2 - (dy / s) * (dp / s))
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            r = p / q
            stpf = stp + r * (sty - stp)
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0.0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    return stx, fx, dx, sty, fy, dy, stpf, brackt


class Dcsrch:
    """dcsrch (MINPACK-2) in reverse communication: start(f, g, stp) then step(f, g) after each
    evaluation; both return (task, stp) with task 'FG' (evaluate at stp), 'CONVERGENCE',
    'WARNING: ...' or 'ERROR: ...'."""

    XTRAPL, XTRAPU = 1.1, 4.0

    def __init__(self, ftol=FTOL, gtol=GTOL, xtol=XTOL, stpmin=0.0, stpmax=BIG):
        self.ftol, self.gtol, self.xtol = ftol, gtol, xtol
        self.stpmin, self.stpmax = stpmin, stpmax

    def start(self, f, g, stp):
        if stp < self.stpmin:
            return "ERROR: STP .LT. STPMIN", stp
        if stp > self.stpmax:
            return "ERROR: STP .GT. STPMAX", stp
        if g >= 0.0:
            return "ERROR: INITIAL G .GE. ZERO", stp
        self.brackt = False
        self.stage = 1
        self.finit, self.ginit = f, g
        self.gtest = self.ftol * g
        self.width = self.stpmax - self.stpmin
        self.width1 = self.width / 0.5
        self.stx, self.fx, self.gx = 0.0, f, g
        self.sty, self.fy, self.gy = 0.0, f, g
        self.stmin = 0.0
        self.stmax = stp + self.XTRAPU * stp
        self.stp = stp
        return "FG", stp

    def step(self, f, g):
        stp = self.stp
        ftest = self.finit + stp * self.gtest
        if self.stage == 1 and f <= ftest and g >= 0.0:
            self.stage = 2
        task = ""
        if self.brackt and (stp <= self.stmin or stp >= self.stmax):
            task = "WARNING: ROUNDING ERRORS PREVENT PROGRESS"
        if self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax:
            task = "WARNING: XTOL TEST SATISFIED"
        if stp == self.stpmax and f <= ftest and g <= self.gtest:
            task = "WARNING: STP = STPMAX"
        if stp == self.stpmin and (f > ftest or g >= self.gtest):
            task = "WARNING: STP = STPMIN"
        if f <= ftest and abs(g) <= self.gtol * (-self.ginit):
            task = "CONVERGENCE"
        if task:
            return task, stp
        if self.stage == 1 and f <= self.fx and f > ftest:
            # the modified function psi(stp) = f - stp * gtest until the first stage ends
            fm, gm = f - stp * self.gtest, g - self.gtest
            fxm, gxm = self.fx - self.stx * self.gtest, self.gx - self.gtest
            fym, gym = self.fy - self.sty * self.gtest, self.gy - self.gtest
            (self.stx, fxm, gxm, self.sty, fym, gym, stp, self.brackt) = dcstep(
                self.stx, fxm, gxm, self.sty, fym, gym, stp, fm, gm, self.brackt, self.stmin,
                self.stmax)
            self.fx, self.gx = fxm + self.stx * self.gtest, gxm + self.gtest
            self.fy, self.gy = fym + self.sty * self.gtest, gym + self.gtest
        else:
            (self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, stp, self.brackt) = dcstep(
                self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, stp, f, g, self.brackt,
                self.stmin, self.stmax)
        if self.brackt:
            # bisect when the interval did not shrink enough over two steps
            if abs(self.sty - self.stx) >= 0.66 * self.width1:
                stp = self.stx + 0.5 * (self.sty - self.stx)
            self.width1 = self.width
            self.width = abs(self.sty - self.stx)
            self.stmin = min(self.stx, self.sty)
            self.stmax = max(self.stx, self.sty)
        else:
            self.stmin = stp + self.XTRAPL * (stp - self.stx)
            self.stmax = stp + self.XTRAPU * (stp - self.stx)
        stp = max(stp, self.stpmin)
        stp = min(stp, self.stpmax)
        if (self.brackt and (stp <= self.stmin or stp >= self.stmax)) or \
                (self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax):
            stp = self.stx
        self.stp = stp
        return "FG", stp


class LBFGSB:
    """setulb for n unbounded variables with m corrections.  Use: opt = LBFGSB(n, m, factr,
    pgtol); task = opt.start(x); while task starts with 'FG' or is 'NEW_X': on 'FG' evaluate
    f, g at x, then task = opt.setulb(x, f, g) (x is updated in place)."""

    def __init__(self, n: int, m: int, factr: float, pgtol: float, maxls: int = 20):
        if n <= 0:
            raise ValueError("ERROR: N .LE. 0")
        if m <= 0:
            raise ValueError("ERROR: M .LE. 0")
        if factr < 0.0:
            raise ValueError("ERROR: FACTR .LT. 0")
        self.n, self.m = n, m
        self.tol = factr * EPSMCH
        self.pgtol = pgtol
        self.maxls = maxls
        self.s: List[np.ndarray] = []
        self.y: List[np.ndarray] = []
        self.rho: List[float] = []
        self.theta = 1.0
        self.iter = 0
        self.nfgv = 0
        self.nskip = 0
        self.task = "START"

    def start(self, x: np.ndarray) -> str:
        self.task = "FG_START"
        return self.task

    # direction of the next iteration
    def _direction(self, g):
        if not self.s:
            return -g / self.theta
        q = g.copy()
        alpha = []
        for s, y, rho in zip(reversed(self.s), reversed(self.y), reversed(self.rho)):
            a = rho * float(np.dot(s, q))
            alpha.append(a)
            q -= a * y
        r = q / self.theta
        for (s, y, rho), a in zip(zip(self.s, self.y, self.rho), reversed(alpha)):
            b = rho * float(np.dot(y, r))
            r += (a - b) * s
        return -r

    def _restart_memory(self):
        self.s.clear()
        self.y.clear()
        self.rho.clear()
        self.theta = 1.0

    def _begin_search(self, x):
        # z = the model minimiser, d = z - x as mainlb forms them (555)
        self.z = x + self._direction(self.g)
        self.d = self.z - x
        self.dtd = float(np.dot(self.d, self.d))
        self.dnorm = math.sqrt(self.dtd)
        self.stp = min(1.0 / self.dnorm, BIG) if self.iter == 0 else 1.0
        self.t = x.copy()
        self.r = self.g.copy()
        self.fold = self.f
        self.ifun = 0
        self.iback = 0
        self.ls = Dcsrch()
        self.first = True

    def _search(self, x, f, g):
        """lnsrlb: returns 'FG_LNSRCH' (x holds the next trial), 'NEW_X' or 'ERROR_ASCENT'."""
        gd = float(np.dot(g, self.d))
        if self.first:
            self.gdold = gd
            if gd >= 0.0:
                return "ERROR_ASCENT"
            task, stp = self.ls.start(f, gd, self.stp)
            self.first = False
        else:
            task, stp = self.ls.step(f, gd)
        self.gd = gd
        self.stp = stp
        if task.startswith("CONV") or task.startswith("WARN"):
            return "NEW_X"
        self.ifun += 1
        self.nfgv += 1
        self.iback = self.ifun - 1
        if stp == 1.0:
            x[:] = self.z
        else:
            x[:] = stp * self.d + self.t
        return "FG_LNSRCH"

    def _iterate(self, x):
        """From an accepted point (self.f, self.g at x): start the next line search; returns the
        task."""
        self._begin_search(x)
        res = self._search(x, self.f, self.g)
        return self._after_search(x, res)

    def _after_search(self, x, res):
        if res == "ERROR_ASCENT" or (res == "FG_LNSRCH" and self.iback >= self.maxls):
            x[:] = self.t
            self.g = self.r.copy()
            self.f = self.fold
            if not self.s:
                self.iter += 1
                self.task = "ABNORMAL_TERMINATION_IN_LNSRCH"
                return self.task
            self._restart_memory()
            return self._iterate(x)
        if res == "FG_LNSRCH":
            self.task = "FG_LNSRCH"
            return self.task
        self.iter += 1
        self.task = "NEW_X"
        return self.task

    def setulb(self, x: np.ndarray, f: float, g: np.ndarray) -> str:
        task = self.task
        if task == "FG_START":
            self.nfgv = 1
            self.f, self.g = float(f), np.array(g, dtype=np.float64)
            if float(np.max(np.abs(self.g))) <= self.pgtol:
                self.task = "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL"
                return self.task
            return self._iterate(x)
        if task == "FG_LNSRCH":
            self.f, self.g = float(f), np.array(g, dtype=np.float64)
            res = self._search(x, self.f, self.g)
            return self._after_search(x, res)
        if task == "NEW_X":
            if float(np.max(np.abs(self.g))) <= self.pgtol:
                self.task = "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL"
                return self.task
            ddum = max(abs(self.fold), abs(self.f), 1.0)
            if (self.fold - self.f) <= self.tol * ddum:
                self.task = "CONVERGENCE: REL_REDUCTION_OF_F_<=_FACTR*EPSMCH"
                return self.task
            y = self.g - self.r
            rr = float(np.dot(y, y))
            if self.stp == 1.0:
                dr = self.gd - self.gdold
                ddum = -self.gdold
                s = self.d
            else:
                dr = (self.gd - self.gdold) * self.stp
                s = self.d * self.stp
                ddum = -self.gdold * self.stp
            if dr <= EPSMCH * ddum:
                self.nskip += 1
            else:
                if len(self.s) == self.m:
                    self.s.pop(0)
                    self.y.pop(0)
                    self.rho.pop(0)
                self.s.append(np.array(s))
                self.y.append(y)
                self.rho.append(1.0 / dr)
                self.theta = rr / dr
            return self._iterate(x)
        raise RuntimeError(f"setulb called in state {task!r}")
