"""Limited-memory BFGS in reverse communication, with the Moré-Thuente line search.

Restates the reference's optimiser, LIB/sor/f/LBFGS/LBFGS.F90 (LBFGS :15-397, MCSRCH :577-867,
MCSTEP :868-1094): the method of Liu and Nocedal (Math. Programming B 45 (1989) 503-528) with
the two-loop recursion of Nocedal (Math. Comp. 35 (1980) 773-782) for -H g, the diagonal scaling
y.s / y.y, and the line search of Moré and Thuente (ACM TOMS 20 (1994) 286-307) that finds a step
with sufficient decrease (FTOL) and curvature (GTOL) inside an interval of uncertainty kept by
the safeguarded cubic/quadratic step of MCSTEP.  The settings are the reference's module
defaults: GTOL = 1e-2, STPMIN = 1e-20, STPMAX = 1e20, FTOL = 1e-4, at most 20 evaluations per
line search, the first trial step 1/|g|.  The safeguard 0.66 of MCSTEP is the single-precision
literal of the Fortran source (0.660000026...), that of MCSRCH the double 0.66.

Use: opt = LBFGS(n, m, eps, xtol); then repeatedly evaluate f, g at x and call
iflag = opt.step(x, f, g) -- x is updated in place; iflag 1 asks for another evaluation at the
new x, 0 means |g| / max(1, |x|) <= eps, a negative value an error (-1 the line search failed,
-3 bad n or m); calling again after an error restarts the method from the current x, as the
reference's computed GO TO does.
"""
from __future__ import annotations

import math

import numpy as np

P66_SINGLE = float(np.float32(0.66))  # the REAL literal 0.66 of MCSTEP


def _mcstep(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    """MCSTEP (LBFGS.F90:868-1094): safeguarded step of the Moré-Thuente search.  Returns
    (stx, fx, dx, sty, fy, dy, stp, brackt, info)."""
    info = 0
    if ((brackt and (stp <= min(stx, sty) or stp >= max(stx, sty))) or dx * (stp - stx) >= 0.0
            or stpmax < stpmin):
        return stx, fx, dx, sty, fy, dy, stp, brackt, info
    sgnd = dp * (dx / abs(dx))
    if fp > fx:
        # case 1: higher function value -- the minimum is bracketed
        info, bound = 1, True
        theta = 3 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        stpc = stx + (p / q) * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2) * (stp - stx)
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2
        brackt = True
    elif sgnd < 0.0:
        # case 2: lower value, derivatives of opposite sign -- bracketed
        info, bound = 2, False
        theta = 3 * (fx - fp) / (stp - stx) + dx + dp
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
        # case 3: lower value, same sign, derivative decreasing in magnitude
        info, bound = 3, True
        theta = 3 * (fx - fp) / (stp - stx) + dx + dp
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
            stpf = stpc if abs(stp - stpc) < abs(stp - stpq) else stpq
        else:
            stpf = stpc if abs(stp - stpc) > abs(stp - stpq) else stpq
    else:
        # case 4: lower value, same sign, derivative not decreasing
        info, bound = 4, False
        if brackt:
            theta = 3 * (fp - fy) / (sty - stp) + dy + dp
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
    # update the interval of uncertainty
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0.0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    stpf = max(stpmin, min(stpmax, stpf))
    stp = stpf
    if brackt and bound:
        if sty > stx:
            stp = min(stx + P66_SINGLE * (sty - stx), stp)
        else:
            stp = max(stx + P66_SINGLE * (sty - stx), stp)
    return stx, fx, dx, sty, fy, dy, stp, brackt, info


class _LineSearch:
    """MCSRCH (LBFGS.F90:577-867) as a resumable object: start() then resume() after each
    evaluation; both set self.info (-1: evaluate at x and resume; 1: done; other: failure)."""

    def __init__(self, gtol, ftol, xtol, stpmin, stpmax, maxfev):
        self.gtol, self.ftol, self.xtol = gtol, ftol, xtol
        self.stpmin, self.stpmax, self.maxfev = stpmin, stpmax, maxfev

    def start(self, x, f, g, s, stp):
        self.info = 0
        self.infoc = 1
        self.s = s
        self.stp = stp
        if (x.size <= 0 or stp <= 0.0 or self.ftol < 0.0 or self.gtol < 0.0 or self.xtol < 0.0
                or self.stpmin < 0.0 or self.stpmax < self.stpmin or self.maxfev <= 0):
            return
        self.dginit = float(np.dot(g, s))
        if self.dginit >= 0.0:  # not a descent direction
            return
        self.brackt = False
        self.stage1 = True
        self.nfev = 0
        self.finit = f
        self.dgtest = self.ftol * self.dginit
        self.width = self.stpmax - self.stpmin
        self.width1 = self.width / 0.5
        self.wa = x.copy()
        self.stx, self.fx, self.dgx = 0.0, self.finit, self.dginit
        self.sty, self.fy, self.dgy = 0.0, self.finit, self.dginit
        self._next_trial(x)

    def _next_trial(self, x):
        if self.brackt:
            self.stmin = min(self.stx, self.sty)
            self.stmax = max(self.stx, self.sty)
        else:
            self.stmin = self.stx
            self.stmax = self.stp + 4.0 * (self.stp - self.stx)
        self.stp = min(max(self.stp, self.stpmin), self.stpmax)
        if ((self.brackt and (self.stp <= self.stmin or self.stp >= self.stmax))
                or self.nfev >= self.maxfev - 1 or self.infoc == 0
                or (self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax)):
            self.stp = self.stx
        x[:] = self.wa + self.stp * self.s
        self.info = -1

    def resume(self, x, f, g):
        self.info = 0
        self.nfev += 1
        dg = float(np.dot(g, self.s))
        ftest1 = self.finit + self.stp * self.dgtest
        if (self.brackt and (self.stp <= self.stmin or self.stp >= self.stmax)) or self.infoc == 0:
            self.info = 6
        if self.stp == self.stpmax and f <= ftest1 and dg <= self.dgtest:
            self.info = 5
        if self.stp == self.stpmin and (f > ftest1 or dg >= self.dgtest):
            self.info = 4
        if self.nfev >= self.maxfev:
            self.info = 3
        if self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax:
            self.info = 2
        if f <= ftest1 and abs(dg) <= self.gtol * (-self.dginit):
            self.info = 1
        if self.info != 0:
            return
        if self.stage1 and f <= ftest1 and dg >= min(self.ftol, self.gtol) * self.dginit:
            self.stage1 = False
        if self.stage1 and f <= self.fx and f > ftest1:
            # the modified function f - stp*dgtest while in the first stage
            fm = f - self.stp * self.dgtest
            fxm = self.fx - self.stx * self.dgtest
            fym = self.fy - self.sty * self.dgtest
            dgm = dg - self.dgtest
            dgxm = self.dgx - self.dgtest
            dgym = self.dgy - self.dgtest
            (self.stx, fxm, dgxm, self.sty, fym, dgym, self.stp, self.brackt,
             self.infoc) = _mcstep(self.stx, fxm, dgxm, self.sty, fym, dgym, self.stp, fm, dgm,
                                   self.brackt, self.stmin, self.stmax)
            self.fx = fxm + self.stx * self.dgtest
            self.fy = fym + self.sty * self.dgtest
            self.dgx = dgxm + self.dgtest
            self.dgy = dgym + self.dgtest
        else:
            (self.stx, self.fx, self.dgx, self.sty, self.fy, self.dgy, self.stp, self.brackt,
             self.infoc) = _mcstep(self.stx, self.fx, self.dgx, self.sty, self.fy, self.dgy,
                                   self.stp, f, dg, self.brackt, self.stmin, self.stmax)
        if self.brackt:
            if abs(self.sty - self.stx) >= 0.66 * self.width1:
                self.stp = self.stx + 0.5 * (self.sty - self.stx)
            self.width1 = self.width
            self.width = abs(self.sty - self.stx)
        self._next_trial(x)


class LBFGS:
    """The reverse-communication driver of LBFGS (LBFGS.F90:15-397)."""

    def __init__(self, n: int, m: int, eps: float, xtol: float, gtol: float = 1.0e-2,
                 ftol: float = 1.0e-4, stpmin: float = 1.0e-20, stpmax: float = 1.0e20,
                 maxfev: int = 20):
        self.n, self.m, self.eps = int(n), int(m), eps
        if gtol <= 1.0e-4:
            gtol = 0.9  # as LBFGS resets a GTOL that is too small
        self.ls = _LineSearch(gtol, ftol, xtol, stpmin, stpmax, maxfev)
        self.iflag = 0
        self.iter = 0
        self.nfun = 0

    def _init(self, g):
        n, m = self.n, self.m
        self.iter = 0
        self.nfun = 1
        self.point = 0
        self.diag = np.ones(n)
        self.s = np.zeros((m, n))    # last m search steps
        self.y = np.zeros((m, n))    # last m gradient differences
        self.rho = np.zeros(m)
        self.alpha = np.zeros(m)
        self.s[0] = -g * self.diag
        self.stp1 = 1.0 / math.sqrt(float(np.dot(g, g)))

    def _direction(self, g):
        """-H g into s[point] by the two-loop recursion (LBFGS.F90 labels 80-145)."""
        m, bound = self.m, min(self.iter - 1, self.m)
        last = self.npt
        ys = float(np.dot(self.y[last], self.s[last]))
        yy = float(np.dot(self.y[last], self.y[last]))
        self.diag[:] = ys / yy
        cp = self.point if self.point != 0 else m
        self.rho[cp - 1] = 1.0 / ys
        q = -g.copy()
        cp = self.point
        for _ in range(bound):
            cp -= 1
            if cp == -1:
                cp = m - 1
            sq = float(np.dot(self.s[cp], q))
            self.alpha[cp] = self.rho[cp] * sq
            q -= self.alpha[cp] * self.y[cp]
        q *= self.diag
        for _ in range(bound):
            yr = float(np.dot(self.y[cp], q))
            beta = self.alpha[cp] - self.rho[cp] * yr
            q += beta * self.s[cp]
            cp += 1
            if cp == m:
                cp = 0
        self.s[self.point] = q

    def _new_iteration(self, x, f, g):
        self.iter += 1
        if self.iter > 1:
            self._direction(g)
        stp = self.stp1 if self.iter == 1 else 1.0
        self.w = g.copy()
        self.ls.start(x, f, g, self.s[self.point], stp)

    def step(self, x: np.ndarray, f: float, g: np.ndarray) -> int:
        if self.iflag != 1:
            # first call, or a restart after an error (the reference falls into its set-up)
            if self.n <= 0 or self.m <= 0:
                self.iflag = -3
                return self.iflag
            self._init(g)
            self._new_iteration(x, f, g)
        else:
            self.ls.resume(x, f, g)
        while True:
            if self.ls.info == -1:
                self.iflag = 1
                return self.iflag
            if self.ls.info != 1:
                self.iflag = -1
                return self.iflag
            self.nfun += self.ls.nfev
            self.npt = self.point
            self.s[self.npt] *= self.ls.stp
            self.y[self.npt] = g - self.w
            self.point = (self.point + 1) % self.m
            gnorm = math.sqrt(float(np.dot(g, g)))
            xnorm = max(1.0, math.sqrt(float(np.dot(x, x))))
            if gnorm / xnorm <= self.eps:
                self.iflag = 0
                return self.iflag
            self._new_iteration(x, f, g)
