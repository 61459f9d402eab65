"""The host L-BFGS optimiser (msmpscu_amd/lbfgs.py, a restatement of the reference's LBFGS with
the Moré-Thuente line search) on problems with known minimisers, and the quench loop around it
(coordinate selection, CGS units, periodic wrap) driven by the oracle's forces.  CPU only."""
import math

import numpy as np
import pytest

import msmpscu_amd.mdlib as md
from msmpscu_amd.lbfgs import LBFGS, P66_SINGLE
from test_oracle_pins import bcc, w_closed, w_table


def _minimise(fg, x0, m, eps, maxeval):
    x = np.array(x0, dtype=np.float64)
    opt = LBFGS(x.size, m, eps, 1e-16)
    fs = []
    for k in range(maxeval):
        f, g = fg(x)
        fs.append(f)
        iflag = opt.step(x, f, g)
        if iflag <= 0:
            return x, iflag, k + 1, opt
    return x, 1, maxeval, opt


def rosenbrock(x):
    """Extended Rosenbrock (pairs (x_{2i-1}, x_{2i})), minimum 0 at all ones."""
    a, b = x[0::2], x[1::2]
    f = float(np.sum(100.0 * (b - a * a) ** 2 + (1.0 - a) ** 2))
    g = np.zeros_like(x)
    g[0::2] = -400.0 * a * (b - a * a) - 2.0 * (1.0 - a)
    g[1::2] = 200.0 * (b - a * a)
    return f, g


def test_rosenbrock_from_the_standard_start():
    x, iflag, nev, opt = _minimise(rosenbrock, [-1.2, 1.0], 5, 1e-10, 500)
    assert iflag == 0 and nev < 200
    assert np.allclose(x, 1.0, atol=1e-8)


def test_extended_rosenbrock_100():
    x0 = np.tile([-1.2, 1.0], 50)
    x, iflag, nev, opt = _minimise(rosenbrock, x0, 7, 1e-9, 3000)
    assert iflag == 0
    assert np.allclose(x, 1.0, atol=1e-6)
    assert opt.iter > 10 and opt.nfun >= opt.iter


def test_quadratic_minimiser_and_monotone_decrease():
    rng = np.random.default_rng(1)
    n = 60
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    a = q @ np.diag(np.linspace(1.0, 100.0, n)) @ q.T
    b = rng.normal(size=n)
    xs = np.linalg.solve(a, b)
    accepted = []

    def fg(x):
        return float(0.5 * x @ a @ x - b @ x), a @ x - b

    x = np.zeros(n)
    # |g| / max(1, |x|) <= 1e-6: far enough above the rounding floor of f near the minimum,
    # where a line search can no longer see a decrease
    opt = LBFGS(n, 7, 1e-6, 1e-16)
    for _ in range(5000):
        f, g = fg(x)
        it0 = opt.iter
        iflag = opt.step(x, f, g)
        if opt.iter != it0 or iflag == 0:
            accepted.append(f)  # f at the point the last line search accepted
        if iflag <= 0:
            break
    assert iflag == 0
    assert np.allclose(x, xs, atol=1e-6)
    # every accepted step decreased f (sufficient-decrease condition of the line search)
    assert all(f1 < f0 for f0, f1 in zip(accepted, accepted[1:]))


def test_errors_and_restart():
    """A gradient of the wrong sign: the line search exhausts its 20 evaluations (iflag -1);
    the next call restarts the method from the current point; n = 0 is an input error."""
    def bad(x):
        return float(np.sum(x * x)), -2.0 * x

    x = np.array([1.0, -2.0])
    opt = LBFGS(2, 3, 1e-10, 1e-16)
    for k in range(25):
        f, g = bad(x)
        iflag = opt.step(x, f, g)
        if iflag != 1:
            break
    assert iflag == -1 and opt.ls.info in (3, 6) and k <= 20
    x = np.array([-1.2, 1.0])
    f, g = rosenbrock(x)
    assert opt.step(x, f, g) == 1 and opt.iter == 1
    assert LBFGS(0, 3, 1e-5, 1e-16).step(np.zeros(0), 0.0, np.zeros(0)) == -3


def test_mcstep_safeguard_constant_is_the_single_precision_literal():
    assert P66_SINGLE == float(np.float32(0.66)) and P66_SINGLE != 0.66


def test_coordinate_selection():
    """IDX/IDY/IDZ: active, not OUTOFBOX, FIXPOS bit of that direction clear."""
    st = np.array([1, 0, 1 | 2, 1 | 4 | 8, 1 | 65536, 1 | 16 | 64, 3 << 1], np.int32)
    ix, iy, iz = md.lbfgs_coordinates(st)
    assert list(ix) == [0, 3, 5] and list(iy) == [0, 2, 5] and list(iz) == [0, 2, 5]


def _oracle_provider(orc, x, st, it, lo, ln, ifpd, t):
    kv, ind = orc.nlist("full", x, it, st, lo, ln, ifpd, np.full((1, 1), 4.7002), 96)
    state = {"x": np.array(x), "n": 0}

    def evaluate():
        state["n"] += 1
        f, _, e = orc.force_full(state["x"], it, st, kv, ind, ln, ifpd, 4.4002 ** 2, t)
        return f, e

    def put(xa):
        state["x"] = np.array(xa)

    return evaluate, put, state, (kv, ind)


def _perfect_w_energy(a):
    shells = [(a * math.sqrt(3) / 2, 8), (a, 6), (a * math.sqrt(2), 12)]
    v, rho = w_closed(np.array([s[0] for s in shells]))
    nsh = np.array([s[1] for s in shells])
    return 0.5 * np.sum(nsh * v) - math.sqrt(np.sum(nsh * rho))


def _converged(f, x1, st):
    """The optimiser's own test |G| <= 1e-4 max(1, |X|), G and X in CGS."""
    ids = md.lbfgs_coordinates(st)
    g = np.concatenate([f[ids[k], k] for k in range(3)]) * md.CP_EVERG * md.CP_CM2A
    xc = np.concatenate([x1[ids[k], k] for k in range(3)]) / md.CP_CM2A
    return np.linalg.norm(g) <= 1e-4 * max(1.0, np.linalg.norm(xc))


def test_quench_relaxes_a_displaced_lattice(orc):
    """The quench of a displaced BCC W lattice (oracle forces, fixed list) stops on the
    optimiser's convergence test in CGS; the energy has then dropped to within 0.5% of the way
    from the start to the closed-form perfect-lattice energy, never below it."""
    a = 3.165
    x, lo, ln = bcc(4, a, 0.04, 11)
    n = len(x)
    t = w_table(orc, 100000)
    it = np.zeros(n, np.int32)
    st = np.ones(n, np.int32)
    ifpd = (1, 1, 1)
    evaluate, put, state, _ = _oracle_provider(orc, x, st, it, lo, ln, ifpd, t)
    lo3, ln3 = np.tile(lo, (n, 1)), np.tile(ln, (n, 1))
    e0 = float(evaluate()[1].sum())
    steps, iflag, x1 = md.lbfgs_quench(x, st, lo3, ln3, ifpd, evaluate, put, 200, 7)
    assert iflag == 0 and 0 < steps < 200
    assert np.array_equal(x1, state["x"])
    f, e = evaluate()
    assert _converged(f, x1, st)
    ep = n * _perfect_w_energy(a)
    assert ep - 1e-9 <= e.sum() < ep + 0.005 * (e0 - ep)
    assert np.all(x1 >= lo) and np.all(x1 <= lo + ln)


def test_quench_keeps_fixed_inactive_and_out_of_box_coordinates(orc):
    """Coordinates with a FIXPOS bit, inactive atoms and OUTOFBOX atoms stay bit-identical; the
    rest relax (lower energy, convergence test met)."""
    a = 3.165
    x, lo, ln = bcc(4, a, 0.04, 11)
    n = len(x)
    t = w_table(orc, 100000)
    it = np.zeros(n, np.int32)
    st = np.ones(n, np.int32)
    st[5] |= 2
    st[9] |= 4 | 8
    st[17] = 0
    st[30] |= 65536
    ifpd = (1, 1, 1)
    evaluate, put, state, _ = _oracle_provider(orc, x, st, it, lo, ln, ifpd, t)
    lo3, ln3 = np.tile(lo, (n, 1)), np.tile(ln, (n, 1))
    e0 = float(evaluate()[1].sum())
    steps, iflag, x1 = md.lbfgs_quench(x, st, lo3, ln3, ifpd, evaluate, put, 200, 7)
    assert iflag == 0
    f, e = evaluate()
    assert _converged(f, x1, st) and e.sum() < e0
    assert x1[5, 0] == x[5, 0] and x1[9, 1] == x[9, 1] and x1[9, 2] == x[9, 2]
    assert x1[5, 1] != x[5, 1] and x1[9, 0] != x[9, 0]
    assert np.array_equal(x1[17], x[17]) and np.array_equal(x1[30], x[30])


def test_first_trial_and_single_wrap(orc):
    """The first trial point is the reference's, X0 - G0/|G0| in CGS (the first L-BFGS step has
    length 1 in X's unit, the centimetre), each coordinate then moved by one box edge if it left
    the box (below: +edge, above: -edge) -- one shift, so a trial far outside stays outside, as in
    the reference; coordinates outside the selection keep their bits; the quench still
    converges from there."""
    a = 3.165
    x, lo, ln = bcc(4, a, 0.04, 12)
    n = len(x)
    t = w_table(orc, 100000)
    it = np.zeros(n, np.int32)
    st = np.ones(n, np.int32)
    st[3] |= 4
    ifpd = (1, 1, 1)
    lo3, ln3 = np.tile(lo, (n, 1)), np.tile(ln, (n, 1))
    evaluate, put, state, _ = _oracle_provider(orc, x, st, it, lo, ln, ifpd, t)
    f0, _ = evaluate()
    state["n"] = 0
    seen = []

    def put_seen(xa):
        seen.append(np.array(xa))
        put(xa)

    steps, iflag, x1 = md.lbfgs_quench(x, st, lo3, ln3, ifpd, evaluate, put_seen, 200, 7)
    assert iflag == 0 and len(seen) == steps == state["n"]
    ids = md.lbfgs_coordinates(st)
    xc = np.concatenate([x[ids[k], k] for k in range(3)]) / md.CP_CM2A
    g0 = -np.concatenate([f0[ids[k], k] for k in range(3)]) * md.CP_EVERG * md.CP_CM2A
    lb = np.concatenate([np.full(len(ids[k]), lo[k]) for k in range(3)]) / md.CP_CM2A
    bs = np.concatenate([np.full(len(ids[k]), ln[k]) for k in range(3)]) / md.CP_CM2A
    xt = xc - g0 / np.linalg.norm(g0)
    assert np.max(np.abs(xt - xc)) > 10 * bs[0]  # far outside the box
    want = np.where(xt < lb, xt + bs, np.where(xt > lb + bs, xt - bs, xt)) * md.CP_CM2A
    got = np.concatenate([seen[0][ids[k], k] for k in range(3)])
    assert np.allclose(got, want, rtol=1e-12, atol=0.0)
    assert seen[0][3, 1] == x[3, 1]
    assert np.count_nonzero((want < -0.5 * ln[0]) | (want > 0.5 * ln[0])) > 0
