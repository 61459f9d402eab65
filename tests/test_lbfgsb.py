"""CPU pins of the L-BFGS-B restatement (msmpscu_amd/lbfgsb.py) for the unbounded problems the
reference's DO_LBFGSB_FORSTEPS_DEV hands SETULB (MD_LBFGSScheme_GPU.F90:177-388, NBD = 0):

- the line search: the same step and outcome as scipy's Python port of MINPACK-2's dcsrch
  (scipy.optimize._dcsrch.DCSRCH) on one-dimensional functions with sufficient-decrease,
  curvature, extrapolation and bracketing cases;
- the whole method: the same evaluation points, to rounding, as scipy's L-BFGS-B 3.0
  (scipy.optimize.fmin_l_bfgs_b, the code the reference vendors as LIB/sor/f/LBFGSB/lbfgsb.f)
  over the first 30 evaluations, the same number of evaluations and the same convergence task
  on Rosenbrock problems, a convex quadratic's minimiser, the projected-gradient stop, and the
  abnormal termination of a search along an ascent-free but non-decreasing function;
- the quench loop: a displaced BCC W box relaxed under the oracle's forces to the closed-form
  perfect-lattice energy."""
import numpy as np
import pytest

from msmpscu_amd.lbfgsb import LBFGSB, Dcsrch

so = pytest.importorskip("scipy.optimize")


def rosen(x):
    f = np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] += -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
    g[1:] += 200 * (x[1:] - x[:-1] ** 2)
    return f, g


def run_mine(fun, x0, m, factr, pgtol):
    x = np.array(x0, dtype=np.float64)
    pts = []
    opt = LBFGSB(x.size, m, factr, pgtol)
    task = opt.start(x)
    f, g = 0.0, None
    while task.startswith("FG") or task == "NEW_X":
        if task.startswith("FG"):
            pts.append(x.copy())
            f, g = fun(x)
        task = opt.setulb(x, f, g)
    return x, pts, task


def run_scipy(fun, x0, m, factr, pgtol):
    pts = []

    def f(x):
        pts.append(x.copy())
        return fun(x)

    xs, fs, info = so.fmin_l_bfgs_b(f, np.array(x0, dtype=np.float64), m=m, factr=factr,
                                    pgtol=pgtol, maxls=20)
    return xs, pts, info["task"]


@pytest.mark.parametrize("n,m", [(2, 5), (10, 7)])
def test_iterates_match_scipy_lbfgsb(n, m):
    x0 = np.linspace(-1.2, 1.0, n)
    xm, pm, tm = run_mine(rosen, x0, m, 1e7, 1e-8)
    xs, ps, ts = run_scipy(rosen, x0, m, 1e7, 1e-8)
    assert len(pm) == len(ps)
    for a, b in zip(pm[:30], ps[:30]):
        assert np.max(np.abs(a - b)) <= 1e-10 * max(1.0, np.max(np.abs(b)))
    assert tm.startswith("CONVERGENCE: REL_REDUCTION_OF_F")
    ts = ts.decode() if isinstance(ts, bytes) else ts
    assert ts.startswith("CONVERGENCE: REL")
    assert np.max(np.abs(xm - xs)) <= 1e-8
    assert np.max(np.abs(xm - 1.0)) <= 1e-3


def test_quadratic_minimiser_and_pgtol_stop():
    rng = np.random.default_rng(5)
    a = rng.normal(size=(30, 30))
    hmat = a @ a.T + 30 * np.eye(30)
    b = rng.normal(size=30)

    def quad(x):
        return 0.5 * x @ hmat @ x - b @ x, hmat @ x - b

    x, pts, task = run_mine(quad, np.zeros(30), 5, 0.0, 1e-9)
    assert task.startswith("CONVERGENCE: NORM_OF_PROJECTED_GRADIENT")
    assert np.max(np.abs(hmat @ x - b)) <= 1e-9
    assert np.allclose(x, np.linalg.solve(hmat, b), rtol=0, atol=1e-9)
    # scipy stops with factr = 0 when rounding leaves f unchanged; the paths agree until then
    xs, ps, _ = run_scipy(quad, np.zeros(30), 5, 0.0, 1e-9)
    for a_, b_ in zip(pts[:20], ps[:20]):
        assert np.max(np.abs(a_ - b_)) <= 1e-12
    assert np.max(np.abs(xs - x)) <= 1e-8
    # a start that is already stationary stops at once
    _, p0, t0 = run_mine(quad, np.linalg.solve(hmat, b), 5, 0.0, 1e-6)
    assert len(p0) == 1 and t0.startswith("CONVERGENCE: NORM")


def test_abnormal_termination_matches_scipy():
    """A gradient inconsistent with the function makes the searches fail: restarts from the
    steepest descent and finally ABNORMAL_TERMINATION_IN_LNSRCH, after the same evaluations as
    scipy's L-BFGS-B."""
    def liar(x):
        return 1.0 + 1e-3 * float(np.sum(np.abs(x))), np.ones_like(x)

    _, pts, task = run_mine(liar, np.ones(4), 5, 0.0, 0.0)
    _, ps, ts = run_scipy(liar, np.ones(4), 5, 0.0, 0.0)
    ts = ts.decode() if isinstance(ts, bytes) else ts
    assert task == "ABNORMAL_TERMINATION_IN_LNSRCH" and ts.startswith("ABNORMAL")
    assert len(pts) == len(ps)
    for a_, b_ in zip(pts, ps):
        assert np.max(np.abs(a_ - b_)) <= 1e-12


@pytest.mark.parametrize("case", ["quad", "quartic", "steep", "flat"])
def test_line_search_matches_minpack2_port(case):
    from scipy.optimize._dcsrch import DCSRCH
    funcs = {
        "quad": (lambda a: (a - 3.0) ** 2, lambda a: 2.0 * (a - 3.0), 1.0),
        "quartic": (lambda a: (a - 0.2) ** 4 - a, lambda a: 4 * (a - 0.2) ** 3 - 1.0, 10.0),
        "steep": (lambda a: np.exp(4.0 * a) - 8.0 * a, lambda a: 4.0 * np.exp(4.0 * a) - 8.0, 2.0),
        "flat": (lambda a: -np.arctan(a), lambda a: -1.0 / (1.0 + a * a), 0.01),
    }
    phi, dphi, a1 = funcs[case]
    ref = DCSRCH(phi, dphi, ftol=1e-3, gtol=0.9, xtol=0.1, stpmin=0.0, stpmax=1e10)
    stp_ref, *_ , task_ref = ref(alpha1=a1, phi0=phi(0.0), derphi0=dphi(0.0), maxiter=20)
    ls = Dcsrch()
    task, stp = ls.start(phi(0.0), dphi(0.0), a1)
    for _ in range(20):
        task, stp = ls.step(phi(stp), dphi(stp))
        if task != "FG":
            break
    task_ref = task_ref.decode() if isinstance(task_ref, bytes) else str(task_ref)
    assert task.split(":")[0] == task_ref.split(":")[0].strip()
    assert stp == pytest.approx(stp_ref, rel=1e-14)


def test_quench_loop_relaxes_a_displaced_lattice(orc):
    """lbfgsb_quench (the DO_LBFGSB_FORSTEPS_DEV loop) under the oracle's forces, in CGS with
    PGTOL 1e-8 dyn (6e-5 eV/A): a displaced BCC W lattice relaxes to the closed-form
    perfect-lattice energy,
    the forces end below PGTOL, the positions stay in the box, and the step count counts every
    SETULB call."""
    import msmpscu_amd.mdlib as md
    from test_lbfgs import _oracle_provider, _perfect_w_energy
    from test_oracle_pins import bcc, w_table
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
    state["n"] = 0
    steps, iflag, x1, task = md.lbfgsb_quench(x, st, lo3, ln3, ifpd, evaluate, put, 400, 7,
                                              0.0, 1e-8)
    assert iflag == 0 and task.startswith("CONVERGENCE: NORM"), task
    assert steps > state["n"] > 1  # NEW_X calls count as steps too
    f, e = evaluate()
    assert np.max(np.abs(f)) * md.CP_EVERG * md.CP_CM2A <= 1e-8
    ep = n * _perfect_w_energy(a)
    assert ep - 1e-9 <= e.sum() < ep + 1e-4 * (e0 - ep)
    assert np.all(x1 >= lo) and np.all(x1 <= lo + ln)
    assert np.array_equal(x1, state["x"])
