"""Pins of the oracle's adjoint (orc_adjoint_full) against central finite differences of the
oracle's forward force/energy (orc_force_full), knot by knot, for FS and EAM tables, with
inactive atoms; and of the table <-> spline-coefficient maps of msmpscu_amd.adjoint against
their own definitions.  CPU only."""
import numpy as np
import pytest

import msmpscu_amd.configs as cfg


def _setup(orc, pot_type, seed=3, inactive=(5, 40)):
    box = cfg.make_bcc_box((4, 4, 4), cfg.W_A, 0.12, 300.0, seed)
    st = box.STATU.copy()
    st[list(inactive)] = 0
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl, pot_type=pot_type)
    t = orc.from_mdlib(ft)
    it = box.ITYP - 1
    kv, ind = orc.nlist("full", box.XP, it, st, box.BOXLOW, box.ZL, ctrl.IFPD, ctrl.NB_RM,
                        ctrl.NB_MXNBS)
    ru2 = float(np.max(ctrl.RU)) ** 2
    return box, st, ctrl, t, it, kv, ind, ru2


def _loss(orc, box, st, ctrl, t, it, kv, ind, ru2, G, g):
    fp, _, ep = orc.force_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t)
    return float(np.sum(G * fp) + g * np.sum(ep))


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
def test_oracle_adjoint_matches_central_differences(orc, pot_type):
    """L = sum G.F + g E with fixed random G, g (so dL/dknot is exactly the adjoint's output):
    every table's gradient agrees with (L(v+h) - L(v-h)) / 2h at knots that carry weight and at
    a few that carry none, to 1e-6 relative."""
    box, st, ctrl, t, it, kv, ind, ru2 = _setup(orc, pot_type)
    rng = np.random.default_rng(11)
    G = rng.normal(size=(box.NPRT, 3))
    g = 0.37
    grads = orc.adjoint_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t, G, g)
    names = ["potr", "fpotr", "potb", "fpotb"] + (["fembd", "dfembd"] if t.pot_type else [])
    for name in names:
        gr = grads[name]
        arr = getattr(t, name)[0]
        hot = np.flatnonzero(np.abs(gr) > 0)
        assert hot.size > 0, name
        pick = list(rng.choice(hot, size=min(12, hot.size), replace=False))
        cold = np.flatnonzero(gr == 0)
        pick += list(rng.choice(cold[1:-1], size=3, replace=False))
        scale = np.max(np.abs(gr))
        for k in pick:
            v = arr[k]
            # the loss carries g E (|E| ~ 1e3 eV): h = 1e-4 keeps the rounding of L below
            # 1e-8 of the difference while the truncation error (h^2 terms) stays near 1e-9
            h = 1e-4 * max(1.0, abs(v))
            arr[k] = v + h
            lp = _loss(orc, box, st, ctrl, t, it, kv, ind, ru2, G, g)
            arr[k] = v - h
            lm = _loss(orc, box, st, ctrl, t, it, kv, ind, ru2, G, g)
            arr[k] = v
            fd = (lp - lm) / (2 * h)
            assert abs(fd - gr[k]) <= 1e-6 * max(abs(gr[k]), 1e-3 * scale), (name, k, fd, gr[k])


def test_oracle_adjoint_is_linear_in_the_loss_weights(orc):
    """The gradient is linear in (G, g): grad(aG1 + bG2, a g1 + b g2) = a grad1 + b grad2."""
    box, st, ctrl, t, it, kv, ind, ru2 = _setup(orc, "EAM_TYPE")
    rng = np.random.default_rng(5)
    G1, G2 = rng.normal(size=(box.NPRT, 3)), rng.normal(size=(box.NPRT, 3))
    a1 = orc.adjoint_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t, G1, 0.5)
    a2 = orc.adjoint_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t, G2, -1.5)
    a3 = orc.adjoint_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t, 2 * G1 - 3 * G2,
                          2 * 0.5 - 3 * (-1.5))
    for k in a1:
        ref = 2 * a1[k] - 3 * a2[k]
        assert np.allclose(a3[k], ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))


def test_inactive_atoms_get_no_gradient_through_their_forces(orc):
    """dL/dF of an inactive atom does not reach the tables (its force is identically zero)."""
    box, st, ctrl, t, it, kv, ind, ru2 = _setup(orc, "FS_TYPE")
    rng = np.random.default_rng(2)
    G = rng.normal(size=(box.NPRT, 3))
    a1 = orc.adjoint_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t, G, 0.0)
    G2 = G.copy()
    G2[[5, 40]] += 100.0
    a2 = orc.adjoint_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t, G2, 0.0)
    for k in a1:
        assert np.array_equal(a1[k], a2[k])


def test_spline_maps_are_transposes():
    """coefficients -> tables is linear; the gradient map is its exact transpose:
    <A c, y> = <c, A^T y> for random c, y (all six tables)."""
    from msmpscu_amd import adjoint as adj
    m = adj.SplineModel(ntab=300, csi=300 / np.sqrt(4.4002), nembd=200, rhod=0.37)
    rng = np.random.default_rng(9)
    c = rng.normal(size=(3, 300))
    c[2, :200] = rng.normal(size=200)
    tabs = m.tables(c)
    ys = {k: rng.normal(size=v.shape) for k, v in tabs.items()}
    lhs = sum(float(np.sum(tabs[k] * ys[k])) for k in tabs)
    rhs = float(np.sum(c * m.coeff_grad(ys)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_spline_fit_reproduces_value_tables():
    """The base coefficients fitted to the W EAM tables reproduce the value tables (POTR, POTB,
    FEMBD) at every knot to 1e-9 relative, and the derivative tables the spline implies agree
    with the analytic ones to the spline's own accuracy (1e-3 relative of the table's range)."""
    from msmpscu_amd import adjoint as adj
    box = cfg.make_bcc_box((4, 4, 4), cfg.W_A, 0.0, 0.0, 1)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl, pot_type="EAM_TYPE")
    m = adj.SplineModel.from_table(ft)
    c = m.fit(ft)
    tabs = m.tables(c)
    for name, ref in (("POTR", ft.POTR[0]), ("POTB", ft.POTB[0]), ("FEMBD", ft.FEMBD[0])):
        assert np.max(np.abs(tabs[name] - ref)) <= 1e-9 * np.max(np.abs(ref)), name
    for name, ref in (("FPOTR", ft.FPOTR[0]), ("FPOTB", ft.FPOTB[0]), ("DFEMBD", ft.DFEMBD[0])):
        sl = slice(20, -20)
        err = np.max(np.abs(tabs[name][sl] - ref[sl]))
        assert err <= 1e-3 * np.max(np.abs(ref[sl])), (name, err)
