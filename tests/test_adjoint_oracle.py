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
        gr = grads[name][0]
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


def _alloy(orc, case):
    """Two atom types: 'asym' -- W lattice with alternating types and four FS kinds with KPAIR
    [[1, 3], [4, 2]] (density and force channels differ by direction); 'cuni' -- a Cu-Ni slab
    (free in z) with three TB-SMA kinds and two EAM embedding kinds."""
    import msmpscu_amd.mdlib as md
    if case == "asym":
        box = cfg.make_bcc_box((4, 4, 4), cfg.W_A, 0.12, 300.0, 8)
        box.NGROUP = 2
        box.ITYP = (np.arange(box.NPRT) % 2 + 1).astype(np.int32)
        box.CM = np.array([box.CM[0], box.CM[0] * 0.5])
        box.PTYPE = np.array([[1, 3], [4, 2]], dtype=np.int32)
        ctrl = md.SimMDCtrl(RU=np.full((2, 2), 4.4002), NB_RM=np.full((2, 2), 4.7002),
                            NB_MXNBS=64, NUMFTABR=2000)
        ft = md.MDForceTable()
        md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
        md.Register_ForceTableProc(2, ft, "msm_pot_ackland_ww_nn", "msm_pot_ackland_ww_ne")
        md.Register_ForceTableProc(3, ft, "msm_pot_wfs_nn", "msm_pot_ackland_ww_ne")
        md.Register_ForceTableProc(4, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
        md.Create_Interaction_ForceTable(box, ctrl, ft)
    else:
        box = cfg.make_cuni_slab((4, 4, 4), 3.57, 300.0, 5)
        ctrl = cfg.cuni_ctrl(ntab=2000)
        ft = cfg.cuni_forcetable(box, ctrl, pot_type="EAM_TYPE")
    st = box.STATU.copy()
    st[[3, 50]] = 0
    t = orc.from_mdlib(ft)
    it = box.ITYP - 1
    kv, ind = orc.nlist("full", box.XP, it, st, box.BOXLOW, box.ZL, ctrl.IFPD, ctrl.NB_RM,
                        ctrl.NB_MXNBS)
    ru2 = float(np.max(ctrl.RU)) ** 2
    return box, st, ctrl, t, it, kv, ind, ru2


@pytest.mark.parametrize("case", ["asym", "cuni"])
def test_oracle_adjoint_of_several_kinds_matches_central_differences(orc, case):
    """Every table kind: knots drawn from all kinds of each table (the asymmetric KPAIR of 'asym'
    sends F'_i and F'_j through different FPOTB kinds; 'cuni' has two embedding kinds) agree with
    central differences of the forward model to 1e-6 relative (plus the rounding of the loss over
    the step), and every kind in use carries gradient."""
    box, st, ctrl, t, it, kv, ind, ru2 = _alloy(orc, case)
    rng = np.random.default_rng(17)
    G = rng.normal(size=(box.NPRT, 3))
    g = -0.23
    grads = orc.adjoint_full(box.XP, it, st, kv, ind, box.ZL, ctrl.IFPD, ru2, t, G, g)
    names = ["potr", "fpotr", "potb", "fpotb"] + (["fembd", "dfembd"] if t.pot_type else [])
    used = set(int(k) for k in np.unique(t.kpair))
    for name in names:
        gr = grads[name]
        arr = getattr(t, name)
        assert gr.shape == arr.shape
        kinds = used if name not in ("fembd", "dfembd") else set(int(k) for k in t.kembd[:2])
        for kd in kinds:
            assert np.any(gr[kd] != 0.0), (name, kd)
        hot = np.argwhere(np.abs(gr) > 0)
        pick = [tuple(hot[i]) for i in rng.choice(len(hot), size=min(16, len(hot)), replace=False)]
        scale = np.max(np.abs(gr))
        l0 = abs(_loss(orc, box, st, ctrl, t, it, kv, ind, ru2, G, g))
        for kd, k in pick:
            v = arr[kd, k]
            # small steps: the piecewise-linear embedding lookup has a kink at every knot of
            # rho / RHOD, which a larger step in POTB can move an atom's density across; the
            # bar then allows the rounding of L over the step (8 ulp of |L| / h)
            h = 1e-6 * max(1.0, abs(v))
            arr[kd, k] = v + h
            lp = _loss(orc, box, st, ctrl, t, it, kv, ind, ru2, G, g)
            arr[kd, k] = v - h
            lm = _loss(orc, box, st, ctrl, t, it, kv, ind, ru2, G, g)
            arr[kd, k] = v
            fd = (lp - lm) / (2 * h)
            bar = 1e-6 * max(abs(gr[kd, k]), 1e-3 * scale) + 8 * 2.2e-16 * l0 / h
            assert abs(fd - gr[kd, k]) <= bar, (name, kd, k, fd, gr[kd, k])


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


def test_spline_model_of_fs_tables_has_no_embedding_row():
    """For FS tables (F = -sqrt(rho) built in) the model has no embedding knots: the third
    coefficient row is zero after the fit, its gradient row is zero, and applying coefficients
    leaves the table's (absent) embedding arrays alone while the value tables are reproduced."""
    from msmpscu_amd import adjoint as adj
    box = cfg.make_bcc_box((4, 4, 4), cfg.W_A, 0.0, 0.0, 1)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl, pot_type="FS_TYPE")
    m = adj.SplineModel.from_table(ft)
    assert m.nembd == 0 and m.shape == (3, int(ft.NTAB))
    c = m.fit(ft)
    assert np.all(c[2] == 0.0)
    potr = np.array(ft.POTR[0], dtype=np.float64)
    m.apply(c, ft)
    assert ft.FEMBD is None and ft.DFEMBD is None
    assert np.max(np.abs(ft.POTR[0] - potr)) <= 1e-9 * np.max(np.abs(potr))
    rng = np.random.default_rng(3)
    g = {k: rng.normal(size=v.shape) for k, v in m.tables(c).items()}
    assert np.all(m.coeff_grad(g)[2] == 0.0)
