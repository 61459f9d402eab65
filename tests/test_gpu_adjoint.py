"""GPU parity of the adjoint (msm_adjoint through the C-ABI): knot gradients against the oracle's
adjoint restatement, and coefficient gradients of the force/energy matching loss against central
finite differences of the device forward path (64 strained 1024-atom W boxes, EAM tables)."""
import sys

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md
from msmpscu_amd import adjoint as adj

pytestmark = pytest.mark.gpu

NAMES = ("POTR", "FPOTR", "POTB", "FPOTB", "FEMBD", "DFEMBD")


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _device(boxes, ctrl, ft):
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(ft.PotType, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc.Device


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
@pytest.mark.parametrize("nbox", [1, 3])
def test_adjoint_matches_oracle(orc, pot_type, nbox):
    """Random dL/dF and dL/dE (per box), inactive atoms, one or several uneven boxes: every
    table's knot gradient equals the oracle's to 1e-10 of its largest entry."""
    ctrl = cfg.w_ctrl(mx=64)
    boxes = [cfg.make_bcc_box(c, cfg.W_A, 0.1, 300.0, 40 + k, n_vac=k, n_dumbbell=k % 2)
             for k, c in enumerate([(6, 6, 6), (5, 7, 6), (6, 5, 5)][:nbox])]
    boxes[0].STATU[[7, 100]] = 0
    ft = cfg.w_forcetable(boxes[0], ctrl, pot_type=pot_type)
    say(f"adjoint vs oracle {pot_type} boxes={nbox}")
    dev = _device(boxes, ctrl, ft)
    dev.build_neighbors()
    rng = np.random.default_rng(8)
    ntot = sum(b.NPRT for b in boxes)
    G = rng.normal(size=(ntot, 3))
    g = rng.normal(size=nbox)
    eam = pot_type == "EAM_TYPE"
    gd = adj.table_grads(dev, G, g, ft.NTAB, ft.NEMBD if eam else 1, eam)
    dev.close()
    t = orc.from_mdlib(ft)
    grads = None
    off = 0
    for k, b in enumerate(boxes):
        it = b.ITYP - 1
        kv, ind = orc.nlist("full", b.XP, it, b.STATU, b.BOXLOW, b.ZL, ctrl.IFPD, ctrl.NB_RM,
                            ctrl.NB_MXNBS)
        grads = orc.adjoint_full(b.XP, it, b.STATU, kv, ind, b.ZL, ctrl.IFPD,
                                 float(np.max(ctrl.RU)) ** 2, t, G[off:off + b.NPRT], g[k],
                                 grads)
        off += b.NPRT
    for name in NAMES[: 6 if eam else 4]:
        ref = grads[name.lower()][0, 1:-1]
        got = gd[name]
        scale = np.max(np.abs(ref))
        assert scale > 0, name
        assert np.max(np.abs(got - ref)) <= 1e-10 * scale, (name, np.max(np.abs(got - ref)),
                                                            scale)


def _alloy(case):
    """Two atom types: 'asym' -- W lattice with alternating types, four FS kinds, KPAIR
    [[1, 3], [4, 2]]; 'cuni' -- a Cu-Ni slab, three TB-SMA kinds and two EAM embedding kinds."""
    if case == "asym":
        box = cfg.make_bcc_box((7, 7, 7), cfg.W_A, 0.1, 300.0, 8)
        box.NGROUP = 2
        box.ITYP = (np.arange(box.NPRT) % 2 + 1).astype(np.int32)
        box.CM = np.array([box.CM[0], box.CM[0] * 0.5])
        box.PTYPE = np.array([[1, 3], [4, 2]], dtype=np.int32)
        ctrl = md.SimMDCtrl(RU=np.full((2, 2), 4.4002), NB_RM=np.full((2, 2), 4.7002),
                            NB_MXNBS=64, NUMFTABR=3000)
        ft = md.MDForceTable()
        md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
        md.Register_ForceTableProc(2, ft, "msm_pot_ackland_ww_nn", "msm_pot_ackland_ww_ne")
        md.Register_ForceTableProc(3, ft, "msm_pot_wfs_nn", "msm_pot_ackland_ww_ne")
        md.Register_ForceTableProc(4, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
        md.Create_Interaction_ForceTable(box, ctrl, ft)
    else:
        box = cfg.make_cuni_slab((5, 5, 5), 3.57, 600.0, 5)
        ctrl = cfg.cuni_ctrl(ntab=3000)
        ft = cfg.cuni_forcetable(box, ctrl, pot_type="EAM_TYPE")
    box.STATU[[3, 50]] = 0
    return box, ctrl, ft


@pytest.mark.parametrize("case", ["asym", "cuni"])
def test_adjoint_of_several_kinds_matches_oracle(orc, case):
    """Two atom types (asymmetric KPAIR: F'_i and F'_j reach different FPOTB kinds; two EAM
    embedding kinds): every kind of every table equals the oracle's adjoint to 1e-10 of that
    table's largest entry, and every kind in use carries gradient."""
    box, ctrl, ft = _alloy(case)
    say(f"adjoint of several kinds: {case}")
    dev = _device([box], ctrl, ft)
    dev.build_neighbors()
    rng = np.random.default_rng(12)
    G = rng.normal(size=(box.NPRT, 3))
    g = np.array([0.31])
    eam = ft.PotType.strip() == "EAM_TYPE"
    t = orc.from_mdlib(ft)
    gd = adj.table_grads(dev, G, g, ft.NTAB, ft.NEMBD if eam else 1, eam, nkind=t.nkind,
                         nkind1=max(t.nkind1, 1))
    dev.close()
    it = box.ITYP - 1
    kv, ind = orc.nlist("full", box.XP, it, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                        ctrl.NB_RM, ctrl.NB_MXNBS)
    grads = orc.adjoint_full(box.XP, it, box.STATU, kv, ind, box.ZL, ctrl.IFPD,
                             float(np.max(ctrl.RU)) ** 2, t, G, g[0])
    used = set(int(k) for k in np.unique(t.kpair))
    for name in NAMES[: 6 if eam else 4]:
        ref = grads[name.lower()][:, 1:-1]
        got = gd[name]
        assert got.shape == ref.shape, name
        kinds = used if name not in ("FEMBD", "DFEMBD") else set(int(k) for k in t.kembd[:2])
        for kd in kinds:
            assert np.any(ref[kd] != 0.0), (name, kd)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= 1e-10 * scale, (name, np.max(np.abs(got - ref)),
                                                            scale)


def test_coefficient_gradient_central_differences():
    """64 boxes x 1024 atoms (diagonal strains up to 2%), targets from coefficients with 1%
    seeded noise, L = mean|dF|^2 + 0.1 mean(dE)^2 (dE per atom of each box): dL/dc for 200
    random coefficients among the 3 x 5000 (drawn from those the data reaches) agrees with
    central differences at h = 1e-6 to 1e-6 relative, up to the differences' own rounding floor
    8 eps L / h (L is evaluated in FP64, so no step can resolve it more finely); coefficients the
    data does not reach have zero gradient both ways."""
    boxes, ctrl, ft = adj.make_d_config()
    fm = adj.ForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    fm.set_targets(adj.perturbed(c0, 0.01, seed=5))
    L, grad = fm.loss_and_grad(c0)
    assert L > 0
    h = 1e-6
    floor = 8.0 * np.finfo(np.float64).eps * L / h
    say(f"D config n={fm.n} boxes={fm.nbox} L={L:.6e} FD rounding floor {floor:.2e}")
    rng = np.random.default_rng(21)
    hot = np.argwhere(np.abs(grad) > 0)
    cold = np.argwhere(grad == 0)
    pick = hot[rng.choice(len(hot), size=200, replace=False)]
    worst_rel, worst_ok = 0.0, 0.0
    for (a, k) in pick:
        c = c0.copy()
        c[a, k] += h
        lp = fm.loss(c)
        c[a, k] -= 2 * h
        lm = fm.loss(c)
        fd = (lp - lm) / (2 * h)
        err = abs(fd - grad[a, k])
        worst_ok = max(worst_ok, err / (1e-6 * abs(grad[a, k]) + floor))
        if abs(grad[a, k]) >= 1e7 * floor:
            worst_rel = max(worst_rel, err / abs(grad[a, k]))
    say(f"worst error / (1e-6 |g| + floor) = {worst_ok:.3e}; worst relative error where "
        f"|g| >= 1e7 floor: {worst_rel:.3e}")
    for (a, k) in cold[rng.choice(len(cold), size=5, replace=False)]:
        c = c0.copy()
        c[a, k] += h
        lp = fm.loss(c)
        c[a, k] -= 2 * h
        assert (lp - fm.loss(c)) == 0.0
    fm.close()
    assert worst_ok <= 1.0
    assert worst_rel <= 1e-6


def test_torch_autograd_function_matches_loss_and_grad():
    """The torch.autograd.Function over the C-ABI gives the same loss and coefficient gradient as
    ForceMatch.loss_and_grad."""
    torch = pytest.importorskip("torch")
    boxes, ctrl, ft = adj.make_d_config(nbox=4)
    fm = adj.ForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    fm.set_targets(adj.perturbed(c0, 0.01, seed=6))
    L, grad = fm.loss_and_grad(c0)
    Fn = adj.make_torch_function()
    c = torch.tensor(c0, dtype=torch.float64, requires_grad=True)
    f, e = Fn.apply(c, fm)
    ft_ = torch.from_numpy(fm.f_target)
    et_ = torch.from_numpy(fm.e_target)
    na = torch.from_numpy(fm.natoms)
    lt = ((f - ft_) ** 2).sum() / fm.n + 0.1 * (((e - et_) / na) ** 2).sum() / fm.nbox
    lt.backward()
    fm.close()
    assert abs(lt.item() - L) <= 1e-12 * abs(L)
    assert np.max(np.abs(c.grad.numpy() - grad)) <= 1e-12 * np.max(np.abs(grad))
