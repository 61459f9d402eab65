"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle on the same seeded
inputs.  Bars: neighbour lists bit-exact as sorted sets; per-atom forces within 1e-10 eV/A
(max-abs); total energy within 1e-12 relative; integrator within rounding of the reference
formulas."""
import sys

import copy

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md
from msmpscu_amd._lib import MSMError

pytestmark = pytest.mark.gpu

F_TOL = 1e-10      # eV/A, max-abs per atom and component
E_REL = 1e-12      # relative, total energy


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def device_for(boxes, ctrl, ft):
    boxes = boxes if isinstance(boxes, list) else [boxes]
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(ft.PotType, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc.Device


def oracle_half(orc, box, ctrl, ft, xp=None, with_vabs=False):
    xp = box.XP if xp is None else xp
    t = orc.from_mdlib(ft)
    it = box.ITYP - 1
    kv, ind = orc.nlist("half", xp, it, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD, ctrl.NB_RM,
                        ctrl.NB_MXNBS)
    return orc.force_fs_half(xp, it, kv, ind, box.ZL, ctrl.IFPD, ctrl.RU, t, with_vabs=with_vabs)


def oracle_full(orc, box, ctrl, ft, xp=None, statu=None):
    xp = box.XP if xp is None else xp
    st = box.STATU if statu is None else statu
    t = orc.from_mdlib(ft)
    it = box.ITYP - 1
    kv, ind = orc.nlist("full", xp, it, st, box.BOXLOW, box.ZL, ctrl.IFPD, ctrl.NB_RM,
                        ctrl.NB_MXNBS)
    return orc.force_full(xp, it, st, kv, ind, box.ZL, ctrl.IFPD, float(np.max(ctrl.RU)) ** 2, t)


def pairset(kv, ind, one_based):
    s = set()
    off = 1 if one_based else 0
    for i in range(len(kv)):
        for j in ind[i, : kv[i]]:
            s.add((i, int(j) - off))
    return s


@pytest.fixture(scope="module")
def _p0_master():
    box = cfg.make_bcc_box((16, 16, 16), cfg.W_A, 0.03, 300.0, 2024)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl)
    return box, ctrl, ft


@pytest.fixture
def p0(_p0_master):
    """The 8192-atom W box, its control block and FS table; every test gets its own copy."""
    return copy.deepcopy(_p0_master)


@pytest.mark.parametrize("case", ["p0", "disordered"])
def test_neighbour_list_bit_exact(orc, p0, case):
    if case == "p0":
        box, ctrl, ft = p0
    else:
        box = cfg.make_bcc_box((10, 10, 10), cfg.W_A, 0.15, 900.0, 77, n_vac=5, n_dumbbell=3)
        ctrl = cfg.w_ctrl(mx=64)
        ft = cfg.w_forcetable(box, ctrl)
    say(f"nlist {case} n={box.NPRT}")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    kv, ind = dev.get_neighbors()
    okv, oind = orc.nlist("full", box.XP, box.ITYP - 1, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                          ctrl.NB_RM, ctrl.NB_MXNBS)
    assert np.array_equal(kv, okv)
    assert pairset(kv, ind, True) == pairset(okv, oind, False)
    tot, mk = dev.list_stats()
    assert tot == int(okv.sum()) and mk == int(okv.max())
    dev.close()


def _list_all_kernels(box, ctrl, ft):
    """Lists from every list kernel (global scan, small bricks, wide bricks, automatic choice)
    and, for the automatic choice, with cells numbered in 4^3 bricks."""
    import msmpscu_amd._lib as L
    out = []
    for mode, cbs in ((1, 1), (2, 1), (3, 1), (0, 1), (0, 4)):
        dev = device_for(box, ctrl, ft)
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_CELL_BRICK, cbs), "msm_set_option")
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_NLIST_KERNEL, mode), "msm_set_option")
        dev.build_neighbors()
        out.append(dev.get_neighbors())
        dev.close()
    return out


@pytest.mark.parametrize("case", ["small_ncell3", "dense_blob", "slab", "w16"])
def test_brick_list_equals_global_scan(orc, case):
    """The LDS-staged brick builds (small bricks with one lane per atom, wide bricks with
    several lanes per atom and ballot placement, including bricks that overflow LDS and take the
    global path) give the identical list, in identical order, to the per-atom global-memory
    scan; so does the brick-ordered cell numbering."""
    if case == "small_ncell3":
        box = cfg.make_bcc_box((4, 4, 4), cfg.W_A, 0.1, 300.0, 3)
        ctrl = cfg.w_ctrl(mx=64)
    elif case == "dense_blob":
        box = cfg.make_bcc_box((12, 12, 12), cfg.W_A, 0.05, 0.0, 4)
        rng = np.random.default_rng(4)
        blob = rng.normal(0.0, 1.2, (600, 3))
        x = np.concatenate([box.XP, blob])
        n = x.shape[0]
        box = md.SimMDBox(NPRT=n, NGROUP=1, ZL=box.ZL, BOXLOW=box.BOXLOW,
                          XP=np.asfortranarray(x), XP1=np.zeros((n, 3)),
                          ITYP=np.ones(n, np.int32), STATU=np.ones(n, np.int32), CM=box.CM,
                          PTYPE=box.PTYPE)
        ctrl = cfg.w_ctrl(mx=800)
    elif case == "slab":
        box = cfg.make_cuni_slab((6, 6, 6), 3.57, 300.0, 9)
        ctrl = cfg.cuni_ctrl()
    else:
        box = cfg.make_bcc_box((16, 16, 16), cfg.W_A, 0.05, 300.0, 12)
        ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl) if case != "slab" else cfg.cuni_forcetable(box, ctrl)
    say(f"brick vs global list: {case} n={box.NPRT}")
    lists = _list_all_kernels(box, ctrl, ft)
    kv0, ind0 = lists[0]
    for kv1, ind1 in lists[1:]:
        assert np.array_equal(kv0, kv1)
        for i in range(len(kv0)):
            assert np.array_equal(ind0[i, : kv0[i]], ind1[i, : kv1[i]])
    okv, oind = orc.nlist("full", box.XP, box.ITYP - 1, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                          ctrl.NB_RM, ctrl.NB_MXNBS)
    assert pairset(kv0, ind0, True) == pairset(okv, oind, False)


@pytest.mark.parametrize("lanes,nlist", [(1, 2), (2, 2), (2, 3)])
def test_brick_table_overflowing_bricks(orc, lanes, nlist):
    """A dense blob of 600 extra atoms (Gaussian, 1.2 A wide) in a W box: the bricks around it
    hold more halo atoms than the LDS-brick passes stage, so they take the global-list groups
    (with the LDS knot window and its below-window fallback) while the other bricks run from
    LDS; F' and forces agree with the oracle to the bar scaled by the largest force.  Small
    list bricks (nlist 2) and wide ones (nlist 3, whose staging capacity the passes follow)."""
    import msmpscu_amd._lib as L
    box = cfg.make_bcc_box((12, 12, 12), cfg.W_A, 0.05, 0.0, 4)
    rng = np.random.default_rng(4)
    x = np.concatenate([box.XP, rng.normal(0.0, 1.2, (600, 3))])
    n = x.shape[0]
    box = md.SimMDBox(NPRT=n, NGROUP=1, ZL=box.ZL, BOXLOW=box.BOXLOW, XP=np.asfortranarray(x),
                      XP1=np.zeros((n, 3)), ITYP=np.ones(n, np.int32),
                      STATU=np.ones(n, np.int32), CM=box.CM, PTYPE=box.PTYPE)
    ctrl = cfg.w_ctrl(mx=800)
    ft = cfg.w_forcetable(box, ctrl)
    say(f"brick table, overflowing bricks, lanes {lanes} list kernel {nlist}")
    dev = device_for(box, ctrl, ft)
    for opt, v in ((L.MSM_OPT_FORCE_PATH, 0), (L.MSM_OPT_TABLE_LDS, 1),
                   (L.MSM_OPT_BRICK_LANES, lanes), (L.MSM_OPT_NLIST_KERNEL, nlist)):
        L.check(L.lib().msm_set_option(dev.h, opt, v), "msm_set_option")
    assert dev.pass_kernel() == "brick_table"
    dev.build_neighbors()
    dev.calc_force()
    _, _, fp, _, _ = dev.get_atoms()
    den = dev.get_den()
    dev.close()
    fo, do, _ = oracle_full(orc, box, ctrl, ft)
    assert np.max(np.abs(fp - fo)) <= F_TOL * max(1.0, np.max(np.abs(fo)) / 10.0)
    assert np.max(np.abs(den - do)) <= 1e-13 * max(1.0, np.max(np.abs(do)))


def test_brick_passes_without_slot_list(orc, p0):
    """A list built for the global-list passes carries no halo slots: switching to the brick
    passes (global or LDS table) without a rebuild sends every brick to the global list, with the
    same results; the next rebuild writes the slots and the staged passes agree again."""
    import msmpscu_amd._lib as L
    box, ctrl, ft = p0
    say("brick passes on a list without slots")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    fh, *_ = oracle_half(orc, box, ctrl, ft)
    for lds in (1, 0):
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_FORCE_PATH, 0), "msm_set_option")
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_TABLE_LDS, lds), "msm_set_option")
        dev.calc_force()
        _, _, fp, _, _ = dev.get_atoms()
        assert np.max(np.abs(fp - fh)) <= F_TOL
    dev.build_neighbors()
    dev.calc_force()
    _, _, fp, _, _ = dev.get_atoms()
    assert np.max(np.abs(fp - fh)) <= F_TOL
    dev.close()


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
def test_force_energy_vs_oracle(orc, p0, pot_type):
    box, ctrl, _ = p0
    ft = cfg.w_forcetable(box, ctrl, pot_type=pot_type)
    say(f"force {pot_type} n={box.NPRT}")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.calc_epot()
    _, _, fp, _, _ = dev.get_atoms()
    den = dev.get_den()
    ep = dev.get_epot()
    ebox = dev.get_box_epot()
    fo, do, eo = oracle_full(orc, box, ctrl, ft)
    assert np.max(np.abs(fp - fo)) <= F_TOL
    assert np.max(np.abs(den - do)) <= 1e-13
    assert abs(ep.sum() - eo.sum()) <= E_REL * abs(eo.sum())
    assert abs(ebox[0] - orc.tree_sum(eo)) <= E_REL * abs(eo.sum())
    if pot_type == "FS_TYPE":
        fh, eh, _, ddh, _ = oracle_half(orc, box, ctrl, ft)
        assert np.max(np.abs(fp - fh)) <= F_TOL
        assert np.max(np.abs(den - ddh)) <= 1e-13
        assert abs(ebox[0] - orc.tree_sum(eh)) <= E_REL * abs(eh.sum())
    dev.close()


def test_two_species_slab(orc):
    box = cfg.make_cuni_slab((6, 6, 6), 3.57, 600.0, 5)
    rng = np.random.default_rng(1)
    box.XP = np.asfortranarray(box.XP + rng.normal(0, 0.05, box.XP.shape))
    box.XP[:, :2] = cfg.wrap_into(box.XP, box.BOXLOW, box.ZL, (1, 1, 0))[:, :2]
    ctrl = cfg.cuni_ctrl()
    ft = cfg.cuni_forcetable(box, ctrl)
    say(f"cu-ni slab n={box.NPRT}")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    kv, ind = dev.get_neighbors()
    okv, oind = orc.nlist("full", box.XP, box.ITYP - 1, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                          ctrl.NB_RM, ctrl.NB_MXNBS)
    assert pairset(kv, ind, True) == pairset(okv, oind, False)
    dev.calc_force()
    dev.calc_epot()
    _, _, fp, _, _ = dev.get_atoms()
    fh, eh, _, ddh, _ = oracle_half(orc, box, ctrl, ft)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    assert abs(dev.get_epot().sum() - eh.sum()) <= E_REL * abs(eh.sum())
    dev.close()


def test_asymmetric_density_channels(orc):
    """KPAIR(i,j) != KPAIR(j,i) with different density tables (the pair term stays symmetric, as
    any energy requires): checks the KTAB/KTAB1 pairing."""
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.08, 300.0, 8)
    box.NGROUP = 2
    box.ITYP = (np.arange(box.NPRT) % 2 + 1).astype(np.int32)
    box.CM = np.array([box.CM[0], box.CM[0] * 0.5])
    box.PTYPE = np.array([[1, 3], [4, 2]], dtype=np.int32)
    ctrl = md.SimMDCtrl(RU=np.full((2, 2), 4.4002), NB_RM=np.full((2, 2), 4.7002), NB_MXNBS=64)
    ft = md.MDForceTable()
    md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
    md.Register_ForceTableProc(2, ft, "msm_pot_ackland_ww_nn", "msm_pot_ackland_ww_ne")
    md.Register_ForceTableProc(3, ft, "msm_pot_wfs_nn", "msm_pot_ackland_ww_ne")
    md.Register_ForceTableProc(4, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
    md.Create_Interaction_ForceTable(box, ctrl, ft)
    say("asymmetric kpair")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    _, _, fp, _, _ = dev.get_atoms()
    fh, eh, _, ddh, _ = oracle_half(orc, box, ctrl, ft)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    assert np.max(np.abs(dev.get_den() - ddh)) <= 1e-13
    dev.close()


def test_inactive_atoms(orc, p0):
    box, ctrl, ft = p0
    st = box.STATU.copy()
    st[[3, 100, 4000]] = 0
    say("inactive atoms")
    dev = device_for(box, ctrl, ft)
    dev.set_atoms(box.XP, box.XP1, box.ITYP, st)
    dev.build_neighbors()
    dev.calc_force()
    dev.calc_epot()
    _, _, fp, _, _ = dev.get_atoms()
    fo, do, eo = oracle_full(orc, box, ctrl, ft, statu=st)
    assert np.max(np.abs(fp - fo)) <= F_TOL
    assert np.all(fp[[3, 100, 4000]] == 0)
    assert abs(dev.get_epot().sum() - eo.sum()) <= E_REL * abs(eo.sum())
    dev.close()


def test_integrator_kernels(orc, p0):
    box, ctrl, ft = p0
    say("integrator")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    x0, v0, f0, st0, d0 = dev.get_atoms()
    dev.predictor(ctrl.H)
    x1, v1, _, st1, d1 = dev.get_atoms()
    ox, ov, od, ost = orc.predictor(x0, v0, f0, d0, st0, box.ITYP - 1, box.CM, ctrl.H,
                                    box.BOXLOW, box.ZL, ctrl.IFPD)
    assert np.max(np.abs(x1 - ox)) <= 1e-13
    assert np.max(np.abs(v1 - ov)) <= 1e-16
    assert np.max(np.abs(d1 - od)) <= 1e-13
    assert np.array_equal(st1 & 0xFFFF, ost & 0xFFFF)
    dev.correction(ctrl.H)
    _, v2, _, st2, _ = dev.get_atoms()
    ov2, _ = orc.correction(v1, f0, st1, box.ITYP - 1, box.CM, ctrl.H)
    assert np.max(np.abs(v2 - ov2)) <= 1e-16
    dev.close()


def test_md_run_matches_oracle(orc, p0):
    """100 NVE steps, rebuild every 10 (the 8192-atom W benchmark case)."""
    box, ctrl, ft = p0
    say("md run 100 steps")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.run(100, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.calc_epot()
    x, v, fp, st, dis = dev.get_atoms()
    ep = dev.get_epot()
    t = orc.from_mdlib(ft)
    ox, ov, of, od, ost, oep, _, _ = orc.md_run(box.XP, box.XP1, box.STATU, box.ITYP - 1, box.CM,
                                                ctrl.H, 100, ctrl.NB_UPTAB, 0, box.BOXLOW, box.ZL,
                                                ctrl.IFPD, ctrl.NB_RM, ctrl.RU, ctrl.NB_MXNBS, t)
    dx = x - ox
    dx -= np.round(dx / box.ZL) * box.ZL
    assert np.max(np.abs(dx)) <= 1e-9
    assert np.max(np.abs(v - ov)) <= 1e-11
    assert np.max(np.abs(fp - of)) <= 1e-8
    assert np.max(np.abs(dis - od)) <= 1e-9
    # per-snapshot parity: the oracle evaluated on the GPU's own final positions
    fh, eh, *_ = oracle_half(orc, box, ctrl, ft, xp=x)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    assert abs(ep.sum() - eh.sum()) <= E_REL * abs(eh.sum())
    # NVE: energy drift small
    ek0 = 0.5 * box.CM[0] * np.sum(box.XP1 ** 2)
    dev2 = device_for(box, ctrl, ft)
    dev2.build_neighbors()
    dev2.calc_force()
    dev2.calc_epot()
    e0 = dev2.get_epot().sum() + ek0
    e1 = ep.sum() + 0.5 * box.CM[0] * np.sum(v ** 2)
    assert abs(e1 - e0) <= 1e-4 * abs(e0)
    dev.close()
    dev2.close()


def test_run_is_deterministic(p0):
    box, ctrl, ft = p0
    say("determinism")
    outs = []
    for _ in range(2):
        dev = device_for(box, ctrl, ft)
        dev.build_neighbors()
        dev.calc_force()
        dev.run(30, ctrl.NB_UPTAB, 0, ctrl.H)
        outs.append(dev.get_atoms())
        dev.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_uneven_multibox_equals_single_boxes(p0):
    _, ctrl, _ = p0
    boxes = [cfg.make_bcc_box(c, cfg.W_A, 0.05, 500.0, 10 + k, n_vac=k, n_dumbbell=k % 2)
             for k, c in enumerate([(5, 6, 7), (9, 9, 9), (6, 4, 5)])]
    ft = cfg.w_forcetable(boxes[0], ctrl)
    say("multibox")
    dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.run(25, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.calc_epot()
    xb, vb, fb, _, _ = dev.get_atoms()
    eb = dev.get_box_epot()
    dev.close()
    off = 0
    for k, b in enumerate(boxes):
        d1 = device_for(b, ctrl, ft)
        d1.build_neighbors()
        d1.calc_force()
        d1.run(25, ctrl.NB_UPTAB, 0, ctrl.H)
        d1.calc_epot()
        x1, v1, f1, _, _ = d1.get_atoms()
        s = slice(off, off + b.NPRT)
        assert np.array_equal(xb[s], x1) and np.array_equal(fb[s], f1)
        assert eb[k] == d1.get_box_epot()[0]
        off += b.NPRT
        d1.close()


def test_errors_are_reported_not_fatal(p0):
    box, ctrl, ft = p0
    say("error paths")
    small = cfg.w_ctrl(mx=8)
    dev = device_for(box, small, ft)
    with pytest.raises(MSMError) as ei:
        dev.build_neighbors()
    assert ei.value.code == 4
    dev.close()
    b2 = cfg.make_bcc_box((6, 6, 6), cfg.W_A, 0.0, 0.0, 1)
    b2.XP[0, 2] = b2.BOXLOW[2] + b2.ZL[2] + 3.0
    c2 = cfg.w_ctrl()
    c2.IFPD = (1, 1, 0)
    dev = device_for(b2, c2, cfg.w_forcetable(b2, c2))
    with pytest.raises(MSMError) as ei:
        dev.build_neighbors()
    assert ei.value.code == 5
    dev.close()


def test_million_atom_snapshot(orc):
    """The 80^3 BCC benchmark box (0.5% vacancies): list sets, forces and energy against the
    oracle at step 0 and after 20 steps (oracle on the GPU's snapshot)."""
    box = cfg.make_bcc_box((80, 80, 80), cfg.W_A, 0.05, 600.0, 1000, vacancy_frac=0.005)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl)
    say(f"1M snapshot n={box.NPRT}")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.calc_epot()
    kv, _ = dev.get_neighbors()
    _, _, fp, _, _ = dev.get_atoms()
    fh, eh, *_ = oracle_half(orc, box, ctrl, ft)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    assert abs(dev.get_epot().sum() - eh.sum()) <= E_REL * abs(eh.sum())
    okv, _ = orc.nlist("full", box.XP, box.ITYP - 1, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                       ctrl.NB_RM, ctrl.NB_MXNBS)
    assert np.array_equal(kv, okv)
    dev.run(20, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.calc_epot()
    x, v, fp, _, _ = dev.get_atoms()
    fh, eh, *_ = oracle_half(orc, box, ctrl, ft, xp=x)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    assert abs(dev.get_epot().sum() - eh.sum()) <= E_REL * abs(eh.sum())
    dev.close()


@pytest.mark.parametrize("path,compact,lds", [(0, 0, 0), (1, 0, 0), (1, 1, 0), (2, 0, 0),
                                              (2, 1, 0), (3, 0, 0), (4, 1, 0), (8, 1, 0),
                                              (1, 0, 1), (2, 0, 1), (4, 0, 1), (0, 0, 1),
                                              (3, 0, 1), (1, 1, 1), (2, 1, 1), (4, 1, 1)])
def test_force_paths_agree(orc, p0, path, compact, lds):
    """Every density/force kernel variant (LDS bricks with the table in global memory or in LDS;
    global list with 1-8 lanes per atom, with and without the density pass's in-cutoff record for
    the force pass -- a compacted list with the global table, per-lane slot masks with the LDS
    table -- and lane-pair gathers) meets the parity bar on the same box, before and after a
    15-step run."""
    import msmpscu_amd._lib as L
    box, ctrl, ft = p0
    say(f"force path {path} compact {compact} lds table {lds}")
    dev = device_for(box, ctrl, ft)
    L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_FORCE_PATH, path), "msm_set_option")
    L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_COMPACT_LIST, compact), "msm_set_option")
    L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_TABLE_LDS, lds), "msm_set_option")
    expect = {0: "brick_table" if lds else "bricks",
              3: "lds_table" if lds else "lane_pair"}.get(path, "lds_table" if lds else "lanes")
    assert dev.pass_kernel() == expect
    dev.build_neighbors()
    dev.calc_force()
    _, _, fp, _, _ = dev.get_atoms()
    fh, eh, _, ddh, _ = oracle_half(orc, box, ctrl, ft)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    assert np.max(np.abs(dev.get_den() - ddh)) <= 1e-13
    dev.run(15, ctrl.NB_UPTAB, 0, ctrl.H)
    x, _, fp, _, _ = dev.get_atoms()
    fh, *_ = oracle_half(orc, box, ctrl, ft, xp=x)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    dev.close()



@pytest.mark.parametrize("path", [1, 2, 4])
@pytest.mark.parametrize("vacancy,seed", [(0.8, 77), (0.95, 1)])
def test_short_lists_at_capacity(orc, vacancy, seed, path):
    """Lists of every length from 0 up to MXKVOIS, with MXKVOIS set to the longest list (11 and
    4 on these sparse W boxes): the LDS-table passes read each atom's first slot indices before
    its length is known, clamping the slot below MXKVOIS, and take the last slots of a list in
    pairs and singles from indices already read.  Forces and F' against the oracle with 1, 2 and
    4 lanes per atom."""
    import msmpscu_amd._lib as L
    box = cfg.make_bcc_box((10, 10, 10), cfg.W_A, 0.05, 300.0, seed, vacancy_frac=vacancy)
    big = cfg.w_ctrl(mx=64)
    kv, _ = orc.nlist("full", box.XP, box.ITYP - 1, box.STATU, box.BOXLOW, box.ZL, big.IFPD,
                      big.NB_RM, big.NB_MXNBS)
    mx = int(kv.max())
    assert kv.min() == 0 and np.all(np.bincount(kv, minlength=mx + 1) > 0)
    ctrl = cfg.w_ctrl(mx=mx)
    ft = cfg.w_forcetable(box, ctrl)
    say(f"short lists vacancy {vacancy} mx {mx} lanes {path}")
    dev = device_for(box, ctrl, ft)
    L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_FORCE_PATH, path), "msm_set_option")
    L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_TABLE_LDS, 1), "msm_set_option")
    assert dev.pass_kernel() == "lds_table"
    dev.build_neighbors()
    dev.calc_force()
    _, _, fp, _, _ = dev.get_atoms()
    fo, do, _ = oracle_full(orc, box, ctrl, ft)
    assert np.max(np.abs(fp - fo)) <= F_TOL
    # an atom whose only neighbours sit near the cutoff has a tiny rho and a large
    # F' = -rho^(-1/2)/2: the bar is relative to max(1, |F'|)
    den = dev.get_den()
    rel = np.abs(den - do) / np.maximum(1.0, np.abs(do))
    say(f"  F' max |F'| {np.max(np.abs(do)):.3e} worst relative {np.max(rel):.3e}")
    assert np.max(rel) <= 1e-13
    dev.close()


def _asym_case(ntab=5000):
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.08, 300.0, 8)
    box.NGROUP = 2
    box.ITYP = (np.arange(box.NPRT) % 2 + 1).astype(np.int32)
    box.CM = np.array([box.CM[0], box.CM[0] * 0.5])
    box.PTYPE = np.array([[1, 3], [4, 2]], dtype=np.int32)
    ctrl = md.SimMDCtrl(RU=np.full((2, 2), 4.4002), NB_RM=np.full((2, 2), 4.7002), NB_MXNBS=64,
                        NUMFTABR=ntab)
    ft = md.MDForceTable()
    md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
    md.Register_ForceTableProc(2, ft, "msm_pot_ackland_ww_nn", "msm_pot_ackland_ww_ne")
    md.Register_ForceTableProc(3, ft, "msm_pot_wfs_nn", "msm_pot_ackland_ww_ne")
    md.Register_ForceTableProc(4, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
    md.Create_Interaction_ForceTable(box, ctrl, ft)
    return [box], ctrl, ft


@pytest.mark.parametrize("variant", ["pair", "lds", "lds_pair", "lds_mask", "lds_f1", "bt1", "bt2",
                                     "bt4"])
@pytest.mark.parametrize("case", ["fs", "eam", "asym", "inactive", "multibox", "slab"])
def test_kernel_variant_matches_two_lane_path(p0, case, variant):
    """The lane-pair gather kernels (force path 3, with the table in global memory or in LDS) and
    the LDS-table kernels split each atom's list over two lanes as the two-lane kernels do (the
    compiler may contract multiply-adds differently, so agreement is to rounding): F' and forces
    to the parity bar, and positions
    after a 12-step run to 1e-12 A, over single and multiple species (the LDS table serves one
    kind; several kinds take the global table), FS and EAM, inactive atoms, uneven boxes and a
    slab.  lds_f1 runs the LDS-table force pass with one lane per atom after a two-lane
    density pass.  The bt variants are the LDS-brick passes with the table in LDS (1, 2, 4 lanes per
    home atom; the two-species cases use 1500-knot tables so that every kind's window fits)."""
    import msmpscu_amd._lib as L
    box, ctrl, ft = p0
    boxes, statu = [box], None
    bt = variant.startswith("bt")
    if case == "eam":
        ft = cfg.w_forcetable(box, ctrl, pot_type="EAM_TYPE")
    elif case == "asym":
        boxes, ctrl, ft = _asym_case(1500 if bt else 5000)
    elif case == "inactive":
        statu = box.STATU.copy()
        statu[[3, 100, 4000, 4001]] = 0
    elif case == "multibox":
        boxes = [cfg.make_bcc_box(c, cfg.W_A, 0.05, 500.0, 30 + k, n_vac=k, n_dumbbell=k % 2)
                 for k, c in enumerate([(5, 6, 7), (9, 9, 9), (3, 4, 5)])]
        ft = cfg.w_forcetable(boxes[0], ctrl)
    elif case == "slab":
        b = cfg.make_cuni_slab((6, 6, 6), 3.57, 600.0, 5)
        ctrl = cfg.cuni_ctrl(ntab=1500 if bt else 5000)
        ft = cfg.cuni_forcetable(b, ctrl)
        boxes = [b]
    say(f"{variant} vs two lanes: {case}")
    outs = []
    for k in range(2):
        dev = device_for(boxes, ctrl, ft)
        path = 3 if (k == 1 and variant in ("pair", "lds_pair")) else 2
        if k == 1 and bt:
            path = 0
            L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_BRICK_LANES, int(variant[2:])),
                    "msm_set_option")
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_FORCE_PATH, path), "msm_set_option")
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_TABLE_LDS,
                                       int(k == 1 and variant in ("lds", "lds_pair", "lds_mask",
                                                                  "lds_f1", "bt1", "bt2", "bt4"))),
                "msm_set_option")
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_FORCE_LANES,
                                       int(k == 1 and variant == "lds_f1")), "msm_set_option")
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_COMPACT_LIST,
                                       int(k == 1 and variant == "lds_mask")), "msm_set_option")
        if k == 1 and bt:
            assert dev.pass_kernel() == "brick_table"
        if k == 1 and variant in ("lds_pair", "lds_mask", "lds_f1") and case not in ("asym",
                                                                                       "slab"):
            assert dev.pass_kernel() == "lds_table"
        if statu is not None:
            dev.set_atoms(box.XP, box.XP1, box.ITYP, statu)
        dev.build_neighbors()
        dev.calc_force()
        r = [dev.get_den()] + list(dev.get_atoms())
        dev.run(12, ctrl.NB_UPTAB, 0, ctrl.H)
        r += list(dev.get_atoms())
        outs.append(r)
        dev.close()
    den2, x2, v2, f2 = outs[0][0], outs[0][1], outs[0][2], outs[0][3]
    den3, x3, v3, f3 = outs[1][0], outs[1][1], outs[1][2], outs[1][3]
    assert np.max(np.abs(den3 - den2)) <= 1e-13
    assert np.max(np.abs(f3 - f2)) <= F_TOL
    assert np.array_equal(x3, x2)
    assert np.max(np.abs(outs[1][6] - outs[0][6])) <= 1e-12
    assert np.max(np.abs(outs[1][8] - outs[0][8])) <= F_TOL


@pytest.mark.parametrize("path", [2, 0])
@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
def test_lds_table_window_fallback(orc, pot_type, path):
    """A 12000-knot table leaves the bottom of both tables outside the LDS windows (force window
    of the global-list passes starts near r = 2.7 A, density window near 1.45 A; the LDS-brick
    passes hold less), and a few atoms are pushed to 1.15-1.4 A from a neighbour: the pairs
    below the windows take the global-table fallback and every result still meets the parity
    bar against the oracle."""
    import msmpscu_amd._lib as L
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.05, 300.0, 77)
    d = np.array([cfg.W_A / 2, cfg.W_A / 2, cfg.W_A / 2])
    for i, frac in [(10, 0.55), (200, 0.5), (555, 0.58)]:
        # atom i toward its neighbour at +d: distance |d| * (1 - frac), 1.15-1.4 A
        box.XP[i] = box.XP[i] + frac * d
    box.XP = np.asfortranarray(cfg.wrap_into(box.XP, box.BOXLOW, box.ZL, (1, 1, 1)))
    ctrl = cfg.w_ctrl(ntab=12000)
    ft = cfg.w_forcetable(box, ctrl, pot_type=pot_type)
    say(f"lds window fallback {pot_type} path {path}")
    outs = []
    for lds in (0, 1):
        dev = device_for(box, ctrl, ft)
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_FORCE_PATH, path), "msm_set_option")
        L.check(L.lib().msm_set_option(dev.h, L.MSM_OPT_TABLE_LDS, lds), "msm_set_option")
        if lds:
            assert dev.pass_kernel() == ("lds_table" if path else "brick_table")
        dev.build_neighbors()
        dev.calc_force()
        _, _, fp, _, _ = dev.get_atoms()
        outs.append((fp, dev.get_den()))
        dev.close()
    fo, do, _ = oracle_full(orc, box, ctrl, ft)
    for fp, den in outs:
        assert np.max(np.abs(fp - fo)) <= F_TOL * max(1.0, np.max(np.abs(fo)) / 10.0)
        assert np.max(np.abs(den - do)) <= 1e-13 * max(1.0, np.max(np.abs(do)))
    assert np.max(np.abs(outs[1][0] - outs[0][0])) <= F_TOL


def test_autotune_keeps_state_and_parity(orc, p0):
    """The autotuner times kernel variants on the live state: positions and velocities stay as
    they were, the chosen variant's forces meet the parity bar, and a run afterwards matches the
    oracle's forces on the resulting snapshot."""
    box, ctrl, ft = p0
    say("autotune")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    x0, v0, _, _, _ = dev.get_atoms()
    choice = dev.autotune(2)
    say(f"autotune choice {choice}")
    x1, v1, f1, _, _ = dev.get_atoms()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    fh, *_ = oracle_half(orc, box, ctrl, ft)
    assert np.max(np.abs(f1 - fh)) <= F_TOL
    dev.run(12, ctrl.NB_UPTAB, 0, ctrl.H)
    x, _, fp, _, _ = dev.get_atoms()
    fh, *_ = oracle_half(orc, box, ctrl, ft, xp=x)
    assert np.max(np.abs(fp - fh)) <= F_TOL
    dev.close()


def test_uneven_batch_matches_single_boxes():
    """A small uneven batch (sizes 500-4000, vacancies, dumbbells, 300-1200 K) run as one
    multibox simulation after autotuning agrees with each box run alone: positions after 20
    steps to 1e-12 A and per-box energies to 1e-12 relative."""
    ctrl = cfg.w_ctrl(mx=64)
    boxes = cfg.make_batch(9, 500, 4000, seed=17)
    ft = cfg.w_forcetable(boxes[0], ctrl)
    say(f"uneven batch n={[b.NPRT for b in boxes]}")
    dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.autotune(2)
    dev.run(20, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.calc_epot()
    xb, _, _, _, _ = dev.get_atoms()
    eb = dev.get_box_epot()
    dev.close()
    off = 0
    for k, b in enumerate(boxes):
        d1 = device_for(b, ctrl, ft)
        d1.build_neighbors()
        d1.calc_force()
        d1.run(20, ctrl.NB_UPTAB, 0, ctrl.H)
        d1.calc_epot()
        x1, _, _, _, _ = d1.get_atoms()
        assert np.max(np.abs(xb[off:off + b.NPRT] - x1)) <= 1e-12
        e1 = d1.get_box_epot()[0]
        assert abs(eb[k] - e1) <= 1e-12 * abs(e1)
        off += b.NPRT
        d1.close()


def test_virial_fs_matches_oracle(orc, p0):
    """Virial tensor of the FS box (full list, FORTOT/2 per ordered pair) equals the oracle's
    half-list VTENSOR (CALFOR, each pair once) and is symmetric.  Each component is a sum of
    about 1e5 pair terms of both signs that largely cancel, so its rounding error scales with
    the sum of the terms' absolute values (reported by the oracle), not with the result: the
    bar is 1e-14 of that sum per component (about 45 ulps of it; a wrong factor, sign or index
    is off by O(1) of the result)."""
    box, ctrl, ft = p0
    say("virial vs oracle")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    v = dev.calc_ptensor()[0]
    _, _, fp, _, _ = dev.get_atoms()
    dev.close()
    fh, _, _, _, vt, vabs = oracle_half(orc, box, ctrl, ft, with_vabs=True)
    assert np.all(np.abs(v - vt) <= 1e-14 * vabs)
    assert np.array_equal(v, v.T)
    assert np.max(np.abs(fp - fh)) <= F_TOL


@pytest.mark.parametrize("case", ["fs", "eam", "asym", "multibox", "inactive"])
def test_atomic_stress_matches_oracle(orc, p0, case):
    """Per-atom stress (msm_calc_atomic_stress, CAL_FS/EAM_AtomicStress_KERNEL) equals the
    oracle's full-list restatement atom by atom, to 1e-13 of the sum of each component's absolute
    pair terms; inactive atoms carry none; summed over the atoms of each box it is twice that
    box's virial from msm_calc_ptensor."""
    box, ctrl, ft = p0
    boxes, statu = [box], None
    if case == "eam":
        ft = cfg.w_forcetable(box, ctrl, pot_type="EAM_TYPE")
    elif case == "asym":
        boxes, ctrl, ft = _asym_case()
    elif case == "multibox":
        boxes = [cfg.make_bcc_box(c, cfg.W_A, 0.05, 500.0, 40 + k, n_vac=k)
                 for k, c in enumerate([(5, 6, 7), (9, 9, 9), (4, 4, 5)])]
        ft = cfg.w_forcetable(boxes[0], ctrl)
    elif case == "inactive":
        statu = box.STATU.copy()
        statu[[3, 100, 4000]] = 0
    say(f"atomic stress {case}")
    dev = device_for(boxes, ctrl, ft)
    if statu is not None:
        dev.set_atoms(box.XP, box.XP1, box.ITYP, statu)
    dev.build_neighbors()
    avp = dev.calc_atomic_stress()
    vt = dev.calc_ptensor()
    dev.close()
    t = orc.from_mdlib(ft)
    off = 0
    for k, b in enumerate(boxes):
        st = b.STATU if statu is None else statu
        it = b.ITYP - 1
        kv, ind = orc.nlist("full", b.XP, it, st, b.BOXLOW, b.ZL, ctrl.IFPD, ctrl.NB_RM,
                            ctrl.NB_MXNBS)
        _, dd, _ = orc.force_full(b.XP, it, st, kv, ind, b.ZL, ctrl.IFPD,
                                  float(np.max(ctrl.RU)) ** 2, t)
        ap, aabs = orc.atomic_stress_full(b.XP, it, st, kv, ind, b.ZL, ctrl.IFPD,
                                          float(np.max(ctrl.RU)) ** 2, t, dd, with_abs=True)
        mine = avp[off:off + b.NPRT]
        assert np.all(np.abs(mine - ap) <= 1e-13 * aabs + 1e-300)
        assert np.all(mine[st == 0] == 0)
        tot = mine.sum(axis=0).reshape(3, 3)
        assert np.all(np.abs(tot - 2.0 * vt[k]) <= 1e-13 * aabs.sum(axis=0).reshape(3, 3))
        off += b.NPRT


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
def test_virial_is_minus_strain_derivative_of_energy(pot_type):
    """W_aa = -dE/d(eps_aa) under a homogeneous diagonal strain of positions and box (central
    differences of the device energy at eps = 1e-6), for FS and EAM tables.  Energies interpolate
    POTR/POTB and forces the separately tabulated FPOTR/FPOTB, so the two agree to the tables'
    interpolation error (about 2e-4 relative here), not to rounding: the bar is 1e-3."""
    base = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.1, 0.0, 21)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(base, ctrl, pot_type=pot_type)
    say(f"virial strain derivative {pot_type}")

    def energy(eps):
        b = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.1, 0.0, 21)
        f = 1.0 + np.asarray(eps)
        b.XP = np.asfortranarray(b.XP * f)
        b.ZL = b.ZL * f
        b.BOXLOW = -0.5 * b.ZL
        dev = device_for(b, ctrl, ft)
        dev.build_neighbors()
        dev.calc_force()
        dev.calc_epot()
        e = dev.get_box_epot()[0]
        v = dev.calc_ptensor()[0]
        dev.close()
        return e, v

    _, v0 = energy([0.0, 0.0, 0.0])
    h = 1e-6
    for a in range(3):
        ep = np.zeros(3)
        ep[a] = h
        e_plus, _ = energy(ep)
        e_minus, _ = energy(-ep)
        dE = (e_plus - e_minus) / (2 * h)
        assert abs(v0[a, a] + dE) <= 1e-3 * abs(v0[a, a]), (a, v0[a, a], dE)


def test_virial_multibox_equals_single_boxes():
    ctrl = cfg.w_ctrl(mx=64)
    boxes = [cfg.make_bcc_box(c, cfg.W_A, 0.08, 0.0, 50 + k, n_vac=k)
             for k, c in enumerate([(6, 6, 6), (7, 5, 6)])]
    ft = cfg.w_forcetable(boxes[0], ctrl)
    say("virial multibox")
    dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    vb = dev.calc_ptensor()
    dev.close()
    for k, b in enumerate(boxes):
        d1 = device_for(b, ctrl, ft)
        d1.build_neighbors()
        v1 = d1.calc_ptensor()[0]
        d1.close()
        assert np.max(np.abs(vb[k] - v1)) <= 1e-12 * np.max(np.abs(v1))


def test_ragged_and_degenerate_boxes(orc):
    """One multibox run over edge cases: the smallest box the cell grid allows (3x3x3 BCC cells,
    where every cell neighbours every other), a box holding a single atom (no neighbours: zero
    force, F(0) = 0), a box whose atoms are all inactive (listed, but no force or energy), and
    an ordinary box.  Each box's list,
    forces and energies equal the oracle's for that box alone, and 20 fused steps keep them
    equal to the boxes run one at a time."""
    ctrl = cfg.w_ctrl(mx=64)
    tiny = cfg.make_bcc_box((3, 3, 3), cfg.W_A, 0.05, 300.0, 61)
    lone = cfg.make_bcc_box((4, 4, 4), cfg.W_A, 0.0, 0.0, 62)
    keep = np.zeros(lone.NPRT, dtype=bool)
    keep[5] = True
    lone.NPRT = 1
    lone.XP = np.asfortranarray(lone.XP[keep])
    lone.XP1 = np.asfortranarray(np.full((1, 3), 0.01))
    lone.ITYP = lone.ITYP[keep]
    lone.STATU = lone.STATU[keep]
    dead = cfg.make_bcc_box((3, 4, 3), cfg.W_A, 0.05, 300.0, 63)
    dead.STATU = np.zeros(dead.NPRT, np.int32)
    plain = cfg.make_bcc_box((5, 4, 6), cfg.W_A, 0.05, 300.0, 64, n_vac=2)
    boxes = [tiny, lone, dead, plain]
    ft = cfg.w_forcetable(plain, ctrl)
    say("ragged and degenerate boxes")
    dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.calc_epot()
    _, _, fp, _, _ = dev.get_atoms()
    ep = dev.get_epot()
    kv, ind = dev.get_neighbors()
    off = 0
    for b in boxes:
        s = slice(off, off + b.NPRT)
        it = b.ITYP - 1
        okv, oind = orc.nlist("full", b.XP, it, b.STATU, b.BOXLOW, b.ZL, ctrl.IFPD, ctrl.NB_RM,
                              ctrl.NB_MXNBS)
        assert np.array_equal(kv[s], okv)
        assert pairset(kv[s], ind[s] - off, True) == pairset(okv, oind, False)
        fo, _, eo = oracle_full(orc, b, ctrl, ft)
        assert np.max(np.abs(fp[s] - fo)) <= F_TOL
        assert np.allclose(ep[s], eo, rtol=0, atol=1e-12 * max(1.0, np.max(np.abs(eo))))
        off += b.NPRT
    assert kv[tiny.NPRT] == 0 and np.all(fp[tiny.NPRT] == 0.0) and ep[tiny.NPRT] == 0.0
    # inactive atoms keep their list entries (the reference's list predicate does not look at
    # STATU) but get no force and no energy
    d0 = tiny.NPRT + 1
    assert np.all(fp[d0:d0 + dead.NPRT] == 0.0) and np.all(ep[d0:d0 + dead.NPRT] == 0.0)
    dev.run(20, ctrl.NB_UPTAB, 0, ctrl.H)
    xb, _, fb, _, _ = dev.get_atoms()
    dev.close()
    off = 0
    for b in boxes:
        d1 = device_for(b, ctrl, ft)
        d1.build_neighbors()
        d1.calc_force()
        d1.run(20, ctrl.NB_UPTAB, 0, ctrl.H)
        x1, _, f1, _, _ = d1.get_atoms()
        s = slice(off, off + b.NPRT)
        assert np.array_equal(xb[s], x1) and np.array_equal(fb[s], f1)
        off += b.NPRT
        d1.close()
