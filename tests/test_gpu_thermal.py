"""GPU parity of the thermalisation and quench entry points (msm_damping, msm_calc_ekin,
msm_scale_velocities, msm_dyn_damp, msm_steepest, msm_cg, and the L-BFGS driver over
msm_put_atoms) against the CPU oracle, through the C-ABI."""
import sys

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md

pytestmark = pytest.mark.gpu

KB = md.CP_KB_EV


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def device_for(boxes, ctrl, ft):
    boxes = boxes if isinstance(boxes, list) else [boxes]
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(ft.PotType, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc, fc.Device


def mixed_statu(n, seed):
    """ACTIVE everywhere; a few atoms with FIXPOS or FIXVEL bits and a few inactive ones."""
    rng = np.random.default_rng(seed)
    st = np.ones(n, np.int32)
    idx = rng.permutation(n)
    st[idx[:6]] |= 2
    st[idx[6:12]] |= 4 | 8
    st[idx[12:18]] |= 16 | 64
    st[idx[18:22]] = 0
    return st


@pytest.fixture
def two_boxes():
    ctrl = cfg.w_ctrl(mx=64)
    boxes = [cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.04, 300.0, 31),
             cfg.make_bcc_box((7, 9, 8), cfg.W_A, 0.04, 900.0, 32, n_vac=3)]
    for k, b in enumerate(boxes):
        b.STATU = mixed_statu(b.NPRT, 40 + k)
    ft = cfg.w_forcetable(boxes[0], ctrl)
    return boxes, ctrl, ft


def test_damping_matches_oracle(orc, two_boxes):
    boxes, ctrl, ft = two_boxes
    say("damping")
    _, dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    x0, v0, f0, st0, _ = dev.get_atoms()
    dev.damping()
    _, v1, _, _, _ = dev.get_atoms()
    assert np.array_equal(v1, orc.damping(v0, f0, st0))
    assert np.count_nonzero(v1 == 0.0) > np.count_nonzero(v0 == 0.0)
    dev.close()


def test_ekin_and_box_temperature_match_oracle(orc, two_boxes):
    boxes, ctrl, ft = two_boxes
    say("ekin")
    fc, dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    ek, tb = dev.calc_ekin(KB)
    _, v, _, st, _ = dev.get_atoms()
    it = np.concatenate([b.ITYP for b in boxes]) - 1
    oek = orc.ekin(v, st, it, boxes[0].CM)
    assert np.max(np.abs(ek - oek) / np.maximum(np.abs(oek), 1e-300)) <= 1e-15
    off = 0
    for k, b in enumerate(boxes):
        e = oek[off:off + b.NPRT]
        tk = 2.0 * e[e >= 0].mean() / (3.0 * KB)
        assert tb[k] == pytest.approx(tk, rel=1e-13)
        off += b.NPRT
    tg = md.Cal_GlobalT_DEV(boxes, ctrl, fc, KB)
    assert tg == pytest.approx(2.0 * oek[oek >= 0].mean() / (3.0 * KB), rel=1e-13)
    dev.close()


def test_scaling_matches_oracle_and_reaches_the_target(orc, two_boxes):
    boxes, ctrl, ft = two_boxes
    say("velocity scaling")
    fc, dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    _, v0, _, st, _ = dev.get_atoms()
    md.VelScaling_DEV(boxes, ctrl, fc, 600.0, KB)
    _, v1, _, _, _ = dev.get_atoms()
    off = 0
    for b in boxes:
        s = slice(off, off + b.NPRT)
        ov = orc.velscale(v0[s], st[s], b.ITYP - 1, b.CM, 600.0, KB)
        assert np.max(np.abs(v1[s] - ov)) <= 1e-15 * np.max(np.abs(ov))
        off += b.NPRT
    # the FIXVEL atoms counted in <EK> before scaling lose their velocity, so the boxes land
    # slightly below 600 K; a second scaling (after the zeroing) lands on it exactly
    md.VelScaling_DEV(boxes, ctrl, fc, 600.0, KB)
    _, tb = dev.calc_ekin(KB)
    assert np.allclose(tb, 600.0, rtol=1e-12)
    dev.close()


@pytest.mark.parametrize("mindele,maxsteps", [(0.0, 40), (1e-3, 200)])
def test_dyn_damp_matches_oracle(orc, mindele, maxsteps):
    """Damped quench of a displaced 1024-atom W box at 300 K on a fixed list against the oracle's
    restatement of the loop: same stopping iteration (41 without convergence; 74 for 1e-3 eV,
    where the last max |dEPOT| sits 5% under the threshold), positions, velocities and forces
    within the bars of the 100-step run."""
    ctrl = cfg.w_ctrl(mx=64)
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.06, 300.0, 77)
    box.STATU = mixed_statu(box.NPRT, 78)
    ft = cfg.w_forcetable(box, ctrl)
    say(f"dyn damp mindele={mindele}")
    fc, dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    x0, v0, f0, st0, d0 = dev.get_atoms()
    kv, ind = dev.get_neighbors()
    ind0 = np.where(ind > 0, ind - 1, 0)
    t = orc.from_mdlib(ft)
    it = box.ITYP - 1
    ru2 = float(np.max(ctrl.RU)) ** 2
    ox, ov, of, od, ost, oep, osteps, ode = orc.dyn_damp(
        x0, v0, f0, d0, st0, it, box.CM, ctrl.H, box.BOXLOW, box.ZL, ctrl.IFPD, kv, ind0, ru2,
        t, maxsteps, mindele)
    if mindele > 0.0:
        assert osteps <= maxsteps
    else:
        assert osteps == maxsteps + 1
    ctrl.STEEPEST_MiDelE = mindele
    steps, de = md.Do_DynDamp_Forsteps_DEV([box], ctrl, fc, maxsteps)
    assert steps == osteps
    assert de == pytest.approx(ode, rel=1e-6, abs=1e-12)
    x, v, f, st, d = dev.get_atoms()
    dx = x - ox
    dx -= np.round(dx / box.ZL) * box.ZL
    assert np.max(np.abs(dx)) <= 1e-9
    assert np.max(np.abs(v - ov)) <= 1e-11
    assert np.max(np.abs(f - of)) <= 1e-8
    assert np.array_equal(st & 0xFFFF, ost & 0xFFFF)
    dev.close()


@pytest.mark.parametrize("meth", [md.CP_DAMPSCHEME_ST, md.CP_DAMPSCHEME_ST | md.CP_DAMPSCHEME_LSEARCH])
def test_steepest_matches_oracle(orc, meth):
    """Steepest-descent quench of a displaced 1024-atom W box (a few inactive atoms) on a fixed
    list against the oracle's restatement: same number of steps, positions to 1e-9 A."""
    ctrl = cfg.w_ctrl(mx=64)
    ctrl.STEEPEST_MiDelE = 1e-6
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.06, 0.0, 91)
    box.STATU = mixed_statu(box.NPRT, 92)
    ft = cfg.w_forcetable(box, ctrl)
    ls = bool(meth & md.CP_DAMPSCHEME_LSEARCH)
    say(f"steepest linesearch={ls}")
    fc, dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    x0, _, _, st0, _ = dev.get_atoms()
    kv, ind = dev.get_neighbors()
    ind0 = np.where(ind > 0, ind - 1, 0)
    t = orc.from_mdlib(ft)
    ru2 = float(np.max(ctrl.RU)) ** 2
    maxsteps = 300
    ox, osteps, ode, omm = orc.steepest(
        x0, st0, box.ITYP - 1, box.BOXLOW, box.ZL, ctrl.IFPD, kv, ind0, ru2, t, maxsteps, ls,
        ctrl.STEEPEST_Alpha, ctrl.STEEPEST_MxStep * box.RR, ctrl.STEEPEST_MiStep * box.RR,
        1e-3 if ls else ctrl.STEEPEST_MiDelE)
    assert 0 < osteps <= maxsteps
    steps, de, mm = md.Do_Steepest_Forsteps_DEV([box], ctrl, fc, maxsteps, meth)
    assert steps == osteps
    assert de == pytest.approx(ode, rel=1e-6, abs=1e-13)
    x, _, _, st, _ = dev.get_atoms()
    dx = x - ox
    dx -= np.round(dx / box.ZL) * box.ZL
    assert np.max(np.abs(dx)) <= 1e-9
    assert np.all(x >= box.BOXLOW) and np.all(x <= box.BOXLOW + box.ZL)
    dev.close()


@pytest.mark.parametrize("meth,maxsteps,mindele", [(md.CP_DAMPSCHEME_CG, 4, 1e-12),
                                                   (md.CP_DAMPSCHEME_CG | md.CP_DAMPSCHEME_LSEARCH,
                                                    300, 1e-5)])
def test_cg_matches_oracle(orc, meth, maxsteps, mindele):
    """Conjugate-gradient quench against the oracle's restatement on a displaced 1024-atom W box
    with a few inactive atoms: the plain variant for four iterations (it does not converge on
    this box: the reference divides its secant step by |F0| twice), the line-search variant to
    1e-5 eV; same step counts, positions to 1e-9 A."""
    ctrl = cfg.w_ctrl(mx=64)
    ctrl.STEEPEST_MiDelE = mindele
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.06, 0.0, 93)
    box.STATU = mixed_statu(box.NPRT, 94)
    ft = cfg.w_forcetable(box, ctrl)
    ls = bool(meth & md.CP_DAMPSCHEME_LSEARCH)
    say(f"cg linesearch={ls}")
    fc, dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    x0, _, _, st0, _ = dev.get_atoms()
    kv, ind = dev.get_neighbors()
    ind0 = np.where(ind > 0, ind - 1, 0)
    t = orc.from_mdlib(ft)
    ru2 = float(np.max(ctrl.RU)) ** 2
    ox, osteps, ode = orc.cg(x0, st0, box.ITYP - 1, box.BOXLOW, box.ZL, ctrl.IFPD, kv, ind0, ru2,
                             t, maxsteps, ls, ctrl.STEEPEST_MxStep * box.RR,
                             ctrl.STEEPEST_MiStep * box.RR, mindele)
    steps, de = md.Do_CG_Forsteps_DEV([box], ctrl, fc, maxsteps, meth)
    assert steps == osteps
    assert de == pytest.approx(ode, rel=1e-6, abs=1e-13)
    x, _, _, _, _ = dev.get_atoms()
    dx = x - ox
    dx -= np.round(dx / box.ZL) * box.ZL
    assert np.max(np.abs(dx)) <= 1e-9
    dev.close()


def test_put_atoms_round_trip_and_forces(orc):
    """msm_put_atoms: positions and velocities written in the caller's order come back bit-exact
    (either may be left out), and the forces at the new positions match the oracle's on the same
    list; coordinates several box lengths out (an L-BFGS trial point) are shifted by one box
    length at most, as the reference's minimum image does."""
    ctrl = cfg.w_ctrl(mx=64)
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.04, 300.0, 95)
    box.STATU = mixed_statu(box.NPRT, 96)
    ft = cfg.w_forcetable(box, ctrl)
    say("put_atoms")
    fc, dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    x0, v0, _, st, _ = dev.get_atoms()
    rng = np.random.default_rng(97)
    x1 = np.asfortranarray(x0 + rng.normal(0.0, 0.02, x0.shape))
    v1 = np.asfortranarray(rng.normal(0.0, 1e-3, v0.shape))
    dev.put_atoms(x1, None)
    x, v, _, _, _ = dev.get_atoms()
    assert np.array_equal(x, x1) and np.array_equal(v, v0)
    dev.put_atoms(None, v1)
    x, v, _, _, _ = dev.get_atoms()
    assert np.array_equal(x, x1) and np.array_equal(v, v1)
    kv, ind = dev.get_neighbors()
    ind0 = np.where(ind > 0, ind - 1, 0)
    t = orc.from_mdlib(ft)
    ru2 = float(np.max(ctrl.RU)) ** 2
    it = box.ITYP - 1
    # 40 atoms, none in another's list, to be sent far out
    far, taken = [], set()
    for i in rng.permutation(box.NPRT):
        nb = set(int(j) for j in ind0[i, :kv[i]])
        if i not in taken and not (nb & set(far)):
            far.append(int(i))
            taken |= nb
        if len(far) == 40:
            break
    far = np.asarray(far)
    for shift in (0.0, 3.0e5):
        xs = np.array(x1, order="F")
        xs[far] += shift * rng.choice([-1.0, 1.0], (40, 3))
        dev.put_atoms(xs)
        dev.calc_force()
        dev.calc_epot()
        _, _, f, _, _ = dev.get_atoms()
        assert np.array_equal(dev.get_forces(), f)
        ep = dev.get_epot()
        of, _, oe = orc.force_full(xs, it, st, kv, ind0, box.ZL, ctrl.IFPD, ru2, t)
        assert np.max(np.abs(f - of)) <= 1e-10
        assert abs(ep.sum() - oe.sum()) <= 1e-12 * abs(oe.sum())
        if shift:
            assert np.all(f[far] == 0.0) and np.all(ep[far] == 0.0)
    dev.close()


def test_lbfgs_matches_oracle(orc):
    """L-BFGS quench (the host optimiser over device forces, msm_put_atoms between evaluations)
    of a displaced 1024-atom W box with FIXPOS, inactive and OUTOFBOX atoms, against the same
    loop over the oracle's forces on the device's list: same evaluation count and flag, positions
    to 1e-8 A, velocities zero at the end."""
    ctrl = cfg.w_ctrl(mx=64)
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.05, 300.0, 98)
    box.STATU = mixed_statu(box.NPRT, 99)
    box.STATU[7] |= 65536
    ft = cfg.w_forcetable(box, ctrl)
    say("lbfgs")
    fc, dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    x0, _, _, st0, _ = dev.get_atoms()
    kv, ind = dev.get_neighbors()
    ind0 = np.where(ind > 0, ind - 1, 0)
    t = orc.from_mdlib(ft)
    ru2 = float(np.max(ctrl.RU)) ** 2
    it = box.ITYP - 1
    state = {"x": np.array(x0)}

    def evaluate():
        f, _, e = orc.force_full(state["x"], it, st0, kv, ind0, box.ZL, ctrl.IFPD, ru2, t)
        return f, e

    def put(xa):
        state["x"] = np.array(xa)

    n = box.NPRT
    lo3, zl3 = np.tile(box.BOXLOW, (n, 1)), np.tile(box.ZL, (n, 1))
    osteps, oflag, ox = md.lbfgs_quench(x0, st0, lo3, zl3, ctrl.IFPD, evaluate, put, 300,
                                        ctrl.LBFGS_Msave)
    assert oflag == 0 and 0 < osteps < 300
    steps, iflag = md.Do_LBFGS_Forsteps_DEV([box], ctrl, fc, 300)
    assert (steps, iflag) == (osteps, oflag)
    x, v, _, st, _ = dev.get_atoms()
    assert np.max(np.abs(x - ox)) <= 1e-8
    assert np.all(v == 0.0) and np.array_equal(st, st0)
    dev.close()


def test_check_timestep_matches_oracle(orc, two_boxes):
    """msm_check_timestep against the oracle at steps around the threshold (the largest step
    displacement of the box at 300/900 K with forces), and the halving search of
    Check_Timestep_DEV / Predictor_DEV with IHDUP < 0 against the oracle's."""
    boxes, ctrl, ft = two_boxes
    say("check timestep")
    fc, dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    _, v, f, st, _ = dev.get_atoms()
    it = np.concatenate([b.ITYP for b in boxes]) - 1
    cm = boxes[0].CM
    for th in (0.5, 1.0, 2.0, 4.0, 8.0, 16.0):
        for mxd2 in (1e-4, 1e-3, 1e-2, 1e-1):
            assert dev.check_timestep(th, mxd2) == orc.check_timestep(v, f, st, it, cm, th, mxd2)
    ctrl.HMX, ctrl.DMX, ctrl.IHDUP, ctrl.IT0 = 16.0, 0.05, -5, 1
    want = orc.timestep_search(v, f, st, it, cm, 16.0, 0.05)
    assert want < 16.0
    assert md.Check_Timestep_DEV(boxes, ctrl, fc) == want and ctrl.H == want
    # through the predictor: mod(ITIME - IT0 + 1, 5) = 0 at ITIME 5 searches, at ITIME 6 not
    ctrl.H = 1.0
    md.Predictor_DEV(5, boxes, ctrl, fc)
    assert ctrl.H == want
    ctrl.H = 1.0
    md.Predictor_DEV(6, boxes, ctrl, fc)
    assert ctrl.H == 1.0
    dev.close()


def test_thermalizing_mc_matches_oracle(orc, two_boxes):
    """Thermalizing_MC_DEV on the device (msm_thermalize) against the oracle's restatement fed by
    the same Philox4x32-10 stream: two uneven boxes with FIXPOS/FIXVEL and inactive atoms, two
    draws in a row; then the velocities keep zero box momentum and the caller-order counter makes
    them independent of the cell order (the same draw on a context whose atoms were sorted from
    other positions gives the same velocities)."""
    boxes, ctrl, ft = two_boxes
    say("thermalizing_mc")
    ctrl.SEED = (12345, 678, 0, 0)
    fc, dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    x0, v0, _, st0, _ = dev.get_atoms()
    natoms = [b.NPRT for b in boxes]
    ityp = np.concatenate([b.ITYP for b in boxes])
    seed = md.thermal_seed(ctrl.SEED)
    vo = v0
    for draw in range(2):
        md.Thermalizing_MC_DEV(boxes, ctrl, fc, 800.0)
        _, v, _, st, _ = dev.get_atoms()
        vo = orc.thermalize_philox(natoms, ityp, st0, boxes[0].CM, 800.0 * KB, seed, draw, vo)
        scale = float(np.max(np.abs(vo)))
        assert np.max(np.abs(v - vo)) <= 1e-13 * scale, draw
        off = 0
        for b in boxes:
            s = slice(off, off + b.NPRT)
            m = b.CM[b.ITYP - 1]
            p = (m[:, None] * v[s]).sum(axis=0)
            assert np.max(np.abs(p)) <= 1e-12 * (m[:, None] * np.abs(v[s])).sum()
            off += b.NPRT
    # a second context with every box shifted by 1.3 A along x (another cell order): draw 0
    import copy
    boxes2 = copy.deepcopy(boxes)
    for b in boxes2:
        b.XP = cfg.wrap_into(b.XP + np.array([1.3, 0.0, 0.0]), b.BOXLOW, b.ZL, ctrl.IFPD)
    fc2, dev2 = device_for(boxes2, ctrl, ft)
    dev2.build_neighbors()
    md.Thermalizing_MC_DEV(boxes2, ctrl, fc2, 800.0)
    _, v2, _, _, _ = dev2.get_atoms()
    v_first = orc.thermalize_philox(natoms, ityp, st0, boxes[0].CM, 800.0 * KB, seed, 0, v0)
    assert np.max(np.abs(v2 - v_first)) <= 1e-13 * float(np.max(np.abs(v_first)))
    fc.Device.close()
    dev2.close()


def test_predictor_damping_window_and_step_search(orc, two_boxes):
    """Predictor_DEV inside the damping window at a search step (damping, then the halving search
    from HMX on the damped velocities, then the Swope predictor at the searched step) and outside
    both (plain predictor at H): positions, velocities and DIS against the oracle's damping,
    timestep_search and predictor on the same state."""
    boxes, ctrl, ft = two_boxes
    say("predictor damping window + step search")
    it_all = np.concatenate([b.ITYP for b in boxes]) - 1
    cm = boxes[0].CM
    for itime, h0 in ((5, 1.0), (8, 0.75)):
        fc, dev = device_for(boxes, ctrl, ft)
        dev.build_neighbors()
        dev.calc_force()
        x0, v0, f0, st0, d0 = dev.get_atoms()
        c = md.SimMDCtrl(RU=ctrl.RU, NB_RM=ctrl.NB_RM, NB_MXNBS=ctrl.NB_MXNBS, H=h0, HMX=16.0,
                         DMX=0.05, IHDUP=-5, IT0=1, DAMPTIME0=3, DAMPTIME1=4)
        md.Predictor_DEV(itime, boxes, c, fc)
        x1, v1, _, st1, d1 = dev.get_atoms()
        damp = 3 <= itime <= 6
        search = (itime - 1 + 1) % 5 == 0
        v = orc.damping(v0, f0, st0) if damp else v0
        h = orc.timestep_search(v, f0, st0, it_all, cm, 16.0, 0.05) if search else h0
        assert c.H == h
        if search:
            assert h < 16.0
        off = 0
        for b in boxes:
            s = slice(off, off + b.NPRT)
            ox, ov, od, ost = orc.predictor(x0[s], v[s], f0[s], d0[s], st0[s], b.ITYP - 1, cm, h,
                                            b.BOXLOW, b.ZL, ctrl.IFPD)
            assert np.max(np.abs(x1[s] - ox)) <= 1e-12
            assert np.max(np.abs(v1[s] - ov)) <= 1e-13
            assert np.max(np.abs(d1[s] - od)) <= 1e-12
            # the device predictor also raises PASSBOUND (bit 19) on a periodic wrap, as the
            # reference's Predictor_KERNEL0 does; the host restatement keeps the low bits only
            assert np.array_equal(st1[s] & 0xFFFF, ost & 0xFFFF)
            off += b.NPRT
        dev.close()


def test_lbfgsb_matches_oracle(orc):
    """L-BFGS-B quench (Do_LBFGSB_Forsteps_DEV: the unbounded SETULB over device forces, the
    positions written to the device after every call) of a displaced 1024-atom W box with
    FIXPOS, inactive and OUTOFBOX atoms, against the same loop over the oracle's forces on the
    device's list.  A 40-step prefix: same steps, flag and task, positions to 1e-8 A.  To
    convergence: both stop on the same test, at positions within 1e-8 A; the step count is not
    compared, because the FACTR test near the minimum turns on energy differences of 1e-11 eV
    (forces perturbed by 1e-13 relative move the oracle's own count between 87 and 107 here)."""
    ctrl = cfg.w_ctrl(mx=64)
    ctrl.LBFGS_PGtol = 1e-8
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.05, 300.0, 98)
    box.STATU = mixed_statu(box.NPRT, 99)
    box.STATU[7] |= 65536
    ft = cfg.w_forcetable(box, ctrl)
    t = orc.from_mdlib(ft)
    ru2 = float(np.max(ctrl.RU)) ** 2
    it = box.ITYP - 1
    n = box.NPRT
    lo3, zl3 = np.tile(box.BOXLOW, (n, 1)), np.tile(box.ZL, (n, 1))
    for maxsteps in (40, 400):
        say(f"lbfgsb maxsteps={maxsteps}")
        fc, dev = device_for(box, ctrl, ft)
        dev.build_neighbors()
        x0, _, _, st0, _ = dev.get_atoms()
        kv, ind = dev.get_neighbors()
        ind0 = np.where(ind > 0, ind - 1, 0)
        state = {"x": np.array(x0)}

        def evaluate():
            f, _, e = orc.force_full(state["x"], it, st0, kv, ind0, box.ZL, ctrl.IFPD, ru2, t)
            return f, e

        def put(xa):
            state["x"] = np.array(xa)

        osteps, oflag, ox, otask = md.lbfgsb_quench(x0, st0, lo3, zl3, ctrl.IFPD, evaluate, put,
                                                    maxsteps, ctrl.LBFGS_Msave, ctrl.LBFGS_Factr,
                                                    ctrl.LBFGS_PGtol)
        steps, iflag, task = md.Do_LBFGSB_Forsteps_DEV([box], ctrl, fc, maxsteps)
        if maxsteps == 40:
            assert (osteps, oflag) == (41, 1)
            assert (steps, iflag, task) == (osteps, oflag, otask)
        else:
            assert oflag == 0 and otask.startswith("CONVERGENCE"), otask
            assert (iflag, task) == (oflag, otask)
            assert steps <= maxsteps
        x, v, _, st, _ = dev.get_atoms()
        assert np.max(np.abs(x - ox)) <= 1e-8
        assert np.all(v == 0.0) and np.array_equal(st, st0)
        dev.close()
