"""The Fortran drop-in layer (msmpscu_amd/fortran: MDLIB-named routines over the C-ABI) driven by
a Fortran program: forces and energy of the W box at the start and after 100 steps (one
per-step call, then a fused run) against the oracle on the same positions, the per-atom
stresses of Cal_AtomicStress summed over the atoms against twice the oracle's virial, the
temperature after VelScaling_DEV, the positions after Do_Steepest_Forsteps_DEV against the
oracle's steepest descent, and the copy-in of positions and velocities."""
import os
import subprocess
import sys

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
from msmpscu_amd.mdlib import SimMDBox

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "msmpscu_amd",
                      "fortran", "fs_w_driver")
F_TOL = 1e-10
E_REL = 1e-12


def _box(xp):
    n = xp.shape[0]
    nc = round((n / 2) ** (1 / 3))
    ln = np.full(3, nc * cfg.W_A)
    return SimMDBox(NPRT=n, NGROUP=1, ZL=ln, BOXLOW=-0.5 * ln, XP=np.asfortranarray(xp),
                    XP1=np.zeros((n, 3), order="F"), ITYP=np.ones(n, dtype=np.int32),
                    STATU=np.ones(n, dtype=np.int32), CM=np.array([cfg.W_MASS * cfg.AMU]),
                    PTYPE=np.array([[1]], dtype=np.int32), SYMB=["W"])


def test_fortran_driver_against_oracle(orc, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("fs_w_driver is not built (msmpscu_amd.build.build_fortran)")
    out = tmp_path / "w.txt"
    print("[gpu-test] fortran driver 100 steps", flush=True)
    sys.stdout.flush()
    r = subprocess.run([DRIVER, str(out), "100", "16"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = out.read_text().splitlines()
    n = int(lines[0].split()[0])
    e0 = float(lines[0].split()[1])
    start = np.array([[float(v) for v in l.split()] for l in lines[1:1 + n]])
    end = np.array([[float(v) for v in l.split()] for l in lines[1 + n:1 + 2 * n]])
    e1 = float(lines[1 + 2 * n].split()[3])
    ssum = np.array([float(v) for v in lines[2 + 2 * n].split()[2:]]).reshape(3, 3)
    ctrl = cfg.w_ctrl()
    for blk, e in ((start, e0), (end, e1)):
        box = _box(blk[:, :3])
        ft = cfg.w_forcetable(box, ctrl)
        t = orc.from_mdlib(ft)
        it = box.ITYP - 1
        kv, ind = orc.nlist("half", box.XP, it, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                            ctrl.NB_RM, ctrl.NB_MXNBS)
        fo, eo, _, _, vt, vabs = orc.force_fs_half(box.XP, it, kv, ind, box.ZL, ctrl.IFPD,
                                                   ctrl.RU, t, with_vabs=True)
        assert np.max(np.abs(blk[:, 3:] - fo)) <= F_TOL
        assert abs(e - eo.sum()) <= E_REL * abs(eo.sum())
    # positions were written to 17 significant digits, so the oracle sees them to ~1e-17 relative
    assert np.all(np.abs(ssum - 2.0 * vt) <= 1e-13 * 2.0 * vabs)
    # the run moved the atoms
    assert np.max(np.abs(end[:, :3] - start[:, :3])) > 1e-3
    # VelScaling_DEV then Cal_GlobalT_DEV: every atom counts, so the box lands on 300 K
    temp = float(lines[3 + 2 * n].split()[2])
    assert temp == pytest.approx(300.0, rel=1e-12)
    # Do_Steepest_Forsteps_DEV from the end positions against the oracle's restatement
    assert lines[4 + 2 * n].strip() == "# steepest"
    quench = np.array([[float(v) for v in l.split()] for l in lines[5 + 2 * n:5 + 3 * n]])
    box = _box(end[:, :3])
    box.RR = cfg.W_A
    ft = cfg.w_forcetable(box, ctrl)
    t = orc.from_mdlib(ft)
    it = box.ITYP - 1
    kv, ind = orc.nlist("full", box.XP, it, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                        ctrl.NB_RM, ctrl.NB_MXNBS)
    ox, osteps, _, _ = orc.steepest(box.XP, box.STATU, it, box.BOXLOW, box.ZL, ctrl.IFPD, kv, ind,
                                    float(np.max(ctrl.RU)) ** 2, t, 200, 0, ctrl.STEEPEST_Alpha,
                                    ctrl.STEEPEST_MxStep * box.RR, ctrl.STEEPEST_MiStep * box.RR,
                                    1e-6)
    assert 0 < osteps <= 200
    assert f"steepest finished after steps: {osteps:8d}" in r.stdout
    dx = quench - ox
    dx -= np.round(dx / box.ZL) * box.ZL
    assert np.max(np.abs(dx)) <= 1e-9
    # copy-in of positions and velocities: the moved coordinate comes back as written, and the
    # zeroed velocities give 0 K
    ci = lines[5 + 3 * n].split()
    assert ci[:2] == ["#", "copy-in"]
    assert float(ci[2]) == quench[0, 0] + 0.01 and float(ci[3]) == 0.0


CGS_DRIVER = os.path.join(os.path.dirname(DRIVER), "cgs_driver")
EVERG = 1.60219e-12   # erg per eV (MSM_Const.F90:83)
KB_CGS = 1.38054e-16  # erg/K (MSM_Const.F90:82)


def _blocks(lines, n, start):
    return np.array([[float(v) for v in l.split()] for l in lines[start:start + n]])


def test_cgs_driver_against_oracle(orc, tmp_path):
    """The reference-style Fortran caller in CGS (cm, erg, g, s, K): its forces (erg/cm), energy
    (erg) and F' of CalEDen_ForceClass against the oracle in eV/A after unit conversion; the
    device temperature of Thermalizing_MC_DEV against the velocities' own; a registered external
    force added exactly; Predictor_DEV's damping and step search against the oracle's in CGS; and
    a restart from the binary records resuming bit for bit."""
    from msmpscu_amd.mdlib import CP_KB_CGS
    if not os.path.exists(CGS_DRIVER):
        pytest.fail("cgs_driver is not built (msmpscu_amd.build.build_fortran)")
    out = tmp_path / "cgs.txt"
    print("[gpu-test] cgs driver", flush=True)
    r = subprocess.run([CGS_DRIVER, str(out), "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = out.read_text().splitlines()
    n = int(lines[0].split()[0])
    e0 = float(lines[0].split()[1])
    blk = _blocks(lines, n, 1)
    x_cm, f_cgs, den_cgs = blk[:, :3], blk[:, 3:6], blk[:, 6]
    # oracle in eV / Angstrom on the same positions
    ang = 1e8
    nc = round((n / 2) ** (1 / 3))
    zl = np.full(3, nc * 3.165e-8) * ang
    box = SimMDBox(NPRT=n, NGROUP=1, ZL=zl, BOXLOW=-0.5 * zl, XP=np.asfortranarray(x_cm * ang),
                   XP1=np.zeros((n, 3), order="F"), ITYP=np.ones(n, dtype=np.int32),
                   STATU=np.ones(n, dtype=np.int32), CM=np.array([cfg.W_MASS * cfg.AMU]),
                   PTYPE=np.array([[1]], dtype=np.int32), SYMB=["W"])
    ctrl = cfg.w_ctrl()
    ft = cfg.w_forcetable(box, ctrl)
    t = orc.from_mdlib(ft)
    kv, ind = orc.nlist("full", box.XP, box.ITYP - 1, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                        ctrl.NB_RM, ctrl.NB_MXNBS)
    fo, do, eo = orc.force_full(box.XP, box.ITYP - 1, box.STATU, kv, ind, box.ZL, ctrl.IFPD,
                                float(np.max(ctrl.RU)) ** 2, t)
    assert np.max(np.abs(f_cgs / EVERG / ang - fo)) <= F_TOL
    assert abs(e0 / EVERG - eo.sum()) <= E_REL * abs(eo.sum())
    assert np.max(np.abs(den_cgs * EVERG - do)) <= 1e-13
    rest = lines[1 + n:]
    tl = [l for l in rest if l.startswith("# temperature")][0].split()
    tdev, thost = float(tl[2]), float(tl[3])
    assert abs(tdev - thost) <= 1e-9 * thost
    assert abs(thost / 600.0 - 1.0) < 5 * np.sqrt(2.0 / (3 * n))
    el = [l for l in rest if l.startswith("# extforce")][0].split()
    f0, maxd = float(el[2]), float(el[3])
    assert f0 > 0 and maxd <= 1e-12 * f0
    k = next(i for i, l in enumerate(rest) if l.startswith("# predictor-in"))
    pin = _blocks(rest, n, k + 1)
    h = float(rest[k + 1 + n].split()[3])
    pout = _blocks(rest, n, k + 2 + n)
    st = np.ones(n, dtype=np.int32)
    cm = np.array([183.84 * 1.66054e-24])
    vd = orc.damping(pin[:, :3], pin[:, 3:6], st)
    hs = orc.timestep_search(vd, pin[:, 3:6], st, np.zeros(n, dtype=np.int32), cm, 2.0e-15,
                             0.002e-8)
    assert h == hs and h < 2.0e-15
    ox, ov, _, _ = orc.predictor(x_cm, vd, pin[:, 3:6], np.zeros((n, 3)), st,
                                 np.zeros(n, dtype=np.int32), cm, h, -0.5 * zl / ang, zl / ang,
                                 (1, 1, 1))
    assert np.max(np.abs(pout[:, :3] - ox)) <= 1e-20
    assert np.max(np.abs(pout[:, 3:] - ov)) <= 1e-12 * np.max(np.abs(ov))
    rl = [l for l in rest if l.startswith("# restart")][0].split()
    assert [float(v) for v in rl[2:]] == [0.0] * 5
    assert CP_KB_CGS == KB_CGS


GLUE_DRIVER = os.path.join(os.path.dirname(CGS_DRIVER), "ref_glue_driver")


def test_reference_type_glue_against_oracle(orc, tmp_path):
    """A caller written against the reference's own type names (SimMDBox, SimMDCtrl, MDForceTable,
    the procedure-pointer MDForceClassGPU; CGS values), bound through the committed glue module
    and driven through the force-class pointers only: forces (dyn), energy (erg) and F' against
    the oracle in eV/A after unit conversion; the virial tensor (erg) against the oracle's
    half-list VTENSOR to 1e-14 of the sum of its absolute pair terms; the per-atom stress traces
    sum to twice the tensor's trace."""
    if not os.path.exists(GLUE_DRIVER):
        pytest.fail("ref_glue_driver is not built (msmpscu_amd.build.build_fortran)")
    out = tmp_path / "glue.txt"
    print("[gpu-test] reference-type glue driver", flush=True)
    r = subprocess.run([GLUE_DRIVER, str(out), "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = out.read_text().splitlines()
    n = int(lines[0].split()[0])
    e0 = float(lines[0].split()[1])
    blk = _blocks(lines, n, 1)
    x_cm, f_cgs, den_cgs, avp_tr = blk[:, :3], blk[:, 3:6], blk[:, 6], blk[:, 7]
    assert lines[1 + n].startswith("# vtensor")
    vt_cgs = np.array([float(v) for v in lines[2 + n].split()]).reshape(3, 3, order="F")
    ang = 1e8
    nc = round((n / 2) ** (1 / 3))
    zl = np.full(3, nc * 3.165e-8) * ang
    box = SimMDBox(NPRT=n, NGROUP=1, ZL=zl, BOXLOW=-0.5 * zl, XP=np.asfortranarray(x_cm * ang),
                   XP1=np.zeros((n, 3), order="F"), ITYP=np.ones(n, dtype=np.int32),
                   STATU=np.ones(n, dtype=np.int32), CM=np.array([cfg.W_MASS * cfg.AMU]),
                   PTYPE=np.array([[1]], dtype=np.int32), SYMB=["W"])
    ctrl = cfg.w_ctrl()
    ft = cfg.w_forcetable(box, ctrl)
    t = orc.from_mdlib(ft)
    it = box.ITYP - 1
    kv, ind = orc.nlist("full", box.XP, it, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                        ctrl.NB_RM, ctrl.NB_MXNBS)
    fo, do, eo = orc.force_full(box.XP, it, box.STATU, kv, ind, box.ZL, ctrl.IFPD,
                                float(np.max(ctrl.RU)) ** 2, t)
    assert np.max(np.abs(f_cgs / EVERG / ang - fo)) <= F_TOL
    assert abs(e0 / EVERG - eo.sum()) <= E_REL * abs(eo.sum())
    assert np.max(np.abs(den_cgs * EVERG - do)) <= 1e-13
    kvh, indh = orc.nlist("half", box.XP, it, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                          ctrl.NB_RM, ctrl.NB_MXNBS)
    _, _, _, _, vt, vabs = orc.force_fs_half(box.XP, it, kvh, indh, box.ZL, ctrl.IFPD, ctrl.RU,
                                             t, with_vabs=True)
    assert np.all(np.abs(vt_cgs / EVERG - vt) <= 1e-14 * vabs + 1e-300)
    assert abs(avp_tr.sum() - 2.0 * np.trace(vt_cgs)) <= 1e-13 * EVERG * np.trace(vabs)
