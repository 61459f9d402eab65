"""CPU tests of the product's host side: the C-ABI library loads and exports every declared
symbol, host table generation / RNG / Maxwell sampler agree with the oracle bit for bit, and the
MDLIB-style host mirror builds the reference's tables and raises where the reference stops."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import msmpscu_amd._lib as L
import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "msm_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(msm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_symbols()
    assert len(names) > 30
    for n in names:
        assert hasattr(lib, n), n
    # and every binding signature refers to a declared symbol
    assert set(L.SIGNATURES) == set(names)


def test_pair_table_bitwise_equals_oracle(orc):
    ctrl = cfg.w_ctrl(ntab=5000)
    box = cfg.make_bcc_box((3, 3, 3), 3.165, 0.0, 0.0, 1)
    ft = cfg.w_forcetable(box, ctrl)
    t = orc.Table(0, 1, 1, 5000, 4.4002, np.zeros((1, 1)))
    t.fill_pair(0, "orc_pot_wfs_nn", "orc_pot_wfs_ne")
    assert ft.CSI == t.csi
    assert np.array_equal(ft.POTR[0], t.potr[0, 1:-1])
    assert np.array_equal(ft.FPOTR[0], t.fpotr[0, 1:-1])
    assert np.array_equal(ft.POTB[0], t.potb[0, 1:-1])
    assert np.array_equal(ft.FPOTB[0], t.fpotb[0, 1:-1])
    assert ft.RHOMX == t.rhomx


def test_mdlib_table_matches_reference_export():
    """Register_ForceTableProc + Create_Interaction_ForceTable with the Ackland W-W form and the
    fixture's grid (NTAB=10000, RU=10 A) reproduce the reference's exported table."""
    rows = np.loadtxt(os.path.join(ROOT, "tests", "golden", "w_ackland_pair_rows.txt"))
    box = cfg.make_bcc_box((3, 3, 3), 3.165, 0.0, 0.0, 1)
    ctrl = md.SimMDCtrl(RU=[[10.0]], NB_RM=[[10.3]], NUMFTABR=10000)
    ft = md.MDForceTable()
    md.Register_ForceTableProc(1, ft, "msm_pot_ackland_ww_nn", "msm_pot_ackland_ww_ne")
    md.Create_Interaction_ForceTable(box, ctrl, ft)
    idx = rows[:, 0].astype(int) - 1
    for col, arr, fac in ((2, ft.POTR, 2.0), (3, ft.FPOTR, 1.0), (4, ft.POTB, 1.0),
                          (5, ft.FPOTB, 1.0)):
        ref = rows[:, col]
        got = arr[0, idx] * fac
        assert np.all(np.abs(got - ref) <= 2e-9 * np.abs(ref) + 1e-12 * np.abs(ref).max())


def test_python_callback_potential_equals_builtin():
    box = cfg.make_bcc_box((3, 3, 3), 3.165, 0.0, 0.0, 1)
    ctrl = cfg.w_ctrl(ntab=2000)
    ft1 = cfg.w_forcetable(box, ctrl)

    def nn(r):
        if r <= 3.25:
            rc = r - 3.25
            poly = 47.1346 + -33.7666 * r + 6.2542 * r * r
            return 0.5 * (rc * rc * poly), -(2.0 * rc * poly + rc * rc * (-33.7666 + 2.0 * 6.2542 * r))
        return 0.0, 0.0

    ft2 = md.MDForceTable()
    md.Register_ForceTableProc(1, ft2, nn, "msm_pot_wfs_ne")
    md.Create_Interaction_ForceTable(box, ctrl, ft2)
    assert np.array_equal(ft1.POTR, ft2.POTR) and np.array_equal(ft1.FPOTR, ft2.FPOTR)


def test_registration_errors_follow_reference():
    ft = md.MDForceTable()
    md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn")
    with pytest.raises(ValueError):
        md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn")
    tid = md.Register_ForceTableProc("Ni", ft, "msm_pot_tbsma_cuni_nn", Tgt="Cu")
    assert tid == 2 and md.SymbID2TypeID_ForceTableProc("NI", "CU", ft) == 2
    box = cfg.make_bcc_box((3, 3, 3), 3.165, 0.0, 0.0, 1)
    box.PTYPE[:] = 7
    with pytest.raises(ValueError):
        md.Create_Interaction_ForceTable(box, cfg.w_ctrl(), ft)


def test_drand32_and_maxwell_equal_oracle(orc):
    s = (C.c_int * 2)(777, 999)
    seq = [L.lib().msm_drand32(s) for _ in range(1000)]
    st = (777, 999)
    for v in seq:
        w, st = orc.drand32(st)
        assert v == w
    it = np.ones(500, np.int32)
    cm = np.array([cfg.W_MASS * cfg.AMU])
    v1 = cfg.maxwell_velocities(it, cm, 300.0, 42)
    seed = (42 % 2147483562 + 1, (42 * 7 + 12345) % 2147483398 + 1)
    v2 = orc.thermalize(it - 1, cm, 300.0, cfg.KB, seed)
    assert np.array_equal(np.asarray(v1), v2)


def test_table_file_roundtrip(tmp_path):
    box = cfg.make_bcc_box((3, 3, 3), 3.165, 0.0, 0.0, 1)
    ctrl = cfg.w_ctrl(ntab=500)
    ft = cfg.w_forcetable(box, ctrl)
    md.Export_ForceTable(str(tmp_path / "w"), ft)
    head, rows = md.read_table_file(str(tmp_path / "w.pair"))
    assert head["&NUMPOINT"] == ["500"]
    assert np.allclose(rows[:, 1], 2 * ft.POTR[0], rtol=1e-9, atol=1e-300)
    assert np.allclose(rows[:, 4], ft.FPOTB[0], rtol=1e-9, atol=1e-300)
    head, rows = md.read_table_file(str(tmp_path / "w.embd"))
    assert np.allclose(rows[1:, 1], -np.sqrt(rows[1:, 0]), rtol=1e-8)


def test_workload_builders():
    b = cfg.make_bcc_box((8, 8, 8), 3.165, 0.05, 600.0, 3, vacancy_frac=0.005, n_dumbbell=2)
    assert b.NPRT == 1024 - 5 + 2
    assert np.all(b.XP >= b.BOXLOW) and np.all(b.XP < b.BOXUP)
    p = (b.CM[0] * np.asarray(b.XP1)).sum(axis=0)
    assert np.max(np.abs(p)) < 1e-9
    s = cfg.make_cuni_slab((4, 4, 4), 3.57, 300.0, 1)
    assert s.NPRT == 256 and (s.ITYP == 1).sum() == 128
    assert np.all(np.diff(s.ITYP) >= 0)
    sizes = cfg.batch_sizes(256, 2000, 60000, 5)
    assert min(sizes) >= 2000 and max(sizes) <= 60000


def test_fortran_bindings_name_header_symbols():
    """Every bind(C, name=...) of the Fortran drop-in layer names a function that
    include/msm_hip.h declares (the Fortran interfaces bind nothing the library lacks)."""
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "msmpscu_amd", "fortran", "MD_ForceLib_HIP.F90")).read()
    hdr = open(os.path.join(root, "include", "msm_hip.h")).read()
    names = set(re.findall(r'bind\(C,\s*name="(\w+)"\)', src))
    assert len(names) >= 20
    declared = set(re.findall(r"\b(msm_\w+)\s*\(", hdr))
    assert names <= declared, sorted(names - declared)


def test_raising_potential_callback_is_reported():
    """An exception inside a Python potential callback cannot pass through the C table routine;
    it comes back as an error of Create_Pairwise_ForceTable / Create_EMBDFUNTable instead of a
    table silently holding zeros or stale entries."""
    box = cfg.make_bcc_box((3, 3, 3), 3.165, 0.0, 0.0, 1)
    ctrl = md.SimMDCtrl(RU=[[4.4002]], NB_RM=[[4.7002]], NUMFTABR=200)

    def bad_nn(r):
        if r > 2.0:
            raise ZeroDivisionError("planted")
        return 0.0, 0.0

    ft = md.MDForceTable()
    md.Register_ForceTableProc(1, ft, bad_nn, "msm_pot_wfs_ne")
    with pytest.raises(RuntimeError, match="ZeroDivisionError: planted"):
        md.Create_Interaction_ForceTable(box, ctrl, ft)

    def bad_embd(rho):
        raise ValueError("planted embedding")

    ft = md.MDForceTable(PotType="EAM_TYPE")
    md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne", EMBDF=bad_embd)
    with pytest.raises(RuntimeError, match="ValueError: planted embedding"):
        md.Create_Interaction_ForceTable(box, ctrl, ft)
    # a well-behaved callback still fills the table
    ft = md.MDForceTable()
    md.Register_ForceTableProc(1, ft, lambda r: (0.5 * r, -0.5), "msm_pot_wfs_ne")
    md.Create_Interaction_ForceTable(box, ctrl, ft)
    assert np.all(np.isfinite(ft.POTR)) and ft.POTR[0, -1] != 0.0
