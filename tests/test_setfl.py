"""setfl potential files (msmpscu_amd/setfl.py): the natural cubic spline against scipy's, the
reader and table ids on synthetic files written here from closed forms (one and two elements,
INF/NAN markers), the closed-form lattice energy through the oracle's EAM path, the export /
import chain, and the setfl file shipped with the reference (Mendelev Cu1) when it is present.
CPU only."""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md
from msmpscu_amd import nist as setfl
from test_oracle_pins import w_closed

CU1 = "/root/reference/examples/NIST_Potentials/Cu_EAM/Cu1.eam.fs.setfl"


def test_spline_is_the_natural_cubic_spline():
    x = np.linspace(0.5, 4.0, 57)
    f = np.sin(2.0 * x) * np.exp(-0.3 * x) + 0.1 * x ** 2
    sp = setfl._Spline(x[0], x[1] - x[0], f)
    cs = CubicSpline(x, f, bc_type="natural")
    y = np.random.default_rng(3).uniform(x[0], x[-1], 200)
    v = np.array([sp(t) for t in y])
    assert np.max(np.abs(v[:, 0] - cs(y))) <= 1e-13
    assert np.max(np.abs(v[:, 1] - cs(y, 1))) <= 1e-12
    # interpolates the samples, zero curvature at the ends
    assert all(abs(sp(t)[0] - ft) <= 1e-14 for t, ft in zip(x, f))
    assert abs(cs(x[0], 2)) < 1e-12 and abs(cs(x[-1], 2)) < 1e-12


def _w_pair(r):
    v, _ = w_closed(np.asarray(r, dtype=np.float64))
    return v


def _w_rho(r):
    _, rho = w_closed(np.asarray(r, dtype=np.float64))
    return rho


def _write_setfl(path, elements, nrho, drho, nr, cutoff, frho, rhor, zr, markers=False):
    """frho[i](rho), rhor[i](r), zr[(i, j)](r) for i >= j; five values a line."""
    r = np.arange(nr) * (cutoff / nr)
    rho = np.arange(nrho) * drho

    def block(vals):
        vals = list(vals)
        return "\n".join(" ".join(f"{v:.16e}" for v in vals[k:k + 5])
                         for k in range(0, len(vals), 5)) + "\n"

    with open(path, "w") as fh:
        fh.write("synthetic setfl\nwritten by tests/test_setfl.py\nfor the reader tests\n")
        fh.write(f"{len(elements)} " + " ".join(elements) + "\n")
        fh.write(f"{nrho} {drho:.16e} {nr} {cutoff / nr:.16e} {cutoff:.16e}\n")
        for i, el in enumerate(elements):
            fh.write(f"{74 + i} 183.84 3.165 bcc\n")
            fh.write(block(frho[i](rho)))
            vals = rhor[i](r)
            if markers:
                fh.write("INF " + " ".join(f"{v:.16e}" for v in vals[1:4]) + "\n")
                fh.write(block(vals[4:]))
            else:
                fh.write(block(vals))
        for i in range(len(elements)):
            for j in range(i + 1):
                vals = zr[(i, j)](r)
                if markers:
                    fh.write("NAN " + " ".join(f"{v:.16e}" for v in vals[1:3]) + "\n")
                    fh.write(block(vals[3:]))
                else:
                    fh.write(block(vals))


def _w_setfl(path, markers=False, nr=4000, nrho=4000):
    _write_setfl(path, ["W"], nrho, 400.0 / nrho, nr, 4.8,
                 [lambda rho: -np.sqrt(rho)], [_w_rho], {(0, 0): lambda r: r * _w_pair(r)},
                 markers=markers)


def _bcc_energy(orc, ft, a, n=4):
    x = cfg.bcc_sites(n, n, n, a)
    ln = np.full(3, n * a)
    lo = -0.5 * ln
    x = x + lo
    npart = len(x)
    it = np.zeros(npart, np.int32)
    st = np.ones(npart, np.int32)
    t = orc.from_mdlib(ft)
    kv, ind = orc.nlist("full", x, it, st, lo, ln, (1, 1, 1), np.full((1, 1), 4.7002), 96)
    f, _, e = orc.force_full(x, it, st, kv, ind, ln, (1, 1, 1), 4.4002 ** 2, t)
    return f, e, npart


@pytest.mark.parametrize("markers", [False, True])
def test_w_setfl_tables_and_lattice_energy(orc, tmp_path, markers):
    """A setfl file of the W Finnis-Sinclair closed forms (r*phi, rho, F = -sqrt(rho)) gives
    tables equal to the closed forms to the spline's accuracy, the file's Rmax and RHOMX, and the
    closed-form perfect-lattice energy through the oracle's EAM path; INF/NAN-opened lists read
    the same as plain ones away from r = 0."""
    path = str(tmp_path / "w.setfl")
    _w_setfl(path, markers)
    ft = md.MDForceTable(PotType="EAM_TYPE", NTAB=5000, NEMBD=5000)
    assert setfl.Register_ForceTableProc_Setfl(path, ft) == ["W"]
    setfl.Generate_NIST_ForceTable(ft)
    assert ft.Rmax == 4.8 and ft.RHOMX == pytest.approx(400.0) and ft.RHOD == pytest.approx(0.08)
    r = (np.arange(1, ft.NTAB + 1) * ft.CSIV) ** 2
    # away from the kinks of the closed forms (phi'' jumps at 3.25 A, rho'' at 4.4002 A) the
    # spline of 4000 samples is exact to rounding for these polynomials
    m = (r > 1.0) & (r < 4.7) & (np.abs(r - 3.25) > 0.01) & (np.abs(r - 4.4002) > 0.01)
    assert np.allclose(ft.POTR[0, m], 0.5 * _w_pair(r[m]) * r[m], rtol=0, atol=1e-9)  # V r / 2
    assert np.allclose(ft.POTB[0, m], _w_rho(r[m]), rtol=0, atol=1e-9)
    assert np.allclose(ft.POTB[0, r > 1.0], _w_rho(r[r > 1.0]), rtol=0, atol=1e-5)
    rho = np.arange(ft.NEMBD) * ft.RHOD
    last = 3999 * 0.1  # the file's last sample; beyond it F holds its value with zero slope
    k = (rho > 2.0) & (rho <= last)
    assert np.allclose(ft.FEMBD[0, k], -np.sqrt(rho[k]), rtol=0, atol=1e-6)
    assert np.allclose(ft.DFEMBD[0, k], -0.5 / np.sqrt(rho[k]), rtol=0, atol=1e-5)
    beyond = rho > last
    assert np.all(ft.FEMBD[0, beyond] == ft.FEMBD[0, beyond][0]) and np.all(ft.DFEMBD[0, beyond] == 0)
    f, e, n = _bcc_energy(orc, ft, 3.165)
    shells = [(3.165 * math.sqrt(3) / 2, 8), (3.165, 6), (3.165 * math.sqrt(2), 12)]
    v, rh = w_closed(np.array([s[0] for s in shells]))
    nsh = np.array([s[1] for s in shells])
    e_atom = 0.5 * np.sum(nsh * v) - math.sqrt(np.sum(nsh * rh))
    assert e.sum() / n == pytest.approx(e_atom, rel=1e-6)
    assert np.max(np.abs(f)) < 1e-9


def test_two_element_table_ids(tmp_path):
    """IT = (I-1) NE + J: pair term of the unordered pair, density of element J (zero when F of
    element I is zero everywhere), embedding only for I = J; KPAIR/KEMBD map group I to element
    I."""
    path = str(tmp_path / "ab.setfl")
    fa, fb = (lambda rho: -np.sqrt(rho)), (lambda rho: 0.0 * rho)
    ra, rb = (lambda r: np.exp(-r)), (lambda r: 2.0 * np.exp(-1.5 * r))
    z = {(0, 0): lambda r: r * np.exp(-2.0 * r), (1, 0): lambda r: 3.0 * r * np.exp(-2.0 * r),
         (1, 1): lambda r: 5.0 * r * np.exp(-2.0 * r)}
    _write_setfl(path, ["A", "B"], 1000, 0.01, 2000, 5.0, [fa, fb], [ra, rb], z)
    ft = md.MDForceTable(PotType="EAM_TYPE", NTAB=3000, NEMBD=1000)
    assert setfl.Register_ForceTableProc_Setfl(path, ft) == ["A", "B"]
    assert sorted(ft.REGPROCS) == [1, 2, 3, 4]
    assert ft.REGPROCS[2].NOTE.startswith("setfl table: A<-B")
    setfl.Generate_NIST_ForceTable(ft)
    r = (np.arange(1, ft.NTAB + 1) * ft.CSIV) ** 2
    m = (r > 0.5) & (r < 4.9)
    ph = np.exp(-2.0 * r[m])
    for it, scale in ((1, 1.0), (2, 3.0), (3, 3.0), (4, 5.0)):
        assert np.allclose(ft.POTR[it - 1, m], 0.5 * scale * ph * r[m], atol=1e-8)
    assert np.allclose(ft.POTB[0, m], np.exp(-r[m]), atol=1e-8)             # A<-A: rho_A
    assert np.allclose(ft.POTB[1, m], 2.0 * np.exp(-1.5 * r[m]), atol=1e-8)  # A<-B: rho_B
    assert np.all(ft.POTB[2] == 0.0) and np.all(ft.POTB[3] == 0.0)         # F_B == 0
    rho = np.arange(ft.NEMBD) * ft.RHOD
    k = rho > 0.5
    assert np.allclose(ft.FEMBD[0, k], -np.sqrt(rho[k]), atol=1e-6)
    assert np.all(ft.FEMBD[1] == 0.0) and np.all(ft.FEMBD[2] == 0.0)
    assert ft.KPAIR.tolist() == [[1, 2], [3, 4]] and ft.KEMBD.tolist() == [1, 4]


def test_export_import_chain(orc, tmp_path):
    """The reference's chain: setfl -> generated tables -> .pair/.embd -> imported onto the
    simulation grid; the lattice energy survives to the files' nine digits."""
    path = str(tmp_path / "w.setfl")
    _w_setfl(path)
    ft = md.MDForceTable(PotType="EAM_TYPE", NTAB=5000, NEMBD=5000)
    setfl.Register_ForceTableProc_Setfl(path, ft)
    setfl.Generate_NIST_ForceTable(ft)
    lib = str(tmp_path / "w_lib")
    md.Export_ForceTable(lib, ft)
    box = cfg.make_bcc_box((4, 4, 4), 3.165, 0.0, 0.0, 1)
    ctrl = cfg.w_ctrl(ntab=5000, nembd=5000)
    ft2 = md.MDForceTable()
    md.Register_Imported_ForceTable(box, ctrl, ft2, lib)
    _, e1, n = _bcc_energy(orc, ft, 3.165)
    _, e2, _ = _bcc_energy(orc, ft2, 3.165)
    assert e2.sum() == pytest.approx(e1.sum(), rel=1e-7)


@pytest.mark.skipif(not os.path.exists(CU1), reason="reference example files not present")
def test_reference_cu_setfl_lattice():
    """Mendelev Cu1 (examples/NIST_Potentials/Cu_EAM/Cu1.eam.fs.setfl, read in place): the file's
    FCC lattice constant 3.639087 A minimises the lattice energy (to 2e-3 A), and the energy per
    atom through the tables and the oracle's EAM path equals the direct sum over the file's
    samples (scipy's natural spline, phi = (r phi)/r, F of the summed density) to the tables'
    interpolation error."""
    import oracle as orc
    ft = md.MDForceTable(PotType="EAM_TYPE", NTAB=10000, NEMBD=10000)
    assert setfl.Register_ForceTableProc_Setfl(CU1, ft) == ["Cu"]
    setfl.Generate_NIST_ForceTable(ft)
    t = orc.from_mdlib(ft)

    def energy(a, n=4):
        x = cfg.fcc_sites(n, n, n, a)
        ln = np.full(3, n * a)
        lo = -0.5 * ln
        x = x + lo
        it = np.zeros(len(x), np.int32)
        st = np.ones(len(x), np.int32)
        kv, ind = orc.nlist("full", x, it, st, lo, ln, (1, 1, 1), np.full((1, 1), 6.0), 200)
        _, _, e = orc.force_full(x, it, st, kv, ind, ln, (1, 1, 1), 6.0 ** 2, t)
        return e.sum() / len(x)

    a = np.linspace(3.60, 3.68, 9)
    e = np.array([energy(v) for v in a])
    c = np.polyfit(a, e, 2)
    assert -c[1] / (2.0 * c[0]) == pytest.approx(3.639087, abs=2e-3)
    d = setfl.read_setfl(CU1)
    rg = np.arange(d.nr) * d.rstep
    z = CubicSpline(rg, d.zr[(0, 0)], bc_type="natural")
    dn = CubicSpline(rg, d.rhor[0], bc_type="natural")
    fe = CubicSpline(np.arange(d.nrho) * d.drho, d.frho[0], bc_type="natural")
    a0 = 3.639087
    x = cfg.fcc_sites(6, 6, 6, a0)
    s = x - x[0]
    s -= np.round(s / (6 * a0)) * (6 * a0)
    r = np.sqrt(np.sum(s * s, axis=1))
    r = r[(r > 0.0) & (r <= rg[-1])]
    e_direct = 0.5 * np.sum(z(r) / r) + fe(np.sum(dn(r)))
    assert energy(a0) == pytest.approx(e_direct, rel=2e-4)


def test_moldy_file(orc, tmp_path):
    """A moldy file (lattice-unit knots and coefficients) whose knots reproduce the W closed forms
    -- rho = A^2 (d - r)^2 is not a cubic, so the test builds a density from two cubic knots and
    checks it in closed form -- gives procedures equal to the closed forms and, through the
    tables, the closed-form lattice energy."""
    a0 = 3.165
    # pair: 0.8 (3.3 - r)^3 - 0.1 (4.1 - r)^3; density: 0.05 (4.4 - r)^3 + 0.02 (3.6 - r)^3
    av, rv = np.array([0.8, -0.1]), np.array([3.3, 4.1])
    ar, rr = np.array([0.05, 0.02]), np.array([4.4, 3.6])
    path = tmp_path / "w.moldy"
    path.write_text("W\n" + " ".join(f"{v * a0 ** 3:.17e}" for v in av) + "\n"
                    + " ".join(f"{v / a0:.17e}" for v in rv) + "\n"
                    + " ".join(f"{v * a0 ** 6:.17e}" for v in ar) + "\n"
                    + " ".join(f"{v / a0:.17e}" for v in rr) + "\n"
                    + f"{a0} 183.84\n")
    ft = md.MDForceTable(PotType="EAM_TYPE", NTAB=8000, NEMBD=8000)
    assert setfl.Register_ForceTableProc_Moldy(str(path), ft) == "W"

    def pair(r):
        return sum(a * np.clip(k - r, 0, None) ** 3 for a, k in zip(av, rv))

    def dens(r):
        return sum(a * np.clip(k - r, 0, None) ** 3 for a, k in zip(ar, rr))

    nn, ne, em = (ft.REGPROCS[1].NNFORCE, ft.REGPROCS[1].NEFORCE, ft.REGPROCS[1].EMBDF)
    for r in (2.2, 2.74, 3.0, 3.9, 4.3, 4.5):
        v, mdv = nn(r)
        assert v == pytest.approx(0.5 * pair(r), rel=1e-12, abs=1e-15)
        h = 1e-6
        assert mdv == pytest.approx(-(pair(r + h) - pair(r - h)) / (2 * h), rel=1e-6, abs=1e-9)
        b, mdb = ne(r)
        assert b == pytest.approx(dens(r), rel=1e-12, abs=1e-15)
        assert mdb == pytest.approx(-(dens(r + h) - dens(r - h)) / (2 * h), rel=1e-6, abs=1e-9)
    assert em(4.0) == (-2.0, -0.25) and em(0.0) == (0.0, 0.0)
    box = cfg.make_bcc_box((4, 4, 4), a0, 0.0, 0.0, 1)
    ctrl = cfg.w_ctrl(ntab=8000, nembd=8000, rc=4.5)
    md.Create_Interaction_ForceTable(box, ctrl, ft)
    t = orc.from_mdlib(ft)
    x = box.XP
    it = np.zeros(box.NPRT, np.int32)
    kv, ind = orc.nlist("full", x, it, box.STATU, box.BOXLOW, box.ZL, (1, 1, 1),
                        np.full((1, 1), 4.8), 96)
    f, _, e = orc.force_full(x, it, box.STATU, kv, ind, box.ZL, (1, 1, 1), 4.5 ** 2, t)
    shells = np.array([a0 * math.sqrt(3) / 2, a0, a0 * math.sqrt(2), a0 * math.sqrt(11) / 2])
    nsh = np.array([8, 6, 12, 24])
    e_atom = 0.5 * np.sum(nsh * pair(shells)) - math.sqrt(np.sum(nsh * dens(shells)))
    assert e.sum() / box.NPRT == pytest.approx(e_atom, rel=2e-6)  # table lerp error
    assert np.max(np.abs(f)) < 1e-9


LSPT = "/root/reference/examples/NIST_Potentials/WHHe_EAM1_LSPT/WHHe_EAM1.lspt"


@pytest.mark.skipif(not os.path.exists(LSPT), reason="reference example files not present")
def test_reference_lspt_matches_its_exported_embedding_table():
    """The W-H-He lspt potential shipped with the reference (read in place) against the
    embedding table the reference exported from it (WHHe_EAM1.lspt.embd, 10000 points, nine
    digits): same rho grid (RHOMX = 10 from the F file), F of W to the printed digits, dF as the
    reference writes it (still in erg, i.e. times 1.60219e-12), zero tables for H and He (no F
    files); and the reader's quirk that every pair of a control-file line reads the line's first
    file (H-H holds the W-H samples)."""
    ft = md.MDForceTable(PotType="EAM_TYPE", NTAB=10000, NEMBD=10000)
    assert setfl.Register_ForceTableProc_SPT(LSPT, ft) == ["W", "H", "He"]
    assert ft.RHOMX == 10.0 and ft.Rmax == pytest.approx(5.4604375)
    setfl.Generate_NIST_ForceTable(ft)
    rows = []
    with open(LSPT + ".embd") as fh:
        for line in fh:
            t = line.split()
            if len(t) == 20 and t[0].isdigit():
                rows.append([float(v) for v in t])
    rows = np.array(rows)
    assert rows.shape == (10000, 20)
    assert np.max(np.abs(rows[:, 1] - np.arange(10000) * ft.RHOD)) < 1e-12
    f, df = rows[:, 2], rows[:, 3]
    assert np.all(np.abs(ft.FEMBD[0] - f) <= 5.01e-9 * np.maximum(np.abs(f), 1.0))
    big = np.abs(df) > 1e-14
    assert np.all(np.abs(ft.DFEMBD[0, big] * md.CP_EVERG - df[big]) <= 1e-8 * np.abs(df[big]))
    assert np.all(rows[:, 4:] == 0.0) and np.all(ft.FEMBD[1:] == 0.0)
    for r in (1.0, 2.0, 3.0):
        assert ft.REGPROCS[5].NNFORCE(r) == ft.REGPROCS[2].NNFORCE(r)
