"""ctypes front end of the CPU oracle (oracle/liborcmd.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, always as the checker.  The product package (msmpscu_amd) never imports it.

Every function takes and returns numpy arrays in the oracle's conventions (0-based groups and
neighbour ids, XP(N,3) in Fortran order, padded kind-major tables); see orc_md.h.

Parity status of each restated function (how it is pinned):
  orc_table_pairwise   pinned: reference-exported table fixture (tests/golden), closed forms
  orc_table_embd       pinned: reference-exported .embd fixture (grid and F column)
  orc_nlist_half/full  pinned: brute-force O(N^2) list, half <-> full symmetry, lattice shells
  orc_force_fs_half    pinned: perfect-lattice F = 0, sum F = 0, F = -dE/dx (central FD),
                       analytic lattice energy, full-list (device-kernel) restatement agreement
  orc_force_full       pinned: as above, plus FS == EAM with F = -sqrt(rho) tabulated
  orc_atomic_stress_full pinned: sum over atoms = 2 x the half-list VTENSOR of orc_force_fs_half
                       (independent loop), equal tensors on every atom of a perfect lattice with
                       cubic symmetry, zero on inactive atoms, per-atom symmetry
  orc_adjoint_full     pinned: central finite differences of orc_force_full (FS and EAM, every
                       table, inactive atoms), linearity of the pair-table gradients
  orc_predictor/corr.  pinned: closed-form free flight and constant-force motion, NVE drift
  orc_damping/ekin/velscale  pinned: component-wise closed forms (v.F < 0 and FIXPOS zeroed,
                       EK = m v^2 / 2), scaled velocities reach the target temperature exactly
                       and leave inactive atoms, FIXVEL/FIXPOS components as specified
  orc_drand32          pinned: exact modular arithmetic of L'Ecuyer's generator
  orc_thermalize_mc    pinned: zero net momentum, temperature statistics
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborcmd.so")

STATU_ACTIVE = 1
STATU_OUTOFBOX = 65536

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
PAIRFN = C.CFUNCTYPE(None, _dp, _dp, _dp)


class _FTable(C.Structure):
    _fields_ = [("pot_type", C.c_int), ("ngroup", C.c_int), ("nkind", C.c_int), ("ntab", C.c_int),
                ("csi", C.c_double), ("kpair", _ip),
                ("potr", _dp), ("fpotr", _dp), ("potb", _dp), ("fpotb", _dp),
                ("nkind1", C.c_int), ("nembd", C.c_int), ("rhod", C.c_double), ("kembd", _ip),
                ("fembd", _dp), ("dfembd", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_drand32.restype = C.c_double
        L.orc_sum.restype = C.c_double
        L.orc_sum.argtypes = [_dp, C.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def fn(name: str):
    """C function pointer of an oracle potential (e.g. "orc_pot_wfs_nn")."""
    return C.cast(getattr(lib(), name), C.c_void_p)


class Table:
    """Padded kind-major oracle table: arrays of shape (nkind, ntab+2), pads zero."""

    def __init__(self, pot_type: int, ngroup: int, nkind: int, ntab: int, rmax: float,
                 kpair: np.ndarray, nkind1: int = 0, nembd: int = 0):
        self.pot_type = pot_type
        self.ngroup = ngroup
        self.nkind = nkind
        self.ntab = ntab
        self.rmax = rmax
        self.csi = float(ntab) / np.sqrt(rmax)
        self.kpair = np.ascontiguousarray(kpair, dtype=np.int32)  # [ti, tj] 0-based kinds
        self.potr = np.zeros((nkind, ntab + 2))
        self.fpotr = np.zeros_like(self.potr)
        self.potb = np.zeros_like(self.potr)
        self.fpotb = np.zeros_like(self.potr)
        self.nkind1 = nkind1
        self.nembd = nembd
        self.rhod = 1.0
        self.rhomx = -1.0
        self.kembd = np.zeros(max(ngroup, 1), dtype=np.int32)
        self.fembd = np.zeros((max(nkind1, 1), nembd + 2))
        self.dfembd = np.zeros_like(self.fembd)

    def fill_pair(self, k: int, nn: str, ne: Optional[str], rhomult: float = 20.0):
        rh = C.c_double(self.rhomx)
        lib().orc_table_pairwise(C.c_int(self.ntab), C.c_double(self.rmax), fn(nn),
                                 fn(ne) if ne else None, C.c_double(rhomult), C.byref(rh),
                                 _d(self.potr[k]), _d(self.fpotr[k]), _d(self.potb[k]),
                                 _d(self.fpotb[k]))
        self.rhomx = rh.value

    def fill_embd(self, k: int, emb: str, rhod: float):
        self.rhod = rhod
        lib().orc_table_embd(C.c_int(self.nembd), C.c_double(rhod), fn(emb), _d(self.fembd[k]),
                             _d(self.dfembd[k]))

    def struct(self) -> _FTable:
        s = _FTable()
        s.pot_type = self.pot_type
        s.ngroup = self.ngroup
        s.nkind = self.nkind
        s.ntab = self.ntab
        s.csi = self.csi
        s.kpair = _i(self.kpair)
        s.potr, s.fpotr, s.potb, s.fpotb = (_d(self.potr), _d(self.fpotr), _d(self.potb),
                                            _d(self.fpotb))
        s.nkind1 = self.nkind1
        s.nembd = self.nembd
        s.rhod = self.rhod
        s.kembd = _i(self.kembd)
        s.fembd, s.dfembd = _d(self.fembd), _d(self.dfembd)
        self._keep = s
        return s


def from_mdlib(ft) -> Table:
    """Oracle table holding the same numbers as a product MDForceTable (1-based KPAIR/KEMBD)."""
    ng = ft.KPAIR.shape[0]
    pt = 1 if ft.PotType.strip() == "EAM_TYPE" else 0
    t = Table(pt, ng, ft.NKIND, ft.NTAB, ft.Rmax, np.asarray(ft.KPAIR) - 1,
              ft.NKIND1 if pt else 0, ft.NEMBD if pt else 0)
    t.csi = ft.CSI
    for k in range(ft.NKIND):
        t.potr[k, 1:-1] = ft.POTR[k]
        t.fpotr[k, 1:-1] = ft.FPOTR[k]
        t.potb[k, 1:-1] = ft.POTB[k]
        t.fpotb[k, 1:-1] = ft.FPOTB[k]
    if pt:
        t.rhod = ft.RHOD
        t.kembd = np.ascontiguousarray(np.asarray(ft.KEMBD) - 1, dtype=np.int32)
        for k in range(ft.NKIND1):
            t.fembd[k, 1:-1] = ft.FEMBD[k]
            t.dfembd[k, 1:-1] = ft.DFEMBD[k]
    return t


def _xp(x):
    """(N, 3) -> component-major flat copy (never a view: several oracle routines update it in
    place, and the caller's array must stay as it was)."""
    return np.asarray(x, dtype=np.float64).flatten(order="F")


def _v3(a, dtype=np.float64):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(3))


def nlist(kind: str, xp, ityp0, statu, lo, ln, ifpd, nb_rm, mx):
    """kind in {"half", "full", "brute"}; returns (kvois, indi[N, mx]) with 0-based ids."""
    n = int(np.asarray(ityp0).shape[0])
    x = _xp(xp)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    nb = np.ascontiguousarray(np.asarray(nb_rm, dtype=np.float64))
    ng = int(round(np.sqrt(nb.size)))
    kv = np.zeros(n, dtype=np.int32)
    ind = np.zeros(n * mx, dtype=np.int32)
    if kind == "brute":
        rc = lib().orc_nlist_brute(n, _d(x), _i(it), _i(st), _d(_v3(ln)), _i(_v3(ifpd, np.int32)),
                                   ng, _d(nb), mx, _i(kv), _i(ind))
    else:
        f = lib().orc_nlist_half if kind == "half" else lib().orc_nlist_full
        rc = f(n, _d(x), _i(it), _i(st), _d(_v3(lo)), _d(_v3(ln)), _i(_v3(ifpd, np.int32)), ng,
               _d(nb), mx, _i(kv), _i(ind))
    if rc != 0:
        raise RuntimeError(f"oracle nlist({kind}) error {rc}")
    return kv, ind.reshape(mx, n).T


def force_fs_half(xp, ityp0, kvois, indi, ln, ifpd, ru, table: Table, with_vabs=False):
    n = int(np.asarray(ityp0).shape[0])
    mx = indi.shape[1]
    x = _xp(xp)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    kv = np.ascontiguousarray(kvois, dtype=np.int32)
    ind = np.ascontiguousarray(np.asarray(indi, dtype=np.int32).T.ravel())
    r = np.ascontiguousarray(np.asarray(ru, dtype=np.float64))
    fp = np.zeros(3 * n)
    ep = np.zeros(n)
    den = np.zeros(n)
    dden = np.zeros(n)
    vt = np.zeros(9)
    va = np.zeros(9)
    rc = lib().orc_force_fs_half(n, _d(x), _i(it), _i(kv), _i(ind), mx, _d(_v3(ln)),
                                 _i(_v3(ifpd, np.int32)), _d(r), C.byref(table.struct()), _d(fp),
                                 _d(ep), _d(den), _d(dden), _d(vt), _d(va) if with_vabs else None)
    if rc != 0:
        raise RuntimeError(f"oracle force error {rc}")
    out = (fp.reshape(3, n).T.copy(), ep, den, dden, vt.reshape(3, 3).T.copy())
    return out + (va.reshape(3, 3).T.copy(),) if with_vabs else out


def force_full(xp, ityp0, statu, kvois, indi, ln, ifpd, ru2max, table: Table):
    n = int(np.asarray(ityp0).shape[0])
    mx = indi.shape[1]
    x = _xp(xp)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    kv = np.ascontiguousarray(kvois, dtype=np.int32)
    ind = np.ascontiguousarray(np.asarray(indi, dtype=np.int32).T.ravel())
    fp = np.zeros(3 * n)
    dden = np.zeros(n)
    ep = np.zeros(n)
    rc = lib().orc_force_full(n, _d(x), _i(it), _i(st), _i(kv), _i(ind), mx, _d(_v3(ln)),
                              _i(_v3(ifpd, np.int32)), C.c_double(ru2max),
                              C.byref(table.struct()), _d(fp), _d(dden), _d(ep))
    if rc != 0:
        raise RuntimeError(f"oracle force_full error {rc}")
    return fp.reshape(3, n).T.copy(), dden, ep


def atomic_stress_full(xp, ityp0, statu, kvois, indi, ln, ifpd, ru2max, table: Table, dden,
                       with_abs=False):
    """Per-atom stress AP (N, 9) of CAL_FS/EAM_AtomicStress_KERNEL given F' (dden from
    force_full); with_abs also returns the per-component sums of |term|."""
    n = int(np.asarray(ityp0).shape[0])
    x = _xp(xp)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    kv = np.ascontiguousarray(kvois, dtype=np.int32)
    ind = np.ascontiguousarray(np.asarray(indi, dtype=np.int32).T.ravel())
    dd = np.ascontiguousarray(dden, dtype=np.float64)
    ap = np.zeros(9 * n)
    aa = np.zeros(9 * n)
    rc = lib().orc_atomic_stress_full(n, _d(x), _i(it), _i(st), _i(kv), _i(ind), _d(_v3(ln)),
                                      _i(_v3(ifpd, np.int32)), C.c_double(ru2max),
                                      C.byref(table.struct()), _d(dd), _d(ap), _d(aa))
    if rc != 0:
        raise RuntimeError(f"oracle atomic_stress_full error {rc}")
    ap = ap.reshape(9, n).T.copy()
    return (ap, aa.reshape(9, n).T.copy()) if with_abs else ap


def adjoint_full(xp, ityp0, statu, kvois, indi, ln, ifpd, ru2max, table: Table, dldf, dlde,
                 grads=None):
    """Reverse mode of force_full: adds dL/d(knot) for G = dL/dFP (N,3) and g = dL/dE (E = sum
    of EPOT) into grads (dict of padded kind-major tables, the shapes of the table's arrays;
    created when None)."""
    n = int(np.asarray(ityp0).shape[0])
    if grads is None:
        grads = {k: np.zeros_like(getattr(table, k)) for k in ("potr", "fpotr", "potb", "fpotb")}
        grads.update({k: np.zeros_like(getattr(table, k)) for k in ("fembd", "dfembd")})
    x = _xp(xp)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    kv = np.ascontiguousarray(kvois, dtype=np.int32)
    ind = np.ascontiguousarray(np.asarray(indi, dtype=np.int32).T.ravel())
    g = _xp(dldf)
    emb = table.pot_type != 0
    rc = lib().orc_adjoint_full(n, _d(x), _i(it), _i(st), _i(kv), _i(ind), _d(_v3(ln)),
                                _i(_v3(ifpd, np.int32)), C.c_double(ru2max),
                                C.byref(table.struct()), _d(g), C.c_double(dlde),
                                _d(grads["potr"]), _d(grads["fpotr"]), _d(grads["potb"]),
                                _d(grads["fpotb"]), _d(grads["fembd"]) if emb else None,
                                _d(grads["dfembd"]) if emb else None)
    if rc != 0:
        raise RuntimeError(f"oracle adjoint error {rc}")
    return grads


def predictor(xp, xp1, fp, dis, statu, ityp0, cm, h, lo, ln, ifpd):
    n = int(np.asarray(ityp0).shape[0])
    x, v, f, d = _xp(xp), _xp(xp1), _xp(fp), _xp(dis)
    st = np.ascontiguousarray(statu, dtype=np.int32).copy()
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    lib().orc_predictor(n, _d(x), _d(v), _d(f), _d(d), _i(st), _i(it), _d(c), C.c_double(h),
                        _d(_v3(lo)), _d(_v3(ln)), _i(_v3(ifpd, np.int32)))
    r = lambda a: a.reshape(3, n).T.copy()  # noqa: E731
    return r(x), r(v), r(d), st


def correction(xp1, fp, statu, ityp0, cm, h):
    n = int(np.asarray(ityp0).shape[0])
    v, f = _xp(xp1), _xp(fp)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    ek = np.zeros(n)
    lib().orc_correction(n, _d(v), _d(f), _i(st), _i(it), _d(c), C.c_double(h), _d(ek))
    return v.reshape(3, n).T.copy(), ek


def damping(xp1, fp, statu):
    n = int(np.asarray(statu).shape[0])
    v, f = _xp(xp1), _xp(fp)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    lib().orc_damping(n, _d(v), _d(f), _i(st))
    return v.reshape(3, n).T.copy()


def ekin(xp1, statu, ityp0, cm):
    n = int(np.asarray(statu).shape[0])
    v = _xp(xp1)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    ek = np.zeros(n)
    lib().orc_ekin(n, _d(v), _i(st), _i(it), _d(c), _d(ek))
    return ek


def velscale(xp1, statu, ityp0, cm, temp, kb):
    n = int(np.asarray(statu).shape[0])
    v = _xp(xp1)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    rc = lib().orc_velscale(n, _d(v), _i(st), _i(it), _d(c), C.c_double(temp), C.c_double(kb))
    if rc != 0:
        raise ValueError("oracle velscale: no kinetic energy")
    return v.reshape(3, n).T.copy()


def dyn_damp(xp, xp1, fp, dis, statu, ityp0, cm, h, lo, ln, ifpd, kvois, indi, ru2max,
             table: Table, maxsteps, mindele):
    """Do_DynDamp_Forsteps_DEV (MD_DiffScheme_GPU.F90:1809-1883) on the fixed full list
    (kvois, indi): per iteration damping (DAMPING_KERNEL :125-185), predictor, force and EPOT,
    exit when max |EPOT - EPOT0| <= mindele (before the corrector), else EPOT0 = EPOT and the
    corrector.  fp is the force of the entry configuration.  Returns (x, v, f, dis, statu, epot,
    iteration it stopped at or maxsteps + 1, last max |dEPOT|)."""
    x, v, f, d = (np.array(a, dtype=np.float64) for a in (xp, xp1, fp, dis))
    st = np.ascontiguousarray(statu, dtype=np.int32).copy()
    _, _, ep0 = force_full(x, ityp0, st, kvois, indi, ln, ifpd, ru2max, table)
    ep = ep0
    it, de = 1, 0.0
    while it <= maxsteps:
        v = damping(v, f, st)
        x, v, d, st = predictor(x, v, f, d, st, ityp0, cm, h, lo, ln, ifpd)
        f, _, ep = force_full(x, ityp0, st, kvois, indi, ln, ifpd, ru2max, table)
        de = float(np.max(np.abs(ep - ep0))) if len(ep) else 0.0
        if de <= mindele:
            break
        ep0 = ep
        v, _ = correction(v, f, st, ityp0, cm, h)
        it += 1
    return x, v, f, d, st, ep, it, de


def check_timestep(xp1, fp, statu, ityp0, cm, th, mxd2) -> bool:
    """CheckTimestep_KERNEL (MD_DiffScheme_GPU.F90:1068-1139): some active atom has
    |th v + th^2/2 F/m|^2 > mxd2 over the components without a FIXPOS bit."""
    v = np.asarray(xp1, dtype=np.float64)
    f = np.asarray(fp, dtype=np.float64)
    st = np.asarray(statu)
    m = np.asarray(cm, dtype=np.float64)[np.asarray(ityp0)]
    h2s2 = th * th * 0.5
    d = th * v + h2s2 * (f / m[:, None])
    for k, bit in enumerate((2, 4, 8)):
        d[(st & bit) != 0, k] = 0.0
    act = (st & 1) == 1
    return bool(np.any(np.sum(d * d, axis=1)[act] > mxd2))


def timestep_search(xp1, fp, statu, ityp0, cm, hmx, dmx) -> float:
    """The halving search of Predictor_DEV (MD_DiffScheme_GPU.F90:633-650)."""
    th = hmx
    while check_timestep(xp1, fp, statu, ityp0, cm, th, dmx * dmx):
        th *= 0.5
    return th


def _move_bounded(x, dx, lo, ln, ifpd):
    """x + dx with the bounds of DevAdd (AddBD_DevVec_DF_KERNEL0, MSM_MultiGPU_Basic.F90:
    5444-5477): in a periodic direction a coordinate above BOXUP moves down by the box length and
    one below BOXLOW up by it."""
    r = x + dx
    for k in range(3):
        if not ifpd[k]:
            continue
        lb, hb = lo[k], lo[k] + ln[k]
        c = r[:, k]
        c = np.where(c > hb, c - (hb - lb), np.where(c < lb, c + (hb - lb), c))
        r[:, k] = c
    return r


def steepest(xp, statu, ityp0, lo, ln, ifpd, kvois, indi, ru2max, table: Table, maxsteps,
             linesearch, alpha, maxdis, mindis, mindele):
    """Do_Steepest0_Forsteps_DEV / Do_Steepest1_Forsteps_DEV (MD_SteepestScheme_GPU.F90:20-152,
    156-260) on the fixed full list (kvois, indi), forces and energies from force_full.  Returns
    (x, steps, last max |dEPOT|, last largest move) with the step conventions of msm_steepest
    (where the reference would divide by zero, the same replacements are taken)."""
    lo, ln = np.asarray(lo, dtype=np.float64), np.asarray(ln, dtype=np.float64)
    x = np.array(xp, dtype=np.float64)

    def ff(xx):
        f, _, e = force_full(xx, ityp0, statu, kvois, indi, ln, ifpd, ru2max, table)
        return f, e

    de = mm = 0.0
    if not linesearch:
        f, _ = ff(x)
        pf = f.copy()
        dx = alpha * f
        mm = float(np.max(np.abs(dx)))
        if mm > maxdis:
            dx = dx * (maxdis / mm)
        elif mm <= mindis:
            return x, 0, 0.0, mm
        _, e0 = ff(x)
        x = _move_bounded(x, dx, lo, ln, ifpd)
        iflag = 0
        for i in range(1, maxsteps + 1):
            f, _ = ff(x)
            g = pf - f
            dxdf, dfdf = float(np.sum(dx * g)), float(np.sum(g * g))
            al = dxdf / dfdf if dfdf > 0.0 else -1.0
            if al < 0.0:
                al = alpha
            dx = al * f
            mm = float(np.max(np.abs(dx)))
            if mm > maxdis:
                dx = dx * (maxdis / mm)
            x = _move_bounded(x, dx, lo, ln, ifpd)
            if mm <= mindis:
                iflag = i
                break
            _, e = ff(x)
            de = float(np.max(np.abs(e - e0)))
            if de <= mindele:
                iflag = i
                break
            e0 = e
            pf = f
        return x, (iflag if iflag else maxsteps + 1), de, mm
    it = 0
    step = maxdis
    while it <= maxsteps:
        f, e0 = ff(x)
        nrm = float(np.sum(f * f))
        if not nrm > 0.0:
            break
        d = f * (1.0 / np.sqrt(nrm))
        step = maxdis
        dx = step * d
        pf0 = float(np.sum(f * d))
        while it <= maxsteps:
            x = _move_bounded(x, dx, lo, ln, ifpd)
            f, _ = ff(x)
            pf1 = float(np.sum(f * d))
            if pf1 != pf0:
                step = -step * pf1 / (pf1 - pf0)
            else:
                step = maxdis if -step * pf1 >= 0.0 else -maxdis
            if abs(step) > maxdis:
                step = maxdis if step > 0.0 else -maxdis
            dx = step * d
            it += 1
            if abs(step) <= mindis:
                break
            pf0 = pf1
        _, e = ff(x)
        de = float(np.max(np.abs(e - e0)))
        mm = abs(step)
        if de <= mindele:
            break
    return x, it, de, mm


def _secant(step, pf0, pf1, maxdis):
    if pf1 != pf0:
        st = -step * pf1 / (pf1 - pf0)
    else:
        st = maxdis if -step * pf1 >= 0.0 else -maxdis
    if abs(st) > maxdis:
        st = maxdis if st > 0.0 else -maxdis
    return st


def cg(xp, statu, ityp0, lo, ln, ifpd, kvois, indi, ru2max, table: Table, maxsteps, linesearch,
       maxdis, mindis, mindele):
    """Do_CG0_Forsteps_DEV / Do_CG1_Forsteps_DEV (MD_CGScheme_GPU.F90:16-291) on the fixed full
    list: Polak-Ribiere directions from d = F, secant steps along them, moves with the bounds of
    DevAdd.  Returns (x, steps, last max |dEPOT|) with the conventions of msm_cg."""
    lo, ln = np.asarray(lo, dtype=np.float64), np.asarray(ln, dtype=np.float64)
    x = np.array(xp, dtype=np.float64)
    eps = 1.0e-64

    def ff(xx):
        f, _, e = force_full(xx, ityp0, statu, kvois, indi, ln, ifpd, ru2max, table)
        return f, e

    f, e0 = ff(x)
    d = f.copy()
    f0 = f.copy()
    f0n = float(np.sum(f0 * f0))
    pf0 = float(np.sum(f * d))
    de = 0.0
    if not f0n > eps:
        return x, 0, de

    def new_direction():
        nonlocal d, f0, f0n, pf0
        f1, _ = ff(x)
        g = float(np.sum(f1 * (f1 - f0))) / f0n
        d = f1 + g * d
        f0 = f1.copy()
        f0n = float(np.sum(f0 * f0))
        pf0 = float(np.sum(f1 * d))

    if not linesearch:
        for i in range(1, maxsteps + 1):
            step = maxdis / np.sqrt(f0n)
            x = _move_bounded(x, step * d, lo, ln, ifpd)
            f, _ = ff(x)
            pf1 = float(np.sum(f * d))
            step = _secant(step, pf0, pf1, maxdis)
            x = _move_bounded(x, (step / np.sqrt(f0n)) * d, lo, ln, ifpd)
            _, e = ff(x)
            de = float(np.max(np.abs(e - e0)))
            if de <= mindele:
                return x, i, de
            e0 = e
            new_direction()
            if f0n <= eps:
                return x, i, de
        return x, maxsteps + 1, de
    it = 0
    while it <= maxsteps:
        step = maxdis / np.sqrt(f0n)
        x = _move_bounded(x, step * d, lo, ln, ifpd)
        f, _ = ff(x)
        pf1 = float(np.sum(f * d))
        it += 1
        while it <= maxsteps:
            step = _secant(step, pf0, pf1, maxdis)
            x = _move_bounded(x, (step / np.sqrt(f0n)) * d, lo, ln, ifpd)
            it += 1
            if abs(step) <= mindis:
                break
            pf0 = pf1
            f, _ = ff(x)
            pf1 = float(np.sum(f * d))
        _, e = ff(x)
        de = float(np.max(np.abs(e - e0)))
        if de <= mindele:
            break
        e0 = e
        new_direction()
        if f0n <= eps:
            break
    return x, it, de


def md_run(xp, xp1, statu, ityp0, cm, h, nsteps, nb_uptab, step0, lo, ln, ifpd, nb_rm, ru, mx,
           table: Table, fp=None, init=True):
    """Half-list FS MD with the reference's step order; returns (xp, xp1, fp, dis, statu, epot,
    kvois, indi)."""
    n = int(np.asarray(ityp0).shape[0])
    x, v = _xp(xp), _xp(xp1)
    f = _xp(fp) if fp is not None else np.zeros(3 * n)
    d = np.zeros(3 * n)
    st = np.ascontiguousarray(statu, dtype=np.int32).copy()
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    nb = np.ascontiguousarray(np.asarray(nb_rm, dtype=np.float64))
    r = np.ascontiguousarray(np.asarray(ru, dtype=np.float64))
    kv = np.zeros(n, dtype=np.int32)
    ind = np.zeros(n * mx, dtype=np.int32)
    ep = np.zeros(n)
    rc = lib().orc_md_run(n, _d(x), _d(v), _d(f), _d(d), _i(st), _i(it), _d(c), C.c_double(h),
                          nsteps, nb_uptab, step0, 1 if init else 0, _d(_v3(lo)), _d(_v3(ln)),
                          _i(_v3(ifpd, np.int32)), _d(nb), _d(r), mx, C.byref(table.struct()),
                          _i(kv), _i(ind), _d(ep))
    if rc != 0:
        raise RuntimeError(f"oracle md_run error {rc}")
    rs = lambda a: a.reshape(3, n).T.copy()  # noqa: E731
    return rs(x), rs(v), rs(f), rs(d), st, ep, kv, ind.reshape(mx, n).T


def drand32(seed):
    s = (C.c_int * 2)(*seed)
    v = lib().orc_drand32(s)
    return v, (s[0], s[1])


def thermalize(ityp0, cm, temp, kb, seed):
    n = int(np.asarray(ityp0).shape[0])
    s = (C.c_int * 2)(*seed)
    v = np.zeros(3 * n)
    it = np.ascontiguousarray(ityp0, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    lib().orc_thermalize_mc(n, _i(it), _d(c), len(c), C.c_double(temp), C.c_double(kb),
                            C.cast(s, _ip), _d(v))
    return v.reshape(3, n).T.copy()


def philox4x32_10(ctr, key):
    c = (C.c_uint * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def thermalize_philox(natoms, ityp, statu, cm, tkb, seed, draw, xp1):
    """Thermalizing_MC_DEV with Philox4x32-10 uniforms over boxes of natoms consecutive atoms;
    xp1 (N, 3) is the velocity before the call (kept on FIX components)."""
    n = int(np.asarray(ityp).shape[0])
    v = np.ascontiguousarray(np.asarray(xp1, dtype=np.float64).T).ravel().copy()
    na = np.ascontiguousarray(natoms, dtype=np.int32)
    it = np.ascontiguousarray(ityp, dtype=np.int32)
    st = np.ascontiguousarray(statu, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    lib().orc_thermalize_philox(n, len(na), _i(na), _i(it), _i(st), _d(c), C.c_double(tkb),
                                C.c_ulonglong(int(seed) & ((1 << 64) - 1)),
                                C.c_ulonglong(int(draw)), _d(v))
    return v.reshape(3, n).T.copy()


def tree_sum(a) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_sum(_d(a), int(a.size))
