"""Configuration files of a box, in the reference's text formats.

The writer produces the &BOXCFG18 layout of Putout_Instance_Config_SimMDBox
(MDLIB/sor/Common/MD_TypeDef_SimBox.F90:2388-2630, with the record stamp of Putout_RecordStamp,
MD_TypeDef_RecordStamp.F90:89-115).  The loader reads what Load_Config_SimMDBox0
(MD_TypeDef_SimBox.F90:1389-1520) reads: &BOXCFG14/15/16 files (fixed columns) and &BOXCFG18 files
(columns named by their &...COL tags, Read_ConfigByCol18B_SimMDBox :3029-3133); the plain table
of Read_Initial_Config_SimMDBox0 (:1823-1979: type and position, optionally velocity, in lattice
units, the box coming from the box set-up) is read by Read_Initial_Config_SimMDBox.

Units.  This layer works in eV, Angstrom and fs.  The files hold lengths in lattice units (LU =
SimBox.RR Angstrom), and &BOXCFG18 velocities in LU/ps, forces in eV/LU and POT = -EPOT in eV.  The
older formats and the plain table hold velocities in cm/s and forces in dyn (the reference's CGS
state, written without conversion); they are converted with the reference's constants
(MSMLIB/sor/Common/MSM_Const.F90:68-85: 1 eV = 1.60219e-12 erg, 1 cm = 1e8 A).
"""
from __future__ import annotations

import datetime
from typing import List, Optional, Sequence

import numpy as np

from .mdlib import CP_CM2A, CP_EVERG, SimMDBox

CM_S_TO_A_FS = CP_CM2A * 1.0e-15            # (cm/s) -> (A/fs)
DYN_TO_EV_A = 1.0 / CP_EVERG / CP_CM2A      # dyn = erg/cm -> eV/A

FORMAT18 = "&BOXCFG18"
OLD_FORMATS = ("&BOXCFG14", "&BOXCFG15", "&BOXCFG16")

# column tags of the &BOXCFG18 writer, in order: (tag, width, type)
_COLS18 = [("&TYPECOL", 1, "I"), ("&XYZCOL", 3, "D"), ("&VELCOL", 3, "D"), ("&STATCOL", 1, "I"),
           ("&FPCOL", 3, "D"), ("&EPOTCOL", 1, "D"), ("&EKINCOL", 1, "D"), ("&DISCOL", 3, "D")]
_TITLE18 = ("&TYPE         "
            "POS(LU)  (x)             (y)              (z)         "
            "VEL(LU/ps)(vx)           (vy)             (vz)        "
            "STATU   "
            "FOR(ev/LU)(fx)           (fy)             (fz)        "
            "POT(ev).          "
            "K.E.(ev)          "
            "DISPLACE(dx )            (dy)               (dz)       ")


def _e(v: float, w: int, d: int) -> str:
    """Fortran 1PEw.d of one value."""
    return f"{v:{w}.{d}E}"


def Putout_Instance_Config_SimMDBox(Fhead: str, SimBox: SimMDBox, ITime: int = 0,
                                    Time: float = 0.0, ICfg: int = -1, AppType: str = "GMD",
                                    Temperature: float = 0.0, ITest: int = 0,
                                    IBox: Sequence[int] = (0, 0), ISect: int = 0) -> str:
    """Write SimBox as a &BOXCFG18 file named Fhead (Fhead.NNNN when ICfg >= 0) and return the
    name.  EKIN is m v^2 / 2 of each atom; absent FP, EPOT and DIS are written as zeros."""
    n = SimBox.NPRT
    rr = float(SimBox.RR)
    fname = f"{Fhead}.{ICfg:04d}" if ICfg >= 0 else Fhead
    xp = np.asarray(SimBox.XP, dtype=np.float64)
    xp1 = np.asarray(SimBox.XP1, dtype=np.float64)
    fp = np.zeros((n, 3)) if SimBox.FP is None else np.asarray(SimBox.FP, dtype=np.float64)
    ep = np.zeros(n) if SimBox.EPOT is None else np.asarray(SimBox.EPOT, dtype=np.float64)
    dis = np.zeros((n, 3)) if SimBox.DIS is None else np.asarray(SimBox.DIS, dtype=np.float64)
    ityp = np.asarray(SimBox.ITYP, dtype=np.int64)
    cm = np.asarray(SimBox.CM, dtype=np.float64)
    ek = 0.5 * cm[ityp - 1] * np.sum(xp1 * xp1, axis=1)
    na = [int(np.count_nonzero(ityp == g + 1)) for g in range(SimBox.NGROUP)]
    now = datetime.datetime.now()
    out: List[str] = [FORMAT18]
    # Putout_RecordStamp
    out.append(f'&APPTYPE        "{AppType}" ')
    out.append("&DATE            " + now.strftime("%Y-%m-%d,%Hh%Mm%Ss"))
    out.append(f"&TESTID          {ITest:12d}")
    out.append(f"&BOXID           {IBox[0]:12d}, {IBox[1]:6d}")
    out.append(f"&CFGID           {ICfg:12d}, {0:6d}")
    out.append(f"&RECID           {0:12d}, {0:6d}")
    out.append(f"&TIMESTEPS       {ITime:12d}")
    out.append(f"&TIMESECT #      {ISect:12d}")
    out.append("&TIME (ps)       " + _e(Time, 12, 5))
    out.append("&SCALTIME (ps)   " + _e(0.0, 12, 5))
    out.append("&TEMPSET (K)     " + _e(Temperature, 12, 5))
    out.append(" ")
    out.append("!--- Table colume information:")
    col = 1
    for tag, width, typ in _COLS18:
        out.append(f"{tag:<12s} {col:4d} {width:3d} \"{typ}\"")
        col += width
    out.append("")
    out.append("&LATT     lattice length (in A):       " + _e(rr, 12, 5) + " ")
    out.append("&BOXLOW   low boundary of box (in LU): " +
               "".join(_e(v, 12, 5) + " " for v in np.asarray(SimBox.BOXLOW) / rr))
    out.append("&BOXSIZE  boxsize (in LU):             " +
               "".join(_e(v, 12, 5) + " " for v in np.asarray(SimBox.ZL) / rr))
    out.append(f"&NATOM    total number of atoms:       {n:8d} ")
    out.append(f"&NGROUP   number of group of atoms:    {SimBox.NGROUP:8d} ")
    out.append("    &NA   number of atoms in groups:   " + "".join(f"{v:8d} " for v in na))
    out.append("&TEMPCAL instant temperature(K):      " + _e(Temperature, 12, 5) + " ")
    out.append("!--- Configure:")
    out.append(_TITLE18.rstrip())
    st = np.asarray(SimBox.STATU, dtype=np.int64)
    v_lu_ps = xp1 / rr * 1000.0
    f_ev_lu = fp * rr
    for i in range(n):
        row = [f"{int(ityp[i]):8d}   "]
        row += [_e(v, 17, 8) + " " for v in xp[i] / rr]
        row += [_e(v, 17, 8) + " " for v in v_lu_ps[i]]
        row.append(f"{int(st[i]):6d}  ")
        row += [_e(v, 17, 8) + " " for v in f_ev_lu[i]]
        row.append(_e(-ep[i], 17, 8) + " ")
        row.append(_e(ek[i], 17, 8) + " ")
        row += [_e(v, 17, 8) + " " for v in dis[i] / rr]
        out.append("".join(row))
    with open(fname, "w") as fh:
        fh.write("\n".join(out) + "\n")
    return fname


def _numbers(line: str) -> List[float]:
    vals = []
    for tok in line.replace(",", " ").split():
        try:
            vals.append(float(tok))
        except ValueError:
            pass
    return vals


def _header(lines: List[str]):
    """Box geometry and counts from the keyword lines of a configuration file
    (Read_ConfigFileHeader_SimMDBox, MD_TypeDef_SimBox.F90:1983-2300); returns the header
    values, the &...COL tags and the index of the first table row."""
    h = {"cols": {}}
    natom_seen = False
    start = None
    for k, raw in enumerate(lines):
        s = raw.strip()
        if not s or s.startswith("!"):
            continue
        if not s.startswith("&"):
            start = k
            break
        key = s.split()[0].upper()
        nums = _numbers(s[len(key):])
        if key == "&LATT":
            h["latt"] = nums[-1]
        elif key == "&BOXLOW":
            h["boxlow"] = np.array(nums[-3:])
        elif key == "&BOXSIZE":
            h["boxsize"] = np.array(nums[-3:])
        elif key == "&NATOM":
            if not natom_seen:
                h["natom"] = int(nums[-1])
                natom_seen = True
            else:  # second &NATOM of the older formats: atoms per group
                h["na"] = [int(v) for v in nums]
        elif key == "&NA":
            h["na"] = [int(v) for v in nums]
        elif key == "&NGROUP":
            h["ngroup"] = int(nums[-1])
        elif key.endswith("COL") and nums:
            h["cols"][key] = int(nums[0])
        elif key == "&TYPE":
            h["title_line"] = k
    if start is None:
        raise ValueError("configuration file without a table")
    return h, start


def Load_Config_SimMDBox(fname: str, SimBox: SimMDBox) -> SimMDBox:
    """Load a &BOXCFG14/15/16/18 file into SimBox (Load_Config_SimMDBox0): RR, BOXLOW, ZL, NPRT,
    NGROUP, ITYP, XP, XP1, STATU, FP, EPOT, DIS (and EKIN in the returned box's .EKIN)."""
    with open(fname) as fh:
        lines = fh.read().split("\n")
    first = next((l.strip() for l in lines if l.strip() and not l.strip().startswith("!")), "")
    kind = first.split()[0].upper() if first else ""
    if kind not in OLD_FORMATS and kind != FORMAT18:
        raise ValueError(f"{fname}: not a &BOXCFG14/15/16/18 file (first keyword {kind!r})")
    h, start = _header(lines)
    n = h["natom"]
    rows = [l.split() for l in lines[start:] if l.strip() and not l.strip().startswith("!")]
    if len(rows) < n:
        raise ValueError(f"{fname}: {len(rows)} atom rows for &NATOM {n}")
    rows = rows[:n]
    rr = float(h["latt"])
    SimBox.RR = rr
    SimBox.BOXLOW = h["boxlow"] * rr
    SimBox.ZL = h["boxsize"] * rr
    SimBox.NPRT = n
    if "ngroup" in h:
        SimBox.NGROUP = h["ngroup"]
    if kind in OLD_FORMATS:
        # TYPE, XP(3), XP1(3), STATU, FP(3), POT, EKIN, [ENERGY,] DIS(3)
        a = np.array([[float(v) for v in r] for r in rows])
        off = 0 if kind == "&BOXCFG14" else 1
        ityp = a[:, 0].astype(np.int32)
        xp = a[:, 1:4] * rr
        xp1 = a[:, 4:7] * CM_S_TO_A_FS
        st = a[:, 7].astype(np.int32)
        fp = a[:, 8:11] * DYN_TO_EV_A
        ep = -a[:, 11]
        ek = a[:, 12]
        dis = a[:, 13 + off:16 + off] * rr
    else:
        c = h["cols"]
        if "&XYZCOL" not in c:
            raise ValueError(f"{fname}: &XYZCOL is required")
        ncol = max(len(r) for r in rows)
        a = np.zeros((n, ncol))
        for i, r in enumerate(rows):
            a[i, :len(r)] = [float(v) for v in r]

        def cols(tag, width):
            k = c.get(tag, 0)
            return None if k <= 0 else a[:, k - 1:k - 1 + width]

        t = cols("&TYPECOL", 1)
        ityp = np.ones(n, np.int32) if t is None else t[:, 0].astype(np.int32)
        xp = cols("&XYZCOL", 3) * rr
        v = cols("&VELCOL", 3)
        xp1 = np.zeros((n, 3)) if v is None else v * rr / 1000.0
        s = cols("&STATCOL", 1)
        if s is None:
            s = cols("&STATUCOL", 1)
        st = np.ones(n, np.int32) if s is None else s[:, 0].astype(np.int32)
        f = cols("&FPCOL", 3)
        fp = np.zeros((n, 3)) if f is None else f / rr
        e = cols("&EPOTCOL", 1)
        ep = np.zeros(n) if e is None else -e[:, 0]
        k = cols("&EKINCOL", 1)
        ek = np.zeros(n) if k is None else k[:, 0]
        d = cols("&DISCOL", 3)
        dis = np.zeros((n, 3)) if d is None else d * rr
    SimBox.ITYP = ityp
    SimBox.XP = np.asfortranarray(xp)
    SimBox.XP1 = np.asfortranarray(xp1)
    SimBox.STATU = st
    SimBox.FP = np.asfortranarray(fp)
    SimBox.EPOT = ep
    SimBox.DIS = np.asfortranarray(dis)
    SimBox.EKIN = ek
    return SimBox


def _first_keyword(fname: str) -> str:
    with open(fname) as fh:
        for line in fh:
            t = line.split("!")[0].strip()
            if t:
                return t.split()[0].upper() if t.startswith("&") else ""
    return ""


def _check_consistent(SimBox: SimMDBox) -> SimMDBox:
    """Atoms outside the box are marked OUTOFBOX; FIXPOS components of the velocity are zero."""
    xp = np.asarray(SimBox.XP)
    st = np.asarray(SimBox.STATU, dtype=np.int32).copy()
    lo = np.asarray(SimBox.BOXLOW, dtype=np.float64)
    hi = lo + np.asarray(SimBox.ZL, dtype=np.float64)
    out = np.any((xp < lo) | (xp > hi), axis=1)
    st[out] = 65536  # CP_STATU_OUTOFBOX
    xp1 = np.array(SimBox.XP1, dtype=np.float64, order="F")
    for k, bit in enumerate((2, 4, 8)):
        xp1[(st & bit) != 0, k] = 0.0
    SimBox.STATU = st
    SimBox.XP1 = xp1
    return SimBox


def Read_Initial_Config_SimMDBox(fname: str, SimBox: SimMDBox, with_velocity: bool = False,
                                 prop: Optional[Sequence[int]] = None) -> SimMDBox:
    """Read_Initial_Config_SimMDBox (MD_TypeDef_SimBox.F90:1287-1385) for the plain table and
    &CFGXYZ files.  A file opening with &CFGXYZ goes to Read_ConfigByCol18_SimMDBox; any other
    leading keyword is an error.  The plain table (Read_Initial_Config_SimMDBox0): SimBox.NPRT rows
    of type and position in lattice units (SimBox.RR), with velocities in cm/s when with_velocity
    (CP_INPUT_VEL); the geometry (RR, BOXLOW, ZL) comes from SimBox.  STATU is prop[type - 1]
    (ACTIVE by default); atoms outside the box are marked OUTOFBOX, and FIXPOS components of the
    velocity are zero."""
    kw = _first_keyword(fname)
    if kw == "&CFGXYZ":
        return Read_ConfigByCol18_SimMDBox(fname, SimBox, prop)
    if kw:
        raise ValueError(f"MDPSCU Error: unknown file format: {kw} in {fname}")
    n = SimBox.NPRT
    rows = []
    with open(fname) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("!"):
                continue
            rows.append(s.split())
            if len(rows) == n:
                break
    if len(rows) < n:
        raise ValueError(f"{fname}: only {len(rows)} atoms for NPRT {n}")
    rr = float(SimBox.RR)
    ityp = np.array([int(float(r[0])) for r in rows], np.int32)
    xp = np.array([[float(v) for v in r[1:4]] for r in rows]) * rr
    xp1 = np.zeros((n, 3))
    if with_velocity:
        xp1 = np.array([[float(v) for v in r[4:7]] for r in rows]) * CM_S_TO_A_FS
    ng = int(ityp.max())
    pr = np.ones(ng, np.int32) if prop is None else np.asarray(prop, np.int32)
    SimBox.ITYP = ityp
    SimBox.XP = np.asfortranarray(xp)
    SimBox.XP1 = np.asfortranarray(xp1)
    SimBox.STATU = pr[ityp - 1].copy()
    return _check_consistent(SimBox)


# column keywords of &CFGXYZ files (Read_ConfigByCol18A_SimMDBox, MD_TypeDef_SimBox.F90:2739-3025)
_COLS = {"&XYZCOL": "xp", "&DISCOL": "dis", "&TYPECOL": "ityp", "&VELCOL": "xp1",
         "&FPCOL": "fp", "&EPOTCOL": "epot", "&EKINCOL": "ekin", "&STATUCOL": "statu",
         "&STATCOL": "statu"}
_WIDTH = {"xp": 3, "dis": 3, "xp1": 3, "fp": 3, "ityp": 1, "epot": 1, "ekin": 1, "statu": 1}


def Read_ConfigByCol18_SimMDBox(fname: str, SimBox: SimMDBox,
                                prop: Optional[Sequence[int]] = None) -> SimMDBox:
    """Read_ConfigByCol18_SimMDBox (MD_TypeDef_SimBox.F90:2634-2735, 2739-3133): a header of
    &KEYWORD lines naming the 1-based column of each quantity (&XYZCOL required; &TYPECOL,
    &VELCOL, &FPCOL, &DISCOL, &EPOTCOL, &EKINCOL, &STATCOL/&STATUCOL optional) and the length
    unit (&LUNIT or &LENTHUNIT "A" | "CM" | "NM" | "LU", times an optional number; required
    unless the header has &BOXCFG18), ended by the first line without a keyword; then SimBox.NPRT
    rows.  Positions and displacements are scaled by the unit; the other columns hold the
    reference's internal units (velocity cm/s, force dyn, energies erg) and are converted to
    eV/A/fs; the potential column holds -EPOT.  Without &TYPECOL every atom is of type 1; without
    a STATU column STATU is prop[type - 1] (ACTIVE by default).  The geometry and NPRT come from
    SimBox; out-of-box atoms are marked OUTOFBOX."""
    with open(fname) as fh:
        lines = fh.read().splitlines()
    cols = {}
    unit = None
    needunit = True
    i = 0
    while i < len(lines):
        t = lines[i].split("!")[0].strip()
        if not t:
            i += 1
            continue
        if not t.startswith("&"):
            break
        kw = t.split()[0].upper()
        rest = t[len(kw):]
        if kw == "&BOXCFG18":
            needunit = False
        elif kw in ("&LUNIT", "&LENTHUNIT"):
            words = rest.replace('"', " ").replace("'", " ").replace(",", " ").split()
            if not words:
                raise ValueError(f"MDPSCU Error: length unit missing after {kw} in {fname}")
            u = words[0].upper()
            if u == "A":
                unit = 1.0
            elif u == "CM":
                unit = CP_CM2A
            elif u == "NM":
                unit = 10.0
            elif u == "LU":
                unit = float(SimBox.RR) if SimBox.RR > 0.0 else None
            else:
                raise ValueError(f"MDPSCU Error: unknown length unit {words[0]} in {fname}")
            num = _numbers(" ".join(words[1:]))
            if num and unit is not None:
                unit *= num[0]
        elif kw in _COLS:
            num = _numbers(rest.replace(",", " "))
            c = int(num[0]) if num else 0
            if c <= 0:
                raise ValueError(f"MDPSCU Error: the column of {kw} is missed in {fname}")
            cols[_COLS[kw]] = c
        i += 1
    if "xp" not in cols:
        raise ValueError(f"MDPSCU Error: keyword &XYZCOL not found in {fname}")
    if needunit and (unit is None or unit <= 0.0):
        raise ValueError(f"MDPSCU Error: the length unit (&LUNIT) is missed in {fname}")
    lunit = unit if needunit else 1.0
    ncol = max(c + _WIDTH[k] - 1 for k, c in cols.items())
    n = SimBox.NPRT
    rows = []
    while i < len(lines) and len(rows) < n:
        t = lines[i].split("!")[0].split()
        i += 1
        if not t:
            continue
        if len(t) < ncol:
            raise ValueError(f"MDPSCU Error: {fname}: a row has {len(t)} of {ncol} columns")
        rows.append(t[:ncol])
    if len(rows) < n:
        raise ValueError(f"MDPSCU Error: {fname}: only {len(rows)} atoms for NPRT {n}")

    def block(key, width):
        c = cols[key] - 1
        return np.array([[float(r[c + k].replace("D", "E")) for k in range(width)] for r in rows])

    SimBox.XP = np.asfortranarray(block("xp", 3) * lunit)
    if "dis" in cols:
        SimBox.DIS = np.asfortranarray(block("dis", 3) * lunit)
    SimBox.XP1 = np.asfortranarray(block("xp1", 3) * CM_S_TO_A_FS) if "xp1" in cols \
        else np.zeros((n, 3), order="F")
    if "fp" in cols:
        SimBox.FP = np.asfortranarray(block("fp", 3) * DYN_TO_EV_A)
    if "epot" in cols:
        SimBox.EPOT = -block("epot", 1)[:, 0] / CP_EVERG
    if "ekin" in cols:
        SimBox.EKIN = block("ekin", 1)[:, 0] / CP_EVERG
    SimBox.ITYP = block("ityp", 1)[:, 0].astype(np.int32) if "ityp" in cols \
        else np.ones(n, np.int32)
    if "statu" in cols:
        SimBox.STATU = block("statu", 1)[:, 0].astype(np.int32)
    else:
        ng = int(SimBox.ITYP.max())
        pr = np.ones(ng, np.int32) if prop is None else np.asarray(prop, np.int32)
        SimBox.STATU = pr[SimBox.ITYP - 1].copy()
    return _check_consistent(SimBox)
