"""Potential files of the reference's EAM_NIST family as force-table procedures: setfl (the
DYNAMO / LAMMPS eam/alloy layout), lspt (a control file naming two-column spline files) and moldy
(Finnis-Sinclair cubic knots).

setfl restates MDLIB/sor/Potentials/EAM_NIST: the reader
(Filedatas_Func_Setfl.F90:119-253), the spline procedures that turn the file's samples into the
callbacks of the force-table builder (:292-496), their registration (:499-541) and the table
generation (NIST_ForceTable.F90:356-426), in this layer's units (eV, Angstrom) instead of CGS.

The file: three comment lines; "NE El1 El2 ..."; "Nrho drho Nr dr cutoff"; per element a line
"Z mass a lattice", F(rho) at rho = (i-1) drho (Nrho values) and rho(r) (Nr values); then
r*phi(r) for every pair I >= J (Nr values each).  A list of values may run over any number of
lines and the rest of its last line is skipped, as a Fortran list-directed read does.  Kept
from the reference:

* the r grid is r_i = (i-1) cutoff/Nr (the file's dr is read but not used);
* a density list starting with "INF" (or a pair list starting with "NAN") takes the numbers that
  follow on that line from i = 2 and copies the second value to r = 0;
* every sample set becomes the cubic spline through the samples with zero second derivative at
  both ends (SLATEC DBINT4 with IBCL = IBCR = 2, FBCL = FBCR = 0, knots at the samples), value and
  slope as DBVALU gives them;
* table id IT = (I-1) NE + J (I the element that feels the term, J its neighbour): pair term
  phi of the unordered pair; density term rho_J -- zero when F of element I is zero everywhere;
  embedding term F_I for I = J only;
* below the first sample a term takes the first sample's value with zero slope; beyond the last
  sample the pair and density terms are zero and the embedding term keeps the last value with
  zero slope;
* the embedding table spans the file's Nrho drho.
"""
from __future__ import annotations

import bisect
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from . import mdlib
from .mdlib import MDForceTable


class _Spline:
    """Cubic spline through (x0 + i h, f_i), i = 0..n-1, with zero second derivative at both
    ends.  The interpolant is unique, so this is the function DBINT4 represents in B-splines
    (IBCL = IBCR = 2, FBCL = FBCR = 0, KNTOPT = 1); value and first derivative on [x0, x_n-1]."""

    def __init__(self, x0: float, h: float, f: np.ndarray):
        f = np.asarray(f, dtype=np.float64)
        n = len(f)
        if n < 3:
            raise ValueError("MDPSCU Error: a setfl list needs at least 3 samples")
        self.x0, self.h, self.n = float(x0), float(h), n
        # second derivatives m: m_0 = m_{n-1} = 0, h/6 m_{i-1} + 2h/3 m_i + h/6 m_{i+1} = slope jump
        rhs = (f[2:] - 2.0 * f[1:-1] + f[:-2]) / h
        k = n - 2
        c = np.zeros(k)
        d = np.zeros(k)
        dia, off = 2.0 * h / 3.0, h / 6.0
        c[0], d[0] = off / dia, rhs[0] / dia
        for i in range(1, k):
            m = dia - off * c[i - 1]
            c[i] = off / m
            d[i] = (rhs[i] - off * d[i - 1]) / m
        w = np.zeros(n)
        w[k] = d[k - 1]
        for i in range(k - 2, -1, -1):
            w[i + 1] = d[i] - c[i] * w[i + 2]
        self.f = f.tolist()
        self.w = w.tolist()
        self.xmin = self.x0
        self.xmax = self.x0 + (n - 1) * self.h

    def __call__(self, y: float) -> Tuple[float, float]:
        h = self.h
        j = min(max(int(math.floor((y - self.x0) / h)), 0), self.n - 2)
        a = self.x0 + j * h
        b = a + h
        f0, f1, w0, w1 = self.f[j], self.f[j + 1], self.w[j], self.w[j + 1]
        v = ((w0 * (b - y) ** 3 + w1 * (y - a) ** 3) / (6.0 * h)
             + (f1 / h - w1 * h / 6.0) * (y - a) + (f0 / h - w0 * h / 6.0) * (b - y))
        dv = ((-w0 * (b - y) ** 2 + w1 * (y - a) ** 2) / (2.0 * h)
              + (f1 - f0) / h - (w1 - w0) * h / 6.0)
        return v, dv


@dataclass
class SetflFile:
    """Contents of a setfl file (get_filedatas_setfl, Filedatas_Func_Setfl.F90:119-253)."""
    comments: List[str]
    elements: List[str]
    nrho: int
    drho: float
    nr: int
    dr: float
    cutoff: float
    number: List[int] = field(default_factory=list)
    mass: List[float] = field(default_factory=list)
    lattice: List[float] = field(default_factory=list)
    frho: List[np.ndarray] = field(default_factory=list)      # per element, at (i-1) drho
    rhor: List[np.ndarray] = field(default_factory=list)      # per element, at (i-1) cutoff/Nr
    zr: Dict[Tuple[int, int], np.ndarray] = field(default_factory=dict)  # r*phi, keys (I, J), I >= J
    frho_zero: List[bool] = field(default_factory=list)

    @property
    def rhomx(self) -> float:
        return self.nrho * self.drho

    @property
    def rstep(self) -> float:
        return self.cutoff / self.nr


class _Lines:
    """Line-oriented token source: a value list may span lines; the rest of its last line is
    dropped (Fortran list-directed read)."""

    def __init__(self, lines: List[str]):
        self.lines = lines
        self.i = 0

    def line(self) -> str:
        if self.i >= len(self.lines):
            raise ValueError("MDPSCU Error: unexpected end of setfl file")
        s = self.lines[self.i]
        self.i += 1
        return s

    def peek(self) -> str:
        while self.i < len(self.lines) and not self.lines[self.i].split():
            self.i += 1
        if self.i >= len(self.lines):
            raise ValueError("MDPSCU Error: unexpected end of setfl file")
        return self.lines[self.i].split()[0]

    def values(self, n: int) -> List[float]:
        out: List[float] = []
        while len(out) < n:
            out += [float(t.replace("D", "E").replace("d", "e")) for t in self.line().split()]
        return out[:n]


def _numbers(tokens: List[str]) -> List[float]:
    out = []
    for t in tokens:
        try:
            out.append(float(t.replace("D", "E").replace("d", "e")))
        except ValueError:
            pass
    return out


def _list_with_marker(src: _Lines, n: int, marker: str) -> np.ndarray:
    """n samples; a list opened by `marker` (INF / NAN at r = 0) starts at the second sample and
    repeats it at the first."""
    if src.peek().upper() == marker:
        head = _numbers(src.line().split()[1:])
        vals = [0.0] + head
        vals += src.values(n - len(vals)) if len(vals) < n else []
        vals = vals[:n]
        vals[0] = vals[1]
        return np.asarray(vals, dtype=np.float64)
    return np.asarray(src.values(n), dtype=np.float64)


def read_setfl(path: str) -> SetflFile:
    with open(path) as fh:
        src = _Lines(fh.read().splitlines())
    comments = [src.line() for _ in range(3)]
    tok = src.line().split()
    ne = int(float(tok[0]))
    elements = tok[1:1 + ne]
    if len(elements) != ne:
        raise ValueError(f"MDPSCU Error: setfl file {path} names {len(elements)} of {ne} elements")
    nrho, drho, nr, dr, cutoff = src.values(5)
    out = SetflFile(comments, elements, int(nrho), drho, int(nr), dr, cutoff)
    for _ in range(ne):
        head = _numbers(src.line().split())
        out.number.append(int(head[0]) if head else 0)
        out.mass.append(head[1] if len(head) > 1 else 0.0)
        out.lattice.append(head[2] if len(head) > 2 else 0.0)
        f = np.asarray(src.values(out.nrho), dtype=np.float64)
        out.frho.append(f)
        out.frho_zero.append(bool(f.max() == 0.0 and f.min() == 0.0))
        out.rhor.append(_list_with_marker(src, out.nr, "INF"))
    for i in range(ne):
        for j in range(i + 1):
            out.zr[(i, j)] = _list_with_marker(src, out.nr, "NAN")
    return out


def _procs(data: SetflFile, i: int, j: int):
    """NN_Spline, RHO_Spline, EMBED_Spline (:440-496) of table (I, J) = (i+1, j+1), in eV and A:
    pair (V/2, -dV/dr) from r*phi, density (rho_J, -drho_J/dr), embedding (F_I, dF_I/drho)."""
    hr = data.rstep
    zs = _Spline(0.0, hr, data.zr[(max(i, j), min(i, j))])

    def nn(r):
        if r > zs.xmax:
            return 0.0, 0.0
        if r < zs.xmin:
            z, dz = zs(zs.xmin)[0], 0.0
        else:
            z, dz = zs(r)
        v = z / r
        dv = (dz - v) / r
        return 0.5 * v, -dv

    if data.frho_zero[i]:
        def ne(r):
            return 0.0, 0.0
    else:
        rs = _Spline(0.0, hr, data.rhor[j])

        def ne(r):
            if r > rs.xmax:
                return 0.0, 0.0
            if r < rs.xmin:
                return rs(rs.xmin)[0], 0.0
            v, dv = rs(r)
            return v, -dv

    if i == j:
        fs = _Spline(0.0, data.drho, data.frho[i])

        def em(rho):
            if rho < fs.xmin:
                return fs(fs.xmin)[0], 0.0
            if rho > fs.xmax:
                return fs(fs.xmax)[0], 0.0
            return fs(rho)
    else:
        def em(rho):
            return 0.0, 0.0
    return nn, ne, em


def Register_ForceTableProc_Setfl(fname: str, FTable: MDForceTable) -> List[str]:
    """Register_ForceTableProc_Setfl (Filedatas_Func_Setfl.F90:499-541): one procedure set per
    ordered element pair, id IT = (I-1) NE + J, note "setfl table: I<-J ( fname )".  Sets the
    table's Rmax to the file's cutoff and RHOMX to Nrho drho (as the reader does); returns the
    element names."""
    data = read_setfl(fname)
    ne = len(data.elements)
    it = 0
    for i in range(ne):
        for j in range(ne):
            it += 1
            nn, nb, em = _procs(data, i, j)
            mdlib.Register_ForceTableProc(
                it, FTable, nn, nb, em,
                NOTE=f"setfl table: {data.elements[i]}<-{data.elements[j]} ( {fname} )")
    FTable.Rmax = float(data.cutoff)
    FTable.RHOMX = float(data.rhomx)
    return list(data.elements)


def Generate_NIST_ForceTable(FTable: MDForceTable):
    """Generate_NIST_ForceTalbe (NIST_ForceTable.F90:356-426): tables of every registered
    procedure set on r_k = (k/CSI)^2, CSI = NTAB/sqrt(Rmax), and on rho_k = (k-1) RHOD with the
    file's RHOMX, RHOD = RHOMX/NEMBD -- the tables Export_ForceTable writes.  With the NE^2 sets
    of one setfl file, KPAIR/KEMBD also map group I to element I, so the table can go to the
    device as it is."""
    if FTable.Rmax <= 0.0:
        raise ValueError("MDPSCU Error: the distance range is missed in generating force table")
    if FTable.NTAB <= 0:
        raise ValueError("MDPSCU Error: the number of grid is less than zero in generating force "
                         "table")
    rhomax = FTable.RHOMX
    rhod = FTable.RHOMX / float(FTable.NEMBD)
    ids = sorted(FTable.REGPROCS)
    num = len(ids)
    nr, ne = FTable.NTAB, FTable.NEMBD
    FTable.RmaxSQRT = math.sqrt(FTable.Rmax)
    FTable.CSI = float(nr) / FTable.RmaxSQRT
    FTable.CSIV = 1.0 / FTable.CSI
    FTable.NKIND = num
    FTable.FPAIR = list(ids)
    FTable.HASREG = {}
    for name in ("POTR", "FPOTR", "POTB", "FPOTB"):
        setattr(FTable, name, np.zeros((num, nr), dtype=np.float64, order="F"))
    FTable.RHOMX = -1.0
    for iforce in ids:
        mdlib.Create_Pairwise_ForceTable(iforce, FTable)
    FTable.RHOMX = rhomax
    FTable.RHOD = rhod
    FTable.NKIND1 = num
    FTable.FPAIR1 = list(ids)
    FTable.HASREG1 = {}
    FTable.FEMBD = np.zeros((num, ne), dtype=np.float64, order="F")
    FTable.DFEMBD = np.zeros((num, ne), dtype=np.float64, order="F")
    for iforce in ids:
        mdlib.Create_EMBDFUNTable(iforce, FTable)
    # groups = elements in file order: KPAIR(I, J) = (I-1) NE + J, KEMBD(I) = KPAIR(I, I)
    n = int(round(math.sqrt(num)))
    if n * n == num and ids == list(range(1, num + 1)):
        FTable.KPAIR = np.asfortranarray(np.arange(1, num + 1, dtype=np.int32).reshape(n, n))
        FTable.KEMBD = np.asarray([k * n + k + 1 for k in range(n)], dtype=np.int32)
    return FTable


# ----------------------------------------------------------------------------------------------
# moldy files (Filedatas_Func_Moldy.F90): Finnis-Sinclair cubic-knot forms
# ----------------------------------------------------------------------------------------------

@dataclass
class MoldyFile:
    """get_filedatas_moldy (Filedatas_Func_Moldy.F90:32-85): a symbol line, then one line each of
    pair coefficients a_k, pair knots r_k, density coefficients A_k, density knots R_k and
    constants (the first the lattice constant a0).  Knots are given in a0, coefficients per a0^3
    (pair) and a0^6 (density); stored here converted to A and eV."""
    symbol: str
    a_v: np.ndarray
    r_v: np.ndarray
    a_rho: np.ndarray
    r_rho: np.ndarray
    const: np.ndarray


def read_moldy(path: str) -> MoldyFile:
    with open(path) as fh:
        lines = [l for l in fh.read().splitlines()]
    if len(lines) < 6:
        raise ValueError(f"MDPSCU Error: moldy file {path} needs six lines")
    symbol = lines[0].split()[0] if lines[0].split() else ""
    rows = [np.asarray(_numbers(l.split()), dtype=np.float64) for l in lines[1:6]]
    a_v, r_v, a_rho, r_rho, const = rows
    if len(a_v) != len(r_v) or len(a_rho) != len(r_rho) or len(const) < 1:
        raise ValueError(f"MDPSCU Error: moldy file {path}: coefficient and knot counts differ")
    a0 = const[0]
    return MoldyFile(symbol, a_v / a0 ** 3, r_v * a0, a_rho / a0 ** 6, r_rho * a0, const)


def _cubic_knots(r: float, a: np.ndarray, knots: np.ndarray):
    """sum_k a_k (R_k - r)^3 H(R_k - r) and 3 sum_k a_k (R_k - r)^2 H(R_k - r), i.e. the value and
    minus its slope (NN_FuncPoly3 / RHO_FuncPoly3, MD_Pot_EAM_Utilities.F90:48-85, 259-290)."""
    v = dv = 0.0
    for ak, rk in zip(a.tolist(), knots.tolist()):
        if rk - r > 0.0:
            d = rk - r
            v += ak * d * d * d
            dv += ak * d * d
    return v, 3.0 * dv


def Register_ForceTableProc_Moldy(fname: str, FTable: MDForceTable) -> str:
    """Register_ForceTableProc_Moldy (Filedatas_Func_Moldy.F90:130-150): table id 1 with pair
    (V/2, -dV/dr), density (rho, -drho/dr) and embedding F = -sqrt(rho) (EMBED_FS_PLOY_Func with
    A = (-1), zero for rho <= 0); returns the symbol."""
    data = read_moldy(fname)

    def nn(r):
        v, mdv = _cubic_knots(r, data.a_v, data.r_v)
        return 0.5 * v, mdv

    def ne(r):
        return _cubic_knots(r, data.a_rho, data.r_rho)

    def em(rho):
        if rho <= 0.0:
            return 0.0, 0.0
        q = math.sqrt(rho)
        return -q, -0.5 / q

    mdlib.Register_ForceTableProc(1, FTable, nn, ne, em, NOTE="Moldy table:" + fname)
    return data.symbol


# ----------------------------------------------------------------------------------------------
# lspt files (Filedatas_Func_Lspt.F90): a control file naming two-column sample files
# ----------------------------------------------------------------------------------------------

class _SplineXY:
    """The natural cubic spline of _Spline on arbitrary increasing abscissae (the .spt samples
    need not be evenly spaced)."""

    def __init__(self, x: np.ndarray, f: np.ndarray):
        x = np.asarray(x, dtype=np.float64)
        f = np.asarray(f, dtype=np.float64)
        n = len(x)
        if n < 3:
            raise ValueError("MDPSCU Error: a spline file needs at least 3 samples")
        h = np.diff(x)
        sl = np.diff(f) / h
        k = n - 2
        sub = h[:-1] / 6.0
        dia = (h[:-1] + h[1:]) / 3.0
        sup = h[1:] / 6.0
        rhs = sl[1:] - sl[:-1]
        c = np.zeros(k)
        d = np.zeros(k)
        c[0], d[0] = sup[0] / dia[0], rhs[0] / dia[0]
        for i in range(1, k):
            m = dia[i] - sub[i] * c[i - 1]
            c[i] = sup[i] / m
            d[i] = (rhs[i] - sub[i] * d[i - 1]) / m
        w = np.zeros(n)
        w[k] = d[k - 1]
        for i in range(k - 2, -1, -1):
            w[i + 1] = d[i] - c[i] * w[i + 2]
        self.x, self.f, self.w = x.tolist(), f.tolist(), w.tolist()
        self.n = n
        self.xmin, self.xmax = float(x.min()), float(x.max())

    def __call__(self, y: float) -> Tuple[float, float]:
        j = min(max(bisect.bisect_right(self.x, y) - 1, 0), self.n - 2)
        a, b = self.x[j], self.x[j + 1]
        h = b - a
        f0, f1, w0, w1 = self.f[j], self.f[j + 1], self.w[j], self.w[j + 1]
        v = ((w0 * (b - y) ** 3 + w1 * (y - a) ** 3) / (6.0 * h)
             + (f1 / h - w1 * h / 6.0) * (y - a) + (f0 / h - w0 * h / 6.0) * (b - y))
        dv = ((-w0 * (b - y) ** 2 + w1 * (y - a) ** 2) / (2.0 * h)
              + (f1 - f0) / h - (w1 - w0) * h / 6.0)
        return v, dv


def _quoted(line: str) -> List[str]:
    return re.findall(r'"([^"]*)"', line)


def _spt(path: str) -> _SplineXY:
    xs, fs = [], []
    with open(path) as fh:
        for line in fh:
            t = line.split("#")[0].split()
            if len(t) >= 2:
                xs.append(float(t[0].replace("D", "E")))
                fs.append(float(t[1].replace("D", "E")))
    return _SplineXY(np.asarray(xs), np.asarray(fs))


@dataclass
class LsptFile:
    """get_filedatas_lspt (Filedatas_Func_Lspt.F90:118-301): elements; per element F(rho) and
    rho(r) splines (None for "NA"); pair splines keyed (I, J), I >= J; RHOMX = the largest rho of
    the F files, RMAX = the largest r of the pair files (A)."""
    elements: List[str]
    frho: List[object]
    rhor: List[object]
    vr: Dict[Tuple[int, int], _SplineXY]
    rhomx: float
    rmax: float


def read_lspt(path: str) -> LsptFile:
    """The control file: a line opening with &LSPT, a line with the number of elements, a line
    with the quoted symbols, per element a quoted F(rho) file and a quoted rho(r) file ("NA" for
    none), then for I = 1..NE a line with the I quoted pair files of (I, 1..I).  Kept from the
    reference: every pair (I, J) of line I reads the line's FIRST file (the reader opens
    SUBSTR(1) for each J, :266), so (2,2) holds the W-H samples in the shipped W-H-He file."""
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as fh:
        lines = [l.split("!")[0].rstrip() for l in fh]
    lines = [l for l in lines if l.strip()]
    if not lines or not lines[0].strip().upper().startswith("&LSPT"):
        raise ValueError(f"MDPSCU Error: {path} does not start with &LSPT")
    nums = _numbers(lines[1].replace("=", " ").split())
    if not nums:
        raise ValueError("MDPSCU Error: the number of groups should be given")
    ne = int(nums[0])
    elements = _quoted(lines[2])
    if len(elements) != ne:
        raise ValueError("MDPSCU Error: the number of elements is different from the given number")
    frho, rhor = [], []
    k = 3
    for _ in range(ne):
        for dst in (frho, rhor):
            q = _quoted(lines[k])
            k += 1
            name = q[0] if q else "NA"
            dst.append(None if name.strip().upper() == "NA" else _spt(os.path.join(base, name)))
    vr: Dict[Tuple[int, int], _SplineXY] = {}
    for i in range(ne):
        q = _quoted(lines[k])
        k += 1
        if len(q) != i + 1:
            raise ValueError(f"MDPSCU Error: {i + 1} pair files expected on line {k} of {path}")
        sp = _spt(os.path.join(base, q[0]))
        for j in range(i + 1):
            vr[(i, j)] = sp
    rhomx = max([f.xmax for f in frho if f is not None] + [0.0])
    rmax = max([v.xmax for v in vr.values()] + [0.0])
    return LsptFile(elements, frho, rhor, vr, rhomx, rmax)


def Register_ForceTableProc_SPT(fname: str, FTable: MDForceTable) -> List[str]:
    """Register_ForceTableProc_SPT (Filedatas_Func_Lspt.F90:560-602): ids IT = (I-1) NE + J with
    pair V of the unordered pair given directly (not r V), (V/2, -dV/dr); density rho_J unless
    element I has no F file or J no rho file; embedding F_I for I = J.  Below a file's first
    sample the value is held with zero slope; above its last, pair and density are zero and the
    embedding holds its value.  Sets Rmax and RHOMX from the files."""
    data = read_lspt(fname)
    ne = len(data.elements)
    it = 0
    for i in range(ne):
        for j in range(ne):
            it += 1
            vs = data.vr[(max(i, j), min(i, j))]

            def nn(r, vs=vs):
                if r > vs.xmax:
                    return 0.0, 0.0
                if r < vs.xmin:
                    return 0.5 * vs(vs.xmin)[0], 0.0
                v, dv = vs(r)
                return 0.5 * v, -dv

            rs = data.rhor[j] if data.frho[i] is not None else None

            def nb(r, rs=rs):
                if rs is None or r > rs.xmax:
                    return 0.0, 0.0
                if r < rs.xmin:
                    return rs(rs.xmin)[0], 0.0
                v, dv = rs(r)
                return v, -dv

            fs = data.frho[i] if i == j else None

            def em(rho, fs=fs):
                if fs is None:
                    return 0.0, 0.0
                if rho < fs.xmin:
                    return fs(fs.xmin)[0], 0.0
                if rho > fs.xmax:
                    return fs(fs.xmax)[0], 0.0
                return fs(rho)

            mdlib.Register_ForceTableProc(
                it, FTable, nn, nb, em,
                NOTE=f"lspt table: {data.elements[i]}<-{data.elements[j]}  ( {fname} )")
    FTable.Rmax = float(data.rmax)
    FTable.RHOMX = float(data.rhomx)
    return list(data.elements)
