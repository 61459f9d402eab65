"""Host-side mirror of the reference MDLIB surface for the force hot path.

Names, argument meaning and error behaviour follow the reference modules so that callers (and the
tests) read like MDLIB code:

* ``MDForceTable``, ``Register_ForceTableProc``, ``New_ForceTable``,
  ``Create_Pairwise_ForceTable``, ``Create_EMBDFUNTable``, ``Create_Interaction_ForceTable``,
  ``Export_ForceTable``, ``Import_ForceTable``, ``Register_Imported_ForceTable``
  -- MDLIB/sor/Common/MD_TypeDef_ForceTable.F90:117-155, 530-1056, 1151-1230, 1315-1855
* ``SimMDBox`` / ``SimMDCtrl`` (the fields this path reads)
  -- MDLIB/sor/Common/MD_TypeDef_SimBox.F90:43-132, MD_TypeDef_SimCtrlParam.F90:56-142
* ``MDForceClassGPU``, ``Register_ForceClass``, ``Init_Forcetable_Dev``, ``CalForce_ForceClass``,
  ``CalEpot_ForceClass`` -- MDLIB/sor/CommonGPU/MD_ForceClass_Register_GPU.F90:248-442, 483-822
* ``Initialize_Globle_Variables_DEV``, ``CopyOut_SimBox_DEV``
  -- MDLIB/sor/CommonGPU/MD_Globle_Variables_GPU.F90:515-654
* ``Cal_NeighBoreList_DEV``, ``Copyout_NeighboreList_DEV``
  -- MDLIB/sor/CommonGPU/MD_NeighborsList_GPU.F90:121-181, 604-693
* ``Predictor_DEV``, ``Correction_DEV`` -- MDLIB/sor/CommonGPU/MD_DiffScheme_GPU.F90:604-847
* ``For_One_Step`` loop -- MDLIB/sor/Appshell/MD_Method_GenericMD_GPU.F90:496-659

All compute goes through libmsm_hip.so; where the reference stops the process, these functions
raise ``MSMError`` (device) or ``ValueError`` (host setup).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Union

import numpy as np

from . import _lib
from .lbfgs import LBFGS
from .lbfgsb import LBFGSB
from ._lib import MSM_EAM_TYPE, MSM_FS_TYPE, check, dptr, iptr, lib

PotFn = Union[str, Callable[[float], tuple], None]

CP_KB_EV = 8.617333262e-5  # Boltzmann constant, eV/K
# the reference's CGS constants (MSMLIB/sor/Common/MSM_Const.F90:82-85)
CP_KB_CGS = 1.38054e-16    # Boltzmann constant, erg/K
CP_EVERG = 1.60219e-12     # erg per eV
CP_CM2A = 1.0e8            # Angstrom per cm
# STATU bits (MD_Const.F90:80-96)
CP_STATU_ACTIVE, CP_STATU_FIXPOSX, CP_STATU_OUTOFBOX = 1, 2, 2 ** 16
# quench schemes (MD_Const.F90:23-27); LSEARCH is a bit added to the scheme
CP_DAMPSCHEME_LBFGS, CP_DAMPSCHEME_DYN, CP_DAMPSCHEME_CG, CP_DAMPSCHEME_ST = 0, 1, 2, 3
CP_DAMPSCHEME_LSEARCH = 2 ** 16



# ----------------------------------------------------------------------------------------------
# force tables
# ----------------------------------------------------------------------------------------------

@dataclass
class _ExtProc:
    POTTypeID: int
    POTSymbID: str
    NNFORCE: PotFn
    NEFORCE: PotFn
    EMBDF: PotFn
    NOTE: str


@dataclass
class MDForceTable:
    """MDForceTable (MD_TypeDef_ForceTable.F90:117-155).

    Table arrays keep the reference's Fortran memory order: ``POTR`` has shape (NKIND, NTAB) in
    Fortran order, i.e. element (k, i) at flat offset k + i*NKIND, which is what the C-ABI takes.
    """
    PotType: str = "FS_TYPE"
    NKIND: int = 0
    NTAB: int = 10000
    Rmax: float = -1.0
    RmaxSQRT: float = 0.0
    CSI: float = 0.0
    CSIV: float = 0.0
    FPAIR: List[int] = field(default_factory=list)
    KPAIR: Optional[np.ndarray] = None
    POTR: Optional[np.ndarray] = None
    FPOTR: Optional[np.ndarray] = None
    POTB: Optional[np.ndarray] = None
    FPOTB: Optional[np.ndarray] = None
    NEMBD: int = 10000
    NKIND1: int = 0
    FPAIR1: List[int] = field(default_factory=list)
    KEMBD: Optional[np.ndarray] = None
    FEMBD: Optional[np.ndarray] = None
    DFEMBD: Optional[np.ndarray] = None
    RHOMX: float = -1.0
    RHOD: float = 1.0
    R: Optional[np.ndarray] = None    # grid of an imported table (Import_ForceTable)
    RHO: Optional[np.ndarray] = None  # embedding grid of an imported table
    RHOMULT: float = 20.0
    REGPROCS: Dict[int, _ExtProc] = field(default_factory=dict)
    HASREG: Dict[int, bool] = field(default_factory=dict)
    HASREG1: Dict[int, bool] = field(default_factory=dict)


class _CallbackErrors(list):
    """Exceptions raised inside Python potential callbacks during one table evaluation.  ctypes
    cannot carry an exception through the C caller, so the trampoline records it here, writes NaN
    to the outputs and the caller re-raises the first one once the C routine has returned."""


def _fn_address(fn: PotFn, keep: list):
    """C callable for a potential: a built-in name ("msm_pot_wfs_nn") or a Python callable
    r -> (value, derivative-term).  ``keep`` holds the ctypes objects alive for the C call; its
    first element, a _CallbackErrors, collects the callbacks' exceptions."""
    if fn is None:
        return None
    if isinstance(fn, str):
        return _lib.builtin_fn(fn)
    if not keep or not isinstance(keep[0], _CallbackErrors):
        keep.insert(0, _CallbackErrors())
    errors = keep[0]

    def tramp(rp, vp, dvp, _f=fn):
        try:
            v, dv = _f(rp[0])
            vp[0] = v
            dvp[0] = dv
        except BaseException as exc:  # noqa: BLE001 -- re-raised by _raise_callback_errors
            vp[0] = float("nan")
            dvp[0] = float("nan")
            errors.append(exc)

    cb = _lib.PAIRFN(tramp)
    keep.append(cb)
    return C.cast(cb, C.c_void_p).value


def _raise_callback_errors(keep: list, what: str):
    if keep and isinstance(keep[0], _CallbackErrors) and keep[0]:
        exc = keep[0][0]
        raise RuntimeError(f"{what}: a potential callback raised {type(exc).__name__}: {exc}") \
            from exc


def Register_ForceTableProc(IFORCE, FTable: MDForceTable, NNFORCE: PotFn, NEFORCE: PotFn = None,
                            EMBDF: PotFn = None, NOTE: str = "", Tgt: Optional[str] = None):
    """RegisterByTypeID_ForceTableProc / RegisterBySymbID_ForceTableProc (:626-852).

    ``IFORCE`` is the integer type id; with ``Tgt`` given, ``IFORCE`` is read as the source symbol
    and the id is assigned as max(existing)+1 with symbol "TGT<-SOR"."""
    if Tgt is not None:
        symb = (Tgt.strip() + "<-" + str(IFORCE).strip()).upper()
        if any(p.POTSymbID == symb for p in FTable.REGPROCS.values()):
            raise ValueError(f"MDPSCU Error: the potential type by symb {symb} has been registed")
        tid = max([0] + list(FTable.REGPROCS)) + 1
        FTable.REGPROCS[tid] = _ExtProc(tid, symb, NNFORCE, NEFORCE, EMBDF, NOTE)
        return tid
    if IFORCE in FTable.REGPROCS:
        raise ValueError(f"MDPSCU Error: the potential type #{IFORCE} has been registed")
    FTable.REGPROCS[IFORCE] = _ExtProc(IFORCE, "", NNFORCE, NEFORCE, EMBDF, NOTE)
    return IFORCE


def SymbID2TypeID_ForceTableProc(Sor: str, Tgt: str, FTable: MDForceTable) -> int:
    symb = (Tgt.strip() + "<-" + Sor.strip()).upper()
    for tid, p in FTable.REGPROCS.items():
        if p.POTSymbID == symb:
            return tid
    return 0


def New_ForceTable(SimBox: "SimMDBox", CtrlParam: "SimMDCtrl", FTable: MDForceTable):
    """New_ForceTable (:530-622): distinct pair types -> kinds, grid constants, zeroed tables."""
    n = SimBox.NGROUP
    fpair: List[int] = []
    for i in range(n):
        for j in range(n):
            t = int(SimBox.PTYPE[i, j])
            if t not in fpair:
                fpair.append(t)
    FTable.FPAIR = fpair
    FTable.NKIND = len(fpair)
    nr = int(CtrlParam.NUMFTABR)
    FTable.NTAB = nr
    FTable.KPAIR = np.zeros((n, n), dtype=np.int32, order="F")
    for name in ("POTR", "FPOTR", "POTB", "FPOTB"):
        setattr(FTable, name, np.zeros((FTable.NKIND, nr), dtype=np.float64, order="F"))
    FTable.Rmax = float(np.max(CtrlParam.RU))
    FTable.RmaxSQRT = float(np.sqrt(FTable.Rmax))
    FTable.CSI = float(nr) / FTable.RmaxSQRT
    FTable.CSIV = 1.0 / FTable.CSI
    FTable.RHOMX = -1.0
    FTable.HASREG = {}
    FTable.HASREG1 = {}
    if FTable.PotType.strip() == "EAM_TYPE":
        fpair1: List[int] = []
        for i in range(n):
            t = int(SimBox.PTYPE[i, i])
            if t not in fpair1:
                fpair1.append(t)
        FTable.FPAIR1 = fpair1
        FTable.NKIND1 = len(fpair1)
        FTable.NEMBD = int(CtrlParam.NUMFTABE)
        FTable.KEMBD = np.zeros(n, dtype=np.int32)
        FTable.FEMBD = np.zeros((FTable.NKIND1, FTable.NEMBD), dtype=np.float64, order="F")
        FTable.DFEMBD = np.zeros((FTable.NKIND1, FTable.NEMBD), dtype=np.float64, order="F")
        FTable.RHOMULT = float(CtrlParam.RHOSCAL)


def _col(a: np.ndarray, k: int):
    """Pointer to element (k, 0) of a Fortran-ordered (NKIND, NTAB) array (stride NKIND)."""
    base = a.ctypes.data + k * a.itemsize
    return C.cast(C.c_void_p(base), C.POINTER(C.c_double))


def Create_Pairwise_ForceTable(IFORCE: int, FTable: MDForceTable):
    """Create_Pairwise_ForceTable (:890-982), evaluated by msm_table_pairwise."""
    if IFORCE not in FTable.FPAIR:
        return
    k = FTable.FPAIR.index(IFORCE)
    if FTable.HASREG.get(k):
        return
    proc = FTable.REGPROCS.get(IFORCE)
    if proc is None:
        raise ValueError(f"MDPSCU Error: the potential type #{IFORCE} is not available")
    keep: list = []
    nn = _fn_address(proc.NNFORCE, keep)
    ne = _fn_address(proc.NEFORCE, keep)
    rhomx = C.c_double(FTable.RHOMX)
    rc = lib().msm_table_pairwise(FTable.NTAB, FTable.Rmax, nn, ne, FTable.RHOMULT,
                                  C.byref(rhomx), _col(FTable.POTR, k), _col(FTable.FPOTR, k),
                                  _col(FTable.POTB, k), _col(FTable.FPOTB, k), FTable.NKIND)
    _raise_callback_errors(keep, f"Create_Pairwise_ForceTable (potential type #{IFORCE})")
    check(rc, "msm_table_pairwise")
    FTable.RHOMX = rhomx.value
    FTable.RHOD = FTable.RHOMX / float(FTable.NEMBD)
    FTable.HASREG[k] = True


def Create_EMBDFUNTable(IFORCE: int, FTable: MDForceTable):
    """Create_EMBDFUNTable (:986-1056), evaluated by msm_table_embd."""
    if IFORCE not in FTable.FPAIR1:
        return
    k = FTable.FPAIR1.index(IFORCE)
    if FTable.HASREG1.get(k):
        return
    proc = FTable.REGPROCS.get(IFORCE)
    if proc is None:
        raise ValueError(f"MDPSCU Error: the embedment function of potential type #{IFORCE} "
                         "is not available")
    keep: list = []
    em = _fn_address(proc.EMBDF, keep)
    rc = lib().msm_table_embd(FTable.NEMBD, FTable.RHOD, em, _col(FTable.FEMBD, k),
                              _col(FTable.DFEMBD, k), FTable.NKIND1)
    _raise_callback_errors(keep, f"Create_EMBDFUNTable (potential type #{IFORCE})")
    check(rc, "msm_table_embd")
    FTable.HASREG1[k] = True


def Create_Interaction_ForceTable(SimBox: "SimMDBox", CtrlParam: "SimMDCtrl",
                                  FTable: MDForceTable):
    """Create_Interaction_ForceTable (:1151-1230)."""
    n = SimBox.NGROUP
    for i in range(n):
        for j in range(n):
            if SimBox.PTYPE[i, j] == 0 and SimBox.SYMB:
                SimBox.PTYPE[i, j] = SymbID2TypeID_ForceTableProc(SimBox.SYMB[j], SimBox.SYMB[i],
                                                                  FTable)
            if SimBox.PTYPE[i, j] == 0:
                raise ValueError("MDPSCU Error: cannot find the registed force table for between "
                                 f"atom groups {i + 1} and {j + 1}")
    New_ForceTable(SimBox, CtrlParam, FTable)
    FTable.KPAIR[:, :] = -1
    for i in range(n):
        for j in range(n):
            iforce = int(SimBox.PTYPE[i, j])
            Create_Pairwise_ForceTable(iforce, FTable)
            FTable.KPAIR[i, j] = FTable.FPAIR.index(iforce) + 1
    if FTable.PotType.strip() == "EAM_TYPE":
        FTable.KEMBD[:] = -1
        for i in range(n):
            iforce = int(SimBox.PTYPE[i, i])
            Create_EMBDFUNTable(iforce, FTable)
            FTable.KEMBD[i] = FTable.FPAIR1.index(iforce) + 1
        if FTable.RHOD <= 1.0e-64:
            FTable.RHOD = 1.0


# ----------------------------------------------------------------------------------------------
# box and control parameters
# ----------------------------------------------------------------------------------------------

@dataclass
class SimMDBox:
    """The SimMDBox fields this path reads or writes (MD_TypeDef_SimBox.F90:43-132).

    XP, XP1, FP, DIS have shape (NPRT, 3) in Fortran order (x column, then y, then z)."""
    NPRT: int
    NGROUP: int
    ZL: np.ndarray
    BOXLOW: np.ndarray
    XP: np.ndarray
    XP1: np.ndarray
    ITYP: np.ndarray
    STATU: np.ndarray
    CM: np.ndarray
    PTYPE: np.ndarray
    SYMB: List[str] = field(default_factory=list)
    FP: Optional[np.ndarray] = None
    DIS: Optional[np.ndarray] = None
    EPOT: Optional[np.ndarray] = None
    VTENSOR: Optional[np.ndarray] = None
    RR: float = 1.0  # lattice unit in this layer's length unit (the reference keeps it in cm)
    EKIN: Optional[np.ndarray] = None

    @property
    def BOXUP(self) -> np.ndarray:
        return self.BOXLOW + self.ZL

    def __post_init__(self):
        self.ZL = np.asarray(self.ZL, dtype=np.float64)
        self.BOXLOW = np.asarray(self.BOXLOW, dtype=np.float64)
        self.XP = np.asfortranarray(self.XP, dtype=np.float64)
        self.XP1 = np.asfortranarray(self.XP1, dtype=np.float64)
        self.ITYP = np.ascontiguousarray(self.ITYP, dtype=np.int32)
        self.STATU = np.ascontiguousarray(self.STATU, dtype=np.int32)
        self.CM = np.ascontiguousarray(self.CM, dtype=np.float64)
        self.PTYPE = np.asarray(self.PTYPE, dtype=np.int32)
        if self.FP is None:
            self.FP = np.zeros((self.NPRT, 3), dtype=np.float64, order="F")
        if self.DIS is None:
            self.DIS = np.zeros((self.NPRT, 3), dtype=np.float64, order="F")
        if self.EPOT is None:
            self.EPOT = np.zeros(self.NPRT, dtype=np.float64)


@dataclass
class SimMDCtrl:
    """The SimMDCtrl fields this path reads (MD_TypeDef_SimCtrlParam.F90:56-142).  RU and NB_RM
    are (NGROUP, NGROUP) in length units (the reference converts from lattice units at input)."""
    RU: np.ndarray
    NB_RM: np.ndarray
    NB_UPTAB: int = 10
    NB_MXNBS: int = 64
    NUMFTABR: int = 5000
    NUMFTABE: int = 5000
    RHOSCAL: float = 20.0
    IFPD: Sequence[int] = (1, 1, 1)
    H: float = 1.0
    MULTIBOX: int = 1
    # quench controls (MD_TypeDef_SimCtrlParam.F90:213-217): alpha of the first steepest step,
    # smallest and largest move (in lattice units RR), energy tolerance (eV)
    STEEPEST_Alpha: float = 0.1
    STEEPEST_MiStep: float = 1.0e-5
    STEEPEST_MxStep: float = 0.1
    STEEPEST_MiDelE: float = 1.0e-3
    # L-BFGS(-B) quench controls (MD_TypeDef_SimCtrlParam.F90:208-210): corrections kept, and
    # the FACTR / PGTOL of SETULB (scale-dependent: the quench hands SETULB CGS values)
    LBFGS_Msave: int = 7
    LBFGS_Factr: float = 0.0
    LBFGS_PGtol: float = 0.0
    # time step control (MD_TypeDef_SimCtrlParam.F90:36-47, 78-79): IHDUP < 0 searches the step
    # from HMX every |IHDUP| steps so that no atom moves more than DMX; damping is applied in
    # the predictor for steps DAMPTIME0 .. DAMPTIME0 + DAMPTIME1 - 1
    IT0: int = 1
    IHDUP: int = 0
    HMI: float = 1.0
    HMX: float = 1.0
    DMX: float = 0.1
    DAMPTIME0: int = 1
    DAMPTIME1: int = 0
    # Boltzmann constant in the caller's energy unit per kelvin: eV/K for eV-Angstrom callers,
    # CP_KB_CGS for the reference's CGS arrays (its CP_KB, MSM_Const.F90:82); every temperature
    # routine of this layer reads it from here
    KB: float = CP_KB_EV
    # random seeds (MD_TypeDef_SimCtrlParam.F90:51); SEED[0], SEED[1] key the device generator of
    # Thermalizing_MC_DEV
    SEED: Sequence[int] = (1, 2, 3, 4)

    def __post_init__(self):
        self.RU = np.asfortranarray(np.atleast_2d(self.RU), dtype=np.float64)
        self.NB_RM = np.asfortranarray(np.atleast_2d(self.NB_RM), dtype=np.float64)


# ----------------------------------------------------------------------------------------------
# device side
# ----------------------------------------------------------------------------------------------

class DeviceMD:
    """One libmsm_hip context: the device copy of the boxes, the neighbour list and the force
    table (dm_WorkSpace, dm_Neighbors, dm_FTableWS of the reference)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().msm_create(C.byref(h), device), "msm_create")
        self.h = h
        self.n = 0
        self.nbox = 0
        self.mx = 0
        self.natoms: List[int] = []

    # ---- setup ----
    def set_force_table(self, FTable: MDForceTable):
        pt = MSM_EAM_TYPE if FTable.PotType.strip() == "EAM_TYPE" else MSM_FS_TYPE
        ng = FTable.KPAIR.shape[0]
        kpair = np.ascontiguousarray(np.asarray(FTable.KPAIR, dtype=np.int32).ravel(order="F"))
        potr = np.ascontiguousarray(FTable.POTR.ravel(order="F"))
        fpotr = np.ascontiguousarray(FTable.FPOTR.ravel(order="F"))
        potb = np.ascontiguousarray(FTable.POTB.ravel(order="F"))
        fpotb = np.ascontiguousarray(FTable.FPOTB.ravel(order="F"))
        if pt == MSM_EAM_TYPE:
            kembd = np.ascontiguousarray(FTable.KEMBD, dtype=np.int32)
            fembd = np.ascontiguousarray(FTable.FEMBD.ravel(order="F"))
            dfembd = np.ascontiguousarray(FTable.DFEMBD.ravel(order="F"))
            rc = lib().msm_set_force_table(self.h, pt, ng, FTable.NKIND, FTable.NTAB, FTable.CSI,
                                           iptr(kpair), dptr(potr), dptr(fpotr), dptr(potb),
                                           dptr(fpotb), FTable.NKIND1, FTable.NEMBD, FTable.RHOD,
                                           iptr(kembd), dptr(fembd), dptr(dfembd))
        else:
            rc = lib().msm_set_force_table(self.h, pt, ng, FTable.NKIND, FTable.NTAB, FTable.CSI,
                                           iptr(kpair), dptr(potr), dptr(fpotr), dptr(potb),
                                           dptr(fpotb), 0, 0, 1.0, None, None, None)
        check(rc, "msm_set_force_table", self.h)

    def set_ctrl(self, CtrlParam: SimMDCtrl, CM: np.ndarray):
        ng = CtrlParam.RU.shape[0]
        ru = np.ascontiguousarray(CtrlParam.RU.ravel(order="F"))
        nb = np.ascontiguousarray(CtrlParam.NB_RM.ravel(order="F"))
        cm = np.ascontiguousarray(CM, dtype=np.float64)
        check(lib().msm_set_ctrl(self.h, ng, dptr(ru), dptr(nb), int(CtrlParam.NB_MXNBS),
                                 dptr(cm)), "msm_set_ctrl", self.h)
        self.mx = int(CtrlParam.NB_MXNBS)

    def set_boxes(self, boxes: Sequence[SimMDBox], ifpd: Sequence[int]):
        nb = len(boxes)
        nat = np.array([b.NPRT for b in boxes], dtype=np.int32)
        lo = np.ascontiguousarray(np.concatenate([b.BOXLOW for b in boxes]), dtype=np.float64)
        ln = np.ascontiguousarray(np.concatenate([b.ZL for b in boxes]), dtype=np.float64)
        pd = np.ascontiguousarray(np.tile(np.asarray(ifpd, dtype=np.int32), nb))
        check(lib().msm_set_boxes(self.h, nb, iptr(nat), dptr(lo), dptr(ln), iptr(pd)),
              "msm_set_boxes", self.h)
        self.nbox = nb
        self.natoms = [int(x) for x in nat]

    def set_atoms(self, xp, xp1, ityp, statu):
        n = int(np.asarray(ityp).shape[0])
        xp = np.ascontiguousarray(np.asarray(xp, dtype=np.float64).ravel(order="F"))
        xp1 = np.ascontiguousarray(np.asarray(xp1, dtype=np.float64).ravel(order="F"))
        ityp = np.ascontiguousarray(ityp, dtype=np.int32)
        statu = np.ascontiguousarray(statu, dtype=np.int32)
        check(lib().msm_set_atoms(self.h, n, dptr(xp), dptr(xp1), iptr(ityp), iptr(statu)),
              "msm_set_atoms", self.h)
        self.n = n

    # ---- compute ----
    def build_neighbors(self):
        check(lib().msm_build_neighbors(self.h), "msm_build_neighbors", self.h)

    def calc_force(self):
        check(lib().msm_calc_force(self.h), "msm_calc_force", self.h)

    def calc_density(self):
        check(lib().msm_calc_density(self.h), "msm_calc_density", self.h)

    def calc_epot(self):
        check(lib().msm_calc_epot(self.h), "msm_calc_epot", self.h)

    def calc_ptensor(self) -> np.ndarray:
        """Virial tensor per box, shape (nbox, 3, 3) (msm_calc_ptensor)."""
        v = np.zeros(9 * self.nbox)
        check(lib().msm_calc_ptensor(self.h, dptr(v)), "msm_calc_ptensor", self.h)
        return v.reshape(self.nbox, 3, 3).transpose(0, 2, 1).copy()

    def calc_atomic_stress(self) -> np.ndarray:
        """Per-atom stress AVP(N, 9) in caller order (msm_calc_atomic_stress): components P11,
        P12, P13, P21, ..., P33 of sum_j s_a s_b FORTOT; Fortran-ordered (N, 9) array."""
        v = np.zeros(9 * self.n)
        check(lib().msm_calc_atomic_stress(self.h, dptr(v)), "msm_calc_atomic_stress", self.h)
        return v.reshape(9, self.n).T.copy(order="F")

    def damping(self):
        check(lib().msm_damping(self.h), "msm_damping", self.h)

    def calc_ekin(self, kb: float = CP_KB_EV):
        """(per-atom EK in caller order, -1e32 for atoms that do not count; temperature per
        box) -- msm_calc_ekin."""
        ek = np.zeros(self.n)
        tb = np.zeros(self.nbox)
        check(lib().msm_calc_ekin(self.h, kb, dptr(ek), dptr(tb)), "msm_calc_ekin", self.h)
        return ek, tb

    def thermalize(self, temp: float, kb: float, seed: int):
        check(lib().msm_thermalize(self.h, float(temp), float(kb), int(seed) & ((1 << 64) - 1)),
              "msm_thermalize", self.h)

    def scale_velocities(self, temp: float, kb: float = CP_KB_EV):
        check(lib().msm_scale_velocities(self.h, temp, kb), "msm_scale_velocities", self.h)

    def dyn_damp(self, maxsteps: int, mindele: float, h: float):
        """Damped dynamics (msm_dyn_damp): returns (iteration it stopped at, last max |dEPOT|)."""
        st = np.zeros(1, dtype=np.int32)
        de = np.zeros(1)
        check(lib().msm_dyn_damp(self.h, int(maxsteps), float(mindele), float(h), iptr(st),
                                 dptr(de)), "msm_dyn_damp", self.h)
        return int(st[0]), float(de[0])

    def steepest(self, maxsteps: int, linesearch: bool, alpha: float, maxdis: float,
                 mindis: float, mindele: float):
        """Steepest-descent quench (msm_steepest): returns (steps, last max |dEPOT|, last largest
        move)."""
        st = np.zeros(1, dtype=np.int32)
        de = np.zeros(1)
        mm = np.zeros(1)
        check(lib().msm_steepest(self.h, int(maxsteps), int(bool(linesearch)), float(alpha),
                                 float(maxdis), float(mindis), float(mindele), iptr(st), dptr(de),
                                 dptr(mm)), "msm_steepest", self.h)
        return int(st[0]), float(de[0]), float(mm[0])

    def cg(self, maxsteps: int, linesearch: bool, maxdis: float, mindis: float, mindele: float):
        """Conjugate-gradient quench (msm_cg): returns (steps, last max |dEPOT|)."""
        st = np.zeros(1, dtype=np.int32)
        de = np.zeros(1)
        check(lib().msm_cg(self.h, int(maxsteps), int(bool(linesearch)), float(maxdis),
                           float(mindis), float(mindele), iptr(st), dptr(de)), "msm_cg", self.h)
        return int(st[0]), float(de[0])

    def put_atoms(self, xp: Optional[np.ndarray] = None, xp1: Optional[np.ndarray] = None):
        """Positions and/or velocities (N, 3) in the caller's order into the device state, no
        list rebuild (msm_put_atoms)."""
        def flat(a):
            if a is None:
                return None
            a = np.asarray(a, dtype=np.float64)
            if a.shape != (self.n, 3):
                raise ValueError(f"expected ({self.n}, 3), got {a.shape}")
            return np.ascontiguousarray(a.T).ravel()
        x, v = flat(xp), flat(xp1)
        check(lib().msm_put_atoms(self.h, None if x is None else dptr(x),
                                  None if v is None else dptr(v)), "msm_put_atoms", self.h)

    def add_forces(self, fext, on_device: bool = False):
        """Add FEXT (N, 3) in the caller's order to the device forces (msm_add_forces); with
        on_device, fext is the device address (an int) of an (N, 3) column-major array on this
        context's device."""
        if on_device:
            check(lib().msm_add_forces(self.h, C.c_void_p(int(fext)), 1), "msm_add_forces", self.h)
            return
        f = np.asarray(fext, dtype=np.float64)
        if f.shape != (self.n, 3):
            raise ValueError(f"expected ({self.n}, 3), got {f.shape}")
        f = np.ascontiguousarray(f.T).ravel()
        check(lib().msm_add_forces(self.h, dptr(f), 0), "msm_add_forces", self.h)

    def add_epot(self, eext, on_device: bool = False):
        """Add EEXT (N) in the caller's order to the per-atom energies (msm_add_epot)."""
        if on_device:
            check(lib().msm_add_epot(self.h, C.c_void_p(int(eext)), 1), "msm_add_epot", self.h)
            return
        e = np.ascontiguousarray(eext, dtype=np.float64)
        if e.shape != (self.n,):
            raise ValueError(f"expected ({self.n},), got {e.shape}")
        check(lib().msm_add_epot(self.h, dptr(e), 0), "msm_add_epot", self.h)

    def put_dis(self, dis):
        """DIS (N, 3) in the caller's order into the device state (msm_put_dis)."""
        d = np.asarray(dis, dtype=np.float64)
        if d.shape != (self.n, 3):
            raise ValueError(f"expected ({self.n}, 3), got {d.shape}")
        d = np.ascontiguousarray(d.T).ravel()
        check(lib().msm_put_dis(self.h, dptr(d)), "msm_put_dis", self.h)

    def check_timestep(self, th: float, mxd2: float) -> bool:
        """msm_check_timestep: True when some active atom would move more than sqrt(mxd2) in a
        step th."""
        fl = np.zeros(1, dtype=np.int32)
        check(lib().msm_check_timestep(self.h, float(th), float(mxd2), iptr(fl)),
              "msm_check_timestep", self.h)
        return bool(fl[0])

    def predictor(self, h: float):
        check(lib().msm_predictor(self.h, h), "msm_predictor", self.h)

    def correction(self, h: float):
        check(lib().msm_correction(self.h, h), "msm_correction", self.h)

    def run(self, nsteps: int, nb_uptab: int, step0: int, h: float):
        check(lib().msm_run(self.h, nsteps, nb_uptab, step0, h), "msm_run", self.h)

    def sync(self):
        check(lib().msm_sync(self.h), "msm_sync", self.h)

    def set_option(self, option: int, value: int):
        check(lib().msm_set_option(self.h, int(option), int(value)), "msm_set_option", self.h)

    def get_option(self, option: int) -> int:
        v = np.zeros(1, dtype=np.int32)
        check(lib().msm_get_option(self.h, int(option), iptr(v)), "msm_get_option", self.h)
        return int(v[0])

    def pass_kernel(self) -> str:
        """Name of the density/force kernels the current options select (MSM_PASS_*)."""
        from ._lib import MSM_OPT_PASS_KERNEL, PASS_KERNELS
        return PASS_KERNELS[self.get_option(MSM_OPT_PASS_KERNEL)]

    def autotune(self, reps: int = 3) -> dict:
        """Times the kernel variants on the current atoms and list and keeps the fastest
        (msm_autotune); returns the choice and the chosen per-pass times."""
        ch = np.zeros(8, dtype=np.int32)
        ms = np.zeros(2)
        check(lib().msm_autotune(self.h, int(reps), iptr(ch), dptr(ms)), "msm_autotune", self.h)
        from ._lib import MSM_OPT_FORCE_LANES
        return {"force_path": int(ch[0]), "table_lds": int(ch[1]),
                "lt_shape": [int(v) for v in ch[2:6]], "nlist_kernel": int(ch[6]),
                "brick_lanes": int(ch[7]), "force_lanes": self.get_option(MSM_OPT_FORCE_LANES),
                "density_force_ms": float(ms[0]), "rebuild_ms": float(ms[1])}

    def kernel_times(self, reset: bool = True) -> dict:
        from ._lib import KCLASS
        ms = np.zeros(len(KCLASS))
        cnt = np.zeros(len(KCLASS), dtype=np.int64)
        check(lib().msm_kernel_times(self.h, dptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_longlong)),
                                     len(KCLASS), 1 if reset else 0), "msm_kernel_times", self.h)
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KCLASS)}

    # ---- copy-out ----
    def get_atoms(self):
        n = self.n
        xp = np.zeros(3 * n)
        xp1 = np.zeros(3 * n)
        fp = np.zeros(3 * n)
        dis = np.zeros(3 * n)
        st = np.zeros(n, dtype=np.int32)
        check(lib().msm_get_atoms(self.h, dptr(xp), dptr(xp1), dptr(fp), iptr(st), dptr(dis)),
              "msm_get_atoms", self.h)
        f = lambda a: a.reshape(3, n).T.copy(order="F")  # noqa: E731
        return f(xp), f(xp1), f(fp), st, f(dis)

    def get_forces(self) -> np.ndarray:
        """FP (N, 3) in the caller's order alone (msm_get_atoms with the other outputs NULL)."""
        n = self.n
        fp = np.zeros(3 * n)
        check(lib().msm_get_atoms(self.h, None, None, dptr(fp), None, None), "msm_get_atoms",
              self.h)
        return fp.reshape(3, n).T.copy(order="F")

    def get_den(self) -> np.ndarray:
        d = np.zeros(self.n)
        check(lib().msm_get_den(self.h, dptr(d)), "msm_get_den", self.h)
        return d

    def get_epot(self) -> np.ndarray:
        d = np.zeros(self.n)
        check(lib().msm_get_epot(self.h, dptr(d)), "msm_get_epot", self.h)
        return d

    def get_box_epot(self) -> np.ndarray:
        d = np.zeros(self.nbox)
        check(lib().msm_get_box_epot(self.h, dptr(d)), "msm_get_box_epot", self.h)
        return d

    def get_neighbors(self):
        n = self.n
        kv = np.zeros(n, dtype=np.int32)
        ind = np.zeros(n * self.mx, dtype=np.int32)
        check(lib().msm_get_neighbors(self.h, iptr(kv), iptr(ind), n), "msm_get_neighbors",
              self.h)
        return kv, ind.reshape(self.mx, n).T  # INDI(N, MXKVOIS), 1-based ids

    def list_stats(self):
        tot = C.c_longlong()
        mk = C.c_int()
        check(lib().msm_list_stats(self.h, C.byref(tot), C.byref(mk)), "msm_list_stats", self.h)
        return tot.value, mk.value

    def close(self):
        if self.h:
            lib().msm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class MDForceClassGPU:
    """MDForceClassGPU (MD_ForceClass_Register_GPU.F90:248-259): the force table, the device
    context it was initialised on and the external-force list (ExtForceList, :240-258): tagged
    (ForceArray, EpotArray) routines called after the force and energy passes."""
    ForceTable: MDForceTable = field(default_factory=MDForceTable)
    Device: Optional[DeviceMD] = None
    ExtForceList: List[tuple] = field(default_factory=list)


def AddExtForce_ForceClass(ForceClass: MDForceClassGPU, Tag: str, ForceArray=None,
                           EpotArray=None):
    """AddExtForce_ForceClass (MD_ForceClass_Register_GPU.F90:982-1180): register routines
    ForceArray(SimBox, CtrlParam, ForceClass) / EpotArray(SimBox, CtrlParam, ForceClass) that add
    their terms through ForceClass.Device.add_forces / add_epot (the reference's routines write
    the device arrays themselves).  A tag already present is replaced."""
    ForceClass.ExtForceList[:] = [e for e in ForceClass.ExtForceList if e[0] != Tag]
    ForceClass.ExtForceList.append((Tag, ForceArray, EpotArray))


def Clear_ExtForce(ForceClass: MDForceClassGPU, Tag: Optional[str] = None):
    """Clear_ExtForce: drop one tagged entry, or all of them."""
    ForceClass.ExtForceList[:] = ([] if Tag is None else
                                  [e for e in ForceClass.ExtForceList if e[0] != Tag])


def Cumulate_ExtForce(SimBox, CtrlParam, ForceClass: MDForceClassGPU):
    """Cumulate_ExtForce (:584-632): every registered ForceArray in registration order."""
    for _, fa, _ in ForceClass.ExtForceList:
        if fa is not None:
            fa(SimBox, CtrlParam, ForceClass)


def Register_ForceClass(classname: str, ForceClass: MDForceClassGPU):
    """Register_ForceClass (:363-442): FS_TYPE or EAM_TYPE."""
    name = classname.strip()
    if name not in ("FS_TYPE", "EAM_TYPE"):
        raise ValueError(f"MDPSCU Error: unsupported type of potential {classname}")
    ForceClass.ForceTable.PotType = name


def _boxes(SimBox) -> List[SimMDBox]:
    return list(SimBox) if isinstance(SimBox, (list, tuple)) else [SimBox]


def Initialize_Globle_Variables_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU,
                                    device: int = 0):
    """Initialize_Globle_Variables_DEV (MD_Globle_Variables_GPU.F90:515-654) generalised to
    boxes of uneven size: creates the device context, uploads control data, boxes and atoms."""
    boxes = _boxes(SimBox)
    if ForceClass.Device is None:
        ForceClass.Device = DeviceMD(device)
    dev = ForceClass.Device
    dev.set_ctrl(CtrlParam, boxes[0].CM)
    dev.set_boxes(boxes, CtrlParam.IFPD)
    xp = np.concatenate([b.XP for b in boxes], axis=0)
    xp1 = np.concatenate([b.XP1 for b in boxes], axis=0)
    ityp = np.concatenate([b.ITYP for b in boxes])
    st = np.concatenate([b.STATU for b in boxes])
    dev.set_atoms(xp, xp1, ityp, st)
    return dev


def Init_Forcetable_Dev(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU):
    """Init_Forcetable_Dev (:483-559): copy the registered table to the device."""
    if ForceClass.Device is None:
        raise ValueError("MDPSCU Error: device workspace not initialised")
    ForceClass.Device.set_force_table(ForceClass.ForceTable)


def Cal_NeighBoreList_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU):
    ForceClass.Device.build_neighbors()


def Copyout_NeighboreList_DEV(ForceClass: MDForceClassGPU):
    return ForceClass.Device.get_neighbors()


def CalForce_ForceClass(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU):
    """CalForce_ForceClass (MD_ForceClass_Register_GPU.F90:636-680): the device passes, then the
    external forces (Cumulate_ExtForce)."""
    ForceClass.Device.calc_force()
    Cumulate_ExtForce(SimBox, CtrlParam, ForceClass)


def CalEDen_ForceClass(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU) -> np.ndarray:
    """The pCalEDen slot (MD_ForceClass_Register_GPU.F90:254, 405, 416):
    CALDEN_FS/EAM_Force_Table2A_DEV (MD_FS_ForceTable_GPU.F90:875-930), the density pass alone.
    Returns F'(rho) per atom (the device's DEN) in the caller's order."""
    ForceClass.Device.calc_density()
    return ForceClass.Device.get_den()


def CalEpot_ForceClass(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU):
    """CalEpot_ForceClass (:778-822): the energy pass, then the external EpotArray routines."""
    ForceClass.Device.calc_epot()
    for _, _, ea in ForceClass.ExtForceList:
        if ea is not None:
            ea(SimBox, CtrlParam, ForceClass)


def CalPTensor_ForceClass(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU):
    """CalPTensor_ForceClass (MD_ForceClass_Register_GPU.F90:684-728): the virial tensor, as
    COPYOUT_VIRIALTENSOR returns it -- the mean over the boxes (MD_FS_ForceTable_GPU.F90:1415-1447)."""
    return ForceClass.Device.calc_ptensor().mean(axis=0)


def Cal_AtomicStress(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU) -> np.ndarray:
    """Cal_AtomicStress (Analysis/AtomicVirialStress_GPU.F90:734-773) through the force class's
    atomic-stress routine (CAL_FS/EAM_AtomicStress_KERNEL): hAVP(N, 9) in the caller's order."""
    return ForceClass.Device.calc_atomic_stress()


def Cal_GlobalT_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU,
                    kb: Optional[float] = None) -> float:
    """Cal_GlobalT_DEV (MD_DiffScheme_GPU.F90:1042-1055): 2 <EK> / (3 kB) over all atoms that
    count (active, no FIXPOS bit); kB is CtrlParam.KB unless given."""
    kb = CtrlParam.KB if kb is None else kb
    ek, _ = ForceClass.Device.calc_ekin(kb)
    m = ek >= 0.0
    return 2.0 * float(ek[m].sum()) / float(m.sum()) / (3.0 * kb)


def VelScaling_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU, DT: float,
                   kb: Optional[float] = None):
    """VelScaling_DEV (MD_DiffScheme_GPU.F90:1390-1448): scale each box to temperature DT
    (kB is CtrlParam.KB unless given)."""
    ForceClass.Device.scale_velocities(DT, CtrlParam.KB if kb is None else kb)


def thermal_seed(SEED: Sequence[int]) -> int:
    """64-bit key of the device generator from the first two control seeds."""
    return ((int(SEED[1]) & 0xFFFFFFFF) << 32) | (int(SEED[0]) & 0xFFFFFFFF)


def Thermalizing_MC_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU, TI: float,
                        kb: Optional[float] = None):
    """Thermalizing_MC_DEV (MD_DiffScheme_GPU.F90:1451-1805): Maxwell velocities at TI on the
    device (components with FIXVEL/FIXPOS kept, inactive atoms at rest), then each box's
    mass-weighted mean velocity removed; the generator is keyed by CtrlParam.SEED[0:2] and
    advances with every call (msm_thermalize)."""
    ForceClass.Device.thermalize(TI, CtrlParam.KB if kb is None else kb,
                                 thermal_seed(CtrlParam.SEED))


def Do_DynDamp_Forsteps_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU,
                            MXNUMSTEPS: int):
    """Do_DynDamp_Forsteps_DEV (MD_DiffScheme_GPU.F90:1809-1883): damped dynamics until the largest
    per-atom energy change is at most CtrlParam.STEEPEST_MiDelE; returns (steps, max |dEPOT|)."""
    it, de = ForceClass.Device.dyn_damp(MXNUMSTEPS, CtrlParam.STEEPEST_MiDelE, CtrlParam.H)
    if it > MXNUMSTEPS:
        print(f" MDPSCU WARNING: DnyDamp finished after max steps:  {it - 1}")
        print(f"                 with max energy uncertainty(ev): {de:13.4E}")
    else:
        print(f" MDPSCU Message: DnyDamp finished after steps:  {it}")
        print(f"                 with max energy uncertainty(ev): {de:13.4E}")
    return it, de


def Do_Steepest_Forsteps_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU,
                             MXNUMSTEPS: int, METH: int = CP_DAMPSCHEME_ST):
    """Do_Steepest_Forsteps_DEV (MD_SteepestScheme_GPU.F90:264-290): steepest descent, with the
    line search when METH has the CP_DAMPSCHEME_LSEARCH bit (whose energy tolerance the reference
    fixes at 1e-3 eV).  Moves are bounded by STEEPEST_MxStep and STEEPEST_MiStep times the lattice
    unit of the first box; returns (steps, max |dEPOT|, largest move)."""
    boxes = SimBox if isinstance(SimBox, (list, tuple)) else [SimBox]
    rr = boxes[0].RR
    ls = bool(METH & CP_DAMPSCHEME_LSEARCH)
    it, de, mm = ForceClass.Device.steepest(
        MXNUMSTEPS, ls, CtrlParam.STEEPEST_Alpha, CtrlParam.STEEPEST_MxStep * rr,
        CtrlParam.STEEPEST_MiStep * rr, 1.0e-3 if ls else CtrlParam.STEEPEST_MiDelE)
    if it > MXNUMSTEPS:
        print(f" MDPSCU WARNING: steepest finished after max steps: {it - 1 if not ls else it}")
    else:
        print(f" MDPSCU Message: steepest finished after steps: {it}")
    print(f"                 with max energy uncertainty(ev): {de:13.4E}")
    return it, de, mm


def Do_CG_Forsteps_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU,
                       MXNUMSTEPS: int, METH: int = CP_DAMPSCHEME_CG):
    """Do_CG_Forsteps_DEV (MD_CGScheme_GPU.F90:296-322): conjugate gradients, with secant line
    searches when METH has the CP_DAMPSCHEME_LSEARCH bit; returns (steps, max |dEPOT|)."""
    boxes = SimBox if isinstance(SimBox, (list, tuple)) else [SimBox]
    rr = boxes[0].RR
    ls = bool(METH & CP_DAMPSCHEME_LSEARCH)
    it, de = ForceClass.Device.cg(MXNUMSTEPS, ls, CtrlParam.STEEPEST_MxStep * rr,
                                  CtrlParam.STEEPEST_MiStep * rr, CtrlParam.STEEPEST_MiDelE)
    name = "CG-LS" if ls else "CG"
    if it > MXNUMSTEPS:
        print(f" MDPSCU WARNING: {name} finished after max steps: {it - 1}")
    else:
        print(f" MDPSCU Message: {name} finished after steps: {it}")
    print(f"                 with max energy uncertainty(ev): {de:13.4E}")
    return it, de


def lbfgs_coordinates(statu: np.ndarray) -> List[np.ndarray]:
    """IDX, IDY, IDZ of DO_LBFGS_FORSTEPS_DEV (MD_LBFGSScheme_GPU.F90:69-97), 0-based: the active
    atoms inside the box whose position in that direction is not fixed."""
    st = np.asarray(statu)
    ok = ((st & CP_STATU_ACTIVE) == CP_STATU_ACTIVE) & ((st & CP_STATU_OUTOFBOX) == 0)
    return [np.flatnonzero(ok & ((st & (CP_STATU_FIXPOSX << k)) == 0)) for k in range(3)]


def lbfgs_quench(xp: np.ndarray, statu: np.ndarray, lo: np.ndarray, zl: np.ndarray,
                 ifpd: Sequence[int], evaluate: Callable[[], tuple],
                 put: Callable[[np.ndarray], None], maxsteps: int, msave: int):
    """The loop of DO_LBFGS_FORSTEPS_DEV (MD_LBFGSScheme_GPU.F90:43-158) over a force provider.

    xp (N, 3) in A are the positions the provider holds; lo, zl (N, 3) the lower corner and edge
    of each atom's box.  Each pass calls evaluate() -> (FP (N, 3) eV/A, EPOT (N) eV) at the
    provider's positions, hands the objective to the optimiser in the reference's CGS units
    (X in cm, G = -FP in dyn, OBJF = sum EPOT in erg; the tolerances EPS = 1e-4 and GTOL = 1e-2
    are scale-dependent, so the units are kept), wraps the new X into the periodic directions
    (below the box: +edge, above: -edge) and gives the positions back through put().  Returns
    (evaluations, iflag, positions); iflag 0 on convergence, 1 when maxsteps ran out, negative
    for an optimiser error (LBFGS flags)."""
    x_a = np.array(xp, dtype=np.float64, order="F")
    ids = lbfgs_coordinates(statu)
    cols = [k for k in range(3) for _ in ids[k]]
    rows = np.concatenate(ids)
    cols = np.asarray(cols, dtype=np.int64)
    to_cm = 1.0 / CP_CM2A
    x = x_a[rows, cols] * to_cm
    lb = np.asarray(lo, dtype=np.float64)[rows, cols] * to_cm
    bs = np.asarray(zl, dtype=np.float64)[rows, cols] * to_cm
    ub = lb + bs
    per = np.asarray([bool(ifpd[k]) for k in range(3)])[cols]
    opt = LBFGS(x.size, msave, 1.0e-4, 1.0e-16)
    g_unit = -CP_EVERG * CP_CM2A  # eV/A -> dyn, with the sign of the gradient
    iflag, it = 1, 0
    for it in range(1, maxsteps + 1):
        fp, ep = evaluate()
        objf = float(np.sum(np.asarray(ep, dtype=np.float64) * CP_EVERG))
        g = np.asarray(fp, dtype=np.float64)[rows, cols] * g_unit
        iflag = opt.step(x, objf, g) if x.size else 0
        low, high = per & (x < lb), per & (x > ub)
        x[low] += bs[low]
        x[high] -= bs[high]
        x_a[rows, cols] = x * CP_CM2A
        put(x_a)
        if iflag == 0:
            break
    return it, iflag, x_a


def Do_LBFGS_Forsteps_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU,
                          MXNUMSTEPS: int):
    """DO_LBFGS_FORSTEPS_DEV (MD_LBFGSScheme_GPU.F90:16-173): L-BFGS quench of all boxes on the
    current neighbour list, at most MXNUMSTEPS force evaluations; the velocities are zeroed at
    the end.  Returns (evaluations, iflag) -- iflag 0 on convergence."""
    boxes = _boxes(SimBox)
    dev = ForceClass.Device
    xp, _, _, st, _ = dev.get_atoms()
    nb = np.asarray([b.NPRT for b in boxes])
    lo = np.repeat(np.asarray([b.BOXLOW for b in boxes]), nb, axis=0)
    zl = np.repeat(np.asarray([b.ZL for b in boxes]), nb, axis=0)

    def evaluate():
        dev.calc_force()
        dev.calc_epot()
        return dev.get_forces(), dev.get_epot()

    it, iflag, _ = lbfgs_quench(xp, st, lo, zl, CtrlParam.IFPD, evaluate, dev.put_atoms,
                                MXNUMSTEPS, CtrlParam.LBFGS_Msave)
    if iflag == 0:
        print(f" MDPSCU Message: LBFG succefully performed for steps {it:6d}")
    dev.put_atoms(None, np.zeros((dev.n, 3)))
    if iflag != 0:
        print(f" MDPSCU Warning: LBFG exit with erro code {iflag:2d}")
        print("                 in O_LBFGS_FORSTEPS_DEV")
    return it, iflag


def lbfgsb_quench(xp: np.ndarray, statu: np.ndarray, lo: np.ndarray, zl: np.ndarray,
                  ifpd: Sequence[int], evaluate: Callable[[], tuple],
                  put: Callable[[np.ndarray], None], maxsteps: int, msave: int, factr: float,
                  pgtol: float):
    """The loop of DO_LBFGSB_FORSTEPS_DEV (MD_LBFGSScheme_GPU.F90:177-388) over a force provider,
    with the reference's units and conventions: X in cm (the coordinates of lbfgs_coordinates,
    no bounds: NBD = 0), G = -FP in dyn, OBJF = sum EPOT in erg, FACTR and PGTOL as given.
    Every SETULB call is one of the maxsteps steps; after each call the positions (X wrapped once
    into the periodic directions) go back through put(); an 'FG' task evaluates the provider,
    'NEW_X' continues, anything else ends the loop.  Returns (steps, iflag, positions, task):
    iflag 1 when the steps ran out, -1 on an 'ERROR' task, else 0."""
    x_a = np.array(xp, dtype=np.float64, order="F")
    ids = lbfgs_coordinates(statu)
    cols = np.asarray([k for k in range(3) for _ in ids[k]], dtype=np.int64)
    rows = np.concatenate(ids)
    to_cm = 1.0 / CP_CM2A
    x = x_a[rows, cols] * to_cm
    lb = np.asarray(lo, dtype=np.float64)[rows, cols] * to_cm
    bs = np.asarray(zl, dtype=np.float64)[rows, cols] * to_cm
    ub = lb + bs
    per = np.asarray([bool(ifpd[k]) for k in range(3)])[cols]
    g_unit = -CP_EVERG * CP_CM2A  # eV/A -> dyn, with the sign of the gradient
    if x.size == 0:
        return 1, 0, x_a, "CONVERGENCE: NO VARIABLES"
    opt = LBFGSB(x.size, msave, factr, pgtol)
    task, f, g = "START", 0.0, np.zeros_like(x)
    steps = 0
    for steps in range(1, maxsteps + 1):
        task = opt.start(x) if task == "START" else opt.setulb(x, f, g)
        xw = x.copy()
        low, high = per & (x < lb), per & (x > ub)
        xw[low] += bs[low]
        xw[high] -= bs[high]
        x_a[rows, cols] = xw * CP_CM2A
        put(x_a)
        if task.startswith("FG"):
            fp, ep = evaluate()
            f = float(np.sum(np.asarray(ep, dtype=np.float64) * CP_EVERG))
            g = np.asarray(fp, dtype=np.float64)[rows, cols] * g_unit
            continue
        if task == "NEW_X":
            continue
        break
    else:
        steps = maxsteps + 1
    iflag = 1 if steps > maxsteps else (-1 if task.startswith("ERROR") else 0)
    return steps, iflag, x_a, task


def Do_LBFGSB_Forsteps_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU,
                           MXNUMSTEPS: int):
    """DO_LBFGSB_FORSTEPS_DEV (MD_LBFGSScheme_GPU.F90:177-388): L-BFGS-B quench of all boxes on
    the current neighbour list (CtrlParam.LBFGS_Factr, LBFGS_PGtol, LBFGS_Msave), positions
    written to the device after every SETULB call, velocities zeroed at the end.  Returns
    (steps, iflag, task)."""
    boxes = _boxes(SimBox)
    dev = ForceClass.Device
    xp, _, _, st, _ = dev.get_atoms()
    nb = np.asarray([b.NPRT for b in boxes])
    lo = np.repeat(np.asarray([b.BOXLOW for b in boxes]), nb, axis=0)
    zl = np.repeat(np.asarray([b.ZL for b in boxes]), nb, axis=0)

    def evaluate():
        dev.calc_force()
        dev.calc_epot()
        return dev.get_forces(), dev.get_epot()

    def put(xa):
        dev.put_atoms(xa, None)

    steps, iflag, _, task = lbfgsb_quench(xp, st, lo, zl, CtrlParam.IFPD, evaluate, put,
                                          MXNUMSTEPS, CtrlParam.LBFGS_Msave,
                                          CtrlParam.LBFGS_Factr, CtrlParam.LBFGS_PGtol)
    print(f" MDPSCU Message: LBFG performed for steps {min(steps, MXNUMSTEPS):6d}")
    print(f"                 exit with code> {task}")
    dev.put_atoms(None, np.zeros((dev.n, 3)))
    if iflag < 0:
        print(f" MDPSCU Warning: LBFG exit with code {iflag:2d}")
    elif steps > MXNUMSTEPS:
        print(f" MDPSCU Warning: LBFG exit with code {iflag:2d} after {MXNUMSTEPS:7d} steps")
    return steps, iflag, task


def Check_Timestep_DEV(SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU) -> float:
    """The variable-step search of Predictor_DEV (MD_DiffScheme_GPU.F90:633-650, the check in
    CheckTimestep_DEV/_KERNEL :1068-1210): from HMX, halve the step while some active atom would
    move more than DMX; the step found is stored in CtrlParam.H and returned."""
    dev = ForceClass.Device
    th = float(CtrlParam.HMX)
    mxd2 = float(CtrlParam.DMX) ** 2
    while dev.check_timestep(th, mxd2):
        th *= 0.5
        if th == 0.0:
            raise ValueError("MDPSCU Error: no time step keeps the displacements under DMX")
    CtrlParam.H = th
    return th


def Predictor_DEV(ITIME: int, SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU):
    """Predictor_DEV (MD_DiffScheme_GPU.F90:604-682): damping when ITIME is in
    [DAMPTIME0, DAMPTIME0 + DAMPTIME1 - 1]; with IHDUP < 0 and mod(ITIME - IT0 + 1, |IHDUP|) = 0
    the step search of Check_Timestep_DEV; then the Swope predictor with CtrlParam.H."""
    dev = ForceClass.Device
    if CtrlParam.DAMPTIME0 <= ITIME <= CtrlParam.DAMPTIME0 + CtrlParam.DAMPTIME1 - 1:
        dev.damping()
    if CtrlParam.IHDUP < 0 and (ITIME - CtrlParam.IT0 + 1) % abs(CtrlParam.IHDUP) == 0:
        Check_Timestep_DEV(SimBox, CtrlParam, ForceClass)
    dev.predictor(CtrlParam.H)


def Correction_DEV(ITIME: int, SimBox, CtrlParam: SimMDCtrl, ForceClass: MDForceClassGPU):
    ForceClass.Device.correction(CtrlParam.H)


def For_One_Steps(ITIME0: int, NSTEPS: int, SimBox, CtrlParam: SimMDCtrl,
                  ForceClass: MDForceClassGPU):
    """NSTEPS consecutive For_One_Step calls (predictor, rebuild when mod(ITIME, NB_UPTAB)=0,
    force, corrector), replayed on the device."""
    ForceClass.Device.run(NSTEPS, CtrlParam.NB_UPTAB, ITIME0, CtrlParam.H)


def CopyOut_SimBox_DEV(SimBox, ForceClass: MDForceClassGPU, with_epot: bool = False):
    """CopyOut_SimBox_DEV: device state back into the SimMDBox objects."""
    boxes = _boxes(SimBox)
    dev = ForceClass.Device
    xp, xp1, fp, st, dis = dev.get_atoms()
    ep = dev.get_epot() if with_epot else None
    off = 0
    for b in boxes:
        s = slice(off, off + b.NPRT)
        b.XP = np.asfortranarray(xp[s])
        b.XP1 = np.asfortranarray(xp1[s])
        b.FP = np.asfortranarray(fp[s])
        b.DIS = np.asfortranarray(dis[s])
        b.STATU = st[s].copy()
        if ep is not None:
            b.EPOT = ep[s].copy()
        off += b.NPRT


# ----------------------------------------------------------------------------------------------
# table files (.pair / .embd), Export_ForceTable (:1315-1459) layout in eV and Angstrom
# ----------------------------------------------------------------------------------------------

def Export_ForceTable(fname: str, FTable: MDForceTable, rho_unit: float = 1.0):
    """Write ``fname.pair`` and ``fname.embd`` in the reference's column layout: r, r*V, -r*dV/dr,
    RHO, -dRHO/dr per table id (POTR is stored as 0.5*V*r, so the r*V column is 2*POTR)."""
    ids = sorted(FTable.FPAIR)
    with open(fname + ".pair", "w") as fh:
        fh.write("&MDPSCU_POTTAB.Pair\n")
        fh.write(f'&POTTYPE "{FTable.PotType.strip()}"\n')
        fh.write(f"&NUMTABLE {len(ids):7d} table IDs: " + "".join(f"{i:5d}" for i in ids) + "\n")
        fh.write(f"&NUMPOINT {FTable.NTAB:7d}\n")
        title = "".join(f"      r*V{i}(r)[eV*A]        -r*dV{i}/dr[eV]            RHO{i}(r)"
                        f"            -dRHO{i}/dr[/A] " for i in ids)
        fh.write(f"&#           Rij(A)      {title}\n")
        for it in range(1, FTable.NTAB + 1):
            t = it * FTable.CSIV
            r = t * t
            cols = []
            for tid in ids:
                k = FTable.FPAIR.index(tid)
                cols += [FTable.POTR[k, it - 1] * 2.0, FTable.FPOTR[k, it - 1],
                         FTable.POTB[k, it - 1] * rho_unit, FTable.FPOTB[k, it - 1] * rho_unit]
            fh.write(f" {it:6d} " + "".join(f"{v:21.9E}" for v in [r] + cols) + "\n")
    with open(fname + ".embd", "w") as fh:
        fh.write("&MDPSCU_POTTAB.Embd\n")
        fh.write(f'&POTTYPE "{FTable.PotType.strip()}"\n')
        ids1 = sorted(FTable.FPAIR1) if FTable.FPAIR1 else [1]
        fh.write(f"&NUMTABLE {len(ids1):7d} table IDs: " + "".join(f"{i:5d}" for i in ids1) + "\n")
        fh.write(f"&NUMPOINT {FTable.NEMBD:7d}\n")
        title = "".join(f"         F{i}(RHO)      dF{i}/d(RHO)" for i in ids1)
        fh.write(f"&#           RHO       {title}\n")
        for it in range(1, FTable.NEMBD + 1):
            rho = (it - 1) * FTable.RHOD
            cols = []
            for tid in ids1:
                if FTable.PotType.strip() == "FS_TYPE" or FTable.FEMBD is None:
                    f = -np.sqrt(rho) if rho > 0 else 0.0
                    df = -0.5 / np.sqrt(rho) if rho > 0 else 0.0
                else:
                    k = FTable.FPAIR1.index(tid)
                    f, df = FTable.FEMBD[k, it - 1], FTable.DFEMBD[k, it - 1]
                cols += [f, df]
            fh.write(f" {it:6d} " + "".join(f"{v:16.8E}" for v in [rho] + cols) + "\n")


def _spline_second_derivatives(x: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Second derivatives of the cubic spline through (x, f) clamped at both ends to the slope of
    the cubic through the four end points -- SPLID1 with IOP = (5, 5)
    (LIB/sor/f/MiniUtilities/DINTF2.F:17-189), solved as one tridiagonal system."""
    n = len(x)
    if n < 4:
        raise ValueError("MDPSCU Error: a table needs at least 4 points to be interpolated")
    a1, a2, a3 = x[0] - x[1], x[0] - x[2], x[0] - x[3]
    a4, a5, a6 = x[1] - x[2], x[1] - x[3], x[2] - x[3]
    d0 = (f[0] * (1.0 / a1 + 1.0 / a2 + 1.0 / a3) - a2 * a3 * f[1] / (a1 * a4 * a5)
          + a1 * a3 * f[2] / (a2 * a4 * a6) - a1 * a2 * f[3] / (a3 * a5 * a6))
    b1, b2, b3 = x[-1] - x[-4], x[-1] - x[-3], x[-1] - x[-2]
    b4, b5, b6 = x[-2] - x[-4], x[-2] - x[-3], x[-3] - x[-4]
    dn = (-b2 * b3 * f[-4] / (b6 * b4 * b1) + b1 * b3 * f[-3] / (b6 * b5 * b2)
          - b1 * b2 * f[-2] / (b4 * b5 * b3) + f[-1] * (1.0 / b1 + 1.0 / b2 + 1.0 / b3))
    h = np.diff(x)
    sl = np.diff(f) / h
    sub = np.zeros(n)
    dia = np.zeros(n)
    sup = np.zeros(n)
    rhs = np.zeros(n)
    dia[0], sup[0], rhs[0] = h[0] / 3.0, h[0] / 6.0, sl[0] - d0
    sub[1:-1], dia[1:-1], sup[1:-1] = h[:-1] / 6.0, (h[:-1] + h[1:]) / 3.0, h[1:] / 6.0
    rhs[1:-1] = sl[1:] - sl[:-1]
    sub[-1], dia[-1], rhs[-1] = h[-1] / 6.0, h[-1] / 3.0, dn - sl[-1]
    # Thomas algorithm (the matrix is diagonally dominant)
    c = np.zeros(n)
    d = np.zeros(n)
    c[0], d[0] = sup[0] / dia[0], rhs[0] / dia[0]
    for i in range(1, n):
        m = dia[i] - sub[i] * c[i - 1]
        c[i] = sup[i] / m
        d[i] = (rhs[i] - sub[i] * d[i - 1]) / m
    w = np.zeros(n)
    w[-1] = d[-1]
    for i in range(n - 2, -1, -1):
        w[i] = d[i] - c[i] * w[i + 1]
    return w


def _spline_eval(x: np.ndarray, f: np.ndarray, w: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Spline value at y (SPLID2, DINTF2.F:191-236): the interval with x[j] <= y < x[j+1], the first
    interval below x[0] and the last at or above x[-1]."""
    y = np.asarray(y, dtype=np.float64)
    j = np.clip(np.searchsorted(x, y, side="right") - 1, 0, len(x) - 2)
    x0, x1 = x[j], x[j + 1]
    hk = x1 - x0
    return ((w[j] * (x1 - y) ** 3 + w[j + 1] * (y - x0) ** 3) / (6.0 * hk)
            + (f[j + 1] / hk - w[j + 1] * hk / 6.0) * (y - x0)
            + (f[j] / hk - hk * w[j] / 6.0) * (x1 - y))


def Import_ForceTable(fname: str, rho_unit: float = 1.0) -> MDForceTable:
    """Import_ForceTable (MD_TypeDef_ForceTable.F90:1463-1608): read ``fname.pair`` and
    ``fname.embd`` as written by Export_ForceTable into a table on the file's own grid: R, and
    POTR = (r V)/2, FPOTR = -r dV/dr, POTB = RHO/rho_unit, FPOTB = -dRHO/dr/rho_unit per imported
    table, RHO and FEMBD, DFEMBD per embedding table (eV and Angstrom)."""
    def parse(path, ncol):
        head, rows = {}, []
        with open(path) as fh:
            in_head = True
            for line in fh:
                s = line.strip()
                if not s or s.startswith("!"):
                    continue
                if in_head:
                    if s.startswith("&#"):
                        in_head = False
                    elif s.startswith("&"):
                        parts = s.split()
                        head[parts[0].upper()] = parts[1:]
                    continue
                rows.append([float(v) for v in s.split()[1:1 + ncol]])
        if in_head:
            raise ValueError(f"MDPSCU Error: no table rows (&# line) in {path}")
        return head, np.array(rows, dtype=np.float64)

    ft = MDForceTable()
    with open(fname + ".pair") as fh:
        for line in fh:
            if line.strip().upper().startswith("&NUMTABLE"):
                nc = int(line.split()[1])
                break
        else:
            raise ValueError(f"MDPSCU Error: no &NUMTABLE in {fname}.pair")
    head, rows = parse(fname + ".pair", 1 + 4 * nc)
    ft.PotType = head.get("&POTTYPE", ['"FS_TYPE"'])[0].strip('"')
    ft.FPAIR = [int(v) for v in head["&NUMTABLE"][3:3 + nc]]
    ft.NKIND = nc
    ft.NTAB = int(head["&NUMPOINT"][0])
    if rows.shape[0] != ft.NTAB:
        raise ValueError(f"MDPSCU Error: {fname}.pair holds {rows.shape[0]} rows, "
                         f"&NUMPOINT says {ft.NTAB}")
    ft.R = rows[:, 0].copy()
    cols = rows[:, 1:].reshape(ft.NTAB, nc, 4)
    ft.POTR = np.asfortranarray(cols[:, :, 0].T * 0.5)
    ft.FPOTR = np.asfortranarray(cols[:, :, 1].T.copy())
    ft.POTB = np.asfortranarray(cols[:, :, 2].T / rho_unit)
    ft.FPOTB = np.asfortranarray(cols[:, :, 3].T / rho_unit)
    with open(fname + ".embd") as fh:
        for line in fh:
            if line.strip().upper().startswith("&NUMTABLE"):
                nc1 = int(line.split()[1])
                break
        else:
            raise ValueError(f"MDPSCU Error: no &NUMTABLE in {fname}.embd")
    head, rows = parse(fname + ".embd", 1 + 2 * nc1)
    ft.FPAIR1 = [int(v) for v in head["&NUMTABLE"][3:3 + nc1]]
    ft.NKIND1 = nc1
    ft.NEMBD = int(head["&NUMPOINT"][0])
    ft.RHO = rows[:, 0] / rho_unit
    cols = rows[:, 1:].reshape(ft.NEMBD, nc1, 2)
    ft.FEMBD = np.asfortranarray(cols[:, :, 0].T.copy())
    ft.DFEMBD = np.asfortranarray(cols[:, :, 1].T.copy())
    return ft


def Register_Imported_ForceTable(SimBox: "SimMDBox", CtrlParam: "SimMDCtrl",
                                 FTable: MDForceTable, fname: str, rho_unit: float = 1.0):
    """Register_Imported_ForceTable (MD_TypeDef_ForceTable.F90:1612-1855): import ``fname`` and
    resample every table the box's PTYPE needs onto the simulation grid (r_k = (k/CSI)^2 from
    CtrlParam) by the clamped cubic spline in r; zero beyond the imported range.  Embedding
    tables are resampled on rho_k = (k-1) RHOD with RHOMX = RHO(NEMBD) NEMBD/(NEMBD-1) of the
    imported grid."""
    src = Import_ForceTable(fname, rho_unit)
    n = SimBox.NGROUP
    for i in range(n):
        for j in range(n):
            if int(SimBox.PTYPE[i, j]) not in src.FPAIR:
                raise ValueError(f"MDPSCU Error: cannot find force table # {int(SimBox.PTYPE[i, j])}"
                                 f" in lib {fname}.pair")
        if int(SimBox.PTYPE[i, i]) not in src.FPAIR1:
            raise ValueError(f"MDPSCU Error: cannot find force table # {int(SimBox.PTYPE[i, i])}"
                             f" in lib {fname}.embd")
    for tid in src.FPAIR:
        if tid > 0 and tid not in FTable.REGPROCS:
            FTable.REGPROCS[tid] = _ExtProc(tid, "", None, None, None, f"Table # {tid}")
    FTable.PotType = src.PotType
    New_ForceTable(SimBox, CtrlParam, FTable)
    FTable.KPAIR[:, :] = -1
    trmax = float(np.max(src.R))
    tt = (np.arange(1, FTable.NTAB + 1) * FTable.CSIV) ** 2
    inside = tt <= trmax
    for i in range(n):
        for j in range(n):
            iforce = int(SimBox.PTYPE[i, j])
            k = FTable.FPAIR.index(iforce)
            if not FTable.HASREG.get(k):
                k0 = src.FPAIR.index(iforce)
                for name in ("POTR", "FPOTR", "POTB", "FPOTB"):
                    fsrc = getattr(src, name)[k0]
                    w = _spline_second_derivatives(src.R, fsrc)
                    out = np.zeros(FTable.NTAB)
                    out[inside] = _spline_eval(src.R, fsrc, w, tt[inside])
                    getattr(FTable, name)[k, :] = out
                FTable.HASREG[k] = True
            FTable.KPAIR[i, j] = k + 1
    nembd_src = len(src.RHO)
    FTable.RHOMX = src.RHO[-1] / (nembd_src - 1) * nembd_src
    FTable.RHOD = FTable.RHOMX / FTable.NEMBD
    if FTable.RHOD <= 1.0e-64:
        FTable.RHOD = 1.0
    if FTable.PotType.strip() == "EAM_TYPE":
        FTable.KEMBD[:] = -1
        rho = np.arange(FTable.NEMBD) * FTable.RHOD
        for i in range(n):
            iforce = int(SimBox.PTYPE[i, i])
            k = FTable.FPAIR1.index(iforce)
            if not FTable.HASREG1.get(k):
                k0 = src.FPAIR1.index(iforce)
                for name in ("FEMBD", "DFEMBD"):
                    fsrc = getattr(src, name)[k0]
                    w = _spline_second_derivatives(src.RHO, fsrc)
                    getattr(FTable, name)[k, :] = _spline_eval(src.RHO, fsrc, w, rho)
                FTable.HASREG1[k] = True
            FTable.KEMBD[i] = k + 1


def read_table_file(path: str):
    """Read a .pair or .embd table (reference layout): returns (header dict, ndarray rows x cols)
    with the leading row index dropped."""
    head = {}
    rows = []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("!"):
                continue
            if s.startswith("&"):
                parts = s.split()
                head[parts[0].upper()] = parts[1:]
                continue
            rows.append([float(x) for x in s.split()[1:]])
    return head, np.array(rows, dtype=np.float64)
