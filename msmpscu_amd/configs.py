"""Workload builders for the FS/EAM MD path: lattices, defects, thermal velocities, force tables
and control blocks, in eV / Angstrom / fs units (masses in eV fs^2 / Angstrom^2).

The box spans [-L/2, L/2) in each direction (the reference's default box placement,
MDLIB/sor/Common/MD_Gvar.F90:922-924).  Velocities come from the reference's Maxwell sampler
(msm_thermalize_mc, a restatement of Thermalize_MC_SimMDBox with DRAND32); displacements,
vacancies and dumbbells use numpy's PCG64 with the given seed.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .mdlib import (CP_KB_EV, MDForceTable, Register_ForceTableProc, SimMDBox, SimMDCtrl,
                    Create_Interaction_ForceTable)

AMU = 1.66053906660e-27 * 1.0e10 / 1.602176634e-19  # eV fs^2 / A^2 per amu (103.6427...)
KB = CP_KB_EV                                         # eV / K
W_MASS = 183.84
CU_MASS = 63.546
NI_MASS = 58.6934

W_A = 3.165
W_RC = 4.4002


def bcc_sites(nx: int, ny: int, nz: int, a: float) -> np.ndarray:
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.reshape(-1, 3).astype(np.float64)
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])
    return ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a


def fcc_sites(nx: int, ny: int, nz: int, a: float) -> np.ndarray:
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.reshape(-1, 3).astype(np.float64)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=np.float64)
    return ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a


def wrap_into(x: np.ndarray, lo: np.ndarray, ln: np.ndarray, ifpd: Sequence[int]) -> np.ndarray:
    x = x.copy()
    for k in range(3):
        if ifpd[k]:
            x[:, k] = lo[k] + np.mod(x[:, k] - lo[k], ln[k])
            # np.mod can return ln for tiny negative offsets; fold back into [lo, lo+ln)
            x[x[:, k] >= lo[k] + ln[k], k] -= ln[k]
    return x


def maxwell_velocities(ityp: np.ndarray, cm: np.ndarray, temp: float, seed: int) -> np.ndarray:
    """Thermalize_MC_SimMDBox via msm_thermalize_mc (DRAND32 seeded with (seed, seed+7))."""
    import ctypes as C
    n = int(ityp.shape[0])
    xp1 = np.zeros(3 * n)
    s = (C.c_int * 2)(int(seed) % 2147483562 + 1, (int(seed) * 7 + 12345) % 2147483398 + 1)
    it = np.ascontiguousarray(ityp, dtype=np.int32)
    c = np.ascontiguousarray(cm, dtype=np.float64)
    rc = _lib.lib().msm_thermalize_mc(n, _lib.iptr(it), _lib.dptr(c), float(temp), KB,
                                      C.cast(s, C.POINTER(C.c_int)), _lib.dptr(xp1))
    _lib.check(rc, "msm_thermalize_mc")
    return xp1.reshape(3, n).T.copy(order="F")


def make_bcc_box(ncell: Sequence[int], a: float, sigma: float, temp: float, seed: int,
                 mass: float = W_MASS, vacancy_frac: float = 0.0, n_vac: int = 0,
                 n_dumbbell: int = 0, strain: Optional[Sequence[float]] = None,
                 ifpd: Sequence[int] = (1, 1, 1)) -> SimMDBox:
    """BCC box of ncell[0] x ncell[1] x ncell[2] conventional cells with seeded Gaussian
    displacements, vacancies (a fraction and/or a count), <110> dumbbells and a diagonal
    homogeneous strain."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(c) for c in ncell)
    sites = bcc_sites(nx, ny, nz, a)
    ln = np.array([nx * a, ny * a, nz * a], dtype=np.float64)
    lo = -0.5 * ln
    x = sites + lo
    n0 = x.shape[0]
    nv = int(round(vacancy_frac * n0)) + int(n_vac)
    keep = np.ones(n0, dtype=bool)
    if nv > 0:
        keep[rng.choice(n0, size=nv, replace=False)] = False
    if n_dumbbell > 0:
        # replace an occupied site by two atoms split along a <110> direction
        cand = np.flatnonzero(keep)
        centers = rng.choice(cand, size=n_dumbbell, replace=False)
        dirs = np.array([[1, 1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]],
                        dtype=np.float64) / math.sqrt(2.0)
        half = 0.3 * a
        extra = []
        for c in centers:
            d = dirs[rng.integers(0, len(dirs))]
            keep[c] = False
            extra.append(x[c] + half * d)
            extra.append(x[c] - half * d)
        x = np.concatenate([x[keep], np.array(extra)], axis=0)
    else:
        x = x[keep]
    if sigma > 0:
        x = x + rng.normal(0.0, sigma, size=x.shape)
    if strain is not None:
        f = 1.0 + np.asarray(strain, dtype=np.float64)
        x = x * f
        ln = ln * f
        lo = -0.5 * ln
    x = wrap_into(x, lo, ln, ifpd)
    n = x.shape[0]
    ityp = np.ones(n, dtype=np.int32)
    cm = np.array([mass * AMU])
    v = maxwell_velocities(ityp, cm, temp, seed) if temp > 0 else np.zeros((n, 3), order="F")
    return SimMDBox(NPRT=n, NGROUP=1, ZL=ln, BOXLOW=lo, XP=np.asfortranarray(x), XP1=v, RR=a,
                    ITYP=ityp, STATU=np.ones(n, dtype=np.int32), CM=cm,
                    PTYPE=np.array([[1]], dtype=np.int32), SYMB=["W"])


def w_ctrl(skin: float = 0.3, rc: float = W_RC, nb_uptab: int = 10, mx: int = 64,
           ntab: int = 5000, nembd: int = 5000, dt: float = 1.0) -> SimMDCtrl:
    return SimMDCtrl(RU=np.array([[rc]]), NB_RM=np.array([[rc + skin]]), NB_UPTAB=nb_uptab,
                     NB_MXNBS=mx, NUMFTABR=ntab, NUMFTABE=nembd, RHOSCAL=20.0, IFPD=(1, 1, 1),
                     H=dt)


def w_forcetable(box: SimMDBox, ctrl: SimMDCtrl, pot_type: str = "FS_TYPE") -> MDForceTable:
    """Benchmark W Finnis-Sinclair table (pair polynomial, rho = A^2 (r - 4.4002)^2; EAM_TYPE
    tabulates F = -sqrt(rho))."""
    ft = MDForceTable(PotType=pot_type)
    Register_ForceTableProc(1, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne", "msm_emb_fs",
                            NOTE="W-W Finnis-Sinclair (benchmark form)")
    Create_Interaction_ForceTable(box, ctrl, ft)
    return ft


def make_cuni_slab(ncell: Sequence[int], a: float, temp: float, seed: int,
                   vacuum: float = 20.0) -> SimMDBox:
    """FCC Cu-Ni random solid solution (50/50) periodic in x, y with free z surfaces and
    `vacuum` Angstrom of empty space on each side.  Group 1 = Cu, 2 = Ni."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(c) for c in ncell)
    sites = fcc_sites(nx, ny, nz, a)
    lz = nz * a + 2.0 * vacuum
    ln = np.array([nx * a, ny * a, lz])
    lo = -0.5 * ln
    x = sites + np.array([lo[0], lo[1], -0.5 * nz * a])
    n = x.shape[0]
    ityp = np.ones(n, dtype=np.int32)
    ityp[rng.permutation(n)[: n // 2]] = 2
    order = np.argsort(ityp, kind="stable")  # atoms grouped by type, as IPA ranges require
    x = x[order]
    ityp = ityp[order]
    x = wrap_into(x, lo, ln, (1, 1, 0))
    cm = np.array([CU_MASS * AMU, NI_MASS * AMU])
    v = maxwell_velocities(ityp, cm, temp, seed) if temp > 0 else np.zeros((n, 3), order="F")
    return SimMDBox(NPRT=n, NGROUP=2, ZL=ln, BOXLOW=lo, XP=np.asfortranarray(x), XP1=v, RR=a,
                    ITYP=ityp, STATU=np.ones(n, dtype=np.int32), CM=cm,
                    PTYPE=np.array([[1, 3], [3, 2]], dtype=np.int32), SYMB=["Cu", "Ni"])


def cuni_ctrl(rc: float = 6.0, skin: float = 0.3, nb_uptab: int = 5, mx: int = 160,
              ntab: int = 5000, dt: float = 2.0) -> SimMDCtrl:
    return SimMDCtrl(RU=np.full((2, 2), rc), NB_RM=np.full((2, 2), rc + skin), NB_UPTAB=nb_uptab,
                     NB_MXNBS=mx, NUMFTABR=ntab, NUMFTABE=5000, RHOSCAL=20.0, IFPD=(1, 1, 0),
                     H=dt)


def cuni_forcetable(box: SimMDBox, ctrl: SimMDCtrl, pot_type: str = "FS_TYPE") -> MDForceTable:
    ft = MDForceTable(PotType=pot_type)
    Register_ForceTableProc(1, ft, "msm_pot_tbsma_cucu_nn", "msm_pot_tbsma_cucu_ne", "msm_emb_fs",
                            NOTE="Cu-Cu TB-SMA")
    Register_ForceTableProc(2, ft, "msm_pot_tbsma_nini_nn", "msm_pot_tbsma_nini_ne", "msm_emb_fs",
                            NOTE="Ni-Ni TB-SMA")
    Register_ForceTableProc(3, ft, "msm_pot_tbsma_cuni_nn", "msm_pot_tbsma_cuni_ne", "msm_emb_fs",
                            NOTE="Cu-Ni TB-SMA")
    Create_Interaction_ForceTable(box, ctrl, ft)
    return ft


def batch_sizes(nbox: int, lo: int, hi: int, seed: int) -> List[int]:
    """Box atom counts drawn log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(v))) for v in rng.uniform(math.log(lo), math.log(hi), size=nbox)]


def bcc_cells_for(natoms: int) -> tuple:
    """Near-cubic BCC cell counts (nx, ny, nz) whose 2*nx*ny*nz is closest to natoms, at least
    4 cells per side."""
    best = None
    c = max(4, int(round((natoms / 2.0) ** (1.0 / 3.0))))
    for nx in range(max(4, c - 2), c + 3):
        for ny in range(max(4, c - 2), c + 3):
            nz = max(4, int(round(natoms / (2.0 * nx * ny))))
            err = abs(2 * nx * ny * nz - natoms)
            if best is None or err < best[0]:
                best = (err, (nx, ny, nz))
    return best[1]


def make_batch(nbox: int = 256, lo: int = 2000, hi: int = 60000, seed: int = 4242,
               sigma: float = 0.05) -> List[SimMDBox]:
    """Uneven batch of W boxes: sizes log-uniform in [lo, hi] atoms (near-cubic BCC cell
    counts), each with 0-5 vacancies, 0-3 <110> dumbbells and a temperature drawn from
    U(300, 1200) K, every box seeded on its own."""
    rng = np.random.default_rng(seed)
    sizes = batch_sizes(nbox, lo, hi, seed)
    boxes = []
    for k, na in enumerate(sizes):
        nv = int(rng.integers(0, 6))
        nd = int(rng.integers(0, 4))
        temp = float(rng.uniform(300.0, 1200.0))
        boxes.append(make_bcc_box(bcc_cells_for(na), W_A, sigma, temp, seed * 1000 + k,
                                  n_vac=nv, n_dumbbell=nd))
    return boxes
