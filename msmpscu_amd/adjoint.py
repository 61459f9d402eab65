"""Gradients of a force/energy matching loss with respect to potential coefficients.

The reference has no derivative code; this module adds reverse mode on top of the same forward
path (the device force and energy passes of MD_FS/EAM_ForceTable_GPU.F90) through the C-ABI entry
point msm_adjoint, which returns dL/d(knot) of every table.

Coefficient model (one atom type; FS tables use the first two rows): three cubic B-splines with
unit knot spacing in the reference's own table coordinates, 3 x max(NTAB, NEMBD) coefficients:
  u(r)   = V(r)/2    in s = sqrt(r) * CSI     (pair)       POTR_i  = u(i) r_i
                                                           FPOTR_i = -2 u'(r_i) r_i = -i du/ds(i)
  rho(r)             in s = sqrt(r) * CSI     (density)    POTB_i  = rho(i)
                                                           FPOTB_i = -drho/dr = -drho/ds(i) CSI^2/(2i)
  F(rho)             in m = rho/RHOD + 1      (embedding)  FEMBD_m = F(m), DFEMBD_m = dF/dm(m)/RHOD
with r_i = (i/CSI)^2 (MD_TypeDef_ForceTable.F90:890-982) and rho_m = (m-1) RHOD (:986-1056).  At
an integer knot a cubic B-spline with coefficients c has value (c[i-1] + 4 c[i] + c[i+1])/6 and
slope (c[i+1] - c[i-1])/2; coefficients outside 1..NTAB are zero.  The tables are linear in the
coefficients, so dL/dc is the transpose of that map applied to msm_adjoint's knot gradients.
CSI and RHOD stay fixed (they are grid constants, not parameters).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._lib import check, dptr, lib
from .mdlib import MDForceTable

TABLES = ("POTR", "FPOTR", "POTB", "FPOTB", "FEMBD", "DFEMBD")


def _val(c: np.ndarray) -> np.ndarray:
    p = np.concatenate([[0.0], c, [0.0]])
    return (p[:-2] + 4.0 * p[1:-1] + p[2:]) / 6.0


def _slope(c: np.ndarray) -> np.ndarray:
    p = np.concatenate([[0.0], c, [0.0]])
    return 0.5 * (p[2:] - p[:-2])


def _val_t(y: np.ndarray) -> np.ndarray:
    q = np.concatenate([[0.0], y, [0.0]])
    return (q[:-2] + 4.0 * q[1:-1] + q[2:]) / 6.0


def _slope_t(y: np.ndarray) -> np.ndarray:
    q = np.concatenate([[0.0], y, [0.0]])
    return 0.5 * (q[:-2] - q[2:])


@dataclass
class SplineModel:
    ntab: int
    csi: float
    nembd: int
    rhod: float

    @classmethod
    def from_table(cls, ft: MDForceTable) -> "SplineModel":
        eam = ft.PotType.strip() == "EAM_TYPE"
        return cls(int(ft.NTAB), float(ft.CSI), int(ft.NEMBD) if eam else 0, float(ft.RHOD))

    @property
    def shape(self):
        return (3, max(self.ntab, self.nembd))

    def _geom(self):
        i = np.arange(1, self.ntab + 1, dtype=np.float64)
        return i, (i / self.csi) ** 2

    def tables(self, c: np.ndarray) -> Dict[str, np.ndarray]:
        """Knot rows 1..NTAB (1..NEMBD) of the six tables for coefficients c (3, n)."""
        i, r = self._geom()
        cp, cq, ce = c[0, : self.ntab], c[1, : self.ntab], c[2, : self.nembd]
        return {
            "POTR": _val(cp) * r,
            "FPOTR": -i * _slope(cp),
            "POTB": _val(cq),
            "FPOTB": -_slope(cq) * (self.csi * self.csi) / (2.0 * i),
            "FEMBD": _val(ce),
            "DFEMBD": _slope(ce) / self.rhod,
        }

    def coeff_grad(self, g: Dict[str, np.ndarray]) -> np.ndarray:
        """dL/dc (3, n) from dL/d(knot) of the six tables (the transpose of tables())."""
        i, r = self._geom()
        out = np.zeros(self.shape)
        out[0, : self.ntab] = _val_t(g["POTR"] * r) + _slope_t(-i * g["FPOTR"])
        out[1, : self.ntab] = _val_t(g["POTB"]) + _slope_t(-g["FPOTB"] * (self.csi * self.csi) /
                                                           (2.0 * i))
        out[2, : self.nembd] = _val_t(g["FEMBD"]) + _slope_t(g["DFEMBD"] / self.rhod)
        return out

    def fit(self, ft: MDForceTable) -> np.ndarray:
        """Coefficients whose value tables interpolate POTR, POTB and FEMBD of ft at every
        knot (one tridiagonal solve each)."""
        from scipy.linalg import solve_banded
        _, r = self._geom()

        def solve(v):
            n = v.shape[0]
            ab = np.zeros((3, n))
            ab[0, 1:] = 1.0 / 6.0
            ab[1, :] = 4.0 / 6.0
            ab[2, :-1] = 1.0 / 6.0
            return solve_banded((1, 1), ab, v)

        c = np.zeros(self.shape)
        c[0, : self.ntab] = solve(np.asarray(ft.POTR[0], dtype=np.float64) / r)
        c[1, : self.ntab] = solve(np.asarray(ft.POTB[0], dtype=np.float64))
        if self.nembd:
            c[2, : self.nembd] = solve(np.asarray(ft.FEMBD[0], dtype=np.float64))
        return c

    def apply(self, c: np.ndarray, ft: MDForceTable) -> MDForceTable:
        """Writes the tables of c into ft (one kind) and returns it."""
        t = self.tables(c)
        ft.POTR = t["POTR"][None, :].copy(order="F")
        ft.FPOTR = t["FPOTR"][None, :].copy(order="F")
        ft.POTB = t["POTB"][None, :].copy(order="F")
        ft.FPOTB = t["FPOTB"][None, :].copy(order="F")
        if self.nembd:
            ft.FEMBD = t["FEMBD"][None, :].copy(order="F")
            ft.DFEMBD = t["DFEMBD"][None, :].copy(order="F")
        return ft


def table_grads(dev, dldf: np.ndarray, dlde: np.ndarray, ntab: int, nembd: int,
                eam: bool = True, nkind: int = 1, nkind1: int = 1) -> Dict[str, np.ndarray]:
    """msm_adjoint on a DeviceMD: dL/d(knot) of the six tables (embedding ones zero for FS).
    With one kind each array has shape (NTAB,) / (NEMBD,); with several, (NKIND, NTAB) and
    (NKIND1, NEMBD), the shapes of MDForceTable's arrays."""
    g = {k: np.zeros(nkind * ntab) for k in ("POTR", "FPOTR", "POTB", "FPOTB")}
    g["FEMBD"] = np.zeros(nkind1 * nembd)
    g["DFEMBD"] = np.zeros(nkind1 * nembd)
    f = np.ascontiguousarray(np.asarray(dldf, dtype=np.float64).ravel(order="F"))
    e = np.ascontiguousarray(np.asarray(dlde, dtype=np.float64))
    check(lib().msm_adjoint(dev.h, dptr(f), dptr(e), dptr(g["POTR"]), dptr(g["FPOTR"]),
                            dptr(g["POTB"]), dptr(g["FPOTB"]),
                            dptr(g["FEMBD"]) if eam else None,
                            dptr(g["DFEMBD"]) if eam else None), "msm_adjoint", dev.h)
    if nkind == 1 and nkind1 == 1:
        return g
    # the C layout is [k + (i-1) * NKIND]
    out = {k: g[k].reshape(ntab, nkind).T.copy() for k in ("POTR", "FPOTR", "POTB", "FPOTB")}
    out.update({k: g[k].reshape(nembd, nkind1).T.copy() for k in ("FEMBD", "DFEMBD")})
    return out


class ForceMatch:
    """Force/energy matching on a fixed set of boxes:
        L(c) = mean_i |F_i(c) - F*_i|^2 + w_e * mean_b ((E_b(c) - E*_b) / N_b)^2
    (energies per atom of each box, so both terms are of the same order), with F, E from the
    device passes and dL/dc from msm_adjoint plus the spline transpose."""

    def __init__(self, boxes, ctrl, ft: MDForceTable, device: int = 0, w_e: float = 0.1):
        from . import mdlib as md
        self.boxes = boxes if isinstance(boxes, list) else [boxes]
        self.ctrl = ctrl
        self.ft = ft
        self.model = SplineModel.from_table(ft)
        self.eam = ft.PotType.strip() == "EAM_TYPE"
        fc = md.MDForceClassGPU(ForceTable=ft)
        md.Register_ForceClass(ft.PotType, fc)
        md.Initialize_Globle_Variables_DEV(self.boxes, ctrl, fc, device=device)
        md.Init_Forcetable_Dev(self.boxes, ctrl, fc)
        self.dev = fc.Device
        self.dev.build_neighbors()
        self.n = self.dev.n
        self.nbox = len(self.boxes)
        self.natoms = np.array([b.NPRT for b in self.boxes], dtype=np.float64)
        self.w_e = w_e
        self.f_target: Optional[np.ndarray] = None
        self.e_target: Optional[np.ndarray] = None

    def forward(self, c: np.ndarray):
        """Forces (N,3) in the caller's atom order and box energies for coefficients c."""
        self.model.apply(c, self.ft)
        self.dev.set_force_table(self.ft)
        self.dev.calc_force()
        self.dev.calc_epot()
        return self.dev.get_forces(), self.dev.get_box_epot()

    def set_targets(self, c_target: np.ndarray):
        self.f_target, self.e_target = self.forward(c_target)

    def loss(self, c: np.ndarray) -> float:
        f, e = self.forward(c)
        return self._loss(f, e)

    def _loss(self, f, e) -> float:
        df = f - self.f_target
        de = (e - self.e_target) / self.natoms
        return float(np.sum(df * df) / self.n + self.w_e * np.sum(de * de) / self.nbox)

    def loss_and_grad(self, c: np.ndarray):
        f, e = self.forward(c)
        L = self._loss(f, e)
        dldf = 2.0 * (f - self.f_target) / self.n
        dlde = 2.0 * self.w_e * (e - self.e_target) / (self.natoms ** 2 * self.nbox)
        g = table_grads(self.dev, dldf, dlde, self.model.ntab, self.model.nembd, self.eam)
        return L, self.model.coeff_grad(g)

    def timing(self, c: np.ndarray, reps: int = 20) -> Dict[str, float]:
        """Wall time per forward (tables, passes, copy-out) and per full gradient."""
        self.loss(c)
        t0 = time.perf_counter()
        for _ in range(reps):
            self.loss(c)
        tf = (time.perf_counter() - t0) / reps
        self.loss_and_grad(c)
        t0 = time.perf_counter()
        for _ in range(reps):
            self.loss_and_grad(c)
        tg = (time.perf_counter() - t0) / reps
        return {"forward_s": tf, "grad_s": tg, "grad_per_s": 1.0 / tg,
                "cost_ratio": tg / tf}

    def device_timing(self, c: np.ndarray, reps: int = 20) -> Dict[str, float]:
        """Device time of the passes alone, from HIP events on the library's stream: per
        gradient, the adjoint passes (msm_adjoint's kernels) against the forward density and
        force passes of the same evaluation (the energy pass and all copies excluded)."""
        from ._lib import MSM_OPT_KERNEL_TIMING
        self.loss_and_grad(c)
        self.dev.set_option(MSM_OPT_KERNEL_TIMING, 1)
        self.dev.kernel_times(reset=True)
        for _ in range(reps):
            self.loss_and_grad(c)
        kt = self.dev.kernel_times(reset=True)
        self.dev.set_option(MSM_OPT_KERNEL_TIMING, 0)
        fwd = (kt["density"][0] + kt["force"][0]) / reps
        adjm = kt["adjoint"][0] / reps
        return {"forward_passes_ms": fwd, "adjoint_passes_ms": adjm,
                "adjoint_vs_forward": adjm / fwd if fwd > 0 else float("nan")}

    def close(self):
        self.dev.close()


class DeviceForceMatch(ForceMatch):
    """ForceMatch with the whole evaluation on the device: msm_set_spline_table builds the
    tables from the coefficients in one kernel, msm_fm_loss_grad runs the forward passes, the
    loss against targets kept on the device, the adjoint passes and the coefficient transpose,
    and only the loss and dL/dc (3 x ld doubles) come back to the host.  Same L and dL/dc as
    ForceMatch (the tables are bit-identical; the loss sums run in another order)."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.ld = self.model.shape[1]

    def _coef(self, c: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(np.asarray(c, dtype=np.float64))
        if c.shape != (3, self.ld):
            raise ValueError(f"coefficients must have shape (3, {self.ld})")
        return c

    def set_tables(self, c: np.ndarray):
        check(lib().msm_set_spline_table(self.dev.h, dptr(self._coef(c)), self.ld, 0),
              "msm_set_spline_table", self.dev.h)

    def forward(self, c: np.ndarray):
        self.set_tables(c)
        self.dev.calc_force()
        self.dev.calc_epot()
        return self.dev.get_forces(), self.dev.get_box_epot()

    def set_targets(self, c_target: np.ndarray):
        f, e = self.forward(c_target)
        self.f_target, self.e_target = f, e
        ff = np.ascontiguousarray(f.ravel(order="F"))
        ee = np.ascontiguousarray(e, dtype=np.float64)
        check(lib().msm_fm_set_targets(self.dev.h, dptr(ff), dptr(ee), float(self.w_e)),
              "msm_fm_set_targets", self.dev.h)

    def loss(self, c: np.ndarray) -> float:
        self.set_tables(c)
        out = np.empty(1)
        check(lib().msm_fm_loss_grad(self.dev.h, dptr(out), None, self.ld, 0),
              "msm_fm_loss_grad", self.dev.h)
        return float(out[0])

    def loss_and_grad(self, c: np.ndarray):
        self.set_tables(c)
        out = np.empty(1)
        g = np.empty((3, self.ld))  # every entry is written
        check(lib().msm_fm_loss_grad(self.dev.h, dptr(out), dptr(g), self.ld, 0),
              "msm_fm_loss_grad", self.dev.h)
        return float(out[0]), g


def make_torch_function():
    """torch.autograd.Function over the C-ABI: coefficients (3, n) float64 tensor -> (forces
    (N,3), box energies (B,)); backward calls msm_adjoint.  Built lazily so that importing this
    module does not need torch."""
    import torch

    class TablesForward(torch.autograd.Function):
        @staticmethod
        def forward(ctx, coeffs, fm: ForceMatch):
            c = coeffs.detach().cpu().numpy()
            f, e = fm.forward(c)
            ctx.fm = fm
            return (torch.from_numpy(f).to(coeffs.device), torch.from_numpy(e).to(coeffs.device))

        @staticmethod
        def backward(ctx, gf, ge):
            fm = ctx.fm
            dldf = gf.detach().cpu().numpy() if gf is not None else np.zeros((fm.n, 3))
            dlde = ge.detach().cpu().numpy() if ge is not None else np.zeros(fm.nbox)
            g = table_grads(fm.dev, dldf, dlde, fm.model.ntab, fm.model.nembd, fm.eam)
            gc = torch.from_numpy(fm.model.coeff_grad(g)).to(gf.device if gf is not None
                                                            else ge.device)
            return gc, None

    return TablesForward


def make_d_config(nbox: int = 64, seed: int = 77, pot_type: str = "EAM_TYPE"):
    """64 W boxes of 8^3 BCC cells (1024 atoms), sigma = 0.1 A, each with its own diagonal
    homogeneous strain drawn from U(-2%, 2%) per component, EAM tables (5000 points)."""
    from . import configs as cfg
    rng = np.random.default_rng(seed)
    boxes = [cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.1, 0.0, seed * 100 + k,
                              strain=rng.uniform(-0.02, 0.02, size=3)) for k in range(nbox)]
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(boxes[0], ctrl, pot_type=pot_type)
    return boxes, ctrl, ft


def perturbed(c: np.ndarray, rel: float = 0.01, seed: int = 5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return c * (1.0 + rel * rng.normal(size=c.shape))
