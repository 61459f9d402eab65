"""A two-element setfl file (written here from closed forms) through the whole chain on the GPU:
reader, generated tables with asymmetric density channels (the table of A<-B holds rho_B, that of
B<-A rho_A), device force, F' and energy of a displaced random bcc alloy against the oracle on
the same tables, through the C-ABI."""
import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md
from msmpscu_amd import nist as setfl
from test_gpu_parity import E_REL, F_TOL, device_for, oracle_full, say
from test_setfl import _w_pair, _w_rho, _write_setfl

pytestmark = pytest.mark.gpu


def test_setfl_alloy_matches_oracle(orc, tmp_path):
    path = str(tmp_path / "ab.setfl")
    _write_setfl(path, ["A", "B"], 4000, 0.1, 4000, 4.8,
                 [lambda rho: -np.sqrt(rho), lambda rho: -1.1 * np.sqrt(rho)],
                 [_w_rho, lambda r: 1.3 * _w_rho(r)],
                 {(0, 0): lambda r: r * _w_pair(r), (1, 0): lambda r: 0.9 * r * _w_pair(r),
                  (1, 1): lambda r: 1.2 * r * _w_pair(r)})
    ft = md.MDForceTable(PotType="EAM_TYPE", NTAB=5000, NEMBD=5000)
    setfl.Register_ForceTableProc_Setfl(path, ft)
    setfl.Generate_NIST_ForceTable(ft)
    box = cfg.make_bcc_box((8, 8, 8), cfg.W_A, 0.05, 300.0, 61)
    rng = np.random.default_rng(62)
    box.NGROUP = 2
    box.ITYP = rng.choice([1, 2], box.NPRT).astype(np.int32)
    box.CM = np.array([box.CM[0], 1.5 * box.CM[0]])
    box.PTYPE = np.array([[1, 2], [3, 4]], np.int32)
    ctrl = cfg.w_ctrl()
    ctrl.RU = np.full((2, 2), 4.7)
    ctrl.NB_RM = np.full((2, 2), 5.0)
    ctrl.NB_MXNBS = 96
    say(f"setfl alloy n={box.NPRT}")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.calc_epot()
    _, _, fp, _, _ = dev.get_atoms()
    den = dev.get_den()
    ep = dev.get_epot()
    fo, do, eo = oracle_full(orc, box, ctrl, ft)
    assert np.max(np.abs(fo)) > 1.0  # the alloy is far from equilibrium
    assert np.max(np.abs(fp - fo)) <= F_TOL
    assert np.max(np.abs(den - do)) <= 1e-13
    assert abs(ep.sum() - eo.sum()) <= E_REL * abs(eo.sum())
    dev.close()
