"""GPU tests of the boundary entry points around the force passes: the density pass alone
(msm_calc_density, the pCalEDen slot), external forces and energies added after the passes
(msm_add_forces / msm_add_epot, host and device arrays, and the Python AddExtForce_ForceClass
list), and the DIS copy-in of a restart (msm_put_dis)."""
import sys

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def device_for(boxes, ctrl, ft):
    boxes = boxes if isinstance(boxes, list) else [boxes]
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(ft.PotType, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc, fc.Device


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
def test_density_pass_alone(orc, pot_type):
    """CalEDen_ForceClass: F' equal bit for bit to the density pass inside msm_calc_force, and to
    the oracle's to 1e-13; the forces are not touched."""
    box = cfg.make_bcc_box((9, 9, 9), cfg.W_A, 0.06, 500.0, 91, n_vac=3)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl, pot_type=pot_type)
    say(f"density alone {pot_type}")
    fc, dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    den = md.CalEDen_ForceClass([box], ctrl, fc)
    f_before = dev.get_forces()
    assert np.all(f_before == 0.0)
    dev.calc_force()
    assert np.array_equal(dev.get_den(), den)
    t = orc.from_mdlib(ft)
    kv, ind = orc.nlist("full", box.XP, box.ITYP - 1, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                        ctrl.NB_RM, ctrl.NB_MXNBS)
    _, do, _ = orc.force_full(box.XP, box.ITYP - 1, box.STATU, kv, ind, box.ZL, ctrl.IFPD,
                              float(np.max(ctrl.RU)) ** 2, t)
    assert np.max(np.abs(den - do)) <= 1e-13
    dev.close()


def _hip_runtime():
    """HIP runtime calls as libmsm_hip.so resolves them: a symbol looked up through the library's
    handle is searched in the library and then its own dependencies, so this is the runtime the
    library uses, whatever else the process has mapped (torch carries a second copy of
    libamdhip64 of its own; memory from one copy is not usable through the other)."""
    from msmpscu_amd._lib import lib
    return lib()


def test_external_forces_and_energies():
    """A registered ForceArray/EpotArray pair adds its terms after the passes (host arrays), the
    same terms from a device array (allocated with the HIP runtime the library uses) give the
    same result, box energies follow, and clearing the list restores the plain forces."""
    import ctypes as C
    hip = _hip_runtime()
    boxes = [cfg.make_bcc_box((6, 6, 6), cfg.W_A, 0.05, 300.0, 5),
             cfg.make_bcc_box((5, 7, 6), cfg.W_A, 0.05, 300.0, 6)]
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(boxes[0], ctrl)
    say("external forces")
    fc, dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    n = dev.n
    rng = np.random.default_rng(3)
    fext = rng.normal(size=(n, 3)) * 0.1
    eext = rng.normal(size=n) * 0.01
    md.CalForce_ForceClass(boxes, ctrl, fc)
    md.CalEpot_ForceClass(boxes, ctrl, fc)
    f0, e0, eb0 = dev.get_forces(), dev.get_epot(), dev.get_box_epot()
    md.AddExtForce_ForceClass(fc, "kick", ForceArray=lambda sb, cp, c: c.Device.add_forces(fext),
                              EpotArray=lambda sb, cp, c: c.Device.add_epot(eext))
    md.CalForce_ForceClass(boxes, ctrl, fc)
    md.CalEpot_ForceClass(boxes, ctrl, fc)
    f1, e1, eb1 = dev.get_forces(), dev.get_epot(), dev.get_box_epot()
    assert np.max(np.abs(f1 - (f0 + fext))) <= 1e-15 * max(1.0, np.max(np.abs(f1)))
    assert np.max(np.abs(e1 - (e0 + eext))) <= 1e-15 * max(1.0, np.max(np.abs(e1)))
    n0 = boxes[0].NPRT
    assert abs(eb1[0] - (eb0[0] + eext[:n0].sum())) <= 1e-11 * abs(eb1[0])
    assert abs(eb1[1] - (eb0[1] + eext[n0:].sum())) <= 1e-11 * abs(eb1[1])
    # a device array: (N, 3) column-major, copied up with hipMemcpy
    dev.calc_force()
    host = np.ascontiguousarray(fext.T)
    dptr_ = C.c_void_p()
    assert hip.hipMalloc(C.byref(dptr_), C.c_size_t(host.nbytes)) == 0
    try:
        assert hip.hipMemcpy(dptr_, host.ctypes.data_as(C.c_void_p), C.c_size_t(host.nbytes),
                             1) == 0  # hipMemcpyHostToDevice
        assert hip.hipDeviceSynchronize() == 0
        dev.add_forces(dptr_.value, on_device=True)
    finally:
        hip.hipFree(dptr_)
    assert np.array_equal(dev.get_forces(), f1)
    md.Clear_ExtForce(fc)
    md.CalForce_ForceClass(boxes, ctrl, fc)
    assert np.array_equal(dev.get_forces(), f0)
    dev.close()


def test_put_dis_restart_resumes_bit_for_bit():
    """30 steps straight against 20 steps, then a new context from the copied-out positions,
    velocities, STATU and DIS (msm_set_atoms + msm_put_dis), a list build, forces, and 10 more
    steps from step 20: identical to the last bit (the rebuild at step 20 sorts the same
    positions into the same order)."""
    box = cfg.make_bcc_box((10, 10, 10), cfg.W_A, 0.05, 800.0, 12)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl)
    say("restart via put_dis")
    fc, dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.run(20, ctrl.NB_UPTAB, 0, ctrl.H)
    x20, v20, _, st20, d20 = dev.get_atoms()
    dev.run(10, ctrl.NB_UPTAB, 20, ctrl.H)
    ref = dev.get_atoms()
    dev.close()
    fc2, dev2 = device_for(box, ctrl, ft)
    dev2.set_atoms(x20, v20, box.ITYP, st20)
    dev2.put_dis(d20)
    dev2.build_neighbors()
    dev2.calc_force()
    dev2.run(10, ctrl.NB_UPTAB, 20, ctrl.H)
    got = dev2.get_atoms()
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    dev2.close()
