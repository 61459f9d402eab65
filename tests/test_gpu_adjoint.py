"""GPU parity of the adjoint (msm_adjoint through the C-ABI): knot gradients against the oracle's
adjoint restatement, and coefficient gradients of the force/energy matching loss against central
finite differences of the device forward path (64 strained 1024-atom W boxes, EAM tables)."""
import sys

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md
from msmpscu_amd import adjoint as adj

pytestmark = pytest.mark.gpu

NAMES = ("POTR", "FPOTR", "POTB", "FPOTB", "FEMBD", "DFEMBD")


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _device(boxes, ctrl, ft):
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(ft.PotType, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc.Device


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
@pytest.mark.parametrize("nbox", [1, 3])
def test_adjoint_matches_oracle(orc, pot_type, nbox):
    """Random dL/dF and dL/dE (per box), inactive atoms, one or several uneven boxes: every
    table's knot gradient equals the oracle's to 1e-10 of its largest entry."""
    ctrl = cfg.w_ctrl(mx=64)
    boxes = [cfg.make_bcc_box(c, cfg.W_A, 0.1, 300.0, 40 + k, n_vac=k, n_dumbbell=k % 2)
             for k, c in enumerate([(6, 6, 6), (5, 7, 6), (6, 5, 5)][:nbox])]
    boxes[0].STATU[[7, 100]] = 0
    ft = cfg.w_forcetable(boxes[0], ctrl, pot_type=pot_type)
    say(f"adjoint vs oracle {pot_type} boxes={nbox}")
    dev = _device(boxes, ctrl, ft)
    dev.build_neighbors()
    rng = np.random.default_rng(8)
    ntot = sum(b.NPRT for b in boxes)
    G = rng.normal(size=(ntot, 3))
    g = rng.normal(size=nbox)
    eam = pot_type == "EAM_TYPE"
    gd = adj.table_grads(dev, G, g, ft.NTAB, ft.NEMBD if eam else 1, eam)
    dev.close()
    t = orc.from_mdlib(ft)
    grads = None
    off = 0
    for k, b in enumerate(boxes):
        it = b.ITYP - 1
        kv, ind = orc.nlist("full", b.XP, it, b.STATU, b.BOXLOW, b.ZL, ctrl.IFPD, ctrl.NB_RM,
                            ctrl.NB_MXNBS)
        grads = orc.adjoint_full(b.XP, it, b.STATU, kv, ind, b.ZL, ctrl.IFPD,
                                 float(np.max(ctrl.RU)) ** 2, t, G[off:off + b.NPRT], g[k],
                                 grads)
        off += b.NPRT
    for name in NAMES[: 6 if eam else 4]:
        ref = grads[name.lower()][0, 1:-1]
        got = gd[name]
        scale = np.max(np.abs(ref))
        assert scale > 0, name
        assert np.max(np.abs(got - ref)) <= 1e-10 * scale, (name, np.max(np.abs(got - ref)),
                                                            scale)


def _alloy(case):
    """Two atom types: 'asym' -- W lattice with alternating types, four FS kinds, KPAIR
    [[1, 3], [4, 2]]; 'cuni' -- a Cu-Ni slab, three TB-SMA kinds and two EAM embedding kinds."""
    if case == "asym":
        box = cfg.make_bcc_box((7, 7, 7), cfg.W_A, 0.1, 300.0, 8)
        box.NGROUP = 2
        box.ITYP = (np.arange(box.NPRT) % 2 + 1).astype(np.int32)
        box.CM = np.array([box.CM[0], box.CM[0] * 0.5])
        box.PTYPE = np.array([[1, 3], [4, 2]], dtype=np.int32)
        ctrl = md.SimMDCtrl(RU=np.full((2, 2), 4.4002), NB_RM=np.full((2, 2), 4.7002),
                            NB_MXNBS=64, NUMFTABR=3000)
        ft = md.MDForceTable()
        md.Register_ForceTableProc(1, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
        md.Register_ForceTableProc(2, ft, "msm_pot_ackland_ww_nn", "msm_pot_ackland_ww_ne")
        md.Register_ForceTableProc(3, ft, "msm_pot_wfs_nn", "msm_pot_ackland_ww_ne")
        md.Register_ForceTableProc(4, ft, "msm_pot_wfs_nn", "msm_pot_wfs_ne")
        md.Create_Interaction_ForceTable(box, ctrl, ft)
    else:
        box = cfg.make_cuni_slab((5, 5, 5), 3.57, 600.0, 5)
        ctrl = cfg.cuni_ctrl(ntab=3000)
        ft = cfg.cuni_forcetable(box, ctrl, pot_type="EAM_TYPE")
    box.STATU[[3, 50]] = 0
    return box, ctrl, ft


@pytest.mark.parametrize("case", ["asym", "cuni"])
def test_adjoint_of_several_kinds_matches_oracle(orc, case):
    """Two atom types (asymmetric KPAIR: F'_i and F'_j reach different FPOTB kinds; two EAM
    embedding kinds): every kind of every table equals the oracle's adjoint to 1e-10 of that
    table's largest entry, and every kind in use carries gradient."""
    box, ctrl, ft = _alloy(case)
    say(f"adjoint of several kinds: {case}")
    dev = _device([box], ctrl, ft)
    dev.build_neighbors()
    rng = np.random.default_rng(12)
    G = rng.normal(size=(box.NPRT, 3))
    g = np.array([0.31])
    eam = ft.PotType.strip() == "EAM_TYPE"
    t = orc.from_mdlib(ft)
    gd = adj.table_grads(dev, G, g, ft.NTAB, ft.NEMBD if eam else 1, eam, nkind=t.nkind,
                         nkind1=max(t.nkind1, 1))
    dev.close()
    it = box.ITYP - 1
    kv, ind = orc.nlist("full", box.XP, it, box.STATU, box.BOXLOW, box.ZL, ctrl.IFPD,
                        ctrl.NB_RM, ctrl.NB_MXNBS)
    grads = orc.adjoint_full(box.XP, it, box.STATU, kv, ind, box.ZL, ctrl.IFPD,
                             float(np.max(ctrl.RU)) ** 2, t, G, g[0])
    used = set(int(k) for k in np.unique(t.kpair))
    for name in NAMES[: 6 if eam else 4]:
        ref = grads[name.lower()][:, 1:-1]
        got = gd[name]
        assert got.shape == ref.shape, name
        kinds = used if name not in ("FEMBD", "DFEMBD") else set(int(k) for k in t.kembd[:2])
        for kd in kinds:
            assert np.any(ref[kd] != 0.0), (name, kd)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= 1e-10 * scale, (name, np.max(np.abs(got - ref)),
                                                            scale)


def test_coefficient_gradient_central_differences():
    """64 boxes x 1024 atoms (diagonal strains up to 2%), targets from coefficients with 1%
    seeded noise, L = mean|dF|^2 + 0.1 mean(dE)^2 (dE per atom of each box): dL/dc for 200
    random coefficients among the 3 x 5000 (drawn from those the data reaches) agrees with
    central differences at h = 1e-6 to 1e-6 relative, up to the differences' own rounding floor
    8 eps L / h (L is evaluated in FP64, so no step can resolve it more finely); coefficients the
    data does not reach have zero gradient both ways."""
    boxes, ctrl, ft = adj.make_d_config()
    fm = adj.ForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    fm.set_targets(adj.perturbed(c0, 0.01, seed=5))
    L, grad = fm.loss_and_grad(c0)
    assert L > 0
    h = 1e-6
    floor = 8.0 * np.finfo(np.float64).eps * L / h
    say(f"D config n={fm.n} boxes={fm.nbox} L={L:.6e} FD rounding floor {floor:.2e}")
    rng = np.random.default_rng(21)
    hot = np.argwhere(np.abs(grad) > 0)
    cold = np.argwhere(grad == 0)
    pick = hot[rng.choice(len(hot), size=200, replace=False)]
    worst_rel, worst_ok = 0.0, 0.0
    for (a, k) in pick:
        c = c0.copy()
        c[a, k] += h
        lp = fm.loss(c)
        c[a, k] -= 2 * h
        lm = fm.loss(c)
        fd = (lp - lm) / (2 * h)
        err = abs(fd - grad[a, k])
        worst_ok = max(worst_ok, err / (1e-6 * abs(grad[a, k]) + floor))
        if abs(grad[a, k]) >= 1e7 * floor:
            worst_rel = max(worst_rel, err / abs(grad[a, k]))
    say(f"worst error / (1e-6 |g| + floor) = {worst_ok:.3e}; worst relative error where "
        f"|g| >= 1e7 floor: {worst_rel:.3e}")
    for (a, k) in cold[rng.choice(len(cold), size=5, replace=False)]:
        c = c0.copy()
        c[a, k] += h
        lp = fm.loss(c)
        c[a, k] -= 2 * h
        assert (lp - fm.loss(c)) == 0.0
    fm.close()
    assert worst_ok <= 1.0
    assert worst_rel <= 1e-6


def test_torch_autograd_function_matches_loss_and_grad():
    """The torch.autograd.Function over the C-ABI gives the same loss and coefficient gradient as
    ForceMatch.loss_and_grad."""
    torch = pytest.importorskip("torch")
    boxes, ctrl, ft = adj.make_d_config(nbox=4)
    fm = adj.ForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    fm.set_targets(adj.perturbed(c0, 0.01, seed=6))
    L, grad = fm.loss_and_grad(c0)
    Fn = adj.make_torch_function()
    c = torch.tensor(c0, dtype=torch.float64, requires_grad=True)
    f, e = Fn.apply(c, fm)
    ft_ = torch.from_numpy(fm.f_target)
    et_ = torch.from_numpy(fm.e_target)
    na = torch.from_numpy(fm.natoms)
    lt = ((f - ft_) ** 2).sum() / fm.n + 0.1 * (((e - et_) / na) ** 2).sum() / fm.nbox
    lt.backward()
    fm.close()
    assert abs(lt.item() - L) <= 1e-12 * abs(L)
    assert np.max(np.abs(c.grad.numpy() - grad)) <= 1e-12 * np.max(np.abs(grad))


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
def test_device_spline_tables_equal_host_tables(pot_type):
    """msm_set_spline_table builds the table bits that the host spline map plus
    msm_set_force_table upload: forces and box energies of 4 strained boxes are bitwise equal
    for two coefficient sets taken in alternation (each build overwrites every knot of the
    other), and the two sets give different forces."""
    say(f"device spline tables vs host tables {pot_type}")
    boxes, ctrl, ft = adj.make_d_config(nbox=4, pot_type=pot_type)
    fm = adj.DeviceForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    c1 = adj.perturbed(c0, 0.02, seed=9)
    fh0, eh0 = adj.ForceMatch.forward(fm, c0)
    fd1, ed1 = fm.forward(c1)
    fh1, eh1 = adj.ForceMatch.forward(fm, c1)
    fd0, ed0 = fm.forward(c0)
    fm.close()
    assert np.array_equal(fd0, fh0) and np.array_equal(ed0, eh0)
    assert np.array_equal(fd1, fh1) and np.array_equal(ed1, eh1)
    assert not np.array_equal(fh0, fh1) and not np.array_equal(eh0, eh1)


@pytest.mark.parametrize("pot_type", ["FS_TYPE", "EAM_TYPE"])
def test_device_force_match_equals_host_path(pot_type):
    """DeviceForceMatch (tables, loss, dL/dF, dL/dE, adjoint passes and coefficient transpose on
    the device) against ForceMatch (host spline map, forces and energies copied out, loss and
    transpose in numpy) on 8 strained boxes: targets bitwise equal, the loss to 1e-13 relative
    (the sums run in another order) and dL/dc to 1e-11 of max|dL/dc|, with the same zero
    pattern, at two coefficient points; loss() alone equals loss_and_grad()'s loss bit for
    bit."""
    say(f"device force matching vs host path {pot_type}")
    hb, hctrl, hft = adj.make_d_config(nbox=8, pot_type=pot_type)
    db, dctrl, dft = adj.make_d_config(nbox=8, pot_type=pot_type)
    c0 = adj.SplineModel.from_table(hft).fit(hft)
    host = adj.ForceMatch(hb, hctrl, hft)
    dev = adj.DeviceForceMatch(db, dctrl, dft)
    ct = adj.perturbed(c0, 0.01, seed=5)
    host.set_targets(ct)
    dev.set_targets(ct)
    try:
        assert np.array_equal(host.f_target, dev.f_target)
        assert np.array_equal(host.e_target, dev.e_target)
        for c in (c0, adj.perturbed(c0, 0.005, seed=8)):
            lh, gh = host.loss_and_grad(c)
            ld, gd = dev.loss_and_grad(c)
            scale = np.max(np.abs(gh))
            say(f"L host {lh:.15e} device {ld:.15e}; max|dg| {np.max(np.abs(gd - gh)):.3e} "
                f"of {scale:.3e}")
            assert gd.shape == gh.shape
            assert lh > 0 and abs(ld - lh) <= 1e-13 * lh
            assert np.max(np.abs(gd - gh)) <= 1e-11 * scale
            assert np.array_equal(gd == 0.0, gh == 0.0)
            assert dev.loss(c) == ld
        if pot_type == "FS_TYPE":
            assert np.all(gd[2] == 0.0)
    finally:
        host.close()
        dev.close()


def _hip_runtime():
    """The HIP runtime libmsm_hip.so is linked against (already loaded in this process, so
    dlopen returns that instance; torch ships a second runtime of its own, whose device pointers
    the library's runtime does not know)."""
    import ctypes as C
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def test_device_force_match_with_device_pointers():
    """Coefficients and gradient as device pointers (hipMalloc of the library's own runtime):
    the same loss bit for bit and dL/dc to 1e-12 of max|dL/dc| (the adjoint's LDS atomics may
    round differently between calls), every gradient entry written over a NaN fill; calls out
    of order or with a short row are refused."""
    import ctypes as C
    from msmpscu_amd._lib import check, dptr, lib
    say("device force matching with device pointers")
    boxes, ctrl, ft = adj.make_d_config(nbox=4)
    fm = adj.DeviceForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    out = np.zeros(1)
    rc = lib().msm_fm_loss_grad(fm.dev.h, dptr(out), None, fm.ld, 0)
    assert rc == 7  # MSM_ERR_STATE: no targets yet
    short = np.zeros((3, fm.ld - 1))
    assert lib().msm_set_spline_table(fm.dev.h, dptr(short), fm.ld - 1, 0) == 1  # MSM_ERR_ARG
    fm.set_targets(adj.perturbed(c0, 0.01, seed=6))
    L, g = fm.loss_and_grad(c0)
    hip = _hip_runtime()
    nb = c0.nbytes
    cd, gd = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(cd), nb) == 0 and hip.hipMalloc(C.byref(gd), nb) == 0
    try:
        c = np.ascontiguousarray(c0)
        nan = np.full(c0.shape, np.nan)
        assert hip.hipMemcpy(cd, c.ctypes.data, nb, 1) == 0
        assert hip.hipMemcpy(gd, nan.ctypes.data, nb, 1) == 0
        check(lib().msm_set_spline_table(fm.dev.h, cd, fm.ld, 1), "msm_set_spline_table",
              fm.dev.h)
        check(lib().msm_fm_loss_grad(fm.dev.h, dptr(out), gd, fm.ld, 1), "msm_fm_loss_grad",
              fm.dev.h)
        got = np.zeros(c0.shape)
        assert hip.hipMemcpy(got.ctypes.data, gd, nb, 2) == 0
    finally:
        hip.hipFree(cd)
        hip.hipFree(gd)
        fm.close()
    assert out[0] == L
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - g)) <= 1e-12 * np.max(np.abs(g))


def test_device_force_match_row_lengths():
    """Coefficient rows longer than the tables (ld beyond max(NTAB, NEMBD), which also grows the
    pinned staging buffer while a table build may be queued): the same loss bit for bit, the
    same dL/dc on the common columns to 1e-12 of its largest entry, zeros beyond."""
    from msmpscu_amd._lib import check, dptr, lib
    say("device force matching with longer coefficient rows")
    boxes, ctrl, ft = adj.make_d_config(nbox=4)
    fm = adj.DeviceForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    fm.set_targets(adj.perturbed(c0, 0.01, seed=7))
    L, g = fm.loss_and_grad(c0)
    ld2 = fm.ld + 777
    c2 = np.zeros((3, ld2))
    c2[:, : fm.ld] = c0
    c2[:, fm.ld:] = 1.0e3  # never read: beyond NTAB and NEMBD
    out = np.zeros(1)
    g2 = np.full((3, ld2), np.nan)
    check(lib().msm_set_spline_table(fm.dev.h, dptr(c2), ld2, 0), "msm_set_spline_table",
          fm.dev.h)
    check(lib().msm_fm_loss_grad(fm.dev.h, dptr(out), dptr(g2), ld2, 0), "msm_fm_loss_grad",
          fm.dev.h)
    fm.close()
    assert out[0] == L
    assert np.all(g2[:, fm.ld:] == 0.0)
    assert np.max(np.abs(g2[:, : fm.ld] - g)) <= 1e-12 * np.max(np.abs(g))


def test_fitting_loop_graph_replays_match_direct_launches():
    """msm_fm_loss_grad replays one captured graph on repeated calls; with kernel timing on it
    launches the same work directly.  Over four coefficient sets, alternating the two modes with
    eager work (a host-path table upload and force evaluation) and synchronisations in between,
    every replay gives the direct launches' loss bit for bit and their dL/dc to 1e-12 of its
    largest entry (the adjoint's LDS atomics may round differently between calls)."""
    from msmpscu_amd._lib import MSM_OPT_KERNEL_TIMING
    say("fitting-loop graph replays against direct launches")
    boxes, ctrl, ft = adj.make_d_config(nbox=4)
    fm = adj.DeviceForceMatch(boxes, ctrl, ft)
    c0 = fm.model.fit(ft)
    fm.set_targets(adj.perturbed(c0, 0.01, seed=11))
    try:
        for k in range(4):
            c = adj.perturbed(c0, 0.004 * (k + 1), seed=20 + k)
            fm.dev.set_option(MSM_OPT_KERNEL_TIMING, 1)
            le, ge = fm.loss_and_grad(c)
            fm.dev.set_option(MSM_OPT_KERNEL_TIMING, 0)
            lg, gg = fm.loss_and_grad(c)
            lg2, gg2 = fm.loss_and_grad(c)
            assert lg == le and lg2 == le, (k, lg, lg2, le)
            scale = np.max(np.abs(ge))
            assert np.max(np.abs(gg - ge)) <= 1e-12 * scale
            assert np.max(np.abs(gg2 - ge)) <= 1e-12 * scale
            assert fm.loss(c) == le
            adj.ForceMatch.forward(fm, c)  # eager work on the same context in between
    finally:
        fm.close()
