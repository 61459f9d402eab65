"""GPU parity at the benchmark configurations themselves, on the kernels the benchmark runs.

- the 1.02M-atom W box (80^3 BCC, 0.5% vacancies, 600 K): neighbour list as sorted (i, j) sets,
  forces and energy against the oracle at step 0 and, after 1000 graph-replayed steps with a
  rebuild every 10, on the GPU's own snapshot (trajectories diverge chaotically, so per-snapshot
  comparisons are the meaningful ones);
- the uneven batch (256 W boxes of 2,000-60,000 atoms, vacancies, <110> dumbbells, 300-1200 K)
  on the autotuned kernels: every box's list as a sorted set, its forces and its energy against
  the oracle at step 0 and after 500 steps;
- the 4M-atom Cu-Ni TB-SMA slab (free z surfaces, three pair and two density kinds) on the
  autotuned kernels: lists, forces and energy at step 0 and after 100 steps (20 rebuilds);
- two time steps on one context: the graphs captured for one h are not replayed for another.

Bars as everywhere: lists bit-exact as sets; forces within 1e-10 eV/A max-abs; total energy
within 1e-12 relative."""
import sys

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md
from msmpscu_amd._lib import MSM_OPT_KERNEL_TIMING

pytestmark = pytest.mark.gpu

F_TOL = 1e-10
E_REL = 1e-12


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def device_for(boxes, ctrl, ft):
    boxes = boxes if isinstance(boxes, list) else [boxes]
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(ft.PotType, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc.Device


def sorted_pairs(kv, ind, one_based):
    """The (i, j) entries of a list (KVOIS, INDI[N, MX]) as sorted int64 keys i * N + j."""
    n, mx = ind.shape
    mask = np.arange(mx)[None, :] < np.asarray(kv)[:, None]
    ii = np.broadcast_to(np.arange(n, dtype=np.int64)[:, None], (n, mx))[mask]
    jj = ind[mask].astype(np.int64) - (1 if one_based else 0)
    return np.sort(ii * n + jj)


def lists_equal_as_sets(dev, orc, xp, ityp, statu, lo, ln, ifpd, nb_rm, mx):
    kv, ind = dev.get_neighbors()
    okv, oind = orc.nlist("full", xp, ityp - 1, statu, lo, ln, ifpd, nb_rm, mx)
    assert np.array_equal(kv, okv)
    assert np.array_equal(sorted_pairs(kv, ind, True), sorted_pairs(okv, oind, False))
    return int(okv.sum())


def oracle_half(orc, xp, ityp, statu, lo, ln, ifpd, nb_rm, ru, mx, ft):
    t = orc.from_mdlib(ft)
    kv, ind = orc.nlist("half", xp, ityp - 1, statu, lo, ln, ifpd, nb_rm, mx)
    fp, ep, *_ = orc.force_fs_half(xp, ityp - 1, kv, ind, ln, ifpd, ru, t)
    return fp, ep


def check_snapshot(dev, orc, box, ctrl, ft, label, xp=None):
    """List (as sets), forces and total energy of the device state against the oracle evaluated
    at the same positions (the device's own positions when xp is None)."""
    dev.calc_epot()
    x, _, fp, st, _ = dev.get_atoms()
    if xp is not None:
        x = xp
    npairs = lists_equal_as_sets(dev, orc, x, box.ITYP, st, box.BOXLOW, box.ZL, ctrl.IFPD,
                                 ctrl.NB_RM, ctrl.NB_MXNBS)
    fo, eo = oracle_half(orc, x, box.ITYP, st, box.BOXLOW, box.ZL, ctrl.IFPD, ctrl.NB_RM,
                         ctrl.RU, ctrl.NB_MXNBS, ft)
    df = float(np.max(np.abs(fp - fo)))
    e, e0 = float(dev.get_epot().sum()), float(eo.sum())
    say(f"{label}: pairs={npairs} max|dF|={df:.3e} |dE|/|E|={abs(e - e0) / abs(e0):.3e}")
    assert df <= F_TOL
    assert abs(e - e0) <= E_REL * abs(e0)
    return x


@pytest.mark.timeout(300)
def test_p1_million_atoms_step0_and_step1000(orc):
    """The 1,018,880-atom W box: step 0, then 1000 steps (dt 1 fs, rebuild every 10) through the
    graph-replayed msm_run; the oracle re-evaluates the GPU's step-1000 snapshot, whose list both
    sides build again on the snapshot's positions."""
    box = cfg.make_bcc_box((80, 80, 80), cfg.W_A, 0.05, 600.0, 1000, vacancy_frac=0.005)
    ctrl = cfg.w_ctrl(mx=32)  # the benchmark's list capacity (longest list 26)
    ft = cfg.w_forcetable(box, ctrl)
    say(f"P1 n={box.NPRT}")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    check_snapshot(dev, orc, box, ctrl, ft, "P1 step 0", xp=box.XP)
    dev.run(1000, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.sync()
    # the step-1000 snapshot: rebuild the list on it so both sides list the same positions
    dev.build_neighbors()
    dev.calc_force()
    x = check_snapshot(dev, orc, box, ctrl, ft, "P1 step 1000")
    # the atoms moved: a snapshot equal to the start would make the check vacuous
    d = x - box.XP
    d -= np.round(d / box.ZL) * box.ZL
    assert np.max(np.abs(d)) > 0.05
    dev.close()


@pytest.mark.timeout(300)
def test_uneven_batch_256_boxes_vs_oracle(orc):
    """The benchmark batch (256 boxes, 2,000-60,000 atoms) on the autotuned kernels: each box's
    list, forces and energy against the oracle at step 0 and after 500 graph-replayed steps."""
    ctrl = cfg.w_ctrl(mx=64)
    boxes = cfg.make_batch(256, 2000, 60000)
    ft = cfg.w_forcetable(boxes[0], ctrl)
    sizes = [b.NPRT for b in boxes]
    say(f"batch 256 boxes, {sum(sizes)} atoms, {min(sizes)}..{max(sizes)} per box")
    dev = device_for(boxes, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    dev.autotune(2)
    dev.calc_force()

    def per_box(label):
        dev.calc_epot()
        x, _, fp, st, _ = dev.get_atoms()
        kv, ind = dev.get_neighbors()
        eb = dev.get_box_epot()
        off = 0
        worst_f, worst_e = 0.0, 0.0
        for k, b in enumerate(boxes):
            s = slice(off, off + b.NPRT)
            xk = np.asfortranarray(x[s])
            okv, oind = orc.nlist("full", xk, b.ITYP - 1, st[s], b.BOXLOW, b.ZL, ctrl.IFPD,
                                  ctrl.NB_RM, ctrl.NB_MXNBS)
            # device ids are 1-based in the caller's order of the whole batch
            assert np.array_equal(kv[s], okv), k
            dind = ind[s].copy()
            dind[dind > 0] -= off
            assert np.array_equal(sorted_pairs(kv[s], dind, True),
                                  sorted_pairs(okv, oind, False)), k
            fo, eo = oracle_half(orc, xk, b.ITYP, st[s], b.BOXLOW, b.ZL, ctrl.IFPD, ctrl.NB_RM,
                                 ctrl.RU, ctrl.NB_MXNBS, ft)
            worst_f = max(worst_f, float(np.max(np.abs(fp[s] - fo))))
            worst_e = max(worst_e, abs(eb[k] - eo.sum()) / abs(eo.sum()))
            off += b.NPRT
        say(f"batch {label}: worst max|dF|={worst_f:.3e} worst |dE|/|E|={worst_e:.3e}")
        assert worst_f <= F_TOL
        assert worst_e <= E_REL
        return x

    per_box("step 0")
    dev.run(500, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.sync()
    dev.build_neighbors()
    dev.calc_force()
    x = per_box("step 500")
    assert not np.array_equal(x, np.concatenate([b.XP for b in boxes]))
    dev.close()


@pytest.mark.timeout(400)
def test_cuni_slab_4m_atoms_vs_oracle(orc):
    """The 4,000,000-atom Cu-Ni slab on its autotuned kernels (wide-brick list, LDS tables of the
    three kinds): lists, forces and energy at step 0 and after 100 steps of dt 2 fs with a
    rebuild every 5."""
    box = cfg.make_cuni_slab((100, 100, 100), 3.57, 1200.0, 31)
    ctrl = cfg.cuni_ctrl()
    ft = cfg.cuni_forcetable(box, ctrl)
    say(f"slab n={box.NPRT}")
    dev = device_for(box, ctrl, ft)
    dev.build_neighbors()
    dev.calc_force()
    choice = dev.autotune(2)
    say(f"slab autotune {choice}")
    dev.calc_force()
    check_snapshot(dev, orc, box, ctrl, ft, "slab step 0", xp=box.XP)
    dev.run(100, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.sync()
    dev.build_neighbors()
    dev.calc_force()
    check_snapshot(dev, orc, box, ctrl, ft, "slab step 100")
    dev.close()


def test_two_time_steps_on_one_context(orc):
    """msm_run with h1, then with h2 on the same context (whose graphs for h1 are cached): equal
    to the eager launches (no graphs) bit for bit, and to the oracle's two runs."""
    box = cfg.make_bcc_box((10, 10, 10), cfg.W_A, 0.05, 900.0, 5)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(box, ctrl)
    h1, h2 = 1.0, 0.5
    say("two time steps on one context")
    outs = []
    for eager in (0, 1):
        dev = device_for(box, ctrl, ft)
        dev.set_option(MSM_OPT_KERNEL_TIMING, eager)
        dev.build_neighbors()
        dev.calc_force()
        dev.run(15, ctrl.NB_UPTAB, 0, h1)
        dev.run(15, ctrl.NB_UPTAB, 15, h2)
        dev.run(15, ctrl.NB_UPTAB, 30, h1)
        outs.append(dev.get_atoms())
        dev.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    t = orc.from_mdlib(ft)
    args = (box.STATU, box.ITYP - 1, box.CM)
    geo = (box.BOXLOW, box.ZL, ctrl.IFPD, ctrl.NB_RM, ctrl.RU, ctrl.NB_MXNBS, t)
    ox, ov, of, *_ = orc.md_run(box.XP, box.XP1, *args, h1, 15, ctrl.NB_UPTAB, 0, *geo)
    # each continuation lists and evaluates the forces at its start again (same forces to
    # rounding as the device's, which keeps its list from the last rebuild)
    ox, ov, of, *_ = orc.md_run(ox, ov, *args, h2, 15, ctrl.NB_UPTAB, 15, *geo)
    ox, ov, of, *_ = orc.md_run(ox, ov, *args, h1, 15, ctrl.NB_UPTAB, 30, *geo)
    x, v, fp, _, _ = outs[0]
    dx = x - ox
    dx -= np.round(dx / box.ZL) * box.ZL
    assert np.max(np.abs(dx)) <= 1e-9
    assert np.max(np.abs(v - ov)) <= 1e-11
    assert np.max(np.abs(fp - of)) <= 1e-8
