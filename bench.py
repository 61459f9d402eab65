#!/usr/bin/env python3
"""Benchmark of the tabulated FS/EAM MD hot path on one MI355X.

Default workload: BCC W, 80x80x80 conventional cells minus 0.5% vacancies = 1,018,880 atoms,
periodic, sigma = 0.05 A displacements, 600 K Maxwell velocities, Finnis-Sinclair tables with
5000 points, FP64 velocity Verlet with dt = 1 fs, neighbour list (skin 0.3 A) rebuilt every 10
steps.  A "step" is one full MD step (predictor, rebuild when due, density, force, corrector)
over the whole box with all data resident in HBM.

Prints one JSON line: value = atom-steps/s over the timed steps, a roofline object for the
dominant kernel (algorithmic bytes per launch / average launch time from HIP events), and a
CPU-baseline object (the oracle's serial half-list MD on a bounded sample of the same box).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(kmean: float, track_dis: bool):
    """Algorithmic (compulsory) HBM bytes per atom for each kernel class.

    density: neighbour count + K indices (4K+4), own x,y,z (24), status (4), F' store (8)
    force:   list (4K+4), own position and F' (32), status (4), velocity load+store (48),
             next-position store (24), DIS load+store (48 when tracked)
    Neighbour positions and F' values are gathers that the model counts as cache-served."""
    dens = 4 * kmean + 4 + 24 + 4 + 8
    force = 4 * kmean + 4 + 32 + 4 + 48 + 24 + (48 if track_dis else 0)
    return {"density": dens, "force": force}


def step_model_bytes(kmean: float, n_rb: int) -> float:
    """Whole-step model of the full-list algorithm: list read twice, own x twice, F' write+read,
    F write, fused VV (read x,v,F; write x,v), amortised rebuild."""
    return 2 * (4 * kmean + 4) + 48 + 16 + 24 + 120 + (100 + 4 * kmean) / n_rb


def profiled_traffic(kernel: str):
    """HBM bytes per launch of a pass of the default workload, as measured with FETCH_SIZE /
    WRITE_SIZE (scaled by a copy kernel of known size) and recorded in profiles/."""
    import re
    path = os.path.join(ROOT, "profiles", "p1_hbm_traffic.md")
    try:
        with open(path) as fh:
            m = re.search(r"traffic_bytes_per_launch (.*)", fh.read())
    except OSError:
        return None
    if not m:
        return None
    vals = dict(kv.split("=") for kv in m.group(1).split())
    return int(vals[kernel]) if kernel in vals else None


def oracle_cflags() -> str:
    """Compiler and flags the oracle library was built with (its recipe, oracle/Makefile)."""
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as fh:
            for line in fh:
                if line.startswith("CFLAGS"):
                    return "gcc " + line.split("=", 1)[1].strip()
    except OSError:
        pass
    return "gcc (flags unknown)"


def cpu_baseline(box, ctrl, ft, nsteps: int):
    import numpy as np

    import oracle
    t = oracle.from_mdlib(ft)
    it = box.ITYP - 1
    # initial list and force outside the timed sample (as on the GPU side)
    x, v, f, d, st, ep, kv, ind = oracle.md_run(box.XP, box.XP1, box.STATU, it, box.CM, ctrl.H, 0,
                                               ctrl.NB_UPTAB, 0, box.BOXLOW, box.ZL, ctrl.IFPD,
                                               ctrl.NB_RM, ctrl.RU, ctrl.NB_MXNBS, t)
    t0 = time.perf_counter()
    oracle.md_run(x, v, st, it, box.CM, ctrl.H, nsteps, ctrl.NB_UPTAB, 0, box.BOXLOW, box.ZL,
                  ctrl.IFPD, ctrl.NB_RM, ctrl.RU, ctrl.NB_MXNBS, t, fp=f, init=False)
    dt = time.perf_counter() - t0
    return box.NPRT * nsteps / dt, dt


def make_device(boxes, ctrl, ft, pot):
    import msmpscu_amd.mdlib as md
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(pot, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc, device=0)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc.Device


def run_batch(nbox: int, steps: int, warmup: int, pot: str, tune: bool):
    """Uneven batch: nbox W boxes (2000-60000 atoms, log-uniform; 0-5 vacancies, 0-3 dumbbells,
    300-1200 K) run as one multibox simulation (one launch sequence for all boxes, kernel
    variant chosen by the autotuner on the batch) against the same boxes run one after another,
    one simulation and launch sequence each, with the library defaults."""
    import msmpscu_amd.configs as cfg
    boxes = cfg.make_batch(nbox)
    ntot = sum(b.NPRT for b in boxes)
    ctrl = cfg.w_ctrl(mx=64)
    ft = cfg.w_forcetable(boxes[0], ctrl, pot_type=pot)
    log(f"batch: {nbox} boxes, {ntot} atoms")
    dev = make_device(boxes, ctrl, ft, pot)
    dev.build_neighbors()
    dev.calc_force()
    choice = dev.autotune(3) if tune else None
    log(f"batch autotune: {choice}")
    dev.run(warmup, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.sync()
    t0 = time.perf_counter()
    dev.run(steps, ctrl.NB_UPTAB, warmup, ctrl.H)
    dev.sync()
    tb = time.perf_counter() - t0
    dev.close()
    rate_b = ntot * steps / tb
    log(f"batched: {rate_b:.4e} atom-steps/s ({1e3 * tb / steps:.4f} ms/step)")
    devs = []
    for b in boxes:
        d = make_device([b], ctrl, ft, pot)
        d.build_neighbors()
        d.calc_force()
        d.run(warmup, ctrl.NB_UPTAB, 0, ctrl.H)
        d.sync()
        devs.append(d)
    t0 = time.perf_counter()
    for d in devs:
        d.run(steps, ctrl.NB_UPTAB, warmup, ctrl.H)
        d.sync()
    ts = time.perf_counter() - t0
    for d in devs:
        d.close()
    rate_s = ntot * steps / ts
    log(f"one box at a time: {rate_s:.4e} atom-steps/s; ratio {rate_b / rate_s:.2f}")
    return {"value": round(rate_b, 1), "unit": "atom-steps/s", "per_box_value": round(rate_s, 1),
            "ratio": round(rate_b / rate_s, 3), "boxes": nbox, "atoms": ntot, "steps": steps,
            "warmup": warmup, "autotune": choice,
            "workload": f"{nbox} W boxes, 2000-60000 atoms log-uniform, 0-5 vacancies, 0-3 "
                        "<110> dumbbells, 300-1200 K, FS 5000-pt tables, rebuild every 10"}


def run_slab(ncell: int, steps: int, warmup: int, tune: bool):
    """Cu-Ni TB-SMA slab: FCC ncell^3 cells (4 ncell^3 atoms), random 50/50, periodic in x and
    y with 20 A of vacuum above and below, 1200 K, cutoff 6.0 A (skin 0.3), dt 2 fs, list
    rebuilt every 5 steps; three pair and three density table kinds (Cu-Cu, Ni-Ni, Cu-Ni)."""
    import msmpscu_amd.configs as cfg
    box = cfg.make_cuni_slab((ncell, ncell, ncell), 3.57, 1200.0, 31)
    ctrl = cfg.cuni_ctrl()
    ft = cfg.cuni_forcetable(box, ctrl)
    log(f"slab: {box.NPRT} atoms")
    dev = make_device([box], ctrl, ft, "FS_TYPE")
    dev.build_neighbors()
    dev.calc_force()
    tot, kmax = dev.list_stats()
    choice = dev.autotune(2) if tune else None
    log(f"slab autotune: {choice}")
    dev.run(warmup, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.sync()
    t0 = time.perf_counter()
    dev.run(steps, ctrl.NB_UPTAB, warmup, ctrl.H)
    dev.sync()
    dt = time.perf_counter() - t0
    dev.close()
    rate = box.NPRT * steps / dt
    log(f"slab: {rate:.4e} atom-steps/s ({1e3 * dt / steps:.4f} ms/step)")
    return {"value": round(rate, 1), "unit": "atom-steps/s", "atoms": box.NPRT, "steps": steps,
            "ms_per_step": round(1e3 * dt / steps, 4), "list_mean": round(tot / box.NPRT, 3),
            "list_max": kmax, "autotune": choice,
            "workload": f"Cu-Ni TB-SMA slab, fcc {ncell}^3 cells, 50/50 random, IFPD (1,1,0) "
                        "with 20 A vacuum, 1200 K, cutoff 6.0 A, skin 0.3 A, dt 2 fs, rebuild "
                        "every 5 steps"}


def run_adjoint(reps: int):
    """Coefficient gradients of the force/energy matching loss on 64 strained 1024-atom W
    boxes with EAM tables (3 x 5000 spline coefficients):
    full gradients per second of the device-resident path (tables built from the coefficients
    on the device, forward passes, loss, adjoint passes and coefficient transpose; only the
    coefficients, the loss and dL/dc cross PCIe), the cost of one gradient relative to one
    loss evaluation, and the same figures for the host-table path (spline map, forces and
    energies copied out, loss and transpose in numpy)."""
    from msmpscu_amd import adjoint as adj
    boxes, ctrl, ft = adj.make_d_config()
    c0 = adj.SplineModel.from_table(ft).fit(ft)
    fm = adj.DeviceForceMatch(boxes, ctrl, ft)
    fm.set_targets(adj.perturbed(c0, 0.01, seed=5))
    tm = fm.timing(c0, reps)
    dt = fm.device_timing(c0, reps)
    fm.close()
    hb, hctrl, hft = adj.make_d_config()
    hm = adj.ForceMatch(hb, hctrl, hft)
    hm.set_targets(adj.perturbed(c0, 0.01, seed=5))
    th = hm.timing(c0, max(2, reps // 4))
    hm.close()
    log(f"adjoint (device path): {tm} device passes: {dt} host-table path: {th}")
    return {"value": round(tm["grad_per_s"], 2), "unit": "grad/s",
            "cost_vs_forward": round(tm["cost_ratio"], 3),
            "forward_ms": round(1e3 * tm["forward_s"], 3), "grad_ms": round(1e3 * tm["grad_s"], 3),
            "device_adjoint_vs_forward": round(dt["adjoint_vs_forward"], 3),
            "device_forward_passes_ms": round(dt["forward_passes_ms"], 4),
            "device_adjoint_passes_ms": round(dt["adjoint_passes_ms"], 4),
            "host_tables_grad_per_s": round(th["grad_per_s"], 2),
            "host_tables_grad_ms": round(1e3 * th["grad_s"], 3),
            "workload": "64 W boxes x 1024 atoms (8^3 bcc, sigma 0.1 A, diagonal strain up to "
                        "2%), EAM 5000-pt tables, gradient of mean|dF|^2 + 0.1 mean(dE/atom)^2 "
                        "wrt 3 x 5000 B-spline coefficients, "
                        "FP64, tables, loss and transpose on the device"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    # 500 steps (about 0.1 s of GPU work): kernel traces show the first tens of milliseconds
    # after the host-side setup running 7-20% slower (clock ramp) on some boxes
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--ncell", type=int, default=80)
    ap.add_argument("--mx", type=int, default=32, help="list capacity MXKVOIS of the P1 leg")
    ap.add_argument("--pot", default="FS_TYPE", choices=["FS_TYPE", "EAM_TYPE"])
    ap.add_argument("--no-dis", action="store_true", help="do not accumulate DIS in the step")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing-pass", action="store_true")
    ap.add_argument("--force-path", type=int, default=None,
                    help="density/force kernels: 0 LDS bricks, 1/2/4/8 global list with that "
                         "many lanes per atom, 3 two lanes with lane-pair gathers (default: "
                         "library default)")
    ap.add_argument("--compact", type=int, default=None,
                    help="1: force pass visits only the pairs the density pass found inside the "
                         "cutoff (global-list paths 1, 2, 4)")
    ap.add_argument("--table-lds", type=int, default=None,
                    help="1: table knots from an LDS window (one table kind, paths 1, 2, 4)")
    ap.add_argument("--lt-shape", default=None,
                    help="LDS-table passes: density block,wgs,force block,wgs (e.g. 1024,1,1024,1)")
    ap.add_argument("--nlist-kernel", type=int, default=None,
                    help="list build: 0 automatic, 1 global scan, 2 small bricks, 3 wide bricks")
    ap.add_argument("--force-lanes", type=int, default=None,
                    help="lanes per atom of the LDS-table force pass (0: as --force-path)")
    ap.add_argument("--chunk-sweep", type=int, default=None,
                    help="1: the workgroups of an XCD take its atom chunks round-robin")
    ap.add_argument("--brick-lanes", type=int, default=None,
                    help="lanes per home atom of the LDS-brick passes with the table in LDS")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch measurement")
    ap.add_argument("--batch-boxes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient measurement")
    ap.add_argument("--no-slab", action="store_true", help="skip the Cu-Ni slab measurement")
    ap.add_argument("--slab-ncell", type=int, default=100)
    ap.add_argument("--slab-steps", type=int, default=500)
    ap.add_argument("--adjoint-reps", type=int, default=20)
    ap.add_argument("--cell-brick", type=int, default=None,
                    help="cell numbering: bricks of N^3 cells contiguous (1 = row-major)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device five times (HBM counter "
                         "calibration for profiling) and exit")
    args = ap.parse_args()

    if args.calib_copy:
        import torch
        nbytes = args.calib_copy << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        src.fill_(1)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": nbytes, "copies": 5}), flush=True)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; multi-GPU sharding is a "
                                   "design note, not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    import ctypes as C

    import numpy as np

    import msmpscu_amd.configs as cfg
    import msmpscu_amd.mdlib as md
    from msmpscu_amd import _lib

    nc = args.ncell
    log(f"building workload: bcc {nc}^3, 0.5% vacancies")
    box = cfg.make_bcc_box((nc, nc, nc), cfg.W_A, 0.05, 600.0, 1000, vacancy_frac=0.005)
    # list capacity MXKVOIS: the longest list of this lattice is 26.  The list's 64-atom blocks
    # are mx rows deep and only the rows below the block's longest list are read, so a deep
    # block leaves gaps whose placement in physical memory decides the pass speed (contexts with
    # mx = 64 ran 0.212-0.244 ms/step, with mx = 32 0.209-0.212); 32 keeps the rows dense.
    ctrl = cfg.w_ctrl(mx=args.mx)
    ft = cfg.w_forcetable(box, ctrl, pot_type=args.pot)
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(args.pot, fc)
    md.Initialize_Globle_Variables_DEV(box, ctrl, fc, device=0)
    md.Init_Forcetable_Dev(box, ctrl, fc)
    dev = fc.Device
    track = not args.no_dis
    _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_TRACK_DIS, 1 if track else 0),
               "msm_set_option")
    if args.force_path is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_FORCE_PATH, args.force_path),
                   "msm_set_option")
    if args.compact is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_COMPACT_LIST, args.compact),
                   "msm_set_option")
    if args.table_lds is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_TABLE_LDS, args.table_lds),
                   "msm_set_option")
    if args.cell_brick is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_CELL_BRICK, args.cell_brick),
                   "msm_set_option")
    if args.nlist_kernel is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_NLIST_KERNEL, args.nlist_kernel),
                   "msm_set_option")
    if args.force_lanes is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_FORCE_LANES, args.force_lanes),
                   "msm_set_option")
    if args.chunk_sweep is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_CHUNK_SWEEP, args.chunk_sweep),
                   "msm_set_option")
    if args.brick_lanes is not None:
        _lib.check(_lib.lib().msm_set_option(dev.h, _lib.MSM_OPT_BRICK_LANES, args.brick_lanes),
                   "msm_set_option")
    if args.lt_shape:
        vals = [int(v) for v in args.lt_shape.split(",")]
        for opt, v in zip((_lib.MSM_OPT_LT_DENSITY_BLOCK, _lib.MSM_OPT_LT_DENSITY_WGS,
                           _lib.MSM_OPT_LT_FORCE_BLOCK, _lib.MSM_OPT_LT_FORCE_WGS), vals):
            _lib.check(_lib.lib().msm_set_option(dev.h, opt, v), "msm_set_option")
    dev.build_neighbors()
    dev.calc_force()
    tot, kmax = dev.list_stats()
    kmean = tot / box.NPRT
    pass_kernels = dev.pass_kernel()
    log(f"n={box.NPRT} list mean={kmean:.3f} max={kmax} pass kernels {pass_kernels}")

    log(f"warmup {args.warmup} steps")
    dev.run(args.warmup, ctrl.NB_UPTAB, 0, ctrl.H)
    dev.sync()
    step0 = args.warmup

    log(f"timed {args.steps} steps (graph replay)")
    dev.sync()
    t0 = time.perf_counter()
    dev.run(args.steps, ctrl.NB_UPTAB, step0, ctrl.H)
    dev.sync()
    t1 = time.perf_counter()
    step0 += args.steps
    wall = t1 - t0
    value = box.NPRT * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    log(f"{value:.4e} atom-steps/s, {ms_per_step:.4f} ms/step")

    # per-kernel device time from HIP events on the library's stream, same step count, right
    # after the timed region and on the same device buffers (the host-boundary run below
    # re-allocates them)
    kt = {}
    if not args.no_timing_pass:
        L = _lib.lib()
        _lib.check(L.msm_set_option(dev.h, _lib.MSM_OPT_KERNEL_TIMING, 1), "msm_set_option")
        _lib.check(L.msm_kernel_times(dev.h, None, None, 0, 1), "msm_kernel_times")
        dev.sync()
        te0 = time.perf_counter()
        dev.run(args.steps, ctrl.NB_UPTAB, step0, ctrl.H)
        dev.sync()
        te1 = time.perf_counter()
        nk = len(_lib.KCLASS)
        ms = (C.c_double * nk)()
        cnt = (C.c_longlong * nk)()
        _lib.check(L.msm_kernel_times(dev.h, ms, cnt, nk, 1), "msm_kernel_times")
        _lib.check(L.msm_set_option(dev.h, _lib.MSM_OPT_KERNEL_TIMING, 0), "msm_set_option")
        for k, name in enumerate(_lib.KCLASS):
            if cnt[k]:
                kt[name] = {"total_ms": ms[k], "launches": cnt[k], "avg_us": 1e3 * ms[k] / cnt[k]}
        kt["_eager_wall_ms_per_step"] = 1e3 * (te1 - te0) / args.steps
        step0 += args.steps
        log(f"kernel times: {kt}")

    # the same steps including the host boundary: upload of the atoms, first list and forces,
    # the steps, and the copy-out of positions, velocities and forces
    xh, vh, _, sth, _ = dev.get_atoms()
    dev.sync()
    tp0 = time.perf_counter()
    dev.set_atoms(xh, vh, box.ITYP, sth)
    dev.build_neighbors()
    dev.calc_force()
    dev.run(args.steps, ctrl.NB_UPTAB, step0, ctrl.H)
    dev.get_atoms()
    tp1 = time.perf_counter()
    step0 += args.steps
    pcie_value = box.NPRT * args.steps / (tp1 - tp0)
    log(f"with host upload/download: {pcie_value:.4e} atom-steps/s")

    kb = kernel_bytes(kmean, track)
    roof = None
    if kt:
        cands = {k: kt[k]["total_ms"] for k in ("density", "force") if k in kt}
        dom = max(cands, key=cands.get)
        avg_s = kt[dom]["avg_us"] * 1e-6
        bytes_launch = kb[dom] * box.NPRT
        achieved = bytes_launch / avg_s / 1e9
        default_p1 = nc == 80 and args.pot == "FS_TYPE"
        traffic = profiled_traffic(dom) if default_p1 else None
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "bytes_per_launch": int(bytes_launch),
                "avg_launch_us": round(kt[dom]["avg_us"], 3),
                "traffic_source": "profiles/p1_hbm_traffic.md" if traffic else None}

    step_bytes = step_model_bytes(kmean, ctrl.NB_UPTAB)
    step_model = {"bytes_per_atom_step": round(step_bytes, 1),
                  "achieved_GBs": round(step_bytes * value / 1e9, 2),
                  "frac": round(step_bytes * value / 1e9 / HBM_PEAK_GBS, 4)}

    dev.close()
    batch = None
    if not args.no_batch:
        batch = run_batch(args.batch_boxes, args.batch_steps, 10, args.pot, not args.no_autotune)

    slab = None
    if not args.no_slab:
        slab = run_slab(args.slab_ncell, args.slab_steps, 10, not args.no_autotune)

    grad = None
    if not args.no_adjoint:
        grad = run_adjoint(args.adjoint_reps)

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: oracle serial half-list MD, {args.cpu_steps} steps on the same box")
        rate, dt = cpu_baseline(box, ctrl, ft, args.cpu_steps)
        model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        cpu = {"value": round(rate, 1), "unit": "atom-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} MD steps (incl. one list rebuild) of the same "
                         f"{box.NPRT}-atom box, {dt:.1f} s, serial C restatement of the "
                         "reference host path (half list, Newton's 3rd law), "
                         f"{oracle_cflags()}, one thread of {os.cpu_count()} on {model}"}
        log(f"cpu {rate:.4e} atom-steps/s")

    out = {
        "metric": "atom-steps/s EAM MD on 1 MI355X, batched vs plain; % HBM roofline; grad/s",
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain MD, bcc W {nc}^3 cells - 0.5% vacancies "
                               f"({box.NPRT} atoms), {args.pot} 5000-pt tables, dt 1 fs, "
                               "rebuild every 10 steps, skin 0.3 A",
                   "atoms": box.NPRT, "mxkvois": args.mx, "list_mean": round(kmean, 3),
                   "list_max": kmax,
                   "dis_tracking": track, "pass_kernels": pass_kernels,
                   "parallelism": "single GPU"},
        "pcie_inclusive_value": round(pcie_value, 1),
        "roofline": roof,
        "step_model": step_model,
        "kernel_times": kt,
        "batched": batch,
        "gradient": grad,
        "slab": slab,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
