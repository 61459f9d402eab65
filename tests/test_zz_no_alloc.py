"""msm_calc_ptensor_noalloc, the entry point that allocates no device memory (its per-box
partial sums sit in a fixed device array), against msm_calc_ptensor on the same contexts, through
the C-ABI: the per-box tensors agree to 1e-12 of each box's largest component (the sums are added
in another fixed order) and the forces it stores bit for bit.  Two calls in a row: a three-box
system with inactive atoms (the out-of-cell bucket blocks), then a smaller single box, which
reads back only what it wrote itself."""
import sys

import numpy as np
import pytest

import msmpscu_amd.configs as cfg
import msmpscu_amd.mdlib as md
from msmpscu_amd._lib import check, dptr, lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _device(boxes, ctrl, ft):
    fc = md.MDForceClassGPU(ForceTable=ft)
    md.Register_ForceClass(ft.PotType, fc)
    md.Initialize_Globle_Variables_DEV(boxes, ctrl, fc)
    md.Init_Forcetable_Dev(boxes, ctrl, fc)
    return fc.Device


def _noalloc(dev):
    v = np.zeros(9 * dev.nbox)
    check(lib().msm_calc_ptensor_noalloc(dev.h, dptr(v)), "msm_calc_ptensor_noalloc", dev.h)
    return v.reshape(dev.nbox, 3, 3).transpose(0, 2, 1).copy()


def _compare(dev):
    dev.build_neighbors()
    v1 = _noalloc(dev)
    f1 = dev.get_forces()
    v0 = dev.calc_ptensor()
    f0 = dev.get_forces()
    assert np.array_equal(f1, f0)
    assert np.any(f0 != 0.0)
    for b in range(dev.nbox):
        scale = np.max(np.abs(v0[b]))
        assert scale > 0.0
        assert np.max(np.abs(v1[b] - v0[b])) <= 1e-12 * scale, b
        assert np.array_equal(v1[b], v1[b].T)
    return v1


def test_noalloc_ptensor_matches_ptensor_on_two_calls():
    ctrl = cfg.w_ctrl(mx=64)
    boxes = [cfg.make_bcc_box(c, cfg.W_A, 0.05, 500.0, 70 + k, n_vac=k)
             for k, c in enumerate([(5, 6, 7), (9, 9, 9), (4, 4, 5)])]
    for b in boxes:
        b.STATU[[1, 7]] = 0
    ft = cfg.w_forcetable(boxes[0], ctrl)
    say("no-alloc ptensor: three boxes with inactive atoms")
    dev = _device(boxes, ctrl, ft)
    try:
        _compare(dev)
    finally:
        dev.close()

    small = cfg.make_bcc_box((6, 6, 6), cfg.W_A, 0.04, 300.0, 77)
    say("no-alloc ptensor: one smaller box")
    dev = _device([small], ctrl, cfg.w_forcetable(small, ctrl))
    try:
        v = _compare(dev)
        assert v.shape == (1, 3, 3)
    finally:
        dev.close()
