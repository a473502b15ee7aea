"""The two convection kernels against each other on the GPU: the line kernel
(a lane per MLT or R line of a (V,K) plane, the line in registers through
every inner QUICKEST step) and the tile kernel (a lane per grid point, four
barrier-separated passes over LDS per step). Both restate the same
operations of src/Convection_2D.cpp:36-169 and
src/Convection_1D_ULTIMATE_QUICKEST6.cpp:22-482 in the same order, and the
tile kernel is pinned bit for bit against the reference's own fixtures
(test_gpu_parity.py, test_gpu_options.py), so the bar here is bit-exact:
the PSD and the lost-PSD accumulator, on the production grid at Kp 6 with
several planes per workgroup (a last, partial tile of planes included) and
on a config-0 grid with convection losses, sources, the lost-PSD
accumulator and FAST_CONVECTION (whose smaller outer sub-step count gives
lines with more than one inner QUICKEST step)."""
import ctypes
import sys

import numpy as np
import pytest

import verb4d
from parity import bit_equal
from verb4d import workload as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def _run(dg, ctx, kernel, tq, lost=None, fast=False):
    """One batched convection call from the grid's initial PSD with the given
    kernel and planes per workgroup; returns (psd, lost) on the host."""
    import torch
    dg.set_tuning(convection_tq=tq, convection_kernel=kernel)
    dg.set_psd(dg.w.psd0)
    lt = None
    ptr = None
    if lost is not None:
        lt = torch.from_numpy(lost.copy()).cuda()
        torch.cuda.synchronize()
        ptr = lt.data_ptr()
    opt = verb4d.ConvectionOptions(ptr, 1 if fast else 0)
    ctx.check(verb4d.lib().verb_convection_batched_ex(ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(),
                                                      ctypes.byref(opt)))
    ctx.check_status()
    return dg.psd_host(), (lt.cpu().numpy() if lt is not None else None)


def test_line_kernel_matches_tile_kernel_production_grid(ctx):
    from verb4d.device import DeviceGrid
    w = W.config1(seed=1001)
    w.set_kp(6.0)
    dg = DeviceGrid(w, ctx=ctx)
    ref, _ = _run(dg, ctx, 1, 1)
    assert not bit_equal(ref, w.psd0)  # the call moved the PSD
    for tq in (4, 8, 16):
        say("production grid: line kernel, %d planes per workgroup" % tq)
        got, _ = _run(dg, ctx, 0, tq)
        assert bit_equal(got, ref), tq


@pytest.mark.parametrize("fast", [False, True])
def test_line_kernel_matches_tile_kernel_losses_options(ctx, fast):
    from verb4d.device import DeviceGrid
    w = W.config0(seed=81, rich=True)
    assert np.any(w.f["losses_conv"] != 0) and np.any(w.f["sources"] != 0)
    # a non-uniform convection weight G_conv (the R sweep runs on PSD * G)
    w.f["G_conv"] = 0.5 + np.random.default_rng(4).random(w.shape)
    dg = DeviceGrid(w, ctx=ctx)
    lost0 = 1e-4 * np.random.default_rng(3).random(w.shape)
    ref, ref_lost = _run(dg, ctx, 1, 2, lost0, fast)
    for tq in (2, 8):
        say("config 0 with losses, fast=%s: line kernel, %d planes per workgroup" % (fast, tq))
        got, got_lost = _run(dg, ctx, 0, tq, lost0, fast)
        assert bit_equal(got, ref), tq
        assert bit_equal(got_lost, ref_lost), tq
