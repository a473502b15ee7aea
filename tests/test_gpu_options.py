"""GPU parity of the reference's compile-time options on this path, against
the reference built with each option (tests/golden/*_lost_*, *_fast_*,
local_lucache_node, written by oracle/_ref/libverbref_{lost,fast,lucache})
and the oracle:
  SAVE_PSD_LOST_CONV  Convection_2D(PSD, PSD_lost, ...) accumulates the PSD
                      the losses take (src/Convection_2D.cpp:150-155);
  FAST_CONVECTION     the smaller outer sub-step count (:80-84);
  LU_CACHING          Diffusion_2D's caller-owned band factors
                      (src/MatrixSolver.cpp:249-277).
Bars: bit-exact without convection losses; within 1e-14 where exp(L dt)
is evaluated by the device math library (one ulp per sub-step); pivots
bit-exact and factors / PSD within 1e-12 for the band LU.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle as O
import verb4d
from parity import bit_equal, rel_err
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


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


@pytest.mark.parametrize("name", ["conv_lost_c0_plane.npz", "conv_lost_c1_plane.npz", "conv_fast_c0_plane.npz",
                                  "conv_fast_c1_plane.npz"])
def test_convection_options_vs_reference_build(ctx, golden_dir, name):
    say("convection option", name)
    d = load(golden_dir, name)
    nP, nR = d["psd"].shape
    fast = bool(int(d["fast"]))
    lost_in = d["lost_in"] if "lost_in" in d.files else None
    t = [int(v) for v in d["types"]]
    out, lost = verb4d.Convection_2D_lost(d["psd"], lost_in, d["P"], d["R"], nP, nR, d["Pl"], d["Pu"], d["Rl"],
                                          d["Ru"], *t, d["VP"], d["VR"], d["S"], d["L"], d["G"], float(d["dt"]),
                                          fast_convection=fast, ctx=ctx)
    exact = not np.any(d["L"] != 0)
    if exact:
        assert bit_equal(out, d["out"])
    else:
        assert rel_err(out, d["out"]) < 1e-14
    if lost_in is not None:
        if exact:
            assert bit_equal(lost, d["lost_out"])
        else:
            # the accumulated differences carry exp()'s last-place difference
            assert rel_err(lost, d["lost_out"]) < 1e-12


@pytest.mark.parametrize("fast", [False, True])
def test_batched_convection_options_vs_oracle(ctx, fast):
    """Every (V,K) plane of a config-0 grid with convection losses, the
    lost-PSD accumulator and (or not) FAST_CONVECTION, against the oracle."""
    import torch
    from verb4d.device import DeviceGrid
    w = W.config0(seed=81, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    rng = np.random.default_rng(3)
    lost0 = 1e-4 * rng.random(w.shape)
    lost = torch.from_numpy(lost0.copy()).cuda()
    torch.cuda.synchronize()
    opt = verb4d.ConvectionOptions(lost.data_ptr(), 1 if fast else 0)
    ctx.check(verb4d.lib().verb_convection_batched_ex(ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(),
                                                      ctypes.byref(opt)))
    ctx.check_status()
    got = dg.psd_host()
    got_lost = lost.cpu().numpy()
    worst = 0.0
    for iV in range(w.nV):
        for iK in range(w.nK):
            sl = (slice(None), slice(None), iV, iK)
            out, lo = O.convection_2d_ex(w.psd0[sl], lost0[sl], w.f["Pg"][sl], w.f["Rg"][sl], w.nP, w.nR,
                                         w.bc["Pl"][:, iV, iK], w.bc["Pu"][:, iV, iK], w.bc["Rl"][:, iV, iK],
                                         w.bc["Ru"][:, iV, iK], (0, 0, 1, 1), w.f["VP"][sl], w.f["VR"][sl],
                                         w.f["sources"][sl], w.f["losses_conv"][sl], w.f["G_conv"][sl], w.dt,
                                         fast=fast)
            worst = max(worst, rel_err(got[sl], out))
            assert rel_err(got[sl], out) < 1e-14, (iV, iK)
            assert rel_err(got_lost[sl], lo) < 1e-12, (iV, iK)
    say("batched convection fast=%s worst plane %.2e" % (fast, worst))


def test_step_ex_carries_convection_options(ctx):
    """verb_step_ex with the lost-PSD accumulator: the PSD equals the plain
    step's, and the accumulator equals the batched convection's."""
    import torch
    from verb4d.device import DeviceGrid
    w = W.config0(seed=82, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    lost = torch.zeros(w.shape, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    opt = verb4d.StepOptions(0, None, None, lost.data_ptr(), 0)
    ctx.check(verb4d.lib().verb_step_ex(ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(), 1, 1, 1, None, None, 1,
                                        ctypes.byref(opt)))
    ctx.check_status()
    with_lost = dg.psd_host()
    lost1 = lost.cpu().numpy()
    dg.set_psd(w.psd0)
    dg.step()
    assert bit_equal(with_lost, dg.psd_host())
    dg.set_psd(w.psd0)
    lost.zero_()
    torch.cuda.synchronize()
    copt = verb4d.ConvectionOptions(lost.data_ptr(), 0)
    ctx.check(verb4d.lib().verb_convection_batched_ex(ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(),
                                                      ctypes.byref(copt)))
    ctx.synchronize()
    assert bit_equal(lost1, lost.cpu().numpy())
    assert np.any(lost1 > 0)


def test_diffusion_2d_lu_caching_vs_reference_build(ctx, golden_dir):
    """Diffusion_2D of an LU_CACHING build: the call that factors fills the
    caller's cache with the reference's pivots (bit-exact) and factors
    (1e-12, OpenBLAS's blocked dgbtrf rounds differently), and the call that
    reuses it gives the reference's PSD."""
    d = load(golden_dir, "local_lucache_node.npz")
    nx, ny = d["psd"].shape
    n = nx * ny
    ldab = 3 * (nx + 1) + 1
    cache = np.zeros(ldab * n)
    ipiv = np.zeros(n, dtype=np.int64)
    t = [int(v) for v in d["types"]]
    args = (d["x"], d["y"], nx, ny, d["xl"], d["xu"], d["yl"], d["yu"], *t, d["Dxx"], d["Dyy"], d["Dxy"], d["Dxy"],
            d["G"], d["S"], d["L"], float(d["dt"]), float(d["sub_dt"]))
    out1 = verb4d.Diffusion_2D(d["psd"], *args, lu_cache=cache, index_cache=ipiv, recompute_lu=True, ctx=ctx)
    assert np.array_equal(ipiv, d["ipiv"])
    assert rel_err(cache, d["cache"]) < 1e-12
    assert rel_err(out1, d["out1"]) < 1e-12
    out2 = verb4d.Diffusion_2D(out1, *args, lu_cache=cache, index_cache=ipiv, recompute_lu=False, ctx=ctx)
    assert rel_err(out2, d["out2"]) < 1e-12
