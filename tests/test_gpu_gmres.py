"""GPU: local diffusion by matrix-free GMRES (FP32 cycles, FP64 residual test,
band-LU fallback) against the plain FP64 band-LU path and the CPU oracle.

Bars: with every node falling back (max_cycles = 0) the result is the plain
path's, bit for bit (same kernels, same inputs). Converged nodes satisfy
max|b - A x| <= tol * max|b| in FP64; the PSD then agrees with the oracle's
direct solve within 1e-8 relative per node (the mixed-precision bar) at
tol = 1e-11.
"""
import sys

import numpy as np
import pytest

import oracle as O
import verb4d
from parity import bit_equal, node_rel_err
from verb4d import workload as W
from verb4d.device import DeviceGrid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def test_gmres_all_fallback_equals_plain_path(ctx):
    say("gmres forced fallback, config0")
    w = W.config0(seed=31, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    dg.local()
    dg.check_status()
    plain = dg.psd_host()
    dg.set_psd(w.psd0)
    st = dg.local_gmres(DeviceGrid.gmres_params(tol=1e-12, restart=8, max_cycles=0))
    dg.check_status()
    say("stats", st)
    assert st["converged"] == 0 and st["fallback"] == st["nodes"] > 0
    assert bit_equal(dg.psd_host(), plain)


def test_gmres_config0_matches_oracle(ctx):
    say("gmres config0 vs oracle")
    w = W.config0(seed=32, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    st = dg.local_gmres(DeviceGrid.gmres_params(tol=1e-12, restart=20, max_cycles=12))
    dg.check_status()
    say("stats", st)
    got = dg.psd_host()
    ref = O.step(w, w.psd0, 8, 0, 0, 1)
    err = node_rel_err(got, ref)
    say("max node rel err %.3e" % err)
    assert st["nodes"] == w.nP * (w.nR - 2) or st["nodes"] == w.nP * w.nR
    assert st["converged"] == st["nodes"], st
    assert err < 1e-9


def test_gmres_production_grid_one_step(ctx):
    """29x24x101x91 at Kp = 6: the local step by GMRES against the oracle."""
    say("gmres production grid")
    w = W.config1(seed=1001)
    w.set_kp(6.0)
    dg = DeviceGrid(w, ctx=ctx)
    st = dg.local_gmres(DeviceGrid.gmres_params(tol=1e-11, restart=20, max_cycles=8))
    dg.check_status()
    say("stats", st, "mean iterations %.1f" % (st["iterations"] / max(1, st["nodes"])))
    got = dg.psd_host()
    ref = O.step(w, w.psd0, 16, 0, 0, 1)
    err = node_rel_err(got, ref)
    say("max node rel err %.3e" % err)
    assert err < 1e-8
    assert st["converged"] + st["fallback"] == st["nodes"]
