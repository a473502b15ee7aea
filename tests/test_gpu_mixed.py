"""GPU: local diffusion with an FP32 band LU and FP64 iterative refinement,
against the CPU oracle's FP64 band LU on the same seeded inputs.

Bars: every solved node meets the FP64 residual test max|b - A x| <=
tol * max|b|; the PSD agrees with the oracle within 1e-10 per node at
tol = 1e-13 on the small grid and within 1e-8 (the mixed-precision bar) on
the production grid and on the stiff 151 x 121 systems (kl = 152), where
the FP32 register window with 11 wavefronts is the only band LU.
"""
import sys

import numpy as np
import pytest

import oracle as O
import verb4d
from parity import node_rel_err
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


def test_mixed_config0_vs_oracle(ctx):
    say("mixed config0")
    w = W.config0(seed=51, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    st = dg.local_mixed(tol=1e-13, max_sweeps=10)
    dg.check_status()
    say("stats", st)
    err = node_rel_err(dg.psd_host(), O.step(w, w.psd0, 8, 0, 0, 1))
    say("max node rel err %.3e" % err)
    assert st["converged"] == st["nodes"] > 0 and st["max_residual"] <= 1e-13
    assert err < 1e-10


def test_mixed_production_grid_kp6(ctx):
    say("mixed production grid, Kp 6")
    w = W.config1(seed=1001)
    w.set_kp(6.0)
    dg = DeviceGrid(w, ctx=ctx)
    st = dg.local_mixed(tol=1e-12, max_sweeps=10)
    dg.check_status()
    say("stats", st)
    err = node_rel_err(dg.psd_host(), O.step(w, w.psd0, 16, 0, 0, 1))
    say("max node rel err %.3e" % err)
    assert st["converged"] == st["nodes"]
    assert err < 1e-8


def test_mixed_stiff_kl152(ctx):
    say("mixed stiff 151x121")
    w = W.config3(seed=3001, nP=2, nR=3)
    dg = DeviceGrid(w, ctx=ctx)
    st = dg.local_mixed(tol=1e-12, max_sweeps=12)
    dg.check_status()
    say("stats", st)
    err = node_rel_err(dg.psd_host(), O.step(w, w.psd0, 8, 0, 0, 1))
    say("max node rel err %.3e" % err)
    assert st["converged"] == st["nodes"] == 6
    assert err < 1e-8
