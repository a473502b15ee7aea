"""GPU: the time loop's pieces around the three operators
(src/VERB4D_Solver.cpp:539-1282) through verb_step_ex: the positivity clamp
PSD.max_of(1e-21) before the step and after each operator (:621, 980, 1072,
1218-1221), the count of negative points (:1223-1238), and the V-lower
boundary taken from the PSD after convection (:965-966).

Bars: the same kernels run in the same order, so each result is
bit-identical to the composition of the single-operator entries with the
clamp / slice applied on the host in between (max_of semantics:
(v < 1e-21) ? 1e-21 : v). The count equals numpy's count of the result.
"""
import sys

import numpy as np
import pytest

import verb4d
from parity import bit_equal
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


def max_of(a, v=1e-21):
    return np.where(a < v, v, a)


def _psd_with_negatives(w, seed):
    rng = np.random.default_rng(seed)
    p = w.psd0.copy()
    flat = p.reshape(-1)
    idx = rng.choice(flat.size, size=flat.size // 50, replace=False)
    flat[idx] = -rng.uniform(1e-30, 1e-20, size=idx.size)
    return p


def test_positive_psd_clamps_match_composition(ctx):
    say("step with positivity clamps, config0")
    w = W.config0(seed=51, rich=True)
    p0 = _psd_with_negatives(w, 1)
    dg = DeviceGrid(w, ctx=ctx)
    dg.set_psd(p0)
    dg.step_ex(positive_psd=True)
    dg.check_status()
    got = dg.psd_host()
    # composition
    dg.set_psd(max_of(p0))
    dg.convection()
    dg.set_psd(max_of(dg.psd_host()))
    dg.radial()
    dg.set_psd(max_of(dg.psd_host()))
    dg.local()
    dg.check_status()
    ref = max_of(dg.psd_host())
    assert got.min() >= 1e-21
    assert bit_equal(got, ref)


def test_negative_count_and_plain_result(ctx):
    say("step with negative count, config0")
    w = W.config0(seed=52, rich=True)
    p0 = _psd_with_negatives(w, 2)
    dg = DeviceGrid(w, ctx=ctx)
    dg.set_psd(p0)
    cnt = dg.step_ex(count_negative=True)
    dg.check_status()
    got = dg.psd_host()
    say("negative points", cnt, "of", got.size)
    assert cnt == int(np.sum(got < 0))
    dg.set_psd(p0)
    dg.step()
    dg.check_status()
    assert bit_equal(got, dg.psd_host())


def test_vl_from_convection(ctx):
    say("V-lower boundary from convection, config0")
    import torch
    w = W.config0(seed=53, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    assert dg.grid.Vl_type == 1
    vl = torch.zeros((w.nP, w.nR, w.nK), dtype=torch.float64, device="cuda")
    dg.set_psd(w.psd0)
    dg.step_ex(vl_from_convection=vl)
    dg.check_status()
    got = dg.psd_host()
    got_vl = vl.cpu().numpy()
    # composition: convection, slice, radial and local with that boundary
    old_vl = dg.grid.Vl
    dg.set_psd(w.psd0)
    dg.convection()
    after = dg.psd_host()
    ref_vl = np.ascontiguousarray(after[:, :, 0, :])
    assert bit_equal(got_vl, ref_vl)
    vl2 = torch.from_numpy(ref_vl).to("cuda")
    dg.grid.Vl = vl2.data_ptr()
    dg.radial()
    dg.local()
    dg.check_status()
    ref = dg.psd_host()
    dg.grid.Vl = old_vl
    assert bit_equal(got, ref)
    # and it differs from the plain step when the slice differs from Vl
    dg.set_psd(w.psd0)
    dg.step()
    dg.check_status()
    plain = dg.psd_host()
    say("max |diff| vs plain step %.3e" % float(np.max(np.abs(plain - got))))
