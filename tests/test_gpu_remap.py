"""The adiabatic remapping on the GPU (remap.hip): verb_cubic1d against the
reference's outputs line by line, verb_remap_batched on the production grid
against the oracle, the reference's own remapping loop with the Cubic1D
drop-in linked in, and the time loop with L* updates against the oracle's.

Bars: bit-exact (remap.hip is compiled -ffp-contract=off and follows the
reference's operation order; division and sqrt are IEEE-rounded on both
sides); the same rejections (VERB_ERR_DOMAIN where the interpolator
throws), with rejected lines left unchanged.
"""
import numpy as np
import pytest

import oracle as O
import timeloop_oracle as OT
import verb4d
from parity import node_rel_err
from verb4d import workload as W

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same(a, b):
    """Bit-identical, NaN matching NaN (x86 and the GPU differ in the sign
    and payload of the NaNs they generate)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(bits(np.where(na, 0.0, a)), bits(np.where(nb, 0.0, b)))


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def test_cubic1d_lines_vs_reference(ctx, golden_dir):
    d = np.load(golden_dir + "/remap_lines.npz")
    off = np.concatenate([[0], np.cumsum(d["n"])])
    ok = rejected = 0
    for j in range(len(d["n"])):
        s = slice(off[j], off[j + 1])
        xo, f, xn, want, st_want = d["x_old"][s], d["f"][s].copy(), d["x_new"][s], d["out"][s], int(d["status"][j])
        if st_want == 0:
            got = verb4d.Cubic1D(xo, f, xn, ctx=ctx)
            assert same(got, want), j
            ok += 1
        else:
            with pytest.raises(verb4d.VerbError) as e:
                verb4d.Cubic1D(xo, f, xn, ctx=ctx)
            assert e.value.status == -8
            rejected += 1
    assert ok > 200 and rejected > 20


def test_remap_production_grid_vs_oracle(ctx):
    import torch
    from verb4d.device import DeviceGrid
    w = W.config1(seed=1201)
    dg = DeviceGrid(w, ctx=ctx)
    dg.step(1, 1, 1)
    start = dg.psd_host()
    lo, ln = W.lstar(w, 0.05), W.lstar(w, 0.12)
    tlo, tln = torch.from_numpy(lo).cuda(), torch.from_numpy(ln).cuda()
    dg.remap(tlo, tln)
    dg.check_status()
    got = dg.psd_host()
    st, want = O.remap(start, lo, ln)
    assert st == 0
    assert np.array_equal(bits(got), bits(want))
    # and back: the remapping is not its own inverse, but it moves the PSD
    assert np.abs(got - start).max() > 1e-3 * np.abs(start).max()


def test_remap_rejects_a_non_finite_line_and_leaves_it(ctx):
    import torch
    from verb4d.device import DeviceGrid
    w = W.config0(seed=1202)
    dg = DeviceGrid(w, ctx=ctx)
    start = dg.psd_host()
    lo, ln = W.lstar(w, 0.0), W.lstar(w, 0.1)
    ln[3, 4, 5, 6] = np.nan
    dg.remap(torch.from_numpy(lo).cuda(), torch.from_numpy(ln).cuda())
    with pytest.raises(verb4d.VerbError) as e:
        dg.check_status()
    assert e.value.status == -8
    got = dg.psd_host()
    assert np.array_equal(bits(got[3, :, 5, 6]), bits(start[3, :, 5, 6]))
    # every other line is remapped
    ln[3, 4, 5, 6] = lo[3, 4, 5, 6]
    st, want = O.remap(start, lo, ln)
    got[3, :, 5, 6] = want[3, :, 5, 6]
    assert np.array_equal(bits(got), bits(want))


def test_reference_loop_with_cubic1d_dropin(ctx):
    """src/VERB4D_Solver.cpp:574-606 as the reference runs it, with Cubic1D
    from the drop-in (oracle/_ref/libverbref_hipremap.so)."""
    if not O.ref_available("hipremap"):
        pytest.skip("hybrid harness not built")
    w = W.Workload(nP=4, nR=10, nV=21, nK=17, seed=3)
    lo, ln = W.lstar(w, 0.05), W.lstar(w, 0.12)
    st, got = O.remap(w.psd0, lo, ln, lib="hipremap")
    assert st == 0
    _, want = O.remap(w.psd0, lo, ln)
    assert np.array_equal(bits(got), bits(want))


def test_time_loop_with_lstar_updates_vs_oracle(ctx):
    import torch
    from verb4d.device import DeviceGrid
    from verb4d.timeloop import TimeLoop, UpdatableField
    wd = W.config0(seed=1203, rich=True)
    wo = W.config0(seed=1203, rich=True)
    dg = DeviceGrid(wd, ctx=ctx)
    dtd = wd.dt / 86400.0
    l0 = W.lstar(wd, 0.04)
    l1, l2 = W.lstar(wd, 0.10), W.lstar(wd, 0.07)
    g1 = 1.0 / l1 ** 2
    lq = dict(files=[(1 * dtd, l1), (3 * dtd, l2)])
    gq = dict(files=[(1 * dtd, g1)])
    lt = torch.from_numpy(l0.copy()).cuda()
    lstar = UpdatableField(dg, "Lstar", tensor=lt, **lq)
    jac = {"G_radial": UpdatableField(dg, "G_radial", **gq)}
    loop = TimeLoop(dg, {}, vl_from_convection=True, lstar=lstar, jacobians=jac)
    ora = {"lstar": OT.UpdatableArray(l0, **lq), "copy": l0.copy(),
           "jacobians": {"G_radial": OT.UpdatableArray(wo.f["G_radial"], **gq)}}
    x = wo.psd0.copy()
    for it in range(5):
        info = loop.step(it)
        x, oinfo = OT.step(wo, x, {}, info["time"], vl_from_convection=True, remap=ora)
        assert info["remapped"] == oinfo["remapped"] == (it in (1, 3)), it
        err = node_rel_err(dg.psd_host(), x)
        assert err < 1e-12, (it, err)
    assert np.array_equal(bits(lt.cpu().numpy()), bits(l2))
