"""The time loop around the operators on the device (verb4d.timeloop:
coefficient and boundary updates with scaling, limiting and replacement,
the stability check, the V boundaries from convection, the LU-recompute
flag, clamps and the negative count) against the oracle's restatement of
the same loop (oracle/timeloop.py) over several steps.

Bars: updated coefficient arrays bit-exact; PSD within 1e-12 normwise per
node after each step (the local band solve's rounding); update and
recompute flags identical; negative counts equal up to values within that
rounding of zero.
"""
import ctypes

import numpy as np
import pytest

import timeloop_oracle as OT
import verb4d
from parity import bit_equal, node_rel_err
from verb4d import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def _queues(w):
    dt_days = w.dt / 86400.0
    rng = np.random.default_rng(17)
    new_loss = -(2e-5 * rng.random(w.shape))
    a_mid = float(w.alpha.mean())
    return {
        # scale DKK x3 from step 2 and back to x1 at step 4
        "DKK": dict(scales=[(2 * dt_days, 3.0), (4 * dt_days, 1.0)]),
        # the mixed term limited to pitch angles below the mean from step 3
        # (K axis; zeroing DVK keeps DVV*DKK - DVK^2 >= 0)
        "DVK": dict(limits={3: [(3 * dt_days, -1e99, a_mid)]}),
        # losses replaced at step 1 (a .lst entry), DLL scaled at step 2
        "losses_local": dict(files=[(1 * dt_days, new_loss)]),
        "DLL": dict(scales=[(2 * dt_days, 0.5)]),
        # the V-upper boundary scaled from step 1 (a [P][R][K] boundary array)
        "Vu": dict(scales=[(1 * dt_days, 0.9)]),
    }


@pytest.mark.parametrize("lu_caching", [False, True])
def test_time_loop_vs_oracle(ctx, lu_caching):
    from verb4d.device import DeviceGrid
    from verb4d.timeloop import TimeLoop, UpdatableField
    wd = W.config0(seed=111, rich=True)
    wo = W.config0(seed=111, rich=True)
    dg = DeviceGrid(wd, ctx=ctx)
    q = _queues(wd)
    dev_fields, ora_fields = {}, {}
    for name, kw in q.items():
        is_bc = name in wd.bc
        dev_fields[name] = UpdatableField(dg, name, axes=() if is_bc else ("Pg", "Rg", "Vg", "Kg"), **kw)
        src = wo.bc[name] if is_bc else wo.f[name]
        grids = () if is_bc else (wo.f["Pg"], wo.f["Rg"], wo.f["Vg"], wo.f["Kg"])
        ora_fields[name] = OT.UpdatableArray(src, grids=grids, **kw)
    loop = TimeLoop(dg, dev_fields, positive_psd=False, vl_from_convection=True, vu_from_convection=False,
                    lu_caching=lu_caching)
    x = wo.psd0.copy()
    for it in range(6):
        info = loop.step(it)
        x, oinfo = OT.step(wo, x, ora_fields, info["time"], vl_from_convection=True)
        assert info["updated"] == oinfo["updated"], it
        for name, f in dev_fields.items():
            assert bit_equal(f.cur.cpu().numpy(), ora_fields[name].current), (it, name)
        want_recompute = it == 0 or any(oinfo["updated"].get(n, False) for n in ("losses_local", "DVV", "DVK", "DKK"))
        if lu_caching:
            assert info["recompute_lu"] == want_recompute, it
        err = node_rel_err(dg.psd_host(), x)
        assert err < 1e-12, (it, err)
        # values within the solve's rounding of zero may take either sign: the
        # count lies between the oracle's counts below -tol and below +tol
        tol = 1e-12 * np.abs(x).max(axis=(2, 3), keepdims=True)
        lo, hi = int((x < -tol).sum()), int((x < tol).sum())
        assert lo <= info["negative_points"] <= hi, (it, info["negative_points"], oinfo["negative_points"], lo, hi)
        assert info["stability_min"] == oinfo["stability_min"], it


def test_time_loop_positive_psd_and_vu_from_convection(ctx):
    from verb4d.device import DeviceGrid
    from verb4d.timeloop import TimeLoop
    wd = W.config0(seed=112, rich=True)
    wo = W.config0(seed=112, rich=True)
    dg = DeviceGrid(wd, ctx=ctx)
    loop = TimeLoop(dg, {}, positive_psd=True, vl_from_convection=True, vu_from_convection=True)
    x = wo.psd0.copy()
    for it in range(3):
        info = loop.step(it)
        x, oinfo = OT.step(wo, x, {}, info["time"], positive_psd=True, vl_from_convection=True,
                           vu_from_convection=True)
        assert info["negative_points"] == 0
        assert node_rel_err(dg.psd_host(), x) < 1e-12, it
        # the V-upper line of the convected PSD: convection's exp-based losses
        # (1e-14) and, from the second step, the local solve's rounding
        vu = loop.vu.cpu().numpy()
        d = np.linalg.norm(vu - wo.bc["Vu"], axis=-1) / np.maximum(np.linalg.norm(wo.bc["Vu"], axis=-1), 1e-300)
        assert d.max() < 1e-10, it


def test_stability_check_stops_the_loop(ctx):
    """DVK scaled so that DVK^2 > DVV*DKK somewhere: the device check reports
    VERB_ERR_UNSTABLE with the oracle's minimum, before any operator runs."""
    from verb4d.device import DeviceGrid
    from verb4d.timeloop import TimeLoop, UpdatableField
    wd = W.config0(seed=113)
    wo = W.config0(seed=113)
    dg = DeviceGrid(wd, ctx=ctx)
    before = dg.psd_host()
    f = UpdatableField(dg, "DVK", scales=[(0.0, 5.0)])
    loop = TimeLoop(dg, {"DVK": f})
    with pytest.raises(verb4d.VerbError) as e:
        loop.step(0)
    assert e.value.status == -7
    ora = OT.UpdatableArray(wo.f["DVK"], scales=[(0.0, 5.0)])
    ora.update(0.0)
    wo.f["DVK"][...] = ora.current
    m = OT.stability_min(wo)
    assert m < 0
    got = ctypes.c_double()
    assert verb4d.lib().verb_check_stability(ctx.handle, ctypes.byref(dg.grid), ctypes.byref(got)) == -7
    assert got.value == m
    assert bit_equal(dg.psd_host(), before)
