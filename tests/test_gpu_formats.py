"""Inputs read from files drive the GPU path: an input folder in the
reference's formats (grid.pltb, L*, the initial PSD, coefficients and
boundaries in .pltb / grid-checked .plt, a .tab field with a .lst update
list and a scale queue) read by verb4d.formats.FileWorkload into a
DeviceGrid, stepped by the time loop on the device, against the oracle's
step and UpdatableMatrix restatement on the same arrays.

Bars: PSD within 1e-12 normwise per node after each step; the .tab field's
device array bit-exact.
"""
import numpy as np
import pytest

import timeloop_oracle as OT
import verb4d
from parity import bit_equal, node_rel_err
from verb4d import formats as F
from verb4d import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


@pytest.mark.parametrize("io", [F.BINARY, F.ASCII])
def test_step_from_input_folder(ctx, tmp_path, io):
    from verb4d.device import DeviceGrid
    w = W.config0(seed=131, rich=True)
    w.f["Lg"] = W.lstar(w, 0.03)
    F.write_inputs(str(tmp_path), w, io_method=io)
    fw = F.FileWorkload(str(tmp_path), dt=w.dt, io_method=io, types=w.types)
    fo = F.FileWorkload(str(tmp_path), dt=w.dt, io_method=io, types=w.types)
    dg = DeviceGrid(fw, ctx=ctx)
    import oracle as O
    x = fo.psd0.copy()
    for it in range(3):
        dg.step()
        x = O.step(fo, x, nthreads=8)
        err = node_rel_err(dg.psd_host(), x)
        assert err < 1e-12, (io, it, err)


def test_tab_driven_time_loop(ctx, tmp_path):
    """DKK from a .tab line: a .lst list (the initial file at time 0, two
    replacement files, .pltb and grid-checked .plt, and a "-" entry) and a
    scale queue file; DLL from a .tab with a constant scale and an L limit."""
    from verb4d.device import DeviceGrid
    from verb4d.timeloop import TimeLoop, UpdatableField
    w = W.config0(seed=132, rich=True)
    F.write_inputs(str(tmp_path), w, io_method=F.BINARY)
    d = str(tmp_path) + "/"
    dtd = w.dt / 86400.0
    g4 = (w.f["Pg"], w.f["Rg"], w.f["Vg"], w.f["Kg"])
    F.write_pltb(d + "DKK_a.pltb", w.f["DKK"] * 1.7)
    F.write_plt_grid(d + "DKK_b.plt", w.f["DKK"] * 0.6, g4)
    with open(d + "DKK.lst", "w") as f:
        f.write("0 DKK.pltb\n%r DKK_a.pltb\n%r -\n%r DKK_b.plt\n" % (1 * dtd, 2 * dtd, 3 * dtd))
    with open(d + "DKK_scale.txt", "w") as f:
        f.write("%r 1.0\n%r 2.5\n" % (0.0, 2 * dtd))
    with open(d + "DKK.tab", "w") as f:
        f.write("DKK.lst DKK_scale.txt\n")
    with open(d + "DLL.tab", "w") as f:
        f.write("DLL.pltb 0.5 - - 2.0 5.0\n")
    fw = F.FileWorkload(d, dt=w.dt, io_method=F.BINARY, types=w.types, cwd=d)
    fo = F.FileWorkload(d, dt=w.dt, io_method=F.BINARY, types=w.types, cwd=d)
    assert set(fw.specs) == {"DKK", "DLL"}
    dg = DeviceGrid(fw, ctx=ctx)
    dev = {n: UpdatableField.from_spec(dg, n, fw.specs[n], g4, axes=("Pg", "Lg" if n == "DLL" else "Rg", "Vg", "Kg"))
           for n in ("DKK", "DLL")}
    ora = {}
    for n in ("DKK", "DLL"):
        s = fo.specs[n]
        gr = (fo.f["Pg"], fo.f["Lg"] if n == "DLL" else fo.f["Rg"], fo.f["Vg"], fo.f["Kg"])
        ora[n] = OT.UpdatableArray(fo.f[n], files=s["files"], scales=s["scales"], scale=s["scale"],
                                   limit_from=s["limit_from"], limit_to=s["limit_to"], limit_init=s["limit_init"],
                                   grids=gr, loader=lambda p: F.read_any(p, g4))
    loop = TimeLoop(dg, dev, lu_caching=True)
    x = fo.psd0.copy()
    seen = []
    for it in range(5):
        info = loop.step(it)
        x, oinfo = OT.step(fo, x, ora, info["time"])
        assert info["updated"] == oinfo["updated"], it
        seen.append(info["updated"]["DKK"])
        for n, f in dev.items():
            assert bit_equal(f.cur.cpu().numpy(), ora[n].current), (it, n)
        err = node_rel_err(dg.psd_host(), x)
        assert err < 1e-12, (it, err)
    # the initial file and scale entry, file a, the scale entry with the "-"
    # entry, file b
    assert seen == [True, True, True, True, False]
