"""GPU: launch shapes, the persistent node queue, the autotuner and the
HIP-graph-captured step. They change speed, not results: every variant is
compared with the default launch of the same step.

Bars: the persistent queue and the graph replay run the same kernels on the
same data, so they are bit-identical to the default launches; other
register-window LU shapes only change which exactly-zero tiles are
skipped, so they agree within 1e-13 per node (and are reported when they
are bit-identical).
"""
import sys

import numpy as np
import pytest

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


def _step(dg, psd0, **tune):
    dg.set_tuning(**tune)
    dg.set_psd(psd0)
    dg.step()
    dg.check_status()
    return dg.psd_host()


def test_persistent_queue_and_lu_shapes_config0(ctx):
    say("persistent queue / LU shapes, config0")
    w = W.config0(seed=41, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    base = _step(dg, w.psd0)
    assert bit_equal(_step(dg, w.psd0, lu_persistent=1), base)
    for rb, cb in ((3, 5), (5, 9), (8, 14), (8, 15)):
        got = _step(dg, w.psd0, lu_rb=rb, lu_cb=cb)
        say("shape", rb, cb, "bit-equal", bit_equal(got, base), "node err %.2e" % node_rel_err(got, base))
        assert node_rel_err(got, base) < 1e-13
        assert bit_equal(_step(dg, w.psd0, lu_rb=rb, lu_cb=cb, lu_persistent=1), got)
    dg.set_tuning()


def test_persistent_queue_uneven_extents(ctx):
    say("persistent queue, uneven extents")
    w = W.config0(seed=42)
    rng = np.random.default_rng(5)
    k0 = rng.integers(0, 12, size=(w.nP, w.nR)).astype(np.int32)
    dg = DeviceGrid(w, ctx=ctx)
    dg.set_tuning()
    dg.step_uneven(k0, 0, 0, 1)
    dg.check_status()
    base = dg.psd_host()
    dg.set_tuning(lu_persistent=1)
    dg.set_psd(w.psd0)
    dg.step_uneven(k0, 0, 0, 1)
    dg.check_status()
    assert bit_equal(dg.psd_host(), base)
    dg.set_tuning()


def test_autotune_keeps_results(ctx):
    say("autotune, config0")
    w = W.config0(seed=43, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    dg.set_tuning()
    base = _step(dg, w.psd0)
    dg.set_psd(w.psd0)
    t = dg.autotune()
    say("tuning", t)
    assert t["convection_tq"] in (1, 2, 4, 8) and t["lu_rb"] >= 3 and t["local_ms"] > 0
    assert bit_equal(dg.psd_host(), w.psd0)  # psd restored
    dg.step()
    dg.check_status()
    assert node_rel_err(dg.psd_host(), base) < 1e-13
    dg.set_tuning()


def test_graph_replay_matches_eager_steps(ctx):
    say("graph replay, config0")
    w = W.config0(seed=44, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    dg.set_tuning(lu_persistent=1)
    dg.set_psd(w.psd0)
    eager = []
    for _ in range(3):
        dg.step()
        dg.check_status()
        eager.append(dg.psd_host())
    dg.set_psd(w.psd0)
    for i in range(3):
        dg.step_graph()  # first call: eager step + capture; then replays
        dg.check_status()
        got = dg.psd_host()
        assert bit_equal(got, eager[i]), i
        # eager work between replays
        dg.set_psd(got)
    dg.set_tuning()


def test_graph_replay_uneven_extents(ctx):
    """The uneven step (loss-cone extents, persistent LU queue) replayed from
    its graph equals the eager uneven step, and the extents table is uploaded
    once."""
    say("graph replay, uneven extents")
    w = W.config0(seed=45, rich=True)
    rng = np.random.default_rng(9)
    k0 = rng.integers(0, 14, size=(w.nP, w.nR)).astype(np.int32)
    dg = DeviceGrid(w, ctx=ctx)
    dg.set_tuning(lu_persistent=1)
    dg.set_psd(w.psd0)
    eager = []
    for _ in range(3):
        dg.step_uneven(k0)
        dg.check_status()
        eager.append(dg.psd_host())
    dg.set_psd(w.psd0)
    for i in range(3):
        dg.step_graph(k0)
        dg.check_status()
        got = dg.psd_host()
        assert bit_equal(got, eager[i]), i
        dg.set_psd(got)
    dg.set_tuning()

