"""Synthetic workloads of the benchmark configurations (CPU)."""
import numpy as np

from verb4d import workload as W


def test_config2_extents_follow_the_loss_cone():
    """Uneven production workload: 4040..9191 unknowns per node, extents
    growing with R, Kp 2 on MLT sectors 0..11 and 6 on 12..23."""
    w, k0 = W.config2(seed=2001)
    n = w.nV * (w.nK - k0)
    assert n.min() == 101 * 40 and n.max() == 101 * 91
    assert k0.shape == (w.nP, w.nR)
    mean_active = (w.nK - k0).mean(axis=0)
    assert mean_active[-1] > mean_active[0] + 30
    ratio = w.f["DKK"] / w.daa
    np.testing.assert_allclose(ratio[: w.nP // 2], 1.0)
    np.testing.assert_allclose(ratio[w.nP // 2:], 10.0)


def test_config3_is_stiff():
    w = W.config3(seed=3001, nP=2, nR=3)
    assert (w.nV, w.nK, w.dt) == (151, 121, 600.0)
    assert w.f["DKK"].max() > 0.1 and w.f["DKK"].min() < 1e-9
