"""The oracle's restatement of UpdatableMatrix::update and of the stability
check (oracle/timeloop_oracle.py), pinned by what the rules fix: a scaled
array is the original times the coefficient exactly (and back when the
coefficient returns to 1), limiting zeroes exactly the points outside the
range, a .lst entry loads once per update time, queues flag an update only
when an entry comes due, and NaN entries are passed over by the minimum."""
import numpy as np

import timeloop_oracle as OT


def test_scale_queue():
    a = np.linspace(1.0, 2.0, 12).reshape(3, 4)
    u = OT.UpdatableArray(a, scales=[(1.0, 3.0), (2.0, 1.0)])
    assert u.update(0.0) is False and np.array_equal(u.current, a)
    assert u.update(1.0) is True and np.array_equal(u.current, a * 3.0)
    assert u.update(1.5) is False and np.array_equal(u.current, a * 3.0)
    assert u.update(2.0) is True and np.array_equal(u.current, a)


def test_limits_zero_outside_range():
    g = np.arange(12, dtype=np.float64).reshape(3, 4)
    a = np.ones((3, 4))
    u = OT.UpdatableArray(a, limits={1: [(0.0, 3.0, 7.0)]}, grids=(None, g))
    assert u.update(0.0) is True
    assert np.array_equal(u.current, ((g >= 3) & (g <= 7)).astype(float))
    # a range covering every coordinate limits nothing
    u = OT.UpdatableArray(a, limits={1: [(0.0, -1.0, 100.0)]}, grids=(None, g))
    u.update(0.0)
    assert np.array_equal(u.current, a)


def test_file_queue_loads_once_and_takes_the_last_due_entry():
    a = np.zeros(5)
    b, c = np.full(5, 2.0), np.full(5, 3.0)
    u = OT.UpdatableArray(a, files=[(1.0, b), (1.5, c)], scales=[(0.0, 2.0)])
    assert u.update(0.0) is True and np.array_equal(u.current, a)  # the scale entry
    assert u.update(2.0) is True and np.array_equal(u.current, c * 2.0)
    assert u.update(3.0) is False and np.array_equal(u.current, c * 2.0)


def test_stability_min_passes_over_nan():
    class Wk:
        pass
    w = Wk()
    w.f = {"DVV": np.array([1.0, np.nan, 4.0]), "DKK": np.array([1.0, 1.0, 1.0]), "DVK": np.array([0.5, 0.0, 3.0])}
    assert OT.stability_min(w) == min(1.0 - 0.25, 4.0 - 9.0)
