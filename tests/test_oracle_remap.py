"""The oracle's restatement of the adiabatic remapping (orc_cubic1d /
orc_remap: Cubic1D, src/Interpolation.cpp:25-51, over the monotone cubic
interpolator of src/MonotCubicInterpolator.cpp, and the loop of
src/VERB4D_Solver.cpp:574-606), pinned by

- the reference's own outputs (tests/golden/remap_lines.npz, 300 lines of
  every kind; remap_grid.npz, a 4x10x21x17 grid), bit-exact, with the same
  rejections;
- what the mathematics fixes: data on a line through the origin are
  reproduced (the extra (0, 0) point lies on it, every slope equals it, no
  limiting), knots map to themselves exactly, monotone data give a monotone
  interpolant (the Fritsch-Carlson limit; without it the steep profile
  below overshoots), values beyond the last key are the last value, and
  between the knots the form is the cubic Hermite spline of the stated
  slopes (scipy's CubicHermiteSpline as the independent evaluator).
"""
import numpy as np
import pytest

import oracle as O


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def split(d):
    off = np.concatenate([[0], np.cumsum(d["n"])])
    for j in range(len(d["n"])):
        s = slice(off[j], off[j + 1])
        yield d["x_old"][s], d["f"][s], d["x_new"][s], d["out"][s], int(d["status"][j])


def test_oracle_matches_reference_lines(golden_dir):
    d = np.load(golden_dir + "/remap_lines.npz")
    kinds = {0: 0, -8: 0}
    for xo, f, xn, want, st_want in split(d):
        st, got = O.cubic1d(xo, f, xn)
        assert st == st_want
        kinds[st] += 1
        if st == 0:
            assert np.array_equal(bits(got), bits(want))
    assert kinds[0] > 200 and kinds[-8] > 20


def test_oracle_matches_reference_grid(golden_dir):
    from verb4d import workload as W
    d = np.load(golden_dir + "/remap_grid.npz")
    w = W.Workload(nP=int(d["nP"]), nR=int(d["nR"]), nV=int(d["nV"]), nK=int(d["nK"]), seed=int(d["seed"]))
    st, got = O.remap(w.psd0, W.lstar(w, float(d["c_old"])), W.lstar(w, float(d["c_new"])))
    assert st == 0
    assert np.array_equal(bits(got), bits(d["out"]))
    assert np.abs(got - w.psd0).max() > 1e-2  # the grid did move


def test_line_through_origin_is_reproduced():
    rng = np.random.default_rng(3)
    xo = np.sort(rng.uniform(1.0, 7.0, 29))
    f = 2.5e-7 * xo
    xn = rng.uniform(0.2, xo[-1], 29)
    st, got = O.cubic1d(xo, f, xn)
    assert st == 0
    assert np.max(np.abs(got - 2.5e-7 * xn)) <= 1e-20


def test_knots_map_to_themselves_and_constant_beyond():
    rng = np.random.default_rng(4)
    xo = np.sort(rng.uniform(1.0, 7.0, 20))
    f = np.exp(-(xo - 4.0) ** 2)
    st, got = O.cubic1d(xo, f, xo)
    assert st == 0 and np.array_equal(bits(got), bits(f))
    st, got = O.cubic1d(xo, f, np.full(20, 9.0))
    assert st == 0 and np.all(got == f[-1])
    st, got = O.cubic1d(xo, f, np.full(20, -1.0))
    assert st == 0 and np.all(got == 0.0)  # below the extra (0, 0) point


def test_monotone_data_give_a_monotone_interpolant():
    xo = np.linspace(1.0, 6.0, 11)
    f = np.where(xo < 3.4, 1e-8, 1e-6) + 1e-9 * xo  # a steep step
    xn = np.linspace(0.5, 6.5, 11)
    for k in range(8):
        xs = xn + 0.07 * k
        st, got = O.cubic1d(xo, f, xs)
        assert st == 0
        order = np.argsort(xs)
        assert np.all(np.diff(got[order]) >= 0)
    # the same evaluation with the slopes left unlimited overshoots
    d = simple_slopes(np.concatenate([[0.0], xo]), np.concatenate([[0.0], f]))
    from scipy.interpolate import CubicHermiteSpline
    s = CubicHermiteSpline(np.concatenate([[0.0], xo]), np.concatenate([[0.0], f]), d)
    dense = s(np.linspace(0.0, 6.0, 2001))
    assert np.any(np.diff(dense) < -1e-12)


def simple_slopes(k, v):
    d = np.empty_like(v)
    d[0] = (v[1] - v[0]) / (k[1] - k[0])
    d[-1] = (v[-1] - v[-2]) / (k[-1] - k[-2])
    d[1:-1] = 0.5 * (v[2:] - v[1:-1]) / (k[2:] - k[1:-1]) + 0.5 * (v[1:-1] - v[:-2]) / (k[1:-1] - k[:-2])
    return d


def test_non_monotone_data_follow_the_hermite_spline():
    rng = np.random.default_rng(5)
    xo = np.sort(rng.uniform(1.0, 7.0, 25))
    f = np.exp(-((xo - 4.5) ** 2) / 0.5) * 1e-6  # a peak: not monotone with the (0, 0) point
    k, v = np.concatenate([[0.0], xo]), np.concatenate([[0.0], f])
    from scipy.interpolate import CubicHermiteSpline
    s = CubicHermiteSpline(k, v, simple_slopes(k, v))
    xn = rng.uniform(0.1, xo[-1], 25)
    st, got = O.cubic1d(xo, f, xn)
    assert st == 0
    assert np.max(np.abs(got - s(xn))) <= 1e-12 * np.abs(f).max()


@pytest.mark.parametrize("case", ["nan_new", "inf_new", "nan_key", "nan_direction"])
def test_rejections(case):
    xo = np.linspace(1.0, 5.0, 6)
    f = np.linspace(1.0, 2.0, 6)
    xn = xo.copy()
    if case == "nan_new":
        xn[2] = np.nan
    elif case == "inf_new":
        xn[4] = np.inf
    elif case == "nan_key":
        xo[3] = np.nan
    else:
        f[0] = np.nan  # (0, 0) and (1, NaN): the pair that sets the direction
    st, _ = O.cubic1d(xo, f, xn)
    assert st == -8
    if O.ref_available():
        st_ref, _ = O.cubic1d(xo, f, xn, lib="ref")
        if case != "nan_key":  # std::map's order is undefined for a NaN key
            assert st_ref == -8
