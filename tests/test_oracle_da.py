"""The oracle's restatement of the data assimilation's Kalman update along L
(oracle/da_oracle.py), pinned by the reference's own code: the model
matrices F and the analysis covariances that getModelMatrixDiffusion1D and
the reference's matrix operations produce (tests/golden/da_lines.npz,
written by oracle/ref_da.cpp) are reproduced bit for bit, the analysis PSD
to 1e-15 (its last product is a dgemv whose summation order numpy's matmul
does not share); and by what the filter fixes: without observations the PSD
stays and Pa = Pf, with an observation error far below the forecast error
the analysis takes the observed values, and F keeps the boundary rows
zero (B is zero there)."""
import numpy as np

import da_oracle as D


def cases(golden_dir):
    d = np.load(golden_dir + "/da_lines.npz")
    for i in range(int(d["count"])):
        yield {k.split("_", 1)[1]: d[k] for k in d.files if k.startswith("%d_" % i)}


def test_oracle_matches_reference_lines(golden_dir):
    n = 0
    for c in cases(golden_dir):
        F = D.model_matrix(c["R"], c["D"], c["G"], int(c["lt"]), int(c["ut"]), float(c["dt"]))
        assert np.array_equal(F, c["F"])
        me, oe = float(c["me"]), float(c["oe"])
        x1, P1, Pf1 = D.kalman_line(c["x0"], np.zeros_like(c["P1"]), F, c["obs"], me, oe)
        x2, P2, Pf2 = D.kalman_line(x1, P1, F, c["obs"], me, oe)
        assert np.array_equal(Pf1, c["Pf1"]) and np.array_equal(P1, c["P1"])
        assert np.array_equal(Pf2, c["Pf2"]) and np.array_equal(P2, c["P2"])
        for got, want in ((x1, c["x1"]), (x2, c["x2"])):
            assert np.abs(got - want).max() <= 1e-15 * np.abs(want).max()
        n += 1
    assert n == 15


def test_filter_limits(golden_dir):
    c = next(cases(golden_dir))
    F = c["F"]
    n = len(c["x0"])
    assert np.all(F[0] == 0) and np.all(F[-1] == 0)
    x = c["x0"]
    # no observations: the forecast stays, Pa = Pf
    xa, Pa, Pf = D.kalman_line(x, np.zeros((n, n)), F, np.full(n, np.nan), 1e-3, 1e-3)
    assert np.array_equal(xa, x) and np.array_equal(Pa, Pf)
    # near-exact observations at every point: the analysis takes them
    obs = x * 1.1 + 1e-9
    xa, Pa, _ = D.kalman_line(x, np.zeros((n, n)), F, obs, 1.0, 1e-14)
    assert np.abs(xa - obs).max() <= 1e-10 * np.abs(obs).max()
    assert np.abs(Pa).max() < 1e-12
