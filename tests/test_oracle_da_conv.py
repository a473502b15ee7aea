"""The oracle's restatement of the convection (MLT, R plane) Kalman filter
(oracle/da_oracle.py: conv_model_matrix, kalman_conv_plane), pinned by the
reference's own code: the model matrices, forecast and analysis
covariances that getModelMatrixConvection2D and runKalmanFilterConvection2D
produce on the reference's matrix operations (tests/golden/da_planes.npz,
written by oracle/ref_da.cpp: ref_da_conv_plane) are reproduced bit for bit,
the analysis PSD to 1e-15 (its last product is a dgemv whose summation
order numpy's matmul does not share); and by what the mathematics fixes:
the matrix power against numpy's, rows of the one-step matrix summing to
1 + dt*loss where no velocity leaves the stencil, the boundary rows equal
to the correlation factor, and an unstable step raising."""
import math

import numpy as np
import pytest

import da_oracle as D


def cases(golden_dir):
    d = np.load(golden_dir + "/da_planes.npz")
    for i in range(int(d["count"])):
        yield {k.split("_", 1)[1]: d[k] for k in d.files if k.startswith("%d_" % i)}


def _model(c):
    P, R = c["P"], c["R"]
    dP, dR = P[1, 0] - P[0, 0], R[0, 1] - R[0, 0]
    return D.conv_model_matrix(c["VP"], c["VR"], c["loss"], c["sat"].astype(bool), float(c["dt"]), dP, dR,
                               bool(c["use_log"]), float(c["corr"]))


def test_oracle_matches_reference_planes(golden_dir):
    n = 0
    nts = set()
    for c in cases(golden_dir):
        M = _model(c)
        assert np.array_equal(M, c["M"]), n
        N = M.shape[0]
        dt = float(c["dt"])
        cp = np.abs(c["VP"]) * dt / (c["P"][1, 0] - c["P"][0, 0])
        cr = np.abs(c["VR"]) * dt / (c["R"][0, 1] - c["R"][0, 0])
        nts.add(int(max(1, math.ceil((cp + cr).max()))))
        args = (c["P"], c["R"], dt, float(c["corr"]), float(c["me"]), float(c["oe"]), bool(c["use_log"]),
                bool(c["growing"]))
        x1, P1, Pf1 = D.kalman_conv_plane(c["x0"], np.zeros((N, N)), M, c["obs"], *args)
        x2, P2, Pf2 = D.kalman_conv_plane(x1, P1, M, c["obs"], *args)
        assert np.array_equal(Pf1, c["Pf1"]) and np.array_equal(P1, c["P1"]), n
        assert np.array_equal(Pf2, c["Pf2"]) and np.array_equal(P2, c["P2"]), n
        for got, want in ((x1, c["x1"]), (x2, c["x2"])):
            assert np.abs(got - want).max() <= 1e-15 * np.abs(want).max(), n
        n += 1
    assert n == 6
    assert nts == {1, 2}


def test_matrix_power_against_numpy():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((7, 7)) / 3
    for p in (1, 2, 3, 5, 8, 13, 21, 64, 100):
        want = np.linalg.matrix_power(A, p)
        assert np.abs(D.matrix_power(A, p) - want).max() <= 1e-13 * max(1.0, np.abs(want).max()), p
    # a permutation's powers are exact
    Pm = np.eye(7)[[3, 0, 6, 1, 5, 2, 4]]
    for p in (1, 2, 3, 7, 12, 29):
        assert np.array_equal(D.matrix_power(Pm, p), np.linalg.matrix_power(Pm, p)), p


def test_step_matrix_structure():
    rng = np.random.default_rng(9)
    nP, nR, dt = 7, 6, 0.1
    dP, dR = 0.5, 0.3
    VP = rng.uniform(-1, 1, (nP, nR)) * dP / dt * 0.4
    VR = rng.uniform(-1, 1, (nP, nR)) * dR / dt * 0.4
    loss = -rng.uniform(0, 0.5, (nP, nR))
    sat = rng.random((nP, nR)) < 0.2
    corr = math.exp(-dt / 2.0)
    M = D.conv_step_matrix(VP, VR, loss, sat, dt, dP, dR, False, corr)
    for iP in range(nP):
        for iR in range(nR):
            row = M[iR + iP * nR]
            if iR in (0, nR - 1):
                assert row[iR + iP * nR] == corr and np.count_nonzero(row) == 1
                continue
            lc = 0.0 if sat[iP, iR] else loss[iP, iR]
            # the upwind stencil conserves what does not decay
            assert abs(row.sum() - (1 + dt * lc)) <= 1e-14
            assert np.count_nonzero(row) <= 3 and (row >= 0).all()
    # log space drops the loss term
    Ml = D.conv_step_matrix(VP, VR, loss, sat, dt, dP, dR, True, corr)
    inner = [iR + iP * nR for iP in range(nP) for iR in range(1, nR - 1)]
    assert np.abs(Ml[inner].sum(axis=1) - 1).max() <= 1e-14
    # a Courant number above one in one sub-step raises
    with pytest.raises(FloatingPointError):
        D.conv_step_matrix(VP * 4, VR * 4, loss, sat, dt, dP, dR, False, corr)
