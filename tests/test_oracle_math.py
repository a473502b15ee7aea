"""Pins of the oracle that do not depend on the golden vectors: closed
forms, invariants, textbook routines and LAPACK itself (scipy's dgbtrf /
dgbtrs, backed by the same OpenBLAS 0.3.29 the reference links). Each check
is chosen so that a plausible slip (a dropped term, a wrong sign or index, a
transposed operand) fails it."""
import numpy as np
import pytest
import scipy.linalg.lapack as LA

import oracle as O
from parity import bit_equal, rel_err
from verb4d import workload as W


def dense_from_diag(diag, nx, ny):
    n = nx * ny
    A = np.zeros((n, n))
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            off = dy * nx + dx
            s = (dy + 1) * 3 + dx + 1
            for i in range(n):
                j = i + off
                if 0 <= j < n:
                    A[i, j] = diag[s][i]
    return A


def node_inputs(w, iP, iR):
    o = (iP, iR)
    return (w.f["Vg"][o], w.f["Kg"][o], w.bc["Vl"][o], w.bc["Vu"][o], w.bc["Kl"][o], w.bc["Ku"][o],
            w.f["DVV"][o], w.f["DKK"][o], w.f["DVK"][o], w.f["G_local"][o], w.f["sources"][o],
            w.f["losses_local"][o])


def test_tridag_matches_dense_solve():
    rng = np.random.default_rng(3)
    for n in (3, 5, 29):
        a, c = rng.random(n), rng.random(n)
        b = 3.0 + rng.random(n)
        r = rng.random(n)
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        u = O.tridag(a, b, c, r)
        np.testing.assert_allclose(u, np.linalg.solve(A, r), rtol=1e-13, atol=1e-14)


def test_tridag_bit_exact_vs_reference_when_present():
    if not O.ref_available():
        pytest.skip("reference harness not built here")
    rng = np.random.default_rng(4)
    a, b, c, r = rng.random(29), 4 + rng.random(29), rng.random(29), rng.random(29)
    assert bit_equal(O.tridag(a, b, c, r), O.tridag(a, b, c, r, lib="ref"))


def test_assembly_row_sums_and_boundary_rows():
    """Every SecondDerivativeApproximation_2D call adds four terms that sum to
    zero, so on interior rows sum_j A[i,j] = 1/dt - L[i]; boundary rows are
    the identity (Dirichlet) or the one-sided difference (Neumann)."""
    w = W.config0(seed=5, rich=True)
    nx, ny = w.nV, w.nK
    x, y, xl, xu, yl, yu, Dxx, Dyy, Dxy, G, S, L = node_inputs(w, 3, 4)
    types = (1, 2, 1, 2)
    dt = 300.0
    diag, B0, C0 = O.assemble_2d(x, y, nx, ny, xl, xu, yl, yu, types, Dxx, Dyy, Dxy, Dxy, G, S, L, dt)
    A = dense_from_diag(diag, nx, ny)
    for ix in range(nx):
        for iy in range(ny):
            i = iy * nx + ix
            row = A[i]
            if ix == 0:  # x-lower Dirichlet: identity row, C = BC
                assert row[i] == 1 and np.count_nonzero(row) == 1 and C0[i] == xl[iy] and B0[i] == 0
            elif ix == nx - 1:  # x-upper Neumann: (f_i - f_{i-1}) / dh
                dh = x[ix, iy] - x[ix - 1, iy]
                assert row[i] == 1 / dh and row[i - 1] == -1 / dh and np.count_nonzero(row) == 2
            elif iy == 0:
                assert row[i] == 1 and np.count_nonzero(row) == 1 and C0[i] == yl[ix]
            elif iy == ny - 1:
                dh = y[ix, iy] - y[ix, iy - 1]
                assert row[i] == 1 / dh and row[i - nx] == -1 / dh
            else:
                scale = np.abs(row).max()
                assert abs(row.sum() - (1 / dt - L[ix, iy])) <= 1e-12 * scale
                assert B0[i] == 1 / dt and C0[i] == S[ix, iy]


def test_assembly_matches_closed_form_laplacian():
    """Uniform unit grid, D constant, G = 1, no mixed term: the operator is
    the 5-point Laplacian with weight D/h^2 (factor 0.5 from each of the two
    one-sided calls), diagonal 1/dt + 4D/h^2."""
    nx = ny = 6
    x = np.repeat(np.arange(nx, dtype=float)[:, None], ny, 1) * 0.5
    y = np.repeat(np.arange(ny, dtype=float)[None, :], nx, 0) * 0.5
    one = np.ones((nx, ny))
    zero = np.zeros((nx, ny))
    D = 0.3
    diag, B0, C0 = O.assemble_2d(x, y, nx, ny, one[0], one[0], one[:, 0], one[:, 0], (1, 1, 1, 1), D * one, D * one,
                                 zero, zero, one, zero, zero, 10.0)
    i = 2 * nx + 3
    w = D / 0.25
    assert diag[4][i] == pytest.approx(0.1 + 4 * w, rel=1e-14)
    for s in (1, 3, 5, 7):
        assert diag[s][i] == pytest.approx(-w, rel=1e-14)
    for s in (0, 2, 6, 8):
        assert diag[s][i] == 0.0


def test_band_packing_matches_dense():
    w = W.config0(seed=6)
    nx, ny = w.nV, w.nK
    x, y, xl, xu, yl, yu, Dxx, Dyy, Dxy, G, S, L = node_inputs(w, 1, 2)
    diag, _, _ = O.assemble_2d(x, y, nx, ny, xl, xu, yl, yu, (1, 1, 1, 2), Dxx, Dyy, Dxy, Dxy, G, S, L, 300.0)
    A = dense_from_diag(diag, nx, ny)
    ab = O.pack_band(nx, ny, diag)
    kl = ku = nx + 1
    n = nx * ny
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            assert ab[kl + ku + i - j, j] == A[i, j]


@pytest.mark.parametrize("seed,kp", [(1, 3.0), (2, 6.0)])
def test_dgbtf2_pivots_match_lapack(seed, kp):
    """Same pivot sequence as LAPACK dgbtrf (OpenBLAS), factors to rounding,
    and the solve agrees with LAPACK dgbtrs and a dense solve."""
    w = W.config0(seed=seed)
    w.set_kp(kp)
    nx, ny = w.nV, w.nK
    x, y, xl, xu, yl, yu, Dxx, Dyy, Dxy, G, S, L = node_inputs(w, 4, 6)
    diag, B0, C0 = O.assemble_2d(x, y, nx, ny, xl, xu, yl, yu, (1, 1, 1, 2), Dxx, Dyy, Dxy, Dxy, G, S, L, 300.0)
    ab = O.pack_band(nx, ny, diag)
    kl = ku = nx + 1
    lu1, piv1, info1 = O.dgbtf2(ab, kl, ku)
    lu2, piv2, info2 = LA.dgbtrf(ab, kl, ku)
    assert info1 == 0 and info2 == 0
    assert np.array_equal(piv1, piv2 + 1)  # scipy returns 0-based pivots
    # at least one genuine row interchange happens (boundary rows pivot)
    assert np.any(piv1 != np.arange(1, len(piv1) + 1))
    assert rel_err(lu1, lu2) < 1e-12
    b = np.random.default_rng(seed).random(nx * ny)
    x1 = O.dgbtrs(lu1, kl, ku, b, piv1)
    x2, _ = LA.dgbtrs(lu2, kl, ku, b, piv2)
    assert rel_err(x1, x2) < 1e-12
    A = dense_from_diag(diag, nx, ny)
    assert rel_err(x1, np.linalg.solve(A, b)) < 1e-11


def test_dgbtf2_reports_zero_pivot():
    """A zero column gives info = its 1-based index, like dgbtrf."""
    n, kl, ku = 6, 1, 1
    ab = np.zeros((2 * kl + ku + 1, n), order="F")
    for j in range(n):
        ab[kl + ku, j] = 2.0
        if j > 0:
            ab[kl + ku - 1, j] = 0.5
        if j < n - 1:
            ab[kl + ku + 1, j] = 0.5
    ab[:, 3] = 0.0
    _, _, info = O.dgbtf2(ab, kl, ku)
    _, _, info2 = LA.dgbtrf(ab, kl, ku)
    assert info == info2 == 4


def test_planted_mistake_in_oracle_is_caught():
    """Flip one multiplier sign in a copy of the oracle's factors: the solve
    must then disagree with LAPACK, i.e. the comparison used above has teeth."""
    w = W.config0(seed=9)
    nx, ny = w.nV, w.nK
    x, y, xl, xu, yl, yu, Dxx, Dyy, Dxy, G, S, L = node_inputs(w, 2, 5)
    diag, _, _ = O.assemble_2d(x, y, nx, ny, xl, xu, yl, yu, (1, 1, 1, 2), Dxx, Dyy, Dxy, Dxy, G, S, L, 300.0)
    ab = O.pack_band(nx, ny, diag)
    kl = ku = nx + 1
    lu1, piv1, _ = O.dgbtf2(ab, kl, ku)
    lu2, piv2, _ = LA.dgbtrf(ab, kl, ku)
    bad = lu1.copy(order="F")
    col = nx * 3 + 4
    k = np.nonzero(bad[kl + ku + 1:, col])[0][0]
    bad[kl + ku + 1 + k, col] *= -1
    b = np.ones(nx * ny)
    xb = O.dgbtrs(bad, kl, ku, b, piv1)
    x2, _ = LA.dgbtrs(lu2, kl, ku, b, piv2)
    assert rel_err(xb, x2) > 1e-6


def test_quickest_courant_one_is_exact_shift():
    """With Courant number exactly 1 every (c^2 - k^2) term vanishes, the face
    value is the upwind value and the update shifts the profile one cell."""
    n = 24
    x = np.arange(n, dtype=float)
    rng = np.random.default_rng(11)
    psd = rng.random(n)
    psd[-1] = psd[0]
    ux = np.ones(n)
    out = O.quickest6(psd, x, n, 0.0, 0.0, 0, 0, ux, 1.0)
    expect = np.empty(n)
    expect[1:n - 1] = psd[0:n - 2]
    expect[0] = psd[n - 2]
    expect[n - 1] = expect[0]
    assert bit_equal(out, expect)
    # and to the left for negative velocity
    out = O.quickest6(psd, x, n, 0.0, 0.0, 0, 0, -ux, 1.0)
    expect = np.empty(n)
    expect[0:n - 2] = psd[1:n - 1]
    expect[n - 2] = psd[0]
    expect[n - 1] = expect[0]
    assert bit_equal(out, expect)


def test_quickest_periodic_conserves_mass_and_bounds():
    n = 24
    x = np.linspace(0, 2 * np.pi, n)
    psd = 1.0 + np.exp(-((x - 3.0) ** 2) / 0.2)
    psd[-1] = psd[0]
    ux = (0.7 + 0.2 * np.sin(x)) * (x[1] - x[0]) / 300.0
    ux[-1] = ux[0]
    out = O.quickest6(psd, x, n, 0.0, 0.0, 0, 0, ux, 300.0)
    assert abs(out[:-1].sum() - psd[:-1].sum()) < 1e-13 * psd.sum()
    # with a uniform velocity the limited scheme creates no new extrema
    ux = np.full(n, 0.7 * (x[1] - x[0]) / 300.0)
    out = O.quickest6(psd, x, n, 0.0, 0.0, 0, 0, ux, 300.0)
    assert out.min() >= psd.min() - 1e-14 and out.max() <= psd.max() + 1e-14
    assert abs(out[:-1].sum() - psd[:-1].sum()) < 1e-13 * psd.sum()


def test_quickest_uniform_field_with_uniform_velocity_is_fixed():
    n = 29
    x = np.linspace(1.5, 6.6, n)
    psd = np.full(n, 0.7)
    ux = np.full(n, 0.6 * (x[1] - x[0]) / 300.0)
    out = O.quickest6(psd, x, n, 0.7, 0.7, 1, 1, ux, 300.0)
    assert bit_equal(out, psd)


def test_quickest_courant_violation_status():
    n = 24
    x = np.linspace(0, 2 * np.pi, n)
    psd = np.ones(n)
    ux = np.full(n, np.nan)
    ux[3] = 1.0
    # a NaN-free velocity gives inner sub-steps; the status path is reached by
    # the convection driver only through NaN/inf velocities, so check inf
    ux = np.full(n, np.inf)
    with pytest.raises(RuntimeError):
        O.quickest6(psd, x, n, 0.0, 0.0, 0, 0, ux, 300.0)


def test_convection_substep_counts_closed_form():
    """num_steps = max(ceil(dt max|VP|/dP/0.25), ceil(dt max|VR|/dR/0.25))
    (Convection_2D.cpp:57-86). With dt*|VR|/dR/0.25 = 2.5 the call runs three
    sub-steps of dt/3, i.e. exactly three calls of dt/3 (each one sub-step)."""
    nP, nR = 24, 29
    P = np.repeat(np.linspace(0, 2 * np.pi, nP)[:, None], nR, 1)
    R = np.repeat(np.linspace(1.5, 6.6, nR)[None, :], nP, 0)
    dR = R[0, 1] - R[0, 0]
    rng = np.random.default_rng(2)
    psd = 1 + rng.random((nP, nR))
    psd[-1] = psd[0]
    VP = np.zeros((nP, nR))
    VR = np.full((nP, nR), 2.5 * 0.25 * dR / 300.0) * np.where(np.arange(nP) < 12, 1.0, -0.5)[:, None]
    z = np.zeros((nP, nR))
    G = 1 + 0.1 * rng.random((nP, nR))
    bcP = np.zeros(nR)
    bcR = psd[:, 0].copy(), psd[:, -1].copy()
    args = (P, R, nP, nR, bcP, bcP, bcR[0], bcR[1], (0, 0, 1, 1), VP, VR, z, z, G)
    full = O.convection_2d(psd, *args, 300.0)
    step = psd
    for _ in range(3):
        step = O.convection_2d(step, *args, 100.0)
    assert bit_equal(full, step)
    two = psd
    for _ in range(2):
        two = O.convection_2d(two, *args, 150.0)
    assert not bit_equal(full, two)


def test_convection_losses_and_sources_closed_form():
    """No velocity: f <- f*exp(L dt) on every point, then + S dt on interior R."""
    nP, nR = 24, 10
    P = np.repeat(np.linspace(0, 2 * np.pi, nP)[:, None], nR, 1)
    R = np.repeat(np.linspace(1.5, 6.6, nR)[None, :], nP, 0)
    rng = np.random.default_rng(3)
    psd = rng.random((nP, nR))
    psd[-1] = psd[0]  # closed periodic MLT grid: point nP-1 duplicates point 0
    L = -1e-4 * rng.random((nP, nR))
    S = 1e-6 * rng.random((nP, nR))
    z = np.zeros((nP, nR))
    one = np.ones((nP, nR))
    out = O.convection_2d(psd, P, R, nP, nR, np.zeros(nR), np.zeros(nR), psd[:, 0], psd[:, -1], (0, 0, 1, 1), z, z,
                          S, L, one, 300.0)
    expect = psd * np.exp(L * 300.0)
    expect[:, 1:-1] = expect[:, 1:-1] + S[:, 1:-1] * 300.0
    # numpy's exp and libm's exp may differ in the last place
    np.testing.assert_allclose(out, expect, rtol=3e-16, atol=0)
    assert not np.allclose(out, psd, rtol=1e-3)


def test_radial_row_sums():
    """Interior rows of the radial operator sum to 1/dt - L."""
    n = 29
    x = np.linspace(1.5, 6.6, n)
    D = 1e-3 * x ** 10 / x.max() ** 10
    G = 1.0 / x ** 2
    L = -1e-5 * np.ones(n)
    dt = 300.0
    # solve for the constant vector: A 1 = (1/dt - L) 1 on interior rows,
    # so with RHS = B f + C (C = S = 0) and f = 1 the interior solution stays
    # 1 only if the operator annihilates constants and L = 0
    psd = np.ones(n)
    out = O.diffusion_1d(psd, x, n, 1.0, 1.0, 1, 1, D, G, np.zeros(n), np.zeros(n), dt)
    assert np.abs(out - 1.0).max() < 1e-13
    out2 = O.diffusion_1d(psd, x, n, 1.0, 1.0, 1, 1, D, G, np.zeros(n), L, dt)
    assert out2[n // 2] < 1.0


def test_workload_coefficients_are_admissible():
    w = W.config0(seed=3)
    chk = w.f["DVV"] * w.f["DKK"] - w.f["DVK"] ** 2
    assert chk.min() > 0  # the time loop's stability check (VERB4D_Solver.cpp:747)
    u = W.splitmix_u01(5, 1, 1000)
    assert np.array_equal(u, W.splitmix_u01(5, 1, 1000))
    assert np.array_equal(u[500:], W.splitmix_u01(5, 1, 500, offset=500))
    assert 0.45 < u.mean() < 0.55
