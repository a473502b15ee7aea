"""CPU restatement of the radial data assimilation (the Kalman update along L
of DataAssimilationManagerDiffusion). TEST INFRASTRUCTURE ONLY: the checker
of verb_da_radial.

  model_matrix   getModelMatrixDiffusion1D (src/DataAssimilationHelper.cpp:
                 168-189): F = trans_solve(A, B) = B A^-1 with A, B of
                 getCalculationMatrix1DDiffusion (src/Diffusion_1D.cpp:91-184),
                 zero sources and losses; trans_solve = dgesv on A^T with the
                 rows of B as right-hand sides (src/MatrixOperations.cpp:761-787)
  observations   convertToObservationSpace (src/DataAssimilationHelper.cpp:
                 271-305): H selects the finite observations; none -> one
                 all-zero row with value 0
  kalman_line    runKalmanFilterDiffusion1D (src/DataAssimilationDiffusion.cpp:
                 151-249) and runKalmanFilter (src/DataAssimilation.cpp:18-73)
  da_radial      the loop over (V,K) lines of one P slice (:100-141)
numpy's products and solves run on the same OpenBLAS the reference links.
"""
import numpy as np


def calc_matrix_1d(x, D, G, lt, ut, dt):
    """(A, B) dense of the implicit 1-D diffusion of one line, zero sources
    and losses, in getCalculationMatrix1DDiffusion's operation order."""
    n = len(x)
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    for j in range(n):
        if j == 0:
            dh = x[1] - x[0]
            if lt == 1:
                A[0, 0], A[0, 1] = 1.0, 0.0
            elif lt == 2:
                A[0, 0], A[0, 1] = 1 / dh, -1 / dh
            else:
                raise ValueError("boundary type")
        elif j == n - 1:
            dh = x[j] - x[j - 1]
            if ut == 1:
                A[j, j], A[j, j - 1] = 1.0, 0.0
            elif ut == 2:
                A[j, j], A[j, j - 1] = 1 / dh, -1 / dh
            else:
                raise ValueError("boundary type")
        else:
            b = 0.0 + 1.0 / dt
            B[j, j] = 0.0 + 1.0 / dt
            b -= 0.0
            a = c = 0.0
            xm, x0, xp = x[j - 1], x[j], x[j + 1]
            for d1, d2 in ((-1, 1), (1, -1)):
                dh1 = x0 - x[j + d1]
                dh21 = x0 - x[j + d2]
                dh22 = x[j + d1] - x0
                cp = -0.5 / G[j] / dh1
                t0 = cp / dh21 * D[j] * G[j]
                t1 = cp / dh22 * D[j + d1] * G[j + d1]
                b += t0
                if d2 == 1:
                    c -= t0
                else:
                    a -= t0
                if d1 == -1:
                    a -= t1
                else:
                    c -= t1
                b += t1
            A[j, j - 1], A[j, j], A[j, j + 1] = a, b, c
    return A, B


def model_matrix(x, D, G, lt, ut, dt):
    A, B = calc_matrix_1d(x, D, G, lt, ut, dt)
    return np.linalg.solve(A.T, B.T).T


def observations(obs):
    fin = np.isfinite(obs)
    n = len(obs)
    if not fin.any():
        return np.zeros(1), np.zeros((1, n))
    idx = np.nonzero(fin)[0]
    H = np.zeros((len(idx), n))
    H[np.arange(len(idx)), idx] = 1.0
    return obs[idx].astype(np.float64), H


def kalman_line(x, Pa, F, obs, model_error, observation_error):
    """(x_analysis, Pa_analysis, Pf); raises FloatingPointError on a NaN
    analysis (the reference throws)."""
    n = len(x)
    data, H = observations(obs)
    Q = np.diag(np.full(n, model_error))
    Robs = np.diag(np.full(len(data), observation_error))
    P = np.array(Pa, dtype=np.float64, copy=True)
    if np.abs(P).max() == 0:
        P[np.arange(n), np.arange(n)] = 5 * model_error
    Pf = (F @ P) @ F.T
    Pf = Pf + Q
    PfHT = Pf @ H.T
    S = H @ PfHT
    S = S + Robs
    K = np.linalg.solve(S.T, PfHT.T).T
    innovation = data - H @ x
    xa = x + K @ innovation
    if np.isnan(xa).any():
        raise FloatingPointError("NaN detected in forecast vector after Kalman update.")
    Pa_new = Pf - (K @ H) @ Pf
    return xa, Pa_new, Pf


def da_radial(w, psd, obs, Pa, iP, dt, model_error, observation_error):
    """Every (V,K) line of P slice iP: returns (psd, Pa) updated copies.
    psd, obs [P][R][V][K]; Pa [V][K][R][R]."""
    psd = np.array(psd, copy=True)
    Pa = np.array(Pa, copy=True)
    L, D, G = w.f["Lg"], w.f["DLL"], w.f["G_radial"]
    for iV in range(w.nV):
        for iK in range(w.nK):
            F = model_matrix(L[iP, :, iV, iK], D[iP, :, iV, iK], G[iP, :, iV, iK], w.types["Ll"], w.types["Lu"], dt)
            xa, pa, _ = kalman_line(psd[iP, :, iV, iK], Pa[iV, iK], F, obs[iP, :, iV, iK], model_error,
                                    observation_error)
            psd[iP, :, iV, iK] = xa
            Pa[iV, iK] = pa
    return psd, Pa
