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
  conv_*, kalman_conv_plane   the convection filter of one (V,K) plane
                 (src/DataAssimilationConvection.cpp:124-238,
                 src/DataAssimilationHelper.cpp:25-148, operator^
                 src/MatrixOperations.cpp:466-620)
numpy's products and solves run on the same OpenBLAS the reference links.
"""
import math

import numpy as np

# the reference's element-wise exp/log are the C library's; numpy's own
# vectorised loops may differ from them in the last place
_log = np.vectorize(lambda v: math.log(v) if v > 0 else (-math.inf if v == 0 else math.nan), otypes=[np.float64])
_pow = np.vectorize(math.pow, otypes=[np.float64])


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


def kalman_line(x, Pa, F, obs, model_error, observation_error, Q=None, R=None, data=None, H=None):
    """(x_analysis, Pa_analysis, Pf) of runKalmanFilter with the radial
    filter's Q and R (or the given ones); raises FloatingPointError on a
    NaN analysis (the reference throws)."""
    n = len(x)
    if data is None:
        data, H = observations(obs)
    Q = np.diag(np.full(n, model_error)) if Q is None else Q
    Robs = np.diag(np.full(len(data), observation_error)) if R is None else R
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


# --------------------------------------------------------------------------
# The convection filter of DataAssimilationManagerConvection: one (V,K)
# plane, state = the plane's nP x nR points (src/DataAssimilationConvection.cpp:
# 124-238; model matrix src/DataAssimilationHelper.cpp:25-148; matrix power
# src/MatrixOperations.cpp:466-620)
# --------------------------------------------------------------------------
def conv_step_matrix(VP, VR, loss, sat, dt, dP, dR, use_log, corr):
    """getModelMatrixConvection2DNoStabilityCheck (:25-102); raises
    FloatingPointError where the reference throws (m_cc < 0)."""
    nP, nR = VP.shape
    N = nP * nR
    M = np.zeros((N, N))
    for iP in range(nP):
        for iR in range(nR):
            row = iR + iP * nR
            if iR == 0 or iR == nR - 1:
                M[row, row] = corr
                continue
            if iP == 0 or iP == nP - 1:
                iPl, iPr = nP - 2, 1
            else:
                iPl, iPr = iP - 1, iP + 1
            vp, vr = VP[iP, iR], VR[iP, iR]
            lc = 0.0 if sat[iP, iR] else loss[iP, iR]
            if use_log:
                mcc = 1. - abs(vp) * dt / dP - abs(vr) * dt / dR
            else:
                mcc = 1. - abs(vp) * dt / dP - abs(vr) * dt / dR + dt * lc
            if mcc < 0:
                raise FloatingPointError("unstable model matrix")
            if vp > 0.:
                mlc, mrc = vp * dt / dP, 0.
            else:
                mlc, mrc = 0., -vp * dt / dP
            if vr > 0.:
                mcl, mcr = vr * dt / dR, 0.
            else:
                mcl, mcr = 0., -vr * dt / dR
            M[row, row] = mcc
            M[row, iR + iPl * nR] = mlc
            M[row, iR + iPr * nR] = mrc
            M[row, iR - 1 + iP * nR] = mcl
            M[row, iR + 1 + iP * nR] = mcr
    return M


def matrix_power(A, p):
    """operator^ (src/MatrixOperations.cpp:466-620): A^2, A^4, ... by
    repeated squaring; the squares of the set bits >= 1 combined in
    increasing order, each next one multiplied from the left (its dgemm
    takes the running product as the right-hand factor in row-major terms),
    then A from the left when p is odd."""
    p = int(p)
    if p == 1:
        return A
    sq, cur, bits = [], A, 1
    while (p >> bits) > 0:
        cur = cur @ cur
        if (p >> bits) & 1:
            sq.append(cur)
        bits += 1
    res = sq[0]
    for s in sq[1:]:
        res = s @ res
    if p & 1:
        res = A @ res
    return res


def conv_model_matrix(VP, VR, loss, sat, dt, dP, dR, use_log, corr_time):
    """getModelMatrixConvection2D (:116-148): the Courant split
    (splitTimeStepCourantCondition :104-114, maximum Courant number 1) and
    the one-step matrix raised to the number of sub-steps."""
    cp = VP * dt / dP
    cr = VR * dt / dR
    cour = np.abs(np.abs(cp) + np.abs(cr)).max()
    nt = 1 if cour <= 1 else int(np.ceil(cour / 1))
    dtn = dt / float(nt)
    step = conv_step_matrix(VP, VR, loss, sat, dtn, dP, dR, use_log, math.exp(-dtn / corr_time))
    return matrix_power(step, nt)


def kalman_conv_plane(forecast, Pa, M, obs, P, R, dt, corr_time, model_error, observation_error, use_log,
                      growing_Q_at_boundary):
    """runKalmanFilterConvection2D after the model matrix (:151-238):
    returns (forecast, Pa, Pf); forecast, obs [nP][nR]."""
    nP, nR = forecast.shape
    N = nP * nR
    x = forecast.reshape(-1).astype(np.float64)
    data, H = observations(obs.reshape(-1))
    if use_log:
        x = _log(x)
        data = _log(data)
    cf = math.exp(-dt / corr_time)
    D = np.full(N, model_error)
    tmp = np.diag(D)
    for iP in range(nP):
        tmp[nR - 1 + iP * nR, nR - 1 + iP * nR] = (1. - (cf * cf)) * model_error
        tmp[iP * nR, iP * nR] = (1. - (cf * cf)) * model_error
    M1 = matrix_power(M, int(np.ceil(3. / 24 / dt)))
    # (transpose(...) is a new matrix in the reference: a plain product)
    Q = (M1 @ tmp) @ np.ascontiguousarray(M1.T) + np.diag(D) / 10
    if not growing_Q_at_boundary:
        for iP in range(nP):
            Q[nR - 1 + iP * nR, nR - 1 + iP * nR] = (1. - (cf * cf)) * model_error
            Q[iP * nR, iP * nR] = (1. - (cf * cf)) * model_error
    Robs = np.diag(np.full(len(data), observation_error)) if use_log else np.diag(data * observation_error)
    xa, Pa_new, Pf = kalman_line(x, Pa, M, None, model_error, observation_error, Q=Q, R=Robs, data=data, H=H)
    if use_log:
        xa = _pow(math.exp(1), xa)
    return xa.reshape(nP, nR), Pa_new, Pf
