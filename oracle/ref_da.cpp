// Reference data-assimilation harness: TEST INFRASTRUCTURE ONLY.
//
// Linked by oracle/Makefile into oracle/_ref/libverbref.so with the
// reference's src/MatrixOperations.cpp (its dgemm_/dgemv_/dgesv_ calls
// renamed to numpy's OpenBLAS, as for MatrixSolver.cpp) and
// src/DataAssimilationHelper.cpp. src/DataAssimilation.cpp itself does not
// link here (its parameter reader needs Parameters.cpp and
// UpdatableMatrix.cpp), so ref_da_line and ref_da_conv_plane below make the calls of
// runKalmanFilterDiffusion1D (src/DataAssimilationDiffusion.cpp:151-249)
// and of runKalmanFilter (src/DataAssimilation.cpp:18-73) in their order,
// on the reference's own matrix operations: every product, solve and sum
// is the reference's code.
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "DataAssimilationHelper.h"
#include "MatrixOperations.h"

namespace {

Matrix1D<double> v1(const double* p, int n) {
    Matrix1D<double> m(n);
    std::memcpy(m.plane_array, p, sizeof(double) * n);
    return m;
}

}  // namespace

extern "C" {

// One line: forecast (n, in/out), Pa (n x n row-major, in/out), F (n x n,
// out, nullable), Pf (n x n, out, nullable). Returns 0, or -8 when the
// reference throws on a NaN forecast.
int ref_da_line(int n, const double* R, const double* DLL, const double* G, double lbc, double ubc, int lt, int ut,
                double dt, const double* obs, double* forecast, double* Pa, double* F_out, double* Pf_out,
                double model_error, double observation_error) {
    Matrix1D<double> Rm = v1(R, n), Dm = v1(DLL, n), Gm = v1(G, n), om = v1(obs, n), x = v1(forecast, n);
    Matrix2D<double> F(n, n);
    data_assimilation::internal::getModelMatrixDiffusion1D(F, Rm, Dm, Gm, lbc, ubc,
                                                           static_cast<BoundaryConditionType>(lt),
                                                           static_cast<BoundaryConditionType>(ut), dt);
    data_assimilation::ObservationSpace os = data_assimilation::internal::convertToObservationSpace(om);
    Matrix1D<double> D(n);
    D = model_error;
    Matrix2D<double> Q = diag(D);
    Matrix1D<double> E(os.data.size_q1);
    E = observation_error;
    Matrix2D<double> Robs = diag(E);
    Matrix2D<double> P(n, n);
    std::memcpy(P.plane_array, Pa, sizeof(double) * n * n);
    if (P.maxabs() == 0)
        for (int i = 0; i < n; i++) P[i][i] = 5 * model_error;
    // runKalmanFilter(x, P, F, Q, os.data, os.H, Robs)
    Matrix2D<double> Pf = abtrans(F * P, F);
    Pf += Q;
    Matrix2D<double> PfHT = abtrans(Pf, os.H);
    Matrix2D<double> S = os.H * PfHT;
    S += Robs;
    Matrix2D<double>& K = trans_solve(S, PfHT);
    Matrix1D<double> innovation = os.data - (os.H * x);
    Matrix1D<double> correction = K * innovation;
    x += correction;
    int st = 0;
    for (int i = 0; i < n; i++)
        if (std::isnan(x[i])) st = -8;
    P = Pf - (K * os.H * Pf);
    std::memcpy(forecast, x.plane_array, sizeof(double) * n);
    std::memcpy(Pa, P.plane_array, sizeof(double) * n * n);
    if (F_out) std::memcpy(F_out, F.plane_array, sizeof(double) * n * n);
    if (Pf_out) std::memcpy(Pf_out, Pf.plane_array, sizeof(double) * n * n);
    return st;
}

}  // extern "C"

extern "C" {

// One (V,K) plane of the convection filter: the calls of
// runKalmanFilterConvection2D (src/DataAssimilationConvection.cpp:124-238)
// in their order on the reference's getModelMatrixConvection2D and matrix
// operations. Arrays [nP][nR] (forecast in/out), Pa and M [N][N] (in/out;
// M is recomputed when have_model == 0). Returns 0, -8 when the reference
// throws (an unstable model matrix, a NaN analysis).
int ref_da_conv_plane(int nP, int nR, const double* P, const double* R, const double* VP, const double* VR,
                      const double* loss_rate, const int* saturated, const double* obs, double* forecast, double* Pa,
                      double* M, int have_model, double time_step, double correlation_time, double model_error,
                      double observation_error, int use_log, int growing_Q_at_boundary, double* Pf_out) {
    const int N = nP * nR;
    auto m2 = [&](const double* p) {
        Matrix2D<double> m(nP, nR);
        std::memcpy(m.plane_array, p, sizeof(double) * N);
        return m;
    };
    try {
        Matrix2D<double> Pm = m2(P), Rm = m2(R), VPm = m2(VP), VRm = m2(VR), Lm = m2(loss_rate), Fc = m2(forecast);
        Matrix2D<bool> sat(nP, nR);
        for (int i = 0; i < nP; i++)
            for (int j = 0; j < nR; j++) sat[i][j] = saturated[i * nR + j] != 0;
        const double dP = Pm[1][0] - Pm[0][0];
        const double dR = Rm[0][1] - Rm[0][0];
        auto forecast1D = toMatrix1D(Fc);
        Matrix2D<double> model(N, N);
        if (!have_model) {
            model = data_assimilation::internal::getModelMatrixConvection2D(VPm, VRm, Lm, sat, time_step, dP, dR,
                                                                            use_log != 0, correlation_time);
        } else {
            std::memcpy(model.plane_array, M, sizeof(double) * N * N);
        }
        auto os = data_assimilation::internal::convertToObservationSpace(m2(obs));
        if (use_log) {
            forecast1D = log(forecast1D);
            os.data = log(os.data);
        }
        Matrix1D<double> D(forecast1D.size_q1);
        D = model_error;
        const double cf = exp(-time_step / correlation_time);
        auto tmp = diag(D);
        for (int iP = 0; iP < nP; iP++) {
            tmp[nR - 1 + iP * nR][nR - 1 + iP * nR] = (1. - (cf * cf)) * model_error;
            tmp[iP * nR][iP * nR] = (1. - (cf * cf)) * model_error;
        }
        auto model_1day = model ^ (std::ceil(3. / 24 / time_step));
        auto Q = model_1day * tmp * transpose(model_1day) + diag(D) / 10;
        if (!growing_Q_at_boundary) {
            for (int iP = 0; iP < nP; iP++) {
                Q[nR - 1 + iP * nR][nR - 1 + iP * nR] = (1. - (cf * cf)) * model_error;
                Q[iP * nR][iP * nR] = (1. - (cf * cf)) * model_error;
            }
        }
        Matrix1D<double> E(os.data.size_q1);
        E = observation_error;
        auto Robs = use_log ? diag(E) : diag(os.data * observation_error);
        Matrix2D<double> Pam(N, N);
        std::memcpy(Pam.plane_array, Pa, sizeof(double) * N * N);
        if (Pam.maxabs() == 0)
            for (int i = 0; i < N; i++) Pam[i][i] = 5 * model_error;
        // runKalmanFilter (src/DataAssimilation.cpp:18-73)
        Matrix2D<double> Pf = abtrans(model * Pam, model);
        Pf += Q;
        Matrix2D<double> PfHT = abtrans(Pf, os.H);
        Matrix2D<double> S = os.H * PfHT;
        S += Robs;
        Matrix2D<double>& K = trans_solve(S, PfHT);
        Matrix1D<double> innovation = os.data - (os.H * forecast1D);
        Matrix1D<double> correction = K * innovation;
        forecast1D += correction;
        for (size_t i = 0; i < forecast1D.size_q1; ++i)
            if (std::isnan(forecast1D[i])) return -8;
        Pam = Pf - (K * os.H * Pf);
        if (use_log) forecast1D = pow(std::exp(1), forecast1D);
        auto out = toMatrix2D(forecast1D, nP, nR);
        std::memcpy(forecast, out.plane_array, sizeof(double) * N);
        std::memcpy(Pa, Pam.plane_array, sizeof(double) * N * N);
        std::memcpy(M, model.plane_array, sizeof(double) * N * N);
        if (Pf_out) std::memcpy(Pf_out, Pf.plane_array, sizeof(double) * N * N);
    } catch (...) {
        return -8;
    }
    return 0;
}

}  // extern "C"
