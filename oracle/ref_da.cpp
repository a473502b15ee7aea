// Reference data-assimilation harness: TEST INFRASTRUCTURE ONLY.
//
// Linked by oracle/Makefile into oracle/_ref/libverbref.so with the
// reference's src/MatrixOperations.cpp (its dgemm_/dgemv_/dgesv_ calls
// renamed to numpy's OpenBLAS, as for MatrixSolver.cpp) and
// src/DataAssimilationHelper.cpp. src/DataAssimilation.cpp itself does not
// link here (its parameter reader needs Parameters.cpp and
// UpdatableMatrix.cpp), so ref_da_line below makes the calls of
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
