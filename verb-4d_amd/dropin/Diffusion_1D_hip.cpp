// Drop-in for src/Diffusion_1D.cpp (src/Diffusion_1D.h:69-87):
// Diffusion_1D runs on the GPU through verb_diffusion_1d;
// getCalculationMatrix1DDiffusion (used by data assimilation, not on the
// step's hot path) assembles the tridiagonal operator on the host.
#include "Diffusion_1D.h"

#include "verb_dropin.h"

bool Diffusion_1D(Matrix1D<double>& psd, const Matrix1D<double>& x, int x_size, double x_LBC, double x_UBC,
                  BoundaryConditionType x_LBC_type, BoundaryConditionType x_UBC_type, const Matrix1D<double>& Dxx,
                  const Matrix1D<double>& G, const Matrix1D<double>& Sources, const Matrix1D<double>& Losses,
                  double dt) {
    int rc = verb_diffusion_1d(verb_dropin::ctx(), psd.plane_array, x.plane_array, x_size, x_LBC, x_UBC,
                               (int)x_LBC_type, (int)x_UBC_type, Dxx.plane_array, G.plane_array, Sources.plane_array,
                               Losses.plane_array, dt);
    verb_dropin::check(rc, "1D_DIFF");
    return true;
}

// Tridiagonal operator of implicit 1-D diffusion: boundary rows (value or
// one-sided derivative) and, in the interior, 1/dt - L on the diagonal plus
// the conservative second difference d/dx (D G d/dx) / G built from the two
// one-sided half-steps, scaled by 1/2 each.
std::tuple<CalculationMatrix, CalculationMatrix, CalculationMatrix> getCalculationMatrix1DDiffusion(
    const Matrix1D<double>& x, int x_size, double x_LBC, double x_UBC, BoundaryConditionType x_LBC_type,
    BoundaryConditionType x_UBC_type, const Matrix1D<double>& Dxx, const Matrix1D<double>& G,
    const Matrix1D<double>& Sources, const Matrix1D<double>& Losses, double dt) {
    CalculationMatrix A(x_size, 1, 1, 1), B(x_size, 1, 1, 0), C(x_size, 1, 1, 0);
    for (auto& kv : A) kv.second = 0;
    for (auto& kv : B) kv.second = 0;
    for (auto& kv : C) kv.second = 0;
    auto half = [&](int i, int d1, int d2) {
        const double h1 = x[i] - x[i + d1];
        const double h21 = x[i] - x[i + d2];
        const double h22 = x[i + d1] - x[i + d1 + d2];
        const double cp = -0.5 / G[i] / h1;
        A[0][i] += +cp / h21 * Dxx[i] * G[i];
        A[d2][i] += -cp / h21 * Dxx[i] * G[i];
        A[d1][i] += -cp / h22 * Dxx[i + d1] * G[i + d1];
        A[d1 + d2][i] += +cp / h22 * Dxx[i + d1] * G[i + d1];
    };
    for (int i = 0; i < x_size; i++) {
        const bool lower = (i == 0 && x_size >= 3), upper = (i == x_size - 1 && x_size >= 3);
        if (lower || upper) {
            const int nb = lower ? +1 : -1;
            const BoundaryConditionType t = lower ? x_LBC_type : x_UBC_type;
            C[0][i] = lower ? x_LBC : x_UBC;
            const double dh = lower ? x[i + 1] - x[i] : x[i] - x[i - 1];
            if (t == BoundaryConditionType::ConstantValue) {
                A[0][i] = 1;
                A[nb][i] = 0;
            } else if (t == BoundaryConditionType::ConstantDerivative) {
                A[0][i] = 1 / dh;
                A[nb][i] = -1 / dh;
            } else {
                Logger::error << "1D_DIFF_BOUNDARY: unknown boundary type: " << t << std::endl;
                exit(EXIT_FAILURE);
            }
            continue;
        }
        A[0][i] += 1.0 / dt;
        B[0][i] += 1.0 / dt;
        C[0][i] += Sources[i];
        A[0][i] -= Losses[i];
        if (x_size >= 3) {
            half(i, -1, +1);
            half(i, +1, -1);
        }
    }
    return std::make_tuple(A, B, C);
}
