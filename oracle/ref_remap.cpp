// Reference remapping harness: TEST INFRASTRUCTURE ONLY.
//
// Linked by oracle/Makefile into oracle/_ref/libverbref.so together with the
// reference's src/Interpolation.cpp and src/MonotCubicInterpolator.cpp,
// compiled from where they lie. Exposes the reference's Cubic1D and the
// remapping loop of src/VERB4D_Solver.cpp:574-606 with plain pointers, for
// tests/golden/make_golden.py. Nothing here is part of the product.
#include <cstring>
#include <stdexcept>

#include "Interpolation.h"
#include "Matrix.h"

namespace {

Matrix1D<double> line(const double* p, int n, size_t stride) {
    Matrix1D<double> m(n);
    for (int i = 0; i < n; i++) m.plane_array[i] = p[i * stride];
    return m;
}

// One call as the time loop makes it: PSD_L = Cubic1D(old, PSD_L, new). The
// returned Matrix1D is empty, so the assignment only marks PSD_L
// uninitialised; its array keeps the values Cubic1D wrote into it.
int cubic1d_line(const double* x_old, double* f, const double* x_new, int n, size_t stride) {
    Matrix1D<double> o = line(x_old, n, stride), v = line(f, n, stride), nw = line(x_new, n, stride);
    double* arr = v.plane_array;
    try {
        v = Cubic1D(o, v, nw);
    } catch (const std::exception&) {
        return -8;
    } catch (...) {
        return -8;
    }
    for (int i = 0; i < n; i++) f[i * stride] = arr[i];
    return 0;
}

}  // namespace

extern "C" {

int ref_cubic1d(int n, const double* x_old, double* values, const double* x_new) {
    return cubic1d_line(x_old, values, x_new, n, 1);
}

// src/VERB4D_Solver.cpp:574-606: every (P,V,K) line of psd [P][R][V][K]
// interpolated from l_old to l_new.
int ref_remap(double* psd, const double* l_old, const double* l_new, int nP, int nR, int nV, int nK) {
    const size_t NQ = (size_t)nV * nK;
    for (int iP = 0; iP < nP; iP++)
        for (size_t q = 0; q < NQ; q++) {
            const size_t base = (size_t)iP * nR * NQ + q;
            int st = cubic1d_line(l_old + base, psd + base, l_new + base, nR, NQ);
            if (st) return st;
        }
    return 0;
}

}  // extern "C"
