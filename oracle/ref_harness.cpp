// Reference harness: TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile together with the reference's own hot-path
// translation units, compiled from where they lie in the PRBEM/VERB-4D tree
// (src/Diffusion_2D.cpp, src/Diffusion_1D.cpp, src/Convection_2D.cpp,
// src/Convection_1D_ULTIMATE_QUICKEST6.cpp, src/MatrixSolver.cpp,
// src/Matrix.cpp, src/Logger.cpp). It exposes plain-pointer entry points so
// that tests/golden/make_golden.py can record input/output vectors and
// bench.py can time the reference CPU path as the baseline. Nothing here
// is part of the product.
//
// The loops in ref_step() mirror the OpenMP regions of
// src/VERB4D_Solver.cpp:860-958 (convection over (V,K) planes), :999-1055
// (radial over (P,V,K) lines) and :1100-1210 (local over (P,R) nodes), with
// the same schedule(dynamic,1).
#include <cstring>
#include <cstdint>
#include <omp.h>

#include "Convection_1D_ULTIMATE_QUICKEST6.h"
#include "Convection_2D.h"
#include "Diffusion_1D.h"
#include "Diffusion_2D.h"
#include "Matrix.h"
#include "MatrixSolver.h"

namespace {

BoundaryConditionType bct(int t) { return static_cast<BoundaryConditionType>(t); }

Matrix1D<double> m1(const double* p, size_t n) {
    Matrix1D<double> m(n);
    std::memcpy(m.plane_array, p, sizeof(double) * n);
    return m;
}

Matrix2D<double> m2(const double* p, size_t n1, size_t n2) {
    Matrix2D<double> m(n1, n2);
    std::memcpy(m.plane_array, p, sizeof(double) * n1 * n2);
    return m;
}

}  // namespace

extern "C" {

#ifndef LU_CACHING
// Diffusion_2D (src/Diffusion_2D.h:86-111, non-LU_CACHING variant). 2-D
// arrays are [nx][ny] row-major like Matrix2D.
int ref_diffusion_2d(double* psd, const double* x, const double* y, int nx, int ny,
                     const double* xl, const double* xu, const double* yl, const double* yu,
                     int xlt, int xut, int ylt, int yut,
                     const double* Dxx, const double* Dyy, const double* Dxy, const double* Dyx,
                     const double* G, const double* S, const double* L, double dt, double sub_dt) {
    Matrix2D<double> f = m2(psd, nx, ny);
    Diffusion_2D(f, m2(x, nx, ny), m2(y, nx, ny), nx, ny, m1(xl, ny), m1(xu, ny), m1(yl, nx), m1(yu, nx),
                 bct(xlt), bct(xut), bct(ylt), bct(yut), m2(Dxx, nx, ny), m2(Dyy, nx, ny), m2(Dxy, nx, ny),
                 m2(Dyx, nx, ny), m2(G, nx, ny), m2(S, nx, ny), m2(L, nx, ny), dt, sub_dt);
    std::memcpy(psd, f.plane_array, sizeof(double) * nx * ny);
    return 0;
}
#else
// Diffusion_2D of an LU_CACHING build (src/Diffusion_2D.h:86-99): the band
// factors are written to lu_cache/ipiv_cache when recompute != 0 and reused
// otherwise (src/MatrixSolver.cpp:249-277).
int ref_diffusion_2d_cached(double* psd, const double* x, const double* y, int nx, int ny,
                            const double* xl, const double* xu, const double* yl, const double* yu,
                            int xlt, int xut, int ylt, int yut,
                            const double* Dxx, const double* Dyy, const double* Dxy, const double* Dyx,
                            const double* G, const double* S, const double* L, double dt, double sub_dt,
                            double* lu_cache, int64_t* ipiv_cache, int recompute) {
    Matrix2D<double> f = m2(psd, nx, ny);
    Diffusion_2D(f, m2(x, nx, ny), m2(y, nx, ny), nx, ny, m1(xl, ny), m1(xu, ny), m1(yl, nx), m1(yu, nx),
                 bct(xlt), bct(xut), bct(ylt), bct(yut), m2(Dxx, nx, ny), m2(Dyy, nx, ny), m2(Dxy, nx, ny),
                 m2(Dyx, nx, ny), m2(G, nx, ny), m2(S, nx, ny), m2(L, nx, ny), dt, sub_dt, lu_cache,
                 reinterpret_cast<long*>(ipiv_cache), recompute != 0);
    std::memcpy(psd, f.plane_array, sizeof(double) * nx * ny);
    return 0;
}
#endif

// The 9 diagonals (slot (dy+1)*3+(dx+1), offset dy*nx+dx) plus B0, C0 that
// Diffusion_2D assembles, obtained by running the reference assembly code
// path (AddBoundaries_2D + SecondDerivativeApproximation_2D) on a
// CalculationMatrix exactly as src/Diffusion_2D.cpp:49-130 does.
int ref_assemble_2d(int nx, int ny, const double* x, const double* y,
                    const double* xl, const double* xu, const double* yl, const double* yu,
                    int xlt, int xut, int ylt, int yut,
                    const double* Dxx, const double* Dyy, const double* Dxy, const double* Dyx,
                    const double* G, const double* S, const double* L, double dt,
                    double* diag, double* B0, double* C0) {
    Matrix2D<double> X = m2(x, nx, ny), Y = m2(y, nx, ny), DXX = m2(Dxx, nx, ny), DYY = m2(Dyy, nx, ny),
                     DXY = m2(Dxy, nx, ny), DYX = m2(Dyx, nx, ny), GG = m2(G, nx, ny), SS = m2(S, nx, ny),
                     LL = m2(L, nx, ny);
    Matrix1D<double> XL = m1(xl, ny), XU = m1(xu, ny), YL = m1(yl, nx), YU = m1(yu, nx);
    CalculationMatrix A(nx, ny, 1, 1), B(nx, ny, 1, 0), C(nx, ny, 1, 0);
    for (auto& kv : A) kv.second = 0;
    for (auto& kv : B) kv.second = 0;
    for (auto& kv : C) kv.second = 0;
    for (int ix = 0; ix < nx; ix++) {
        for (int iy = 0; iy < ny; iy++) {
            int in = A.index1d(ix, iy);
            if ((ix == 0 && nx >= 3) || (ix == nx - 1 && nx >= 3) || (iy == 0 && ny >= 3) ||
                (iy == ny - 1 && ny >= 3)) {
                AddBoundaries_2D(A, C, X, Y, nx, ny, XL, XU, YL, YU, bct(xlt), bct(xut), bct(ylt), bct(yut), ix,
                                 iy, in);
            } else {
                A[0][in] += 1.0 / dt;
                B[0][in] += 1.0 / dt;
                C[0][in] += SS[ix][iy];
                A[0][in] -= LL[ix][iy];
                SecondDerivativeApproximation_2D(A, ix, iy, "x_left", "x_right", X, Y, DXX, GG, -0.5);
                SecondDerivativeApproximation_2D(A, ix, iy, "x_right", "x_left", X, Y, DXX, GG, -0.5);
                SecondDerivativeApproximation_2D(A, ix, iy, "y_left", "y_right", X, Y, DYY, GG, -0.5);
                SecondDerivativeApproximation_2D(A, ix, iy, "y_right", "y_left", X, Y, DYY, GG, -0.5);
                SecondDerivativeApproximation_2D(A, ix, iy, "y_left", "x_right", X, Y, DXY, GG, -0.25);
                SecondDerivativeApproximation_2D(A, ix, iy, "y_right", "x_left", X, Y, DXY, GG, -0.25);
                SecondDerivativeApproximation_2D(A, ix, iy, "y_left", "x_left", X, Y, DXY, GG, -0.25);
                SecondDerivativeApproximation_2D(A, ix, iy, "y_right", "x_right", X, Y, DXY, GG, -0.25);
                SecondDerivativeApproximation_2D(A, ix, iy, "x_left", "y_right", X, Y, DYX, GG, -0.25);
                SecondDerivativeApproximation_2D(A, ix, iy, "x_right", "y_left", X, Y, DYX, GG, -0.25);
                SecondDerivativeApproximation_2D(A, ix, iy, "x_left", "y_left", X, Y, DYX, GG, -0.25);
                SecondDerivativeApproximation_2D(A, ix, iy, "x_right", "y_right", X, Y, DYX, GG, -0.25);
            }
        }
    }
    const int n = nx * ny;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            int off = A.index1d(dx, dy);
            std::memcpy(diag + (size_t)((dy + 1) * 3 + (dx + 1)) * n, A[off].plane_array, sizeof(double) * n);
        }
    std::memcpy(B0, B[0].plane_array, sizeof(double) * n);
    std::memcpy(C0, C[0].plane_array, sizeof(double) * n);
    return 0;
}

int ref_diffusion_1d(double* psd, const double* x, int n, double lbc, double ubc, int lt, int ut,
                     const double* D, const double* G, const double* S, const double* L, double dt) {
    Matrix1D<double> f = m1(psd, n);
    Diffusion_1D(f, m1(x, n), n, lbc, ubc, bct(lt), bct(ut), m1(D, n), m1(G, n), m1(S, n), m1(L, n), dt);
    std::memcpy(psd, f.plane_array, sizeof(double) * n);
    return 0;
}

int ref_tridag(const double* a, const double* b, const double* c, const double* r, double* u, int n) {
    Matrix1D<double> A = m1(a, n), B = m1(b, n), C = m1(c, n), R = m1(r, n);
    tridag(A.plane_array, B.plane_array, C.plane_array, R.plane_array, u, (size_t)n);
    return 0;
}

int ref_quickest6(double* psd, const double* x, int n, double lbc, double ubc, int lt, int ut, const double* ux,
                  double dt) {
    Matrix1D<double> f = m1(psd, n);
    Convection_1D_ULTIMATE_QUICKEST6(f, m1(x, n), n, lbc, ubc, bct(lt), bct(ut), m1(ux, n), dt);
    std::memcpy(psd, f.plane_array, sizeof(double) * n);
    return 0;
}

#ifdef SAVE_PSD_LOST_CONV
// Convection_2D of a SAVE_PSD_LOST_CONV build (src/Convection_2D.h:84-96):
// lost [nP][nR] is read and accumulated.
int ref_convection_2d_lost(double* psd, double* lost, const double* Pg, const double* Rg, int nP, int nR,
                           const double* Pl, const double* Pu, const double* Rl, const double* Ru, int plt, int put,
                           int rlt, int rut, const double* VP, const double* VR, const double* S, const double* L,
                           const double* G, double dt) {
    Matrix2D<double> f = m2(psd, nP, nR), lo = m2(lost, nP, nR);
    Convection_2D(f, lo, m2(Pg, nP, nR), m2(Rg, nP, nR), nP, nR, m1(Pl, nR), m1(Pu, nR), m1(Rl, nP), m1(Ru, nP),
                  bct(plt), bct(put), bct(rlt), bct(rut), m2(VP, nP, nR), m2(VR, nP, nR), m2(S, nP, nR),
                  m2(L, nP, nR), m2(G, nP, nR), dt);
    std::memcpy(psd, f.plane_array, sizeof(double) * nP * nR);
    std::memcpy(lost, lo.plane_array, sizeof(double) * nP * nR);
    return 0;
}
#else
int ref_convection_2d(double* psd, const double* Pg, const double* Rg, int nP, int nR, const double* Pl,
                      const double* Pu, const double* Rl, const double* Ru, int plt, int put, int rlt, int rut,
                      const double* VP, const double* VR, const double* S, const double* L, const double* G,
                      double dt) {
    Matrix2D<double> f = m2(psd, nP, nR);
    Convection_2D(f, m2(Pg, nP, nR), m2(Rg, nP, nR), nP, nR, m1(Pl, nR), m1(Pu, nR), m1(Rl, nP), m1(Ru, nP),
                  bct(plt), bct(put), bct(rlt), bct(rut), m2(VP, nP, nR), m2(VR, nP, nR), m2(S, nP, nR),
                  m2(L, nP, nR), m2(G, nP, nR), dt);
    std::memcpy(psd, f.plane_array, sizeof(double) * nP * nR);
    return 0;
}
#endif

// Full-grid arrays handed over by pointer; the layout of the struct is the
// one the oracle uses (oracle/verb_oracle.c, orc_grid).
typedef struct {
    int nP, nR, nV, nK;
    const double *Pg, *Rg, *Vg, *Kg, *Lg;
    const double *DLL, *DVV, *DKK, *DVK, *VP, *VR;
    const double *Gloc, *Grad, *Gconv, *Src, *Lloc, *Lrad, *Lconv;
    const double *Pl, *Pu, *Rl, *Ru, *Ll, *Lu, *Vl, *Vu, *Kl, *Ku;
    int Pl_t, Pu_t, Rl_t, Ru_t, Ll_t, Lu_t, Vl_t, Vu_t, Kl_t, Ku_t;
    double dt, sub_dt;
    double radial_sources, local_sources;
    int include_boundary;
} ref_grid;

static inline size_t idx4(const ref_grid* g, int p, int r, int v, int k) {
    return (((size_t)p * g->nR + r) * g->nV + v) * g->nK + k;
}

#if !defined(LU_CACHING) && !defined(SAVE_PSD_LOST_CONV)
static void conv_plane(const ref_grid* g, double* psd, int iV, int iK) {
    const int nP = g->nP, nR = g->nR;
    Matrix2D<double> f(nP, nR), Pg(nP, nR), Rg(nP, nR), VP(nP, nR), VR(nP, nR), S(nP, nR), L(nP, nR), G(nP, nR);
    Matrix1D<double> Pl(nR), Pu(nR), Rl(nP), Ru(nP);
    for (int p = 0; p < nP; p++)
        for (int r = 0; r < nR; r++) {
            size_t i = idx4(g, p, r, iV, iK);
            f[p][r] = psd[i]; Pg[p][r] = g->Pg[i]; Rg[p][r] = g->Rg[i]; VP[p][r] = g->VP[i];
            VR[p][r] = g->VR[i]; S[p][r] = g->Src[i]; L[p][r] = g->Lconv[i]; G[p][r] = g->Gconv[i];
        }
    for (int r = 0; r < nR; r++) {
        Pl[r] = g->Pl[((size_t)r * g->nV + iV) * g->nK + iK];
        Pu[r] = g->Pu[((size_t)r * g->nV + iV) * g->nK + iK];
    }
    for (int p = 0; p < nP; p++) {
        Rl[p] = g->Rl[((size_t)p * g->nV + iV) * g->nK + iK];
        Ru[p] = g->Ru[((size_t)p * g->nV + iV) * g->nK + iK];
    }
    Convection_2D(f, Pg, Rg, nP, nR, Pl, Pu, Rl, Ru, bct(g->Pl_t), bct(g->Pu_t), bct(g->Rl_t), bct(g->Ru_t), VP, VR,
                  S, L, G, g->dt);
    for (int p = 0; p < nP; p++)
        for (int r = 0; r < nR; r++) psd[idx4(g, p, r, iV, iK)] = f[p][r];
}

static void radial_line(const ref_grid* g, double* psd, int iP, int iV, int iK) {
    const int nR = g->nR;
    Matrix1D<double> f(nR), x(nR), D(nR), S(nR), G(nR), L(nR);
    for (int r = 0; r < nR; r++) {
        size_t i = idx4(g, iP, r, iV, iK);
        f[r] = psd[i]; x[r] = g->Lg[i]; D[r] = g->DLL[i]; S[r] = g->Src[i]; G[r] = g->Grad[i]; L[r] = g->Lrad[i];
    }
    S *= g->radial_sources;
    size_t bi = ((size_t)iP * g->nV + iV) * g->nK + iK;
    Diffusion_1D(f, x, nR, g->Ll[bi], g->Lu[bi], bct(g->Ll_t), bct(g->Lu_t), D, G, S, L, g->dt);
    for (int r = 0; r < nR; r++) psd[idx4(g, iP, r, iV, iK)] = f[r];
}

static void local_node(const ref_grid* g, double* psd, int iP, int iR) {
    const int nV = g->nV, nK = g->nK, n = nV * nK;
    const size_t o = idx4(g, iP, iR, 0, 0);
    Matrix2D<double> f = m2(psd + o, nV, nK);
    const double* Vl = g->Vl + ((size_t)iP * g->nR + iR) * nK;
    const double* Vu = g->Vu + ((size_t)iP * g->nR + iR) * nK;
    const double* Kl = g->Kl + ((size_t)iP * g->nR + iR) * nV;
    const double* Ku = g->Ku + ((size_t)iP * g->nR + iR) * nV;
    Diffusion_2D(f, m2(g->Vg + o, nV, nK), m2(g->Kg + o, nV, nK), nV, nK, m1(Vl, nK), m1(Vu, nK), m1(Kl, nV),
                 m1(Ku, nV), bct(g->Vl_t), bct(g->Vu_t), bct(g->Kl_t), bct(g->Ku_t), m2(g->DVV + o, nV, nK),
                 m2(g->DKK + o, nV, nK), m2(g->DVK + o, nV, nK), m2(g->DVK + o, nV, nK), m2(g->Gloc + o, nV, nK),
                 m2(g->Src + o, nV, nK) * g->local_sources, m2(g->Lloc + o, nV, nK), g->dt, g->sub_dt);
    std::memcpy(psd + o, f.plane_array, sizeof(double) * n);
}

// One full step, or a bounded sample of it: planes [q0,q1), lines [l0,l1),
// nodes [n0,n1) (pass -1 for q1/l1/n1 to run every slice).
int ref_step_sample(const ref_grid* g, double* psd, int nthreads, int run_convection, int run_radial,
                    int run_local, long q0, long q1, long l0, long l1, long n0, long n1) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    if (q1 < 0) q1 = (long)g->nV * g->nK;
    if (l1 < 0) l1 = (long)g->nP * g->nV * g->nK;
    if (n1 < 0) n1 = (long)g->nP * g->nR;
    if ((3 < g->nP || 3 < g->nR) && run_convection) {
#pragma omp parallel for schedule(dynamic, 1)
        for (long q = q0; q < q1; q++) conv_plane(g, psd, (int)(q / g->nK), (int)(q % g->nK));
    }
    if (g->nR >= 3 && run_radial) {
#pragma omp parallel for schedule(dynamic, 1)
        for (long l = l0; l < l1; l++)
            radial_line(g, psd, (int)(l / ((long)g->nV * g->nK)), (int)((l / g->nK) % g->nV), (int)(l % g->nK));
    }
    if (g->nV >= 3 && g->nK >= 3 && run_local) {
#pragma omp parallel for schedule(dynamic, 1)
        for (long q = n0; q < n1; q++) {
            const int iR = (int)(q % g->nR);
            if (!g->include_boundary && (iR == 0 || iR == g->nR - 1)) continue;
            local_node(g, psd, (int)(q / g->nR), iR);
        }
    }
    return 0;
}

#endif

#ifdef LU_CACHING
// Local diffusion of an LU_CACHING build over nodes [n0,n1) (n1 = -1: all),
// each with its own cache slice lu_all + (node-n0)*(2kl+ku+1)*n, ipiv_all +
// (node-n0)*n (src/VERB4D_Solver.cpp:340-365, 1100-1210).
int ref_step_local_cached(const ref_grid* g, double* psd, int nthreads, double* lu_all, int64_t* ipiv_all,
                          int recompute, long n0, long n1) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    if (n1 < 0) n1 = (long)g->nP * g->nR;
    const int nV = g->nV, nK = g->nK, n = nV * nK;
    const size_t ldab = 3 * (size_t)(nV + 1) + 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (long q = n0; q < n1; q++) {
        const int iP = (int)(q / g->nR), iR = (int)(q % g->nR);
        if (!g->include_boundary && (iR == 0 || iR == g->nR - 1)) continue;
        const size_t o = idx4(g, iP, iR, 0, 0);
        Matrix2D<double> f = m2(psd + o, nV, nK);
        const double* Vl = g->Vl + ((size_t)iP * g->nR + iR) * nK;
        const double* Vu = g->Vu + ((size_t)iP * g->nR + iR) * nK;
        const double* Kl = g->Kl + ((size_t)iP * g->nR + iR) * nV;
        const double* Ku = g->Ku + ((size_t)iP * g->nR + iR) * nV;
        Diffusion_2D(f, m2(g->Vg + o, nV, nK), m2(g->Kg + o, nV, nK), nV, nK, m1(Vl, nK), m1(Vu, nK), m1(Kl, nV),
                     m1(Ku, nV), bct(g->Vl_t), bct(g->Vu_t), bct(g->Kl_t), bct(g->Ku_t), m2(g->DVV + o, nV, nK),
                     m2(g->DKK + o, nV, nK), m2(g->DVK + o, nV, nK), m2(g->DVK + o, nV, nK),
                     m2(g->Gloc + o, nV, nK), m2(g->Src + o, nV, nK) * g->local_sources, m2(g->Lloc + o, nV, nK),
                     g->dt, g->sub_dt, lu_all + (size_t)(q - n0) * ldab * n,
                     reinterpret_cast<long*>(ipiv_all + (size_t)(q - n0) * n), recompute != 0);
        std::memcpy(psd + o, f.plane_array, sizeof(double) * n);
    }
    return 0;
}
#endif

size_t ref_grid_struct_size(void) { return sizeof(ref_grid); }

}  // extern "C"
