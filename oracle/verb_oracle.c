/*
 * CPU oracle for the VERB-4D operator-split Fokker-Planck step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is a plain-C restatement of the
 * reference arithmetic; it is loaded by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg as the checker. The product path
 * (verb-4d_amd/) never links, loads or calls it.
 *
 * Every function follows the reference expression by expression (same
 * operand order, no FMA contraction: build with -ffp-contract=off) so that
 * the integer bookkeeping and the limiter branch selection agree bit-for-bit.
 *
 * Reference files (paths relative to the PRBEM/VERB-4D source tree):
 *   local diffusion   src/Diffusion_2D.cpp:43-153, src/MatrixSolver.cpp:118-340,439-494
 *   band LU / solve   LAPACK dgbtf2 / dgbtrs (netlib 3.x published algorithm) as
 *                     called from src/MatrixSolver.cpp:272,302,321
 *   radial diffusion  src/Diffusion_1D.cpp:43-184, src/MatrixSolver.cpp:353-404,689-720
 *   convection        src/Convection_2D.cpp:36-169,
 *                     src/Convection_1D_ULTIMATE_QUICKEST6.cpp:22-482
 *   composition       src/VERB4D_Solver.cpp:851-1216
 *
 * Boundary type codes follow src/BoundaryConditionType.hpp:53-88:
 *   0 = Periodic, 1 = ConstantValue, 2 = ConstantDerivative.
 *
 * Status codes: 0 ok; >0 LAPACK-style info (singular pivot column, 1-based);
 * -1 Courant number above one (the reference exits, QUICKEST6.cpp:48-52);
 * -2 unknown boundary type (the reference exits, MatrixSolver.cpp:57-58);
 * -3 unsupported size.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BC_PERIODIC 0
#define BC_VALUE 1
#define BC_DERIV 2

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* Local (E, alpha) diffusion assembly                                       */
/* ------------------------------------------------------------------------ */

/* Diagonal slot for offset (dx, dy): slot = (dy+1)*3 + (dx+1); the matrix
 * offset of that diagonal is dy*nx + dx because energy (x) is the fast index
 * of the unknown vector, in = iy*nx + ix (src/Matrix.cpp:6479-6481, 6447-6456). */
#define SLOT(dx, dy) (((dy) + 1) * 3 + ((dx) + 1))

/* One SecondDerivativeApproximation_2D call (src/MatrixSolver.cpp:439-494).
 * x, y, D, G are [nx][ny] row-major (index ix*ny + iy). */
static void sda2(double* diag, int n, int nx, int ny, int ix, int iy,
                 int dx1, int dy1, int dx2, int dy2,
                 const double* x, const double* y, const double* D, const double* G,
                 double mult) {
    (void)nx;
#define AT(a, i, j) (a)[(size_t)(i) * ny + (j)]
    const int in = iy * nx + ix;
    double dh1 = (AT(x, ix, iy) + AT(y, ix, iy)) - (AT(x, ix + dx1, iy + dy1) + AT(y, ix + dx1, iy + dy1));
    double dh21 = (AT(x, ix, iy) + AT(y, ix, iy)) - (AT(x, ix + dx2, iy + dy2) + AT(y, ix + dx2, iy + dy2));
    double dh22 = (AT(x, ix + dx1, iy + dy1) + AT(y, ix + dx1, iy + dy1)) -
                  (AT(x, ix + dx1 + dx2, iy + dy1 + dy2) + AT(y, ix + dx1 + dx2, iy + dy1 + dy2));
    double common_part = mult / AT(G, ix, iy) / dh1;
    diag[(size_t)SLOT(0, 0) * n + in] += +common_part / dh21 * AT(D, ix, iy) * AT(G, ix, iy);
    diag[(size_t)SLOT(dx2, dy2) * n + in] += -common_part / dh21 * AT(D, ix, iy) * AT(G, ix, iy);
    diag[(size_t)SLOT(dx1, dy1) * n + in] += -common_part / dh22 * AT(D, ix + dx1, iy + dy1) * AT(G, ix + dx1, iy + dy1);
    diag[(size_t)SLOT(dx1 + dx2, dy1 + dy2) * n + in] +=
        +common_part / dh22 * AT(D, ix + dx1, iy + dy1) * AT(G, ix + dx1, iy + dy1);
#undef AT
}

/* Boundary row (src/MatrixSolver.cpp:118-213): x-lower, x-upper, y-lower,
 * y-upper precedence; ConstantValue -> A0=1, A[id1]=0; ConstantDerivative ->
 * A0=1/dh, A[id1]=-1/dh. */
static int boundary_row(double* diag, double* C0, int n, int nx, int ny, int ix, int iy,
                        const double* x, const double* y,
                        const double* xl, const double* xu, const double* yl, const double* yu,
                        int xlt, int xut, int ylt, int yut) {
    const int in = iy * nx + ix;
    int type, s1;
    double dh;
    if (ix == 0 && nx >= 3) {
        C0[in] = xl[iy];
        s1 = SLOT(1, 0);
        dh = x[(size_t)(ix + 1) * ny + iy] - x[(size_t)ix * ny + iy];
        type = xlt;
    } else if (ix == nx - 1 && nx >= 3) {
        C0[in] = xu[iy];
        s1 = SLOT(-1, 0);
        dh = x[(size_t)ix * ny + iy] - x[(size_t)(ix - 1) * ny + iy];
        type = xut;
    } else if (iy == 0 && ny >= 3) {
        C0[in] = yl[ix];
        s1 = SLOT(0, 1);
        dh = y[(size_t)ix * ny + iy + 1] - y[(size_t)ix * ny + iy];
        type = ylt;
    } else if (iy == ny - 1 && ny >= 3) {
        C0[in] = yu[ix];
        s1 = SLOT(0, -1);
        dh = y[(size_t)ix * ny + iy] - y[(size_t)ix * ny + iy - 1];
        type = yut;
    } else {
        return 0;
    }
    if (type == BC_VALUE) {
        diag[(size_t)SLOT(0, 0) * n + in] = 1;
        diag[(size_t)s1 * n + in] = 0;
    } else if (type == BC_DERIV) {
        diag[(size_t)SLOT(0, 0) * n + in] = 1 / dh;
        diag[(size_t)s1 * n + in] = -1 / dh;
    } else {
        return -2;
    }
    return 0;
}

/* Assembly of A (9 diagonals), B0 and C0 for one (P,R) node following
 * src/Diffusion_2D.cpp:49-130. dt is the sub-step dt (already divided).
 * Requires nx >= 3 and ny >= 3 (the only case the time loop calls,
 * src/VERB4D_Solver.cpp:1081). diag is [9][n], n = nx*ny. */
int orc_assemble_2d(int nx, int ny, const double* x, const double* y,
                    const double* xl, const double* xu, const double* yl, const double* yu,
                    int xlt, int xut, int ylt, int yut,
                    const double* Dxx, const double* Dyy, const double* Dxy, const double* Dyx,
                    const double* G, const double* S, const double* L, double dt,
                    double* diag, double* B0, double* C0) {
    if (nx < 3 || ny < 3) return -3;
    const int n = nx * ny;
    memset(diag, 0, sizeof(double) * 9 * (size_t)n);
    memset(B0, 0, sizeof(double) * (size_t)n);
    memset(C0, 0, sizeof(double) * (size_t)n);
    for (int ix = 0; ix < nx; ix++) {
        for (int iy = 0; iy < ny; iy++) {
            const int in = iy * nx + ix;
            if (ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1) {
                int st = boundary_row(diag, C0, n, nx, ny, ix, iy, x, y, xl, xu, yl, yu, xlt, xut, ylt, yut);
                if (st) return st;
            } else {
                const size_t p = (size_t)ix * ny + iy;
                diag[(size_t)SLOT(0, 0) * n + in] += 1.0 / dt;
                B0[in] += 1.0 / dt;
                C0[in] += S[p];
                diag[(size_t)SLOT(0, 0) * n + in] -= L[p];
                sda2(diag, n, nx, ny, ix, iy, -1, 0, +1, 0, x, y, Dxx, G, -0.5);
                sda2(diag, n, nx, ny, ix, iy, +1, 0, -1, 0, x, y, Dxx, G, -0.5);
                sda2(diag, n, nx, ny, ix, iy, 0, -1, 0, +1, x, y, Dyy, G, -0.5);
                sda2(diag, n, nx, ny, ix, iy, 0, +1, 0, -1, x, y, Dyy, G, -0.5);
                sda2(diag, n, nx, ny, ix, iy, 0, -1, +1, 0, x, y, Dxy, G, -0.25);
                sda2(diag, n, nx, ny, ix, iy, 0, +1, -1, 0, x, y, Dxy, G, -0.25);
                sda2(diag, n, nx, ny, ix, iy, 0, -1, -1, 0, x, y, Dxy, G, -0.25);
                sda2(diag, n, nx, ny, ix, iy, 0, +1, +1, 0, x, y, Dxy, G, -0.25);
                sda2(diag, n, nx, ny, ix, iy, -1, 0, 0, +1, x, y, Dyx, G, -0.25);
                sda2(diag, n, nx, ny, ix, iy, +1, 0, 0, -1, x, y, Dyx, G, -0.25);
                sda2(diag, n, nx, ny, ix, iy, -1, 0, 0, -1, x, y, Dyx, G, -0.25);
                sda2(diag, n, nx, ny, ix, iy, +1, 0, 0, +1, x, y, Dyx, G, -0.25);
            }
        }
    }
    return 0;
}

/* Pack the 9 diagonals into LAPACK band storage (src/MatrixSolver.cpp:283-299):
 * ab[col*ldab + kl+ku+row-col] = A(row,col), kl = ku = nx+1, ldab = 2kl+ku+1. */
void orc_pack_band(int nx, int ny, const double* diag, double* ab) {
    const int n = nx * ny, kl = nx + 1, ku = nx + 1, ldab = 2 * kl + ku + 1;
    memset(ab, 0, sizeof(double) * (size_t)ldab * n);
    for (int in = 0; in < n; in++) {
        /* The reference iterates its std::map of diagonals in increasing offset
         * order; every offset lands in a distinct band slot, so order is immaterial. */
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int off = dy * nx + dx;
                int col = in + off;
                if (col >= 0 && col < n)
                    ab[(size_t)col * ldab + (kl + ku - off)] = diag[(size_t)SLOT(dx, dy) * n + in];
            }
    }
}

/* ------------------------------------------------------------------------ */
/* LAPACK band LU with partial pivoting: restatement of netlib dgbtf2 (the   */
/* published unblocked algorithm; dgbtrf's blocked form selects the same    */
/* pivots in exact arithmetic). Column-major band, 0-based indices here,    */
/* ipiv returned 1-based like LAPACK.                                        */
/* ------------------------------------------------------------------------ */
int orc_dgbtf2(int m, int n, int kl, int ku, double* ab, int ldab, int64_t* ipiv) {
    const int kv = ku + kl;
    int info = 0;
#define AB(i, j) ab[(size_t)(j) * ldab + (i)]
    /* zero the fill-in rows of columns ku+1 .. min(kv, n)-1 */
    for (int j = ku + 1; j <= imin(kv, n) - 1; j++)
        for (int i = kv - j; i < kl; i++) AB(i, j) = 0.0;
    int ju = 0;
    for (int j = 0; j < imin(m, n); j++) {
        if (j + kv < n)
            for (int i = 0; i < kl; i++) AB(i, j + kv) = 0.0;
        const int km = imin(kl, m - 1 - j);
        /* idamax: first index of the largest |value| */
        int jp = 0;
        double best = fabs(AB(kv, j));
        for (int i = 1; i <= km; i++) {
            double v = fabs(AB(kv + i, j));
            if (v > best) { best = v; jp = i; }
        }
        ipiv[j] = (int64_t)(jp + j + 1);
        if (AB(kv + jp, j) != 0.0) {
            ju = imax(ju, imin(j + ku + jp, n - 1));
            if (jp != 0) {
                /* swap rows j and j+jp across columns j..ju */
                for (int c = j; c <= ju; c++) {
                    double* a = &AB(kv + jp + j - c, c);
                    double* b = &AB(kv + j - c, c);
                    double t = *a; *a = *b; *b = t;
                }
            }
            if (km > 0) {
                const double r = 1.0 / AB(kv, j);
                for (int i = 1; i <= km; i++) AB(kv + i, j) *= r;
                /* dger: A(j+1+i, c) += L(i) * (-U(j, c)) for c in j+1..ju */
                for (int c = j + 1; c <= ju; c++) {
                    const double temp = -AB(kv + j - c, c);
                    if (temp != 0.0)
                        for (int i = 1; i <= km; i++) AB(kv + j + i - c, c) += AB(kv + i, j) * temp;
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
    }
#undef AB
    return info;
}

/* dgbtrs('N') restatement: forward elimination with the row interchanges, then
 * dtbsv upper / no-transpose / non-unit (column-oriented back substitution). */
void orc_dgbtrs(int n, int kl, int ku, const double* ab, int ldab, const int64_t* ipiv, double* b) {
    const int kd = ku + kl;
#define AB(i, j) ab[(size_t)(j) * ldab + (i)]
    if (kl > 0) {
        for (int j = 0; j < n - 1; j++) {
            const int lm = imin(kl, n - 1 - j);
            const int l = (int)ipiv[j] - 1;
            if (l != j) { double t = b[l]; b[l] = b[j]; b[j] = t; }
            const double temp = -b[j];
            if (temp != 0.0)
                for (int i = 1; i <= lm; i++) b[j + i] += AB(kd + i, j) * temp;
        }
    }
    const int k = kl + ku;
    for (int j = n - 1; j >= 0; j--) {
        if (b[j] != 0.0) {
            b[j] = b[j] / AB(k, j);
            const double temp = b[j];
            for (int i = j - 1; i >= imax(0, j - k); i--) b[i] = b[i] - temp * AB(k + i - j, j);
        }
    }
#undef AB
}

/* Lapack() of src/MatrixSolver.cpp:223-340. psd is [nx][ny]; the unknown
 * vector is energy-fast (in = iy*nx + ix). When lu != NULL the factors are
 * read from / written to the caller's cache (LU_CACHING build): recompute
 * selects whether to factor first. */
int orc_lapack(int nx, int ny, const double* diag, const double* B0, const double* C0,
               double* psd, int num_substeps, double* lu, int64_t* ipiv, int recompute) {
    const int n = nx * ny, kl = nx + 1, ku = nx + 1, ldab = 2 * kl + ku + 1;
    double* ab = lu;
    int64_t* piv = ipiv;
    int own = 0;
    if (!ab) {
        ab = (double*)malloc(sizeof(double) * (size_t)ldab * n);
        piv = (int64_t*)malloc(sizeof(int64_t) * n);
        own = 1;
        recompute = 1;
    }
    if (recompute) {
        orc_pack_band(nx, ny, diag, ab);
        int info = orc_dgbtf2(n, n, kl, ku, ab, ldab, piv);
        if (info != 0) {
            if (own) { free(ab); free(piv); }
            return info;
        }
    }
    double* rhs = (double*)malloc(sizeof(double) * n);
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) rhs[iy * nx + ix] = psd[(size_t)ix * ny + iy];
    for (int s = 0; s < num_substeps; s++) {
        for (int i = 0; i < n; i++) rhs[i] *= B0[i];
        for (int i = 0; i < n; i++) rhs[i] += C0[i];
        orc_dgbtrs(n, kl, ku, ab, ldab, piv, rhs);
    }
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) psd[(size_t)ix * ny + iy] = rhs[iy * nx + ix];
    free(rhs);
    if (own) { free(ab); free(piv); }
    return 0;
}

/* Diffusion_2D (src/Diffusion_2D.cpp:19-153). All 2-D arrays are [nx][ny]. */
int orc_diffusion_2d(double* psd, const double* x, const double* y, int nx, int ny,
                     const double* xl, const double* xu, const double* yl, const double* yu,
                     int xlt, int xut, int ylt, int yut,
                     const double* Dxx, const double* Dyy, const double* Dxy, const double* Dyx,
                     const double* G, const double* S, const double* L, double dt, double sub_dt,
                     double* lu_cache, int64_t* ipiv_cache, int recompute) {
    if (nx < 3 || ny < 3) return -3;
    const int num_substeps = imax(1, (int)(dt / sub_dt));
    dt /= num_substeps;
    const int n = nx * ny;
    double* diag = (double*)malloc(sizeof(double) * 9 * (size_t)n);
    double* B0 = (double*)malloc(sizeof(double) * n);
    double* C0 = (double*)malloc(sizeof(double) * n);
    /* the reference assembles A, B and C on every call, cached or not */
    int st = orc_assemble_2d(nx, ny, x, y, xl, xu, yl, yu, xlt, xut, ylt, yut, Dxx, Dyy, Dxy, Dyx, G, S, L, dt,
                             diag, B0, C0);
    if (st == 0) st = orc_lapack(nx, ny, diag, B0, C0, psd, num_substeps, lu_cache, ipiv_cache, recompute);
    free(diag);
    free(B0);
    free(C0);
    return st;
}

/* ------------------------------------------------------------------------ */
/* Radial diffusion                                                          */
/* ------------------------------------------------------------------------ */

/* tridag (src/MatrixSolver.cpp:689-720): Thomas algorithm, no pivoting. */
void orc_tridag(const double* a, const double* b, const double* c, const double* r, double* u, int n) {
    double* gam = (double*)malloc(sizeof(double) * n);
    double bet;
    u[0] = r[0] / (bet = b[0]);
    for (int j = 1; j <= n - 1; j++) {
        gam[j] = c[j - 1] / bet;
        bet = b[j] - a[j] * gam[j];
        u[j] = (r[j] - a[j] * u[j - 1]) / bet;
    }
    for (int j = n - 1; j >= 1; j--) u[j - 1] -= gam[j] * u[j];
    free(gam);
}

/* SecondDerivativeApproximation_1D (src/MatrixSolver.cpp:353-404); a, b, c
 * are the -1, 0, +1 diagonals. */
static void sda1(double* a, double* b, double* c, int ix, int dx1, int dx2,
                 const double* x, const double* D, const double* G, double mult) {
    double* dg[3] = {a, b, c};
    double dh1 = (x[ix]) - (x[ix + dx1]);
    double dh21 = (x[ix]) - (x[ix + dx2]);
    double dh22 = (x[ix + dx1]) - (x[ix + dx1 + dx2]);
    double common_part = mult / G[ix] / dh1;
    dg[1][ix] += +common_part / dh21 * D[ix] * G[ix];
    dg[1 + dx2][ix] += -common_part / dh21 * D[ix] * G[ix];
    dg[1 + dx1][ix] += -common_part / dh22 * D[ix + dx1] * G[ix + dx1];
    dg[1 + dx1 + dx2][ix] += +common_part / dh22 * D[ix + dx1] * G[ix + dx1];
}

/* Diffusion_1D + getCalculationMatrix1DDiffusion (src/Diffusion_1D.cpp:43-184). */
int orc_diffusion_1d(double* psd, const double* x, int n, double lbc, double ubc, int lt, int ut,
                     const double* D, const double* G, const double* S, const double* L, double dt) {
    double* a = (double*)calloc(n, sizeof(double));
    double* b = (double*)calloc(n, sizeof(double));
    double* c = (double*)calloc(n, sizeof(double));
    double* B0 = (double*)calloc(n, sizeof(double));
    double* C0 = (double*)calloc(n, sizeof(double));
    double* rhs = (double*)calloc(n, sizeof(double));
    int st = 0;
    for (int ix = 0; ix < n; ix++) {
        if (ix == 0 && n >= 3) {
            C0[ix] = lbc;
            double dh = x[ix + 1] - x[ix];
            if (lt == BC_VALUE) { b[ix] = 1; c[ix] = 0; }
            else if (lt == BC_DERIV) { b[ix] = 1 / dh; c[ix] = -1 / dh; }
            else { st = -2; }
        } else if (ix == n - 1 && n >= 3) {
            C0[ix] = ubc;
            double dh = x[ix] - x[ix - 1];
            if (ut == BC_VALUE) { b[ix] = 1; a[ix] = 0; }
            else if (ut == BC_DERIV) { b[ix] = 1 / dh; a[ix] = -1 / dh; }
            else { st = -2; }
        } else {
            b[ix] += 1.0 / dt;
            B0[ix] += 1.0 / dt;
            C0[ix] += S[ix];
            b[ix] -= L[ix];
            if (n >= 3) {
                sda1(a, b, c, ix, -1, +1, x, D, G, -0.5);
                sda1(a, b, c, ix, +1, -1, x, D, G, -0.5);
            }
        }
    }
    if (st == 0) {
        for (int ix = 0; ix < n; ix++) {
            rhs[ix] = C0[ix];
            rhs[ix] += B0[ix] * psd[ix];
        }
        orc_tridag(a, b, c, rhs, psd, n);
    }
    free(a); free(b); free(c); free(B0); free(C0); free(rhs);
    return st;
}

/* ------------------------------------------------------------------------ */
/* ULTIMATE-QUICKEST advection (src/Convection_1D_ULTIMATE_QUICKEST6.cpp)   */
/* ------------------------------------------------------------------------ */
#define GST 5

static int fill_ghosts(double* t, const double* src_lo, const double* src_hi, int n,
                       double lbc, double ubc, int lt, int ut, const double* x) {
    /* t holds n+2*GST values; t[GST+i] is point i. src_lo / src_hi are the
     * arrays the lower / upper ghosts are taken from (they differ in the
     * limiter pass, src/Convection_1D_ULTIMATE_QUICKEST6.cpp:283-330). */
    double dx;
    switch (lt) {
        case BC_PERIODIC:
            for (int ig = 1; ig <= GST; ig++) t[GST - ig] = src_lo[(n - 1) - ig];
            break;
        case BC_VALUE:
            for (int ig = 1; ig <= GST; ig++) t[GST - ig] = lbc;
            break;
        case BC_DERIV:
            dx = (x[1] - x[0]);
            t[GST - 1] = src_lo[0] - lbc * dx;
            for (int ig = 1; ig <= GST; ig++) t[GST - ig] = t[GST - ig + 1] - lbc * dx;
            break;
        default:
            return -2;
    }
    switch (ut) {
        case BC_PERIODIC:
            for (int ig = 1; ig <= GST; ig++) t[GST + (n - 1) + ig] = src_hi[ig];
            break;
        case BC_VALUE:
            for (int ig = 1; ig <= GST; ig++) t[GST + (n - 1) + ig] = ubc;
            break;
        case BC_DERIV:
            dx = (x[n - 1] - x[n - 2]);
            t[GST + (n - 1) + 1] = src_hi[n - 1] + ubc * dx;
            for (int ig = 1; ig <= GST; ig++) t[GST + (n - 1) + ig] = t[GST + (n - 1) + ig - 1] + ubc * dx;
            break;
        default:
            return -2;
    }
    return 0;
}

int orc_quickest6(double* PSD, const double* x, int n, double lbc, double ubc, int lt, int ut,
                  const double* Ux, double dt_total) {
    if (n < 2) return -3;
    double dx = (x[1] - x[0]);
    double ux_max = 0;
    for (int i = 0; i < n; i++) {
        double cur = fabs(Ux[i]);
        ux_max = ux_max > cur ? ux_max : cur; /* std::max(tmp, current) keeps tmp on ties/NaN */
    }
    double steps_double = (double)dt_total * ux_max / dx / 1.0;
    int num_steps = (steps_double <= 1) ? 1 : (int)ceil(steps_double);
    double dt = dt_total / num_steps;
    double* Cn = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) Cn[i] = Ux[i] / dx * dt;
    const double courant_max = fabs(ux_max / dx * dt);
    if (courant_max > 1) { free(Cn); return -1; }
    const double f3 = 6.0, f5 = 120.0, f7 = 5040.0, f9 = 362880.0;
    double* PSD_t = (double*)calloc(n + 2 * GST, sizeof(double));
    double* Unl = (double*)calloc(n, sizeof(double));
    double* Unl_t = (double*)calloc(n + 2 * GST, sizeof(double));
    double* Faces = (double*)calloc(n, sizeof(double));
    int st = 0;
    for (int it = 0; it < num_steps && st == 0; it++) {
        for (int ix = 0; ix <= n - 1; ix++) PSD_t[GST + ix] = PSD[ix];
        st = fill_ghosts(PSD_t, PSD, PSD, n, lbc, ubc, lt, ut, x);
        if (st) break;
        double psd_face_left = 0.0, psd_face_right = 0.0, courant_left = 0.0, courant_right = 0.0;
        for (int ix = 0; ix <= n - 1; ix++) {
            if (ix == 0) {
                if (lt == BC_PERIODIC) courant_right = (Cn[n - 2] + Cn[0]) / 2;
                else continue;
            } else {
                courant_right = (Cn[ix - 1] + Cn[ix]) / 2;
            }
            const double cs = courant_right * courant_right;
            const double ca = fabs(courant_right);
            double U[5], D[6], C;
            if (courant_right >= 0) {
                for (int ig = 1; ig <= 4; ig++) U[ig] = PSD_t[GST + ix - 1 - ig];
                C = PSD_t[GST + ix - 1];
                for (int ig = 1; ig <= 5; ig++) D[ig] = PSD_t[GST + ix - 1 + ig];
            } else {
                for (int ig = 1; ig <= 5; ig++) D[ig] = PSD_t[GST + ix - ig];
                C = PSD_t[GST + ix];
                for (int ig = 1; ig <= 4; ig++) U[ig] = PSD_t[GST + ix + ig];
            }
            psd_face_right = 0.5 * (C + D[1]);
            const double d1 = C - D[1];
            psd_face_right += 0.5 * ca * d1;
            const double D2 = (U[1] - C - D[1] + D[2]) / 2;
            const double d3 = U[1] - 3 * C + 3 * D[1] - D[2];
            psd_face_right += (cs - 1) / f3 * (D2 + ca / 4 * d3);
            const double D4 = (U[2] - 3 * U[1] + 2 * C + 2 * D[1] - 3 * D[2] + D[3]) / 2;
            const double d5 = U[2] - 5 * U[1] + 10 * C - 10 * D[1] + 5 * D[2] - D[3];
            psd_face_right += (cs - 1) * (cs - 4) / f5 * (D4 + ca / 6 * d5);
            const double D6 = (U[3] - 5 * U[2] + 9 * U[1] - 5 * C - 5 * D[1] + 9 * D[2] - 5 * D[3] + D[4]) / 2;
            const double d7 = U[3] - 7 * U[2] + 21 * U[1] - 35 * C + 35 * D[1] - 21 * D[2] + 7 * D[3] - D[4];
            psd_face_right += (cs - 1) * (cs - 4) * (cs - 9) / f7 * (D6 + ca / 8 * d7);
            const double D8 = (U[4] - 7 * U[3] + 20 * U[2] - 28 * U[1] + 14 * C + 14 * D[1] - 28 * D[2] +
                               20 * D[3] - 7 * D[4] + D[5]) / 2;
            const double d9 = U[4] - 9 * U[3] + 36 * U[2] - 84 * U[1] + 126 * C - 126 * D[1] + 84 * D[2] -
                              36 * D[3] + 9 * D[4] - D[5];
            psd_face_right += (cs - 1) * (cs - 4) * (cs - 9) * (cs - 16) / f9 * (D8 + d9 / 2);
            if (ix > 0) Unl[ix - 1] = PSD[ix - 1] - (courant_right * psd_face_right - courant_left * psd_face_left);
            Faces[ix] = psd_face_right;
            psd_face_left = psd_face_right;
            courant_left = courant_right;
        }
        if (lt == BC_PERIODIC) Unl[n - 1] = Unl[0];
        else if (lt == BC_VALUE) Unl[0] = lbc;
        else if (lt == BC_DERIV) Unl[0] = Unl[1];
        if (ut == BC_VALUE) Unl[n - 1] = ubc;
        else if (ut == BC_DERIV) Unl[n - 1] = Unl[n - 2];

        for (int ix = 0; ix <= n - 1; ix++) Unl_t[GST + ix] = Unl[ix];
        /* lower ghosts from the pre-step PSD, upper ghosts from the unlimited
         * solution (src/Convection_1D_ULTIMATE_QUICKEST6.cpp:283-330) */
        st = fill_ghosts(Unl_t, PSD, Unl, n, lbc, ubc, lt, ut, x);
        if (st) break;
        psd_face_left = 0.0; courant_left = 0.0; courant_right = 0.0;
        for (int ix = 0; ix <= n - 1; ix++) {
            if (ix == 0) {
                if (lt == BC_PERIODIC) courant_right = (Cn[n - 2] + Cn[0]) / 2;
                else continue;
            } else {
                courant_right = (Cn[ix - 1] + Cn[ix]) / 2;
            }
            const double ca = fabs(courant_right);
            double ou, od, oc, UU[5], UD[6], UC;
            if (courant_right >= 0) {
                for (int ig = 1; ig <= 4; ig++) UU[ig] = Unl_t[GST + ix - 1 - ig];
                ou = PSD_t[GST + ix - 2];
                UC = Unl_t[GST + ix - 1];
                oc = PSD_t[GST + ix - 1];
                for (int ig = 1; ig <= 5; ig++) UD[ig] = Unl_t[GST + ix - 1 + ig];
                od = PSD_t[GST + ix];
            } else {
                for (int ig = 1; ig <= 5; ig++) UD[ig] = Unl_t[GST + ix - ig];
                od = PSD_t[GST + ix - 1];
                UC = Unl_t[GST + ix];
                oc = PSD_t[GST + ix];
                for (int ig = 1; ig <= 4; ig++) UU[ig] = Unl_t[GST + ix + ig];
                ou = PSD_t[GST + ix + 1];
            }
            const double DEL = od - ou;
            const double CURV = od - 2 * oc + ou;
            const double ADEL = fabs(DEL);
            const double ACURV = fabs(CURV);
            const double dif1 = UD[2] - UD[3];
            const double dif2 = UD[1] - UD[2];
            const double dif3 = UC - UD[1];
            const double dif4 = UU[1] - UC;
            const double dif5 = UU[2] - UU[1];
            const double dif6 = UU[3] - UU[2];
            const double th = 0;
            const int peak_d = dif1 > th && dif2 > th && dif3 < -th && dif4 < -th && dif2 < dif1 - th && dif3 > dif4 + th;
            const int peak_c = dif2 > th && dif3 > th && dif4 < -th && dif5 < -th && dif3 < dif2 - th && dif4 > dif5 + th;
            const int peak_u = dif3 > th && dif4 > th && dif5 < -th && dif6 < -th && dif4 < dif3 - th && dif5 > dif6 + th;
            double fr = Faces[ix];
            if (!(peak_d || peak_c || peak_u)) {
                if (ACURV >= ADEL) {
                    fr = oc;
                } else if (ca > 1e-31) {
                    const double PSD_REF = ou + (oc - ou) / ca;
                    if (DEL > 0) {
                        fr = fmax(fr, oc);
                        fr = fmin(fr, fmin(PSD_REF, od));
                    } else if (DEL < 0) {
                        fr = fmin(fr, oc);
                        fr = fmax(fr, fmax(PSD_REF, od));
                    }
                }
            }
            if (ix > 0) PSD[ix - 1] -= (courant_right * fr - courant_left * psd_face_left);
            psd_face_left = fr;
            courant_left = courant_right;
        }
        if (lt == BC_PERIODIC) PSD[n - 1] = PSD[0];
        else if (lt == BC_VALUE) PSD[0] = lbc;
        else if (lt == BC_DERIV) PSD[0] = PSD[1];
        if (ut == BC_VALUE) PSD[n - 1] = ubc;
        else if (ut == BC_DERIV) PSD[n - 1] = PSD[n - 2];
    }
    free(Cn); free(PSD_t); free(Unl); free(Unl_t); free(Faces);
    return st;
}

static double maxabs(const double* a, int n) {
    double t = 0;
    for (int i = 0; i < n; i++) {
        double cur = fabs(a[i]);
        t = (t < cur) ? cur : t; /* std::max(t, cur) */
    }
    return t;
}

/* Convection_2D (src/Convection_2D.cpp:22-169). 2-D arrays are [nP][nR].
 * lost (nullable): the SAVE_PSD_LOST_CONV accumulator (:150-155); fast:
 * FAST_CONVECTION's min of the two sub-step counts (:80-84). */
int orc_convection_2d_ex(double* psd, double* lost, const double* Pg, const double* Rg, int nP, int nR,
                         const double* Plbc, const double* Pubc, const double* Rlbc, const double* Rubc,
                         int plt, int put, int rlt, int rut,
                         const double* VP, const double* VR, const double* S, const double* L, const double* Gc,
                         double dt_total, int fast) {
    const double maxCour = 0.25;
    double nsP, nsR;
    const int np = nP * nR;
    if (nP >= 3) {
        double dP = (Pg[1 * nR + 0] - Pg[0]);
        double nd = (double)dt_total * maxabs(VP, np) / dP / maxCour;
        nsP = nd <= 1 ? 1 : ceil(nd);
    } else nsP = 1;
    if (nR >= 3) {
        double dR = (Rg[1] - Rg[0]);
        double nd = (double)dt_total * maxabs(VR, np) / dR / maxCour;
        nsR = nd <= 1 ? 1 : ceil(nd);
    } else nsR = 1;
    double num_steps = fast ? (nsR < nsP ? nsR : nsP)  /* std::min */
                            : (nsP < nsR ? nsR : nsP); /* std::max */
    double dt = dt_total / num_steps;
    double* lexp = (double*)malloc(sizeof(double) * np);
    for (int i = 0; i < np; i++) lexp[i] = exp(L[i] * dt);
    double* line = (double*)malloc(sizeof(double) * (nP > nR ? nP : nR));
    double* gx = (double*)malloc(sizeof(double) * (nP > nR ? nP : nR));
    double* gv = (double*)malloc(sizeof(double) * (nP > nR ? nP : nR));
    double* gG = (double*)malloc(sizeof(double) * nR);
    double* PG = (double*)malloc(sizeof(double) * np);
    int st = 0;
    for (int it = 0; it < num_steps && st == 0; it++) {
        if (nP >= 3) {
            for (int iR = 0; iR < nR - 1 && st == 0; iR++) {
                for (int iP = 0; iP < nP; iP++) {
                    line[iP] = psd[iP * nR + iR];
                    gx[iP] = Pg[iP * nR + iR];
                    gv[iP] = VP[iP * nR + iR];
                }
                st = orc_quickest6(line, gx, nP, Plbc[iR], Pubc[iR], plt, put, gv, dt);
                for (int iP = 0; iP < nP; iP++) psd[iP * nR + iR] = line[iP];
            }
        }
        if (st) break;
        for (int i = 0; i < np; i++) PG[i] = psd[i] * Gc[i];
        if (nR >= 3) {
            for (int iP = 0; iP < nP && st == 0; iP++) {
                for (int iR = 0; iR < nR; iR++) {
                    line[iR] = PG[iP * nR + iR];
                    gx[iR] = Rg[iP * nR + iR];
                    gv[iR] = VR[iP * nR + iR];
                    gG[iR] = Gc[iP * nR + iR];
                }
                st = orc_quickest6(line, gx, nR, Rlbc[iP] * gG[0], Rubc[iP] * gG[nR - 1], rlt, rut, gv, dt);
                for (int iR = 0; iR < nR; iR++) psd[iP * nR + iR] = line[iR] / gG[iR];
            }
        }
        if (st) break;
        for (int i = 0; i < np; i++) {
            const double before = psd[i];
            psd[i] = psd[i] * lexp[i];
            if (lost) lost[i] += (before - psd[i]);
        }
        for (int iP = 0; iP < nP; iP++)
            for (int iR = 1; iR < nR - 1; iR++) psd[iP * nR + iR] = psd[iP * nR + iR] + S[iP * nR + iR] * dt;
    }
    free(lexp); free(line); free(gx); free(gv); free(gG); free(PG);
    return st;
}

int orc_convection_2d(double* psd, const double* Pg, const double* Rg, int nP, int nR,
                      const double* Plbc, const double* Pubc, const double* Rlbc, const double* Rubc,
                      int plt, int put, int rlt, int rut,
                      const double* VP, const double* VR, const double* S, const double* L, const double* Gc,
                      double dt_total) {
    return orc_convection_2d_ex(psd, 0, Pg, Rg, nP, nR, Plbc, Pubc, Rlbc, Rubc, plt, put, rlt, rut, VP, VR, S, L,
                                Gc, dt_total, 0);
}

/* ------------------------------------------------------------------------ */
/* Full operator-split step over a [P][R][V][K] grid                        */
/* (src/VERB4D_Solver.cpp:851-1216, LU caching off, FAST_CONVECTION off).   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nP, nR, nV, nK;
    /* 4-D fields [P][R][V][K] */
    const double *Pg, *Rg, *Vg, *Kg, *Lg;
    const double *DLL, *DVV, *DKK, *DVK, *VP, *VR;
    const double *Gloc, *Grad, *Gconv, *Src, *Lloc, *Lrad, *Lconv;
    /* boundary arrays */
    const double *Pl, *Pu;   /* [R][V][K] */
    const double *Rl, *Ru;   /* [P][V][K] (convection) */
    const double *Ll, *Lu;   /* [P][V][K] (radial diffusion) */
    const double *Vl, *Vu;   /* [P][R][K] */
    const double *Kl, *Ku;   /* [P][R][V] */
    int Pl_t, Pu_t, Rl_t, Ru_t, Ll_t, Lu_t, Vl_t, Vu_t, Kl_t, Ku_t;
    double dt, sub_dt;
    double radial_sources, local_sources;
    int include_boundary;
} orc_grid;

#define IDX4(g, p, r, v, k) ((((size_t)(p) * (g)->nR + (r)) * (g)->nV + (v)) * (g)->nK + (k))

int orc_convection_plane(const orc_grid* g, double* psd, int iV, int iK) {
    const int nP = g->nP, nR = g->nR, np = nP * nR;
    double* buf = (double*)malloc(sizeof(double) * (size_t)np * 9);
    double *f = buf, *Pg = buf + np, *Rg = buf + 2 * np, *VP = buf + 3 * np, *VR = buf + 4 * np,
           *S = buf + 5 * np, *L = buf + 6 * np, *G = buf + 7 * np;
    double* bc = buf + 8 * np; /* Pl[R], Pu[R], Rl[P], Ru[P] fit: 2R + 2P <= np for P,R >= 4 */
    double *Pl = (double*)malloc(sizeof(double) * nR), *Pu = (double*)malloc(sizeof(double) * nR);
    double *Rl = (double*)malloc(sizeof(double) * nP), *Ru = (double*)malloc(sizeof(double) * nP);
    (void)bc;
    for (int p = 0; p < nP; p++)
        for (int r = 0; r < nR; r++) {
            size_t i = IDX4(g, p, r, iV, iK);
            int j = p * nR + r;
            f[j] = psd[i]; Pg[j] = g->Pg[i]; Rg[j] = g->Rg[i]; VP[j] = g->VP[i]; VR[j] = g->VR[i];
            S[j] = g->Src[i]; L[j] = g->Lconv[i]; G[j] = g->Gconv[i];
        }
    for (int r = 0; r < nR; r++) {
        Pl[r] = g->Pl[((size_t)r * g->nV + iV) * g->nK + iK];
        Pu[r] = g->Pu[((size_t)r * g->nV + iV) * g->nK + iK];
    }
    for (int p = 0; p < nP; p++) {
        Rl[p] = g->Rl[((size_t)p * g->nV + iV) * g->nK + iK];
        Ru[p] = g->Ru[((size_t)p * g->nV + iV) * g->nK + iK];
    }
    int st = orc_convection_2d(f, Pg, Rg, nP, nR, Pl, Pu, Rl, Ru, g->Pl_t, g->Pu_t, g->Rl_t, g->Ru_t, VP, VR, S, L,
                               G, g->dt);
    for (int p = 0; p < nP; p++)
        for (int r = 0; r < nR; r++) psd[IDX4(g, p, r, iV, iK)] = f[p * nR + r];
    free(buf); free(Pl); free(Pu); free(Rl); free(Ru);
    return st;
}

int orc_radial_line(const orc_grid* g, double* psd, int iP, int iV, int iK) {
    const int nR = g->nR;
    double* b = (double*)malloc(sizeof(double) * nR * 6);
    double *f = b, *x = b + nR, *D = b + 2 * nR, *S = b + 3 * nR, *G = b + 4 * nR, *L = b + 5 * nR;
    for (int r = 0; r < nR; r++) {
        size_t i = IDX4(g, iP, r, iV, iK);
        f[r] = psd[i]; x[r] = g->Lg[i]; D[r] = g->DLL[i]; S[r] = g->Src[i] * g->radial_sources;
        G[r] = g->Grad[i]; L[r] = g->Lrad[i];
    }
    size_t bi = ((size_t)iP * g->nV + iV) * g->nK + iK;
    int st = orc_diffusion_1d(f, x, nR, g->Ll[bi], g->Lu[bi], g->Ll_t, g->Lu_t, D, G, S, L, g->dt);
    for (int r = 0; r < nR; r++) psd[IDX4(g, iP, r, iV, iK)] = f[r];
    free(b);
    return st;
}

int orc_local_node(const orc_grid* g, double* psd, int iP, int iR) {
    const int nV = g->nV, nK = g->nK, n = nV * nK;
    const size_t o = IDX4(g, iP, iR, 0, 0);
    double* S = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) S[i] = g->Src[o + i] * g->local_sources;
    const double* Vl = g->Vl + ((size_t)iP * g->nR + iR) * nK;
    const double* Vu = g->Vu + ((size_t)iP * g->nR + iR) * nK;
    const double* Kl = g->Kl + ((size_t)iP * g->nR + iR) * nV;
    const double* Ku = g->Ku + ((size_t)iP * g->nR + iR) * nV;
    int st = orc_diffusion_2d(psd + o, g->Vg + o, g->Kg + o, nV, nK, Vl, Vu, Kl, Ku, g->Vl_t, g->Vu_t, g->Kl_t,
                              g->Ku_t, g->DVV + o, g->DKK + o, g->DVK + o, g->DVK + o, g->Gloc + o, S,
                              g->Lloc + o, g->dt, g->sub_dt, NULL, NULL, 1);
    free(S);
    return st;
}

/* One operator-split step. nthreads <= 0 leaves the OpenMP default. The
 * three loops mirror the reference's OpenMP schedule(dynamic,1) regions.
 * Returns the first nonzero status encountered (0 if none). */
int orc_step(const orc_grid* g, double* psd, int nthreads, int run_convection, int run_radial, int run_local) {
    int status = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if ((3 < g->nP || 3 < g->nR) && run_convection) {
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
        for (int iV = g->nV - 1; iV >= 0; iV--)
            for (int iK = 0; iK < g->nK; iK++) {
                int st = orc_convection_plane(g, psd, iV, iK);
                if (st) {
#pragma omp critical
                    if (!status) status = st;
                }
            }
    }
    if (status) return status;
    if (g->nR >= 3 && run_radial) {
#pragma omp parallel for schedule(dynamic, 1) collapse(3)
        for (int iP = 0; iP < g->nP; iP++)
            for (int iV = 0; iV < g->nV; iV++)
                for (int iK = 0; iK < g->nK; iK++) {
                    int st = orc_radial_line(g, psd, iP, iV, iK);
                    if (st) {
#pragma omp critical
                        if (!status) status = st;
                    }
                }
    }
    if (status) return status;
    if (g->nV >= 3 && g->nK >= 3 && run_local) {
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
        for (int iP = 0; iP < g->nP; iP++)
            for (int iR = 0; iR < g->nR; iR++) {
                if (!g->include_boundary && (iR == 0 || iR == g->nR - 1)) continue;
                int st = orc_local_node(g, psd, iP, iR);
                if (st) {
#pragma omp critical
                    if (!status) status = st;
                }
            }
    }
    return status;
}

size_t orc_grid_struct_size(void) { return sizeof(orc_grid); }

/* ------------------------------------------------------------------------
 * Adiabatic remapping: Cubic1D (src/Interpolation.cpp:25-51) over
 * Opm::MonotCubicInterpolator (src/MonotCubicInterpolator.cpp), and the
 * loop over (P,V,K) lines of src/VERB4D_Solver.cpp:574-606.
 *
 * Cubic1D builds its x and f vectors with n leading zero entries (sized n,
 * then push_back, :33-40), so the interpolator holds the point (0, 0)
 * followed by the old grid. The constructor (:92-106) stores the points in
 * a std::map: keys sorted and unique, a repeated key takes the value
 * inserted last. computeInternalFunctionData (:378-497) decides
 * monotonicity, computeSimpleDerivatives (:612-664) takes one-sided slopes
 * at the ends and the mean of the two half-slopes inside,
 * adjustDerivativesForMonotoneness (:668-704) applies the Fritsch-Carlson
 * limit interval by interval (a derivative changed for interval s is the
 * left one of interval s+1), and evaluate (:175-233) extrapolates
 * constants and evaluates the cubic Hermite form with the bases of
 * src/MonotCubicInterpolator.hpp:522-537.
 * Status: 0; -8 where the reference throws (a non-finite new-grid value,
 * NaN in the pair that fixes the direction) or where std::map's ordering is
 * undefined (a NaN old-grid value).
 * ---------------------------------------------------------------------- */
#define ORC_ERR_DOMAIN (-8)

static double h00(double t) { return 2 * t * t * t - 3 * t * t + 1; }
static double h10(double t) { return t * t * t - 2 * t * t + t; }
static double h01(double t) { return -2 * t * t * t + 3 * t * t; }
static double h11(double t) { return t * t * t - t * t; }

/* values (length n) is replaced by the interpolant at x_new. */
int orc_cubic1d(int n, const double* x_old, double* values, const double* x_new) {
    if (n <= 0) return 0;
    for (int i = 0; i < n; i++)
        if (isnan(x_old[i]) || !isfinite(x_new[i])) return ORC_ERR_DOMAIN;
    double* k = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    double* v = k + n + 1;
    double* d = v + n + 1;
    int m = 1;
    k[0] = 0.0;
    v[0] = 0.0;
    for (int i = 0; i < n; i++) {
        const double x = x_old[i], f = values[i];
        int pos = m;
        while (pos > 0 && x < k[pos - 1]) pos--;
        if (pos > 0 && k[pos - 1] == x) {
            v[pos - 1] = f;
        } else {
            for (int s = m; s > pos; s--) {
                k[s] = k[s - 1];
                v[s] = v[s - 1];
            }
            k[pos] = x;
            v[pos] = f;
            m++;
        }
    }
    if (m >= 2) {
        int mono = 1, s = 0;
        while (s + 1 < m && v[s] == v[s + 1]) s++;
        if (s + 1 < m) {
            if (v[s] > v[s + 1]) {
                for (s++; s + 1 < m; s++)
                    if (v[s] < v[s + 1]) { mono = 0; break; }
            } else if (v[s] < v[s + 1]) {
                for (s++; s + 1 < m; s++)
                    if (v[s] > v[s + 1]) { mono = 0; break; }
            } else {
                free(k);
                return ORC_ERR_DOMAIN; /* "Programming logic error." */
            }
        }
        d[0] = (v[1] - v[0]) / (k[1] - k[0]);
        d[m - 1] = (v[m - 1] - v[m - 2]) / (k[m - 1] - k[m - 2]);
        for (s = 1; s < m - 1; s++)
            d[s] = (v[s + 1] - v[s]) / (2 * (k[s + 1] - k[s])) + (v[s] - v[s - 1]) / (2 * (k[s] - k[s - 1]));
        if (mono) {
            for (s = 0; s < m - 1; s++) {
                const double delta = (v[s + 1] - v[s]) / (k[s + 1] - k[s]);
                if (fabs(delta) < 1e-14) {
                    d[s] = 0.0;
                    d[s + 1] = 0.0;
                } else {
                    const double alpha = d[s] / delta, beta = d[s + 1] / delta;
                    if (!((alpha * alpha + beta * beta) <= 9)) {
                        const double tau = 3 / sqrt(alpha * alpha + beta * beta);
                        d[s] = tau * alpha * delta;
                        d[s + 1] = tau * beta * delta;
                    }
                }
            }
        }
    }
    for (int i = 0; i < n; i++) {
        const double x = x_new[i];
        int lo = 0, hi = m;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (k[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        double r;
        if (lo == 0) r = v[0];
        else if (lo == m) r = v[m - 1];
        else {
            const double x1 = k[lo - 1], x2 = k[lo];
            const double t = (x - x1) / (x2 - x1);
            const double h = x2 - x1;
            r = v[lo - 1] * h00(t) + d[lo - 1] * h10(t) * h + v[lo] * h01(t) + d[lo] * h11(t) * h;
        }
        values[i] = r;
    }
    free(k);
    return 0;
}

/* The remapping loop: every (P,V,K) line of psd [P][R][V][K] is
 * interpolated from l_old to l_new (same layout). Returns the first nonzero
 * status. */
int orc_remap(double* psd, const double* l_old, const double* l_new, int nP, int nR, int nV, int nK) {
    const size_t NQ = (size_t)nV * nK;
    int status = 0;
    double* buf = (double*)malloc(sizeof(double) * 3 * (size_t)(nR > 0 ? nR : 1));
    for (int iP = 0; iP < nP && !status; iP++)
        for (size_t q = 0; q < NQ && !status; q++) {
            double *xo = buf, *f = buf + nR, *xn = buf + 2 * nR;
            const size_t base = (size_t)iP * nR * NQ + q;
            for (int i = 0; i < nR; i++) {
                xo[i] = l_old[base + i * NQ];
                f[i] = psd[base + i * NQ];
                xn[i] = l_new[base + i * NQ];
            }
            status = orc_cubic1d(nR, xo, f, xn);
            if (!status)
                for (int i = 0; i < nR; i++) psd[base + i * NQ] = f[i];
        }
    free(buf);
    return status;
}
